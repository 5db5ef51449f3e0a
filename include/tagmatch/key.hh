// Key type of the TagMatch API (reference key.hh:6).
#ifndef TAGMATCH_AMD_KEY_HH_INCLUDED
#define TAGMATCH_AMD_KEY_HH_INCLUDED

#include <cstdint>

typedef uint32_t tagmatch_key_t;

#endif
