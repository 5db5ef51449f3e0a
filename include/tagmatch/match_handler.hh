// Asynchronous result handler (reference match_handler.hh:15-22): the
// engine calls match_done() from one of its threads, which by default calls
// process_results().
#ifndef TAGMATCH_AMD_MATCH_HANDLER_HH_INCLUDED
#define TAGMATCH_AMD_MATCH_HANDLER_HH_INCLUDED

#include "query.hh"

class match_handler {
public:
	virtual void match_done(query * q) { process_results(q); }
	virtual void process_results(query * q) = 0;
	virtual ~match_handler() {}
};

#endif
