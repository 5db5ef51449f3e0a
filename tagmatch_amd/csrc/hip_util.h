// Shared host/device helpers of the HIP library: error reporting, scans and
// fill kernels.  Everything here is internal to libtagmatch_amd.so.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdarg>
#include <cstdio>
#include <string>

#include "../../include/tagmatch_amd.h"

namespace tmx {

// Thread-local text of the last error (tm_last_error()).
std::string & last_error();
int set_error(int code, const char * fmt, ...);
void * pinned_alloc(size_t bytes);   // hipHostMalloc; NULL on failure
void pinned_free(void * p);
int hip_error(hipError_t e, const char * what, const char * file, int line);

#define TM_HIP(expr)                                                         \
	do {                                                                     \
		hipError_t e_ = (expr);                                              \
		if (e_ != hipSuccess) return ::tmx::hip_error(e_, #expr, __FILE__, __LINE__); \
	} while (0)

#define TM_CHECK(expr)                                                       \
	do {                                                                     \
		int rc_ = (expr);                                                    \
		if (rc_ != TM_OK) return rc_;                                        \
	} while (0)

// Device state shared by the library's translation units.
struct device_state {
	int device = -1;
	int num_cu = 0;
	int max_lds = 0;
	char arch[64] = {0};
};
device_state & dev();
int require_device();

// Device-wide exclusive scan of n uint32 counts into n+1 uint64 offsets
// (out[n] = total), on stream s.  tmp must hold scan_tmp_words(n) uint64.
size_t scan_tmp_words(uint64_t n);
void scan_u32_to_u64(const uint32_t * in, uint64_t * out, uint64_t n, uint64_t * tmp, hipStream_t s);

// Fill kernels (used instead of hipMemsetAsync so that steps can be
// captured in a hipGraph and replayed).
void fill_u32(uint32_t * p, uint32_t v, uint64_t n, hipStream_t s);
void fill_u64(uint64_t * p, uint64_t v, uint64_t n, hipStream_t s);
void copy_u64(uint64_t * dst, const uint64_t * src, uint64_t n, hipStream_t s);

template <typename T>
int dalloc(T ** p, size_t count) {
	*p = nullptr;
	if (count == 0) count = 1;
	hipError_t e = hipMalloc((void **)p, count * sizeof(T));
	if (e != hipSuccess) {
		*p = nullptr;
		(void)hipGetLastError();
		return set_error(TM_ERR_NOMEM, "hipMalloc of %zu bytes failed: %s", count * sizeof(T),
						 hipGetErrorString(e));
	}
	return TM_OK;
}

template <typename T>
void dfree(T *& p) {
	if (p) (void)hipFree((void *)p);
	p = nullptr;
}

} // namespace tmx
