// Error reporting, device lifecycle, scan and fill kernels.
#include "hip_util.h"

#include <cstring>
#include <mutex>

namespace tmx {

std::string & last_error() {
	static thread_local std::string s;
	return s;
}

int set_error(int code, const char * fmt, ...) {
	char buf[1024];
	va_list ap;
	va_start(ap, fmt);
	std::vsnprintf(buf, sizeof(buf), fmt, ap);
	va_end(ap);
	last_error() = buf;
	return code;
}

int hip_error(hipError_t e, const char * what, const char * file, int line) {
	(void)hipGetLastError();
	int code = (e == hipErrorOutOfMemory) ? TM_ERR_NOMEM : TM_ERR_HIP;
	return set_error(code, "%s failed: %s (%s:%d)", what, hipGetErrorString(e), file, line);
}

device_state & dev() {
	static device_state d;
	return d;
}

int require_device() {
	if (dev().device < 0)
		return set_error(TM_ERR_STATE, "no device: call tm_dev_init() first");
	TM_HIP(hipSetDevice(dev().device));
	return TM_OK;
}

// ---------------------------------------------------------------- scans
//
// Three-phase exclusive scan: per-block sums, a single-block scan of the
// sums (looping with a carry), then per-block rescans plus the block offset.
// SCAN_ITEMS elements per block, 4 per thread.

static constexpr int SCAN_THREADS = 256;
static constexpr int SCAN_ITEMS = SCAN_THREADS * 4;

__device__ inline uint64_t block_exclusive_scan(uint64_t v, uint64_t * s_warp, uint64_t * total) {
	// wave-level inclusive scan with shuffles, then across the 4 waves
	const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
	uint64_t x = v;
	for (int o = 1; o < 64; o <<= 1) {
		uint64_t y = __shfl_up(x, o, 64);
		if (lane >= o) x += y;
	}
	if (lane == 63) s_warp[w] = x;
	__syncthreads();
	uint64_t base = 0, all = 0;
	for (int i = 0; i < SCAN_THREADS / 64; ++i) {
		if (i < w) base += s_warp[i];
		all += s_warp[i];
	}
	__syncthreads();
	*total = all;
	return base + x - v;
}

__global__ void __launch_bounds__(SCAN_THREADS)
k_scan_reduce(const uint32_t * __restrict__ in, uint64_t n, uint64_t * __restrict__ sums) {
	__shared__ uint64_t s_warp[SCAN_THREADS / 64];
	uint64_t base = (uint64_t)blockIdx.x * SCAN_ITEMS;
	uint64_t acc = 0;
	for (int k = 0; k < 4; ++k) {
		uint64_t i = base + (uint64_t)k * SCAN_THREADS + threadIdx.x;
		if (i < n) acc += in[i];
	}
	uint64_t total;
	block_exclusive_scan(acc, s_warp, &total);
	if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

__global__ void __launch_bounds__(SCAN_THREADS)
k_scan_sums(uint64_t * __restrict__ sums, uint64_t nsums) {
	__shared__ uint64_t s_warp[SCAN_THREADS / 64];
	uint64_t carry = 0;
	for (uint64_t b = 0; b < nsums; b += SCAN_THREADS) {
		uint64_t i = b + threadIdx.x;
		uint64_t v = i < nsums ? sums[i] : 0;
		uint64_t total;
		uint64_t ex = block_exclusive_scan(v, s_warp, &total);
		if (i < nsums) sums[i] = carry + ex;
		carry += total;
		__syncthreads();
	}
	if (threadIdx.x == 0) sums[nsums] = carry;
}

__global__ void __launch_bounds__(SCAN_THREADS)
k_scan_apply(const uint32_t * __restrict__ in, uint64_t n, const uint64_t * __restrict__ sums,
			 uint64_t * __restrict__ out, uint64_t nblocks) {
	__shared__ uint64_t s_warp[SCAN_THREADS / 64];
	uint64_t base = (uint64_t)blockIdx.x * SCAN_ITEMS;
	// each thread owns 4 consecutive elements
	uint64_t i0 = base + (uint64_t)threadIdx.x * 4;
	uint32_t v[4];
	uint64_t acc = 0;
	for (int k = 0; k < 4; ++k) {
		v[k] = (i0 + k < n) ? in[i0 + k] : 0;
		acc += v[k];
	}
	uint64_t total;
	uint64_t ex = block_exclusive_scan(acc, s_warp, &total) + sums[blockIdx.x];
	for (int k = 0; k < 4; ++k) {
		if (i0 + k < n) out[i0 + k] = ex;
		ex += v[k];
	}
	if (blockIdx.x == nblocks - 1 && threadIdx.x == 0) out[n] = sums[nblocks];
}

size_t scan_tmp_words(uint64_t n) {
	return (n + SCAN_ITEMS - 1) / SCAN_ITEMS + 2;
}

void scan_u32_to_u64(const uint32_t * in, uint64_t * out, uint64_t n, uint64_t * tmp, hipStream_t s) {
	if (n == 0) {
		fill_u64(out, 0, 1, s);
		return;
	}
	uint64_t nb = (n + SCAN_ITEMS - 1) / SCAN_ITEMS;
	hipLaunchKernelGGL(k_scan_reduce, dim3((uint32_t)nb), dim3(SCAN_THREADS), 0, s, in, n, tmp);
	hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(SCAN_THREADS), 0, s, tmp, nb);
	hipLaunchKernelGGL(k_scan_apply, dim3((uint32_t)nb), dim3(SCAN_THREADS), 0, s, in, n, tmp, out, nb);
}

// ---------------------------------------------------------------- fills

template <typename T>
__global__ void __launch_bounds__(256) k_fill(T * __restrict__ p, T v, uint64_t n) {
	uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
	for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
		p[i] = v;
}

__global__ void __launch_bounds__(256)
k_copy_u64(uint64_t * __restrict__ d, const uint64_t * __restrict__ s, uint64_t n) {
	uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
	for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
		d[i] = s[i];
}

static uint32_t fill_grid(uint64_t n) {
	uint64_t g = (n + 255) / 256;
	if (g > 4096) g = 4096;
	if (g == 0) g = 1;
	return (uint32_t)g;
}

void fill_u32(uint32_t * p, uint32_t v, uint64_t n, hipStream_t s) {
	if (n) hipLaunchKernelGGL(k_fill<uint32_t>, dim3(fill_grid(n)), dim3(256), 0, s, p, v, n);
}

void fill_u64(uint64_t * p, uint64_t v, uint64_t n, hipStream_t s) {
	if (n) hipLaunchKernelGGL(k_fill<uint64_t>, dim3(fill_grid(n)), dim3(256), 0, s, p, v, n);
}

void copy_u64(uint64_t * d, const uint64_t * src, uint64_t n, hipStream_t s) {
	if (n) hipLaunchKernelGGL(k_copy_u64, dim3(fill_grid(n)), dim3(256), 0, s, d, src, n);
}

// Pinned host memory (hipHostMalloc) for host code compiled without the HIP
// headers: result buffers that device copies fill by DMA.
void * pinned_alloc(size_t bytes) {
	void * p = nullptr;
	return hipHostMalloc(&p, bytes ? bytes : 1, hipHostMallocDefault) == hipSuccess ? p : nullptr;
}
void pinned_free(void * p) {
	if (p) (void)hipHostFree(p);
}

} // namespace tmx

using namespace tmx;

extern "C" {

const char * tm_last_error(void) { return last_error().c_str(); }

const char * tm_version(void) { return "tagmatch_amd 0.1 (gfx950)"; }

int tm_dev_count(int * count) {
	if (!count) return set_error(TM_ERR_INVALID, "count is NULL");
	int n = 0;
	hipError_t e = hipGetDeviceCount(&n);
	if (e != hipSuccess) {
		(void)hipGetLastError();
		*count = 0;
		return set_error(TM_ERR_NODEV, "hipGetDeviceCount: %s", hipGetErrorString(e));
	}
	*count = n;
	return TM_OK;
}

int tm_dev_init(int device) {
	int n = 0;
	hipError_t e = hipGetDeviceCount(&n);
	if (e != hipSuccess || n == 0) {
		(void)hipGetLastError();
		return set_error(TM_ERR_NODEV, "no HIP device available (%s)",
						 e != hipSuccess ? hipGetErrorString(e) : "count 0");
	}
	if (device < 0 || device >= n)
		return set_error(TM_ERR_INVALID, "device %d out of range [0,%d)", device, n);
	TM_HIP(hipSetDevice(device));
	hipDeviceProp_t prop;
	TM_HIP(hipGetDeviceProperties(&prop, device));
	device_state & d = dev();
	d.device = device;
	d.num_cu = prop.multiProcessorCount;
	d.max_lds = (int)prop.sharedMemPerBlock;
	std::snprintf(d.arch, sizeof(d.arch), "%s", prop.gcnArchName);
	if (std::strncmp(d.arch, "gfx950", 6) != 0) {
		d.device = -1;
		return set_error(TM_ERR_NODEV, "device %d is %s; this library is built for gfx950 only",
						 device, prop.gcnArchName);
	}
	return TM_OK;
}

int tm_dev_shutdown(void) {
	if (dev().device < 0) return TM_OK;
	TM_HIP(hipSetDevice(dev().device));
	TM_HIP(hipDeviceSynchronize());
	dev().device = -1;
	return TM_OK;
}

int tm_mem_info(size_t * free_bytes, size_t * total_bytes) {
	TM_CHECK(require_device());
	size_t f = 0, t = 0;
	TM_HIP(hipMemGetInfo(&f, &t));
	if (free_bytes) *free_bytes = f;
	if (total_bytes) *total_bytes = t;
	return TM_OK;
}

int tm_dev_synchronize(void) {
	TM_CHECK(require_device());
	TM_HIP(hipDeviceSynchronize());
	return TM_OK;
}

int tm_dev_arch(char * buf, size_t len) {
	TM_CHECK(require_device());
	if (!buf || len == 0) return set_error(TM_ERR_INVALID, "buffer is NULL");
	std::snprintf(buf, len, "%s", dev().arch);
	return TM_OK;
}

int tm_dev_alloc(void ** ptr, uint64_t bytes) {
	if (!ptr) return set_error(TM_ERR_INVALID, "ptr is NULL");
	*ptr = nullptr;
	TM_CHECK(require_device());
	void * p = nullptr;
	if (hipMalloc(&p, bytes ? bytes : 1) != hipSuccess)
		return set_error(TM_ERR_NOMEM, "hipMalloc of %llu bytes failed", (unsigned long long)bytes);
	*ptr = p;
	return TM_OK;
}

int tm_dev_free(void * ptr) {
	if (ptr) TM_HIP(hipFree(ptr));
	return TM_OK;
}

int tm_dev_upload(void * dev_dst, const void * host_src, uint64_t bytes) {
	if (bytes == 0) return TM_OK;
	if (!dev_dst || !host_src) return set_error(TM_ERR_INVALID, "NULL argument");
	TM_CHECK(require_device());
	TM_HIP(hipMemcpy(dev_dst, host_src, bytes, hipMemcpyHostToDevice));
	return TM_OK;
}

int tm_dev_download(void * host_dst, const void * dev_src, uint64_t bytes) {
	if (bytes == 0) return TM_OK;
	if (!host_dst || !dev_src) return set_error(TM_ERR_INVALID, "NULL argument");
	TM_CHECK(require_device());
	TM_HIP(hipMemcpy(host_dst, dev_src, bytes, hipMemcpyDeviceToHost));
	return TM_OK;
}

int tm_host_alloc(void ** ptr, uint64_t bytes) {
	if (!ptr) return set_error(TM_ERR_INVALID, "ptr is NULL");
	*ptr = nullptr;
	TM_CHECK(require_device());
	void * p = pinned_alloc(bytes);
	if (!p) return set_error(TM_ERR_NOMEM, "pinned host allocation of %llu bytes failed", (unsigned long long)bytes);
	*ptr = p;
	return TM_OK;
}

int tm_host_free(void * ptr) {
	pinned_free(ptr);
	return TM_OK;
}

} // extern "C"
