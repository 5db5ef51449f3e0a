// GPU balanced partitioning of a filter table.
//
// Restates the reference's recursive balanced partitioning
// (reference partitioning.cc:230-411) level-synchronously: every pending
// candidate ("segment") of the current level is processed at once.
//
//   * A BALANCE segment (size > maxp) is split on the bit whose frequency is
//     nearest half its size, searching from its first unused pivot bit
//     (balancing_pivot, partitioning.cc:262-278).  The 0-side keeps the
//     segment's pivot mask, the 1-side adds the pivot (:307-312).  A child
//     larger than maxp stays BALANCE; a smaller child is finished if its
//     pivot mask is non-empty, else it goes to NONZERO (:314-357).
//   * A NONZERO segment is split on its most frequent bit until all its
//     filters share a bit (nonzero_mask_partitioning, :251-271): the 1-side
//     is finished, the 0-side continues.  A segment of all-zero filters is
//     finished with an empty mask (the reference does not terminate there).
//   * Splits are stable: zeros first, then ones, each in input order
//     (std::stable_partition, :165).
//
// Because every decision depends only on the members of a segment, this
// produces the same partitions (as sets, with the same member order) as the
// reference's thread pool, whatever order the reference finishes them in.
//
// Per level: a bit histogram per tile of 1024 filters (LDS atomics), a
// per-segment sum, one device-to-host copy of the segment histograms, the
// pivot decisions on the host (a few hundred segments at most), per-tile
// split offsets, and a stable scatter (ballot + mbcnt) of all positions into
// the other buffer.
#include "hip_util.h"

#include <algorithm>
#include <cstring>
#include <vector>

using namespace tmx;

namespace {

constexpr uint32_t TILE = 1024;
constexpr int NO_PIVOT = -1;

enum seg_mode : uint8_t { BALANCE = 0, NONZERO = 1, DONE = 2 };

struct seg_t {
	uint32_t begin, end;
	uint32_t used[6];       // pivot bits used on the path (BALANCE only)
	uint8_t mode;
	uint8_t mask_nonempty;  // some 1-side pivot on the path
};

// tile: [begin, end) positions of one segment
struct tile_t {
	uint32_t begin, end, seg, pad;
};

// Each filter record: 6 words, input index, pad (32 bytes).
__global__ void __launch_bounds__(256)
k_pack(const uint32_t * __restrict__ w, uint32_t n, uint32_t * __restrict__ rec) {
	uint32_t i = blockIdx.x * 256 + threadIdx.x;
	if (i >= n) return;
	uint32_t * r = rec + (size_t)i * 8;
	for (int j = 0; j < 6; ++j) r[j] = w[(size_t)i * 6 + j];
	r[6] = i;
	r[7] = 0;
}

__global__ void __launch_bounds__(256)
k_tile_hist(const uint32_t * __restrict__ rec, const tile_t * __restrict__ tiles,
			uint32_t * __restrict__ tile_hist) {
	__shared__ uint32_t h[192];
	const tile_t t = tiles[blockIdx.x];
	for (uint32_t b = threadIdx.x; b < 192; b += 256) h[b] = 0;
	__syncthreads();
	for (uint32_t i = t.begin + threadIdx.x; i < t.end; i += 256) {
		const uint32_t * r = rec + (size_t)i * 8;
		for (int j = 0; j < 6; ++j) {
			uint32_t w = r[j];
			while (w) {
				uint32_t b = __builtin_ctz(w);
				w &= w - 1;
				atomicAdd(&h[j * 32 + b], 1u);
			}
		}
	}
	__syncthreads();
	for (uint32_t b = threadIdx.x; b < 192; b += 256) tile_hist[(size_t)blockIdx.x * 192 + b] = h[b];
}

// freq[s][b] = sum of tile_hist over the tiles [tb[s], tb[s+1]) of segment s
__global__ void __launch_bounds__(192)
k_seg_sum(const uint32_t * __restrict__ tile_hist, const uint32_t * __restrict__ tb,
		  uint32_t * __restrict__ freq) {
	const uint32_t s = blockIdx.x, b = threadIdx.x;
	uint32_t acc = 0;
	for (uint32_t t = tb[s]; t < tb[s + 1]; ++t) acc += tile_hist[(size_t)t * 192 + b];
	freq[(size_t)s * 192 + b] = acc;
}

// One thread per split segment: zero/one base of each of its tiles.
// piv[s] is the pivot of split segment s; its tiles are [tb[s], tb[s+1]).
// hidx[t] is the histogram row of tile t.
__global__ void __launch_bounds__(64)
k_tile_bases(const uint32_t * __restrict__ tile_hist, const uint32_t * __restrict__ hidx,
			 const tile_t * __restrict__ tiles,
			 const uint32_t * __restrict__ tb, const int32_t * __restrict__ piv, uint32_t nseg,
			 const uint32_t * __restrict__ seg_zeros, uint32_t * __restrict__ zbase,
			 uint32_t * __restrict__ obase) {
	uint32_t s = blockIdx.x * 64 + threadIdx.x;
	if (s >= nseg) return;
	const int32_t p = piv[s];
	if (p < 0) return;
	uint32_t z = 0, o = 0;
	const uint32_t first = tiles[tb[s]].begin;
	for (uint32_t t = tb[s]; t < tb[s + 1]; ++t) {
		uint32_t ones = tile_hist[(size_t)hidx[t] * 192 + p];
		uint32_t sz = tiles[t].end - tiles[t].begin;
		zbase[t] = first + z;
		obase[t] = first + seg_zeros[s] + o;
		z += sz - ones;
		o += ones;
	}
}

// One wave per tile: stable split on the tile's pivot, or a plain copy.
__global__ void __launch_bounds__(64)
k_scatter(const uint32_t * __restrict__ src, uint32_t * __restrict__ dst,
		  const tile_t * __restrict__ tiles, const int32_t * __restrict__ tile_piv,
		  const uint32_t * __restrict__ zbase, const uint32_t * __restrict__ obase) {
	const tile_t t = tiles[blockIdx.x];
	const int32_t p = tile_piv[blockIdx.x];
	const uint32_t lane = threadIdx.x;
	if (p < 0) {
		for (uint32_t i = t.begin + lane; i < t.end; i += 64) {
			const uint4 * a = reinterpret_cast<const uint4 *>(src + (size_t)i * 8);
			uint4 * d = reinterpret_cast<uint4 *>(dst + (size_t)i * 8);
			d[0] = a[0];
			d[1] = a[1];
		}
		return;
	}
	uint32_t zpos = zbase[blockIdx.x], opos = obase[blockIdx.x];
	const uint32_t pw = (uint32_t)p >> 5, pb = (uint32_t)p & 31;
	for (uint32_t c = t.begin; c < t.end; c += 64) {
		const uint32_t i = c + lane;
		const bool valid = i < t.end;
		uint4 a0 = make_uint4(0, 0, 0, 0), a1 = a0;
		bool one = false;
		if (valid) {
			const uint4 * a = reinterpret_cast<const uint4 *>(src + (size_t)i * 8);
			a0 = a[0];
			a1 = a[1];
			uint32_t w = pw == 0 ? a0.x : pw == 1 ? a0.y : pw == 2 ? a0.z : pw == 3 ? a0.w : pw == 4 ? a1.x : a1.y;
			one = (w >> pb) & 1;
		}
		const uint64_t zm = __ballot(valid && !one), om = __ballot(valid && one);
		const uint32_t zr = __builtin_amdgcn_mbcnt_hi((uint32_t)(zm >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)zm, 0u));
		const uint32_t orr = __builtin_amdgcn_mbcnt_hi((uint32_t)(om >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)om, 0u));
		if (valid) {
			uint32_t dpos = one ? opos + orr : zpos + zr;
			uint4 * d = reinterpret_cast<uint4 *>(dst + (size_t)dpos * 8);
			d[0] = a0;
			d[1] = a1;
		}
		zpos += (uint32_t)__popcll(zm);
		opos += (uint32_t)__popcll(om);
	}
}

// AND of the members of each finished partition; all-ones if empty.
__global__ void __launch_bounds__(256)
k_part_masks(const uint32_t * __restrict__ rec, const uint32_t * __restrict__ ranges,
			 uint32_t * __restrict__ masks) {
	__shared__ uint32_t s[256][6];
	const uint32_t p = blockIdx.x, b = ranges[p], e = ranges[p + 1];
	uint32_t m[6] = {~0u, ~0u, ~0u, ~0u, ~0u, ~0u};
	for (uint32_t i = b + threadIdx.x; i < e; i += 256)
		for (int j = 0; j < 6; ++j) m[j] &= rec[(size_t)i * 8 + j];
	for (int j = 0; j < 6; ++j) s[threadIdx.x][j] = m[j];
	__syncthreads();
	for (uint32_t o = 128; o > 0; o >>= 1) {
		if (threadIdx.x < o)
			for (int j = 0; j < 6; ++j) s[threadIdx.x][j] &= s[threadIdx.x + o][j];
		__syncthreads();
	}
	if (threadIdx.x < 6) masks[(size_t)p * 6 + threadIdx.x] = s[0][threadIdx.x];
}

__global__ void __launch_bounds__(256)
k_extract_perm(const uint32_t * __restrict__ rec, uint32_t n, uint32_t * __restrict__ perm) {
	uint32_t i = blockIdx.x * 256 + threadIdx.x;
	if (i < n) perm[i] = rec[(size_t)i * 8 + 6];
}

inline bool used_bit(const seg_t & s, int b) { return (s.used[b >> 5] >> (b & 31)) & 1; }

uint32_t dist(uint32_t a, uint32_t b) { return a > b ? a - b : b - a; }

// balancing_pivot (reference partitioning.cc:262-278)
int balancing_pivot(const seg_t & s, const uint32_t * f) {
	const uint32_t size = s.end - s.begin, half = size / 2;
	int pivot = 0;
	while (pivot < 192 && used_bit(s, pivot)) ++pivot;
	if (pivot >= 192) return NO_PIVOT;
	uint32_t min_d = dist(half, f[pivot]);
	for (int b = pivot + 1; b < 192; ++b) {
		uint32_t d = dist(half, f[b]);
		if (d < min_d) {
			pivot = b;
			if (d == 0) break;
			min_d = d;
		}
	}
	return pivot;
}

struct dev_buf {
	void * p = nullptr;
	~dev_buf() { if (p) (void)hipFree(p); }
};

template <typename T>
int ensure(dev_buf & b, size_t & cap, size_t need) {
	if (need <= cap && b.p) return TM_OK;
	if (b.p) (void)hipFree(b.p);
	b.p = nullptr;
	size_t n = std::max<size_t>(need, 1) * 2;
	TM_HIP(hipMalloc(&b.p, n * sizeof(T)));
	cap = n;
	return TM_OK;
}

} // namespace

extern "C" int tm_partition(const uint32_t * words, uint32_t n, uint32_t maxp, uint32_t * perm,
							uint32_t * part_offsets, uint32_t * part_masks, uint32_t max_parts,
							uint32_t * n_parts) {
	if (!n_parts || !part_offsets) return set_error(TM_ERR_INVALID, "NULL output argument");
	if (n && (!words || !perm)) return set_error(TM_ERR_INVALID, "NULL input/perm");
	if (maxp == 0) return set_error(TM_ERR_INVALID, "maxp must be > 0");
	*n_parts = 0;
	part_offsets[0] = 0;
	if (n == 0) return TM_OK;
	TM_CHECK(require_device());

	hipStream_t s;
	TM_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
	struct stream_guard {
		hipStream_t s;
		~stream_guard() { (void)hipStreamDestroy(s); }
	} sg{s};

	dev_buf d_words, d_rec[2], d_tiles, d_hist, d_tb, d_freq, d_piv, d_tpiv, d_zeros, d_zb, d_ob, d_rng, d_mask, d_hidx;
	size_t c_tiles = 0, c_hist = 0, c_tb = 0, c_freq = 0, c_piv = 0, c_tpiv = 0, c_zeros = 0, c_zb = 0, c_ob = 0, c_hidx = 0;
	TM_HIP(hipMalloc(&d_words.p, (size_t)n * 24));
	TM_HIP(hipMalloc(&d_rec[0].p, (size_t)n * 32));
	TM_HIP(hipMalloc(&d_rec[1].p, (size_t)n * 32));
	TM_HIP(hipMemcpyAsync(d_words.p, words, (size_t)n * 24, hipMemcpyHostToDevice, s));
	hipLaunchKernelGGL(k_pack, dim3((n + 255) / 256), dim3(256), 0, s, (const uint32_t *)d_words.p, n,
					   (uint32_t *)d_rec[0].p);
	int cur = 0;

	std::vector<seg_t> segs(1);
	std::memset(&segs[0], 0, sizeof(seg_t));
	segs[0].begin = 0;
	segs[0].end = n;
	segs[0].mode = n > maxp ? BALANCE : NONZERO;

	std::vector<tile_t> tiles;
	std::vector<uint32_t> tb, freq, seg_zeros;
	std::vector<int32_t> piv, tpiv;
	for (int level = 0;; ++level) {
		// active segments (need a decision this level)
		std::vector<uint32_t> act;
		for (uint32_t k = 0; k < segs.size(); ++k)
			if (segs[k].mode != DONE) act.push_back(k);
		if (act.empty()) break;

		// tiles of active segments, contiguous per segment
		tiles.clear();
		tb.assign(1, 0);
		for (uint32_t k : act) {
			const seg_t & g = segs[k];
			for (uint32_t b = g.begin; b < g.end; b += TILE)
				tiles.push_back(tile_t{b, std::min(g.end, b + TILE), k, 0});
			if (g.begin == g.end) tiles.push_back(tile_t{g.begin, g.end, k, 0});
			tb.push_back((uint32_t)tiles.size());
		}
		const uint32_t nt = (uint32_t)tiles.size(), na = (uint32_t)act.size();
		TM_CHECK(ensure<tile_t>(d_tiles, c_tiles, nt));
		TM_CHECK(ensure<uint32_t>(d_hist, c_hist, (size_t)nt * 192));
		TM_CHECK(ensure<uint32_t>(d_tb, c_tb, na + 1));
		TM_CHECK(ensure<uint32_t>(d_freq, c_freq, (size_t)na * 192));
		TM_HIP(hipMemcpyAsync(d_tiles.p, tiles.data(), nt * sizeof(tile_t), hipMemcpyHostToDevice, s));
		TM_HIP(hipMemcpyAsync(d_tb.p, tb.data(), (na + 1) * sizeof(uint32_t), hipMemcpyHostToDevice, s));
		const uint32_t * rec = (const uint32_t *)d_rec[cur].p;
		hipLaunchKernelGGL(k_tile_hist, dim3(nt), dim3(256), 0, s, rec, (const tile_t *)d_tiles.p,
						   (uint32_t *)d_hist.p);
		hipLaunchKernelGGL(k_seg_sum, dim3(na), dim3(192), 0, s, (const uint32_t *)d_hist.p,
						   (const uint32_t *)d_tb.p, (uint32_t *)d_freq.p);
		freq.resize((size_t)na * 192);
		TM_HIP(hipMemcpyAsync(freq.data(), d_freq.p, freq.size() * 4, hipMemcpyDeviceToHost, s));
		TM_HIP(hipStreamSynchronize(s));

		// decisions
		piv.assign(na, NO_PIVOT);
		seg_zeros.assign(na, 0);
		std::vector<seg_t> next;
		next.reserve(segs.size() + na);
		// keep segment order by position: rebuild the list
		std::vector<int> act_index(segs.size(), -1);
		for (uint32_t a = 0; a < na; ++a) act_index[act[a]] = (int)a;
		for (uint32_t k = 0; k < segs.size(); ++k) {
			const seg_t g = segs[k];
			int a = act_index[k];
			if (a < 0) {
				next.push_back(g);
				continue;
			}
			const uint32_t * f = &freq[(size_t)a * 192];
			const uint32_t size = g.end - g.begin;
			if (g.mode == NONZERO) {
				uint32_t mf = f[0];
				int pv = 0;
				for (int b = 1; b < 192; ++b)
					if (f[b] > mf) { mf = f[b]; pv = b; }
				if (mf == size || mf == 0) {
					seg_t d = g;
					d.mode = DONE;
					next.push_back(d);
					continue;
				}
				piv[a] = pv;
				seg_zeros[a] = size - mf;
				seg_t z = g, o = g;
				z.end = g.begin + (size - mf);
				o.begin = z.end;
				z.mode = NONZERO;
				o.mode = DONE;
				next.push_back(z);
				next.push_back(o);
			} else {  // BALANCE
				int pv = balancing_pivot(g, f);
				if (pv == NO_PIVOT) {
					seg_t d = g;
					d.mode = DONE;
					next.push_back(d);
					continue;
				}
				piv[a] = pv;
				seg_zeros[a] = size - f[pv];
				seg_t z = g, o = g;
				z.end = g.begin + (size - f[pv]);
				o.begin = z.end;
				z.used[pv >> 5] |= 1u << (pv & 31);
				o.used[pv >> 5] |= 1u << (pv & 31);
				o.mask_nonempty = 1;
				for (seg_t * c : {&z, &o}) {
					uint32_t cs = c->end - c->begin;
					c->mode = cs > maxp ? BALANCE : (c->mask_nonempty ? DONE : NONZERO);
				}
				next.push_back(z);
				next.push_back(o);
			}
		}

		// split or copy every position of the current buffer
		std::vector<tile_t> all;
		tpiv.clear();
		std::vector<uint32_t> split_tb(1, 0);
		std::vector<int32_t> split_piv;
		std::vector<uint32_t> split_zeros;
		// split tiles first (contiguous per split segment) so tile bases can be
		// computed per segment, then copy tiles for everything else
		for (uint32_t a = 0; a < na; ++a) {
			if (piv[a] == NO_PIVOT) continue;
			for (uint32_t t = tb[a]; t < tb[a + 1]; ++t) {
				all.push_back(tiles[t]);
				tpiv.push_back(piv[a]);
			}
			split_tb.push_back((uint32_t)all.size());
			split_piv.push_back(piv[a]);
			split_zeros.push_back(seg_zeros[a]);
		}
		const uint32_t n_split_tiles = (uint32_t)all.size();
		{
			// copy tiles cover every position not in a split segment
			uint32_t pos = 0;
			std::vector<std::pair<uint32_t, uint32_t>> split_ranges;
			for (uint32_t a = 0; a < na; ++a)
				if (piv[a] != NO_PIVOT) split_ranges.push_back({segs[act[a]].begin, segs[act[a]].end});
			std::sort(split_ranges.begin(), split_ranges.end());
			for (auto & r : split_ranges) {
				for (uint32_t b = pos; b < r.first; b += TILE * 16) {
					all.push_back(tile_t{b, std::min(r.first, b + TILE * 16), 0, 0});
					tpiv.push_back(NO_PIVOT);
				}
				pos = r.second;
			}
			for (uint32_t b = pos; b < n; b += TILE * 16) {
				all.push_back(tile_t{b, std::min(n, b + TILE * 16), 0, 0});
				tpiv.push_back(NO_PIVOT);
			}
		}
		const uint32_t nall = (uint32_t)all.size(), nsplit = (uint32_t)split_piv.size();
		// histogram row of each split tile (rows are in `tiles` order)
		std::vector<uint32_t> hist_index;
		for (uint32_t a = 0; a < na; ++a)
			if (piv[a] != NO_PIVOT)
				for (uint32_t t = tb[a]; t < tb[a + 1]; ++t) hist_index.push_back(t);

		TM_CHECK(ensure<tile_t>(d_tiles, c_tiles, std::max(nall, nt)));
		TM_CHECK(ensure<int32_t>(d_tpiv, c_tpiv, nall));
		TM_CHECK(ensure<uint32_t>(d_zb, c_zb, nall));
		TM_CHECK(ensure<uint32_t>(d_ob, c_ob, nall));
		TM_CHECK(ensure<int32_t>(d_piv, c_piv, nsplit + 1));
		TM_CHECK(ensure<uint32_t>(d_zeros, c_zeros, nsplit + 1));
		TM_CHECK(ensure<uint32_t>(d_hidx, c_hidx, n_split_tiles + 1));
		if (nsplit) {
			TM_HIP(hipMemcpyAsync(d_hidx.p, hist_index.data(), n_split_tiles * 4, hipMemcpyHostToDevice, s));
			TM_HIP(hipMemcpyAsync(d_tb.p, split_tb.data(), (nsplit + 1) * 4, hipMemcpyHostToDevice, s));
			TM_HIP(hipMemcpyAsync(d_piv.p, split_piv.data(), nsplit * 4, hipMemcpyHostToDevice, s));
			TM_HIP(hipMemcpyAsync(d_zeros.p, split_zeros.data(), nsplit * 4, hipMemcpyHostToDevice, s));
		}
		TM_HIP(hipMemcpyAsync(d_tiles.p, all.data(), nall * sizeof(tile_t), hipMemcpyHostToDevice, s));
		TM_HIP(hipMemcpyAsync(d_tpiv.p, tpiv.data(), nall * 4, hipMemcpyHostToDevice, s));
		if (nsplit)
			hipLaunchKernelGGL(k_tile_bases, dim3((nsplit + 63) / 64), dim3(64), 0, s, (const uint32_t *)d_hist.p,
							   (const uint32_t *)d_hidx.p, (const tile_t *)d_tiles.p, (const uint32_t *)d_tb.p, (const int32_t *)d_piv.p,
							   nsplit, (const uint32_t *)d_zeros.p, (uint32_t *)d_zb.p, (uint32_t *)d_ob.p);
		if (nall)
			hipLaunchKernelGGL(k_scatter, dim3(nall), dim3(64), 0, s, (const uint32_t *)d_rec[cur].p,
							   (uint32_t *)d_rec[cur ^ 1].p, (const tile_t *)d_tiles.p, (const int32_t *)d_tpiv.p,
							   (const uint32_t *)d_zb.p, (const uint32_t *)d_ob.p);
		TM_HIP(hipGetLastError());
		cur ^= 1;
		segs.swap(next);
		if (level > 100000) return set_error(TM_ERR_HIP, "partitioner did not converge");
	}

	// finished partitions in position order (segments never overlap)
	std::sort(segs.begin(), segs.end(), [](const seg_t & a, const seg_t & b) {
		return a.begin != b.begin ? a.begin < b.begin : a.end < b.end;
	});
	const uint32_t np = (uint32_t)segs.size();
	if (np > max_parts) {
		*n_parts = np;
		return set_error(TM_ERR_OVERFLOW, "%u partitions exceed max_parts %u", np, max_parts);
	}
	std::vector<uint32_t> ranges(np + 1);
	for (uint32_t p = 0; p < np; ++p) ranges[p] = segs[p].begin;
	ranges[np] = n;
	for (uint32_t p = 0; p <= np; ++p) part_offsets[p] = ranges[p];
	TM_HIP(hipMalloc(&d_rng.p, (np + 1) * 4));
	TM_HIP(hipMalloc(&d_mask.p, (size_t)np * 24 + 4));
	TM_HIP(hipMemcpyAsync(d_rng.p, ranges.data(), (np + 1) * 4, hipMemcpyHostToDevice, s));
	hipLaunchKernelGGL(k_part_masks, dim3(np), dim3(256), 0, s, (const uint32_t *)d_rec[cur].p,
					   (const uint32_t *)d_rng.p, (uint32_t *)d_mask.p);
	uint32_t * d_perm = (uint32_t *)d_words.p;  // reuse: n x 6 words >= n
	hipLaunchKernelGGL(k_extract_perm, dim3((n + 255) / 256), dim3(256), 0, s, (const uint32_t *)d_rec[cur].p, n, d_perm);
	TM_HIP(hipGetLastError());
	if (part_masks) TM_HIP(hipMemcpyAsync(part_masks, d_mask.p, (size_t)np * 24, hipMemcpyDeviceToHost, s));
	TM_HIP(hipMemcpyAsync(perm, d_perm, (size_t)n * 4, hipMemcpyDeviceToHost, s));
	TM_HIP(hipStreamSynchronize(s));
	*n_parts = np;
	return TM_OK;
}
