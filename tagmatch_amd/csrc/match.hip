// Query -> partition lookup, the batched subset-match kernel, key expansion
// and per-query dedup, all on the device; rings, schedules and hipGraph
// replay on the host side.
//
// One submission of nq queries runs as a fixed sequence of kernels on the
// ring's stream, with no host synchronisation in the persistent schedule:
//
//   k_lookup_count / scan / k_lookup_fill
//        query -> partitions with mask & ~q == 0, grouped per partition
//        (reference front_end.cc:465-477 do_match, on the device);
//   k_build_items
//        cut each partition's query list into batches of 64*TILES queries
//        and its blocks into chunks: one work item per (batch, chunk)
//        (reference front_end.cc:350-376 batching of 256 queries);
//   k_match
//        per work item: for each 64-filter block, a candidate pass with one
//        query per lane tests the block's AND mask against 64 queries at
//        once; for each surviving query, one filter per lane tests
//        f & ~q == 0 (reference gpu.cu:60-607 phases 2 and 3).  Hits are
//        expanded to dense key ids and OR-ed into per-query key bitmaps
//        (staged in LDS), or appended as (query, filter) pairs;
//   k_bm_count / scan / k_bm_write   (KEYS_UNIQUE)
//        bitmap -> sorted distinct keys (reference tagmatch_query.hh:99-119
//        sort + unique, and back_end.cc:288-301 key expansion);
//   k_hits_count / scan / k_hits_write  (HITS, KEYS_ALL)
//        group hits per query and expand them to ids or keys.
#include "cost_model.h"
#include "fib.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <numeric>
#include <vector>

using namespace tmx;

namespace {

constexpr int NQUEUE = 8;        // work queues, one per XCD of an MI355X
constexpr int QPAD = 32;         // queue cursors 256 B apart: atomics on one line serialize
constexpr uint32_t SEG_MAX_BLOCKS = 1024;   // persistent: largest block segment (measured at 10M)
constexpr uint32_t SEG_MAX_BLOCKS_EXIT = 4096;   // the same with the full-key-set exit: most items
                                                 // skip their blocks, and fewer items cost less
constexpr uint32_t SEG_MAX_BLOCKS_COMPACT = 1024;   // compacted steps: the tiles of an XCD share a
                                                   // segment small enough for its L2 (2.8 ms at
                                                   // 10M filters, 1M queries, against 3.45 with 4096)
constexpr uint32_t WSEG_BASE = 64;          // block weight: fixed part, in units of one expected hit
constexpr uint32_t WSEG_MAX_TILES = 256;    // persistent: tiles up to which segments are cut by weight
constexpr uint64_t COMPACT_MIN_TILES = 128;   // persistent: tiles above which steps are compacted
constexpr uint32_t PRICE_SAMPLE_TILES = 1024;  // cost model: tiles priced when only totals are needed
constexpr uint32_t NO_QUERY = 0xFFFFFFFFu;
constexpr size_t MAX_GRAPHS = 16;             // captured step sizes kept per ring

__device__ inline uint32_t mbcnt64(uint64_t m) {
	return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// ------------------------------------------------------------- lookup

__global__ void __launch_bounds__(256)
k_lookup_count(const uint32_t * __restrict__ q, uint32_t nq, const uint32_t * __restrict__ part_info,
			   uint32_t P, uint32_t * __restrict__ pcount) {
	const uint32_t i = blockIdx.x * 256 + threadIdx.x;
	const bool valid = i < nq;
	uint32_t n0 = ~0u, n1 = ~0u, n2 = ~0u, n3 = ~0u, n4 = ~0u, n5 = ~0u;
	if (valid) {
		uint4 a = *reinterpret_cast<const uint4 *>(q + (size_t)i * 8);
		uint2 b = *reinterpret_cast<const uint2 *>(q + (size_t)i * 8 + 4);
		n0 = ~a.x; n1 = ~a.y; n2 = ~a.z; n3 = ~a.w; n4 = ~b.x; n5 = ~b.y;
	}
	for (uint32_t p = 0; p < P; ++p) {
		const uint32_t * m = part_info + (size_t)p * 8;
		uint32_t t = (m[0] & n0) | (m[1] & n1) | (m[2] & n2) | (m[3] & n3) | (m[4] & n4) | (m[5] & n5);
		uint64_t b = __ballot(valid && t == 0);
		if (b != 0 && (threadIdx.x & 63) == 0) atomicAdd(&pcount[p], (uint32_t)__popcll(b));
	}
}

__global__ void __launch_bounds__(256)
k_lookup_fill(const uint32_t * __restrict__ q, uint32_t nq, const uint32_t * __restrict__ part_info,
			  uint32_t P, unsigned long long * __restrict__ pcur, uint32_t * __restrict__ list,
			  uint64_t list_cap) {
	const uint32_t i = blockIdx.x * 256 + threadIdx.x;
	const uint32_t lane = threadIdx.x & 63;
	const bool valid = i < nq;
	uint32_t n0 = ~0u, n1 = ~0u, n2 = ~0u, n3 = ~0u, n4 = ~0u, n5 = ~0u;
	if (valid) {
		uint4 a = *reinterpret_cast<const uint4 *>(q + (size_t)i * 8);
		uint2 b = *reinterpret_cast<const uint2 *>(q + (size_t)i * 8 + 4);
		n0 = ~a.x; n1 = ~a.y; n2 = ~a.z; n3 = ~a.w; n4 = ~b.x; n5 = ~b.y;
	}
	for (uint32_t p = 0; p < P; ++p) {
		const uint32_t * m = part_info + (size_t)p * 8;
		uint32_t t = (m[0] & n0) | (m[1] & n1) | (m[2] & n2) | (m[3] & n3) | (m[4] & n4) | (m[5] & n5);
		const bool sel = valid && t == 0;
		uint64_t b = __ballot(sel);
		if (b != 0) {
			unsigned long long base = 0;
			if (lane == 0) base = atomicAdd(&pcur[p], (unsigned long long)__popcll(b));
			base = __shfl(base, 0);
			if (sel) {
				unsigned long long pos = base + mbcnt64(b);
				if (pos < list_cap) list[pos] = i;
			}
		}
	}
}

// ------------------------------------------------------------- work items

// Single workgroup.  For each schedule rank r (partition porder[r]): clamp
// its query count to the list capacity, cut it into batches of bq queries
// and its blocks into chunks of `chunk`; items of rank r are
// [item_prefix[r], item_prefix[r+1]), ordered chunk-major so that items
// taken one after another share filter blocks.  The item range is then cut
// into NQUEUE contiguous queues.
__global__ void __launch_bounds__(1024)
k_build_items(const uint32_t * __restrict__ pcount, const uint64_t * __restrict__ pbase,
			  uint64_t list_cap, const uint32_t * __restrict__ part_info, uint32_t P,
			  const uint32_t * __restrict__ porder, uint32_t bq, uint32_t chunk,
			  uint4 * __restrict__ sched, uint64_t * __restrict__ item_prefix,
			  uint64_t * __restrict__ qbeg, unsigned long long * __restrict__ qnext) {
	__shared__ uint64_t s_w[16];
	const uint32_t tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
	uint64_t carry = 0;
	for (uint32_t base = 0; base < P; base += 1024) {
		uint32_t r = base + tid;
		uint64_t items = 0;
		if (r < P) {
			uint32_t p = porder ? porder[r] : r;
			uint64_t c = pcount[p], b = pbase[p];
			if (b >= list_cap) c = 0;
			else if (b + c > list_cap) c = list_cap - b;
			uint32_t nbat = (uint32_t)((c + bq - 1) / bq);
			uint32_t nblk = part_info[(size_t)p * 8 + 7];
			uint32_t nch = (nblk + chunk - 1) / chunk;
			items = (uint64_t)nbat * nch;
			sched[r] = make_uint4(p, nbat, nch, (uint32_t)c);
		}
		uint64_t x = items;
		for (int o = 1; o < 64; o <<= 1) {
			uint64_t y = __shfl_up(x, o, 64);
			if (lane >= o) x += y;
		}
		if (lane == 63) s_w[w] = x;
		__syncthreads();
		uint64_t pre = 0, tot = 0;
		for (int k = 0; k < 16; ++k) {
			if (k < (int)w) pre += s_w[k];
			tot += s_w[k];
		}
		if (r < P) item_prefix[r] = carry + pre + x - items;
		carry += tot;
		__syncthreads();
	}
	if (tid == 0) item_prefix[P] = carry;
	if (tid <= NQUEUE) qbeg[tid] = carry * tid / NQUEUE;
	if (tid < NQUEUE) qnext[tid * QPAD] = carry * tid / NQUEUE;
}

// ------------------------------------------------------------- match

// The overflow list a slot points to, bounded by the table's overflow words:
// whatever a slot holds, nothing outside the table is read (n = 0 then).
__device__ inline uint32_t ovf_list(uint64_t kv, const uint32_t * __restrict__ ovf, uint32_t ovf_n, uint32_t & off) {
	off = (uint32_t)(kv >> 16);
	if (off >= ovf_n) return 0;
	const uint32_t n = ovf[off];
	return n < ovf_n - off ? n : 0;
}

// Number of keys in a slot (inline ids or an overflow list).
__device__ inline uint32_t key_count(uint64_t kv, const uint32_t * __restrict__ ovf, uint32_t ovf_n) {
	uint32_t off;
	if ((kv & 0xFFFFu) == KEY_NONE) return ovf_list(kv, ovf, ovf_n, off);
	uint32_t n = 0;
	for (int x = 0; x < 4; ++x) n += ((kv >> (16 * x)) & 0xFFFFu) != KEY_NONE;
	return n;
}

// ------------------------------------------------------------- query-major match

// Persistent schedule.  A work item is a group of NQ = 64*TILES consecutive
// queries of the submission, owned by one workgroup for the whole match:
//
//   for each partition p (schedule order):
//     lookup, fused: qsel = queries of the group whose bits cover p's mask
//       (reference front_end.cc:465-477), one 64-lane test per tile;
//     for each 64-filter block of p (blocks split over the waves):
//       candidate pass: cand = qsel & (block AND-mask covered by query);
//       subset test of each candidate query against the 64 filters, two
//       candidates per iteration;
//       hits: OR the filters' dense key ids into the group's key bitmaps in
//       LDS (KEYS_UNIQUE), or stage (query, slot) pairs per wave (HITS).
//
// Because a group's results never leave its workgroup until the end, the
// bitmaps are written out once with plain stores: no global atomics on
// results.  The reference instead merges partial results on the host under
// a per-query mutex (back_end.cc:288-301, tagmatch_query.hh:80-93).

struct mq_args {
	dev_fib_view f;
	const uint32_t * __restrict__ q;          // nq x 8 words
	const uint32_t * __restrict__ qinv;       // nq x 8 inverted words (VAR 1)
	uint32_t nq;
	const uint32_t * __restrict__ porder;     // partition schedule order (NULL: 0..P-1)
	const uint32_t * __restrict__ iorder;     // query tile of each work item (NULL: identity; nseg 1 only)
	const uint32_t * __restrict__ qmap;       // query of each local index (NULL: identity)
	const uint32_t * __restrict__ units;      // explicit units, 8 words each: partition,
	                                          // first query, queries, tiles, segments,
	                                          // blocks per segment, first block, end block
	const uint32_t * __restrict__ uprefix;    // n_units + 1: first work item of each unit
	uint32_t n_units;
	uint32_t blk_lo, blk_hi;                  // blocks the launch covers
	int32_t log_unit;                         // >= 0: add hits, keys, ticks into this log row
	uint32_t ntiles;                          // query tiles
	uint32_t nseg;                            // block segments per tile: item = seg * ntiles + tile
	uint32_t seg_blocks;                      // blocks per segment (nseg > 1)
	unsigned long long * qnext;               // NQUEUE queue cursors, QPAD apart
	const uint64_t * __restrict__ qbeg;       // NQUEUE + 1 queue bounds (items)
	uint32_t * pcount;                        // per partition: queries selecting it
	uint32_t * gbm;                           // KEYS_UNIQUE: nq x kw bitmap words
	uint2 * hits;                             // HITS: (query, block slot)
	uint64_t hit_cap;
	unsigned long long * counters;            // [0] hit slots, [1] hits, [2] block visits,
	                                          // [3] subset tests, [4] tests with hits
	uint32_t * log;                           // per item: hits, 0, 10 ns ticks
	uint64_t log_cap;
	const uint32_t * __restrict__ slot_epoch; // lane insert epochs (NULL: every valid lane counts)
	uint32_t epoch;                           // lanes with a later epoch are not seen
	uint32_t full_exit;                       // KEYS_UNIQUE + LDS bitmaps: skip queries whose
	                                          // bitmap holds every key
	uint32_t groups;                          // walk blocks in groups, skipping a group whose
	                                          // AND mask no query holds (compacted steps)
	uint32_t last_mask;                       // valid bits of a bitmap row's last word
	uint32_t * qfull;                         // nseg > 1 with full_exit: per query, set once its
	                                          // merged gbm row holds every key
	const uint32_t * __restrict__ segb;       // nseg + 1 segment bounds (NULL: seg_blocks each)
	const uint32_t * __restrict__ nq_dev;     // device query count (NULL: nq), with
	const uint32_t * __restrict__ ntiles_dev; // device tile count, segments, blocks per segment
	                                          // (NULL: ntiles, nseg, seg_blocks)
	uint32_t map_rows;                        // KEYS_UNIQUE: local query i is row and flag qmap[i]
	const uint32_t * __restrict__ step_vals;  // replayed steps: [0] epoch, [1] valid overflow words,
	                                          // written before each replay (NULL: epoch, f.ovf_n)
};

constexpr int HBUF = 256;                     // staged hits per wave (HITS mode)
constexpr uint32_t COOP_KEYS = 16;            // overflow lists longer than this: walked by the whole wave

// Dense key ids of a hit filter into a bitmap row; `kv` inline form has up to
// four 16-bit ids, KEY_NONE for an unused field (ORs 0 into word 0).
template <bool LDS_BM>
__device__ inline void or_keys(uint32_t * row, uint64_t kv, const uint32_t * __restrict__ ovf, uint32_t ovf_n,
							   uint32_t n_keys) {
	if ((kv & 0xFFFFu) != KEY_NONE) {
#pragma unroll
		for (int i = 0; i < 4; ++i) {
			const uint32_t id = (uint32_t)(kv >> (16 * i)) & 0xFFFFu;
			const bool ok = id != KEY_NONE;
			const uint32_t w = ok ? (id >> 5) : 0u, bit = ok ? (1u << (id & 31)) : 0u;
			if (LDS_BM) atomicOr(&row[w], bit);
			else __hip_atomic_fetch_or(&row[w], bit, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
		}
	} else {
		uint32_t off;
		const uint32_t n = ovf_list(kv, ovf, ovf_n, off);
		for (uint32_t i = 0; i < n; ++i) {
			const uint32_t id = ovf[off + 1 + i];
			if (id >= n_keys) continue;
			if (LDS_BM) atomicOr(&row[id >> 5], 1u << (id & 31));
			else __hip_atomic_fetch_or(&row[id >> 5], 1u << (id & 31), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
		}
	}
}

// Uniform loads through the constant address space compile to scalar
// loads (s_load) into SGPRs; used for data that no thread writes during the
// kernel (FIB block info, partition info, inverted query words).
typedef const __attribute__((address_space(4))) uint32_t cu32_t;

__device__ inline uint4 sld4(const uint32_t * p) {
	cu32_t * c = (cu32_t *)(p);
	return make_uint4(c[0], c[1], c[2], c[3]);
}
__device__ inline uint2 sld2(const uint32_t * p) {
	cu32_t * c = (cu32_t *)(p);
	return make_uint2(c[0], c[1]);
}
__device__ inline uint32_t sld1(const uint32_t * p) { return *(cu32_t *)(p); }

// (f0 & q0) | ... | (f5 & q5): one v_and_b32 and five v_and_or_b32 in one
// asm block.  Left to itself the compiler balances the OR tree into
// v_or3_b32 and plain ANDs, eight instructions for the same value, in the
// subset tests that run for every candidate query of every visited block;
// separate asm statements would each be followed by a wait state.
__device__ inline uint32_t test6(uint32_t f0, uint32_t f1, uint32_t f2, uint32_t f3, uint32_t f4, uint32_t f5,
								 uint32_t q0, uint32_t q1, uint32_t q2, uint32_t q3, uint32_t q4, uint32_t q5) {
	uint32_t acc;
	asm("v_and_b32 %0, %1, %7\n\t"
		"v_and_or_b32 %0, %2, %8, %0\n\t"
		"v_and_or_b32 %0, %3, %9, %0\n\t"
		"v_and_or_b32 %0, %4, %10, %0\n\t"
		"v_and_or_b32 %0, %5, %11, %0\n\t"
		"v_and_or_b32 %0, %6, %12, %0"
		: "=&v"(acc)
		: "v"(f0), "v"(f1), "v"(f2), "v"(f3), "v"(f4), "v"(f5), "v"(q0), "v"(q1), "v"(q2), "v"(q3), "v"(q4), "v"(q5));
	return acc;
}
// The same with uniform masks (scalar registers) against per-lane words.
__device__ inline uint32_t test6s(const uint4 & ma, const uint4 & mb, const uint32_t (&q)[6]) {
	uint32_t acc;
	asm("v_and_b32 %0, %1, %7\n\t"
		"v_and_or_b32 %0, %2, %8, %0\n\t"
		"v_and_or_b32 %0, %3, %9, %0\n\t"
		"v_and_or_b32 %0, %4, %10, %0\n\t"
		"v_and_or_b32 %0, %5, %11, %0\n\t"
		"v_and_or_b32 %0, %6, %12, %0"
		: "=&v"(acc)
		: "s"(ma.x), "s"(ma.y), "s"(ma.z), "s"(ma.w), "s"(mb.x), "s"(mb.y), "v"(q[0]), "v"(q[1]), "v"(q[2]),
		  "v"(q[3]), "v"(q[4]), "v"(q[5]));
	return acc;
}

// True when a bitmap row in LDS holds every key.  Each lane starts at word
// lane % kw: rows are kw words apart, so lanes that started at word 0 would
// all read one LDS bank.
__device__ inline bool row_full(const uint32_t * row, uint32_t kw, uint32_t lane, uint32_t last_mask) {
	uint32_t w = lane % kw;
	for (uint32_t i = 0; i < kw; ++i) {
		uint32_t v = row[w];
		if (w + 1 == kw) v |= ~last_mask;
		if (v != 0xFFFFFFFFu) return false;
		w = w + 1 == kw ? 0 : w + 1;
	}
	return true;
}

// Per-wave hit sink of the query-major kernel.
//
// KEYS_UNIQUE with KLOG: a hit appends (bitmap row, packed key ids) to this
// wave's 64-entry log in LDS with one plain store per hitting lane; a full
// log is drained with one entry per lane, so each LDS OR instruction carries
// up to 64 lanes instead of the two or three lanes of a single test.
template <int MODE, bool LDS_BM, bool KLOG>
struct hit_sink {
	uint32_t * bm_base;       // MODE 0: bitmap rows of the group's queries
	uint32_t kw;
	uint2 * hb;               // MODE 1: this wave's staging buffer
	uint4 * klog;             // KLOG: this wave's key log (64 entries)
	uint32_t hcnt;
	const uint32_t * ovf;
	uint32_t ovf_n, n_keys;
	uint32_t last_mask;       // valid bits of a row's last word
	unsigned long long * counter;
	uint2 * hits;
	uint64_t cap;

	// Inline ids and short overflow lists: one entry per lane.  Long lists
	// (filters that merged many duplicates carry up to every key of the
	// table) would keep one lane busy for thousands of iterations while the
	// rest of the wave waits: the wave walks them together, 64 ids at a time.
	__device__ inline void drain(uint32_t lane) {
		uint4 e = make_uint4(0, 0, 0, 0);
		bool coop = false;
		if (lane < hcnt) {
			e = klog[lane];
			const uint64_t kv = (uint64_t)e.y | ((uint64_t)e.z << 32);
			uint32_t off;
			coop = (kv & 0xFFFFu) == KEY_NONE && ovf_list(kv, ovf, ovf_n, off) > COOP_KEYS;
			if (!coop) or_keys<LDS_BM>(bm_base + (size_t)e.x * kw, kv, ovf, ovf_n, n_keys);
		}
		uint64_t m = __ballot(coop);
		while (m) {
			const uint32_t j = (uint32_t)__builtin_ctzll(m);
			m &= m - 1;
			const uint32_t jrow = __shfl(e.x, j), jlo = __shfl(e.y, j), jhi = __shfl(e.z, j);
			uint32_t off;
			const uint32_t n = ovf_list((uint64_t)jlo | ((uint64_t)jhi << 32), ovf, ovf_n, off);
			uint32_t * row = bm_base + (size_t)jrow * kw;
			// a row that already holds every key gains nothing (rows only grow)
			uint32_t all = 0xFFFFFFFFu;
			for (uint32_t w = lane; w < kw; w += 64) all &= row[w] | (w + 1 == kw ? ~last_mask : 0u);
			if (__ballot(all != 0xFFFFFFFFu) == 0) continue;
#pragma unroll 4
			for (uint32_t i = lane; i < n; i += 64) {
				const uint32_t id = ovf[off + 1 + i];
				if (id < n_keys) {
					if (LDS_BM) atomicOr(&row[id >> 5], 1u << (id & 31));
					else __hip_atomic_fetch_or(&row[id >> 5], 1u << (id & 31), __ATOMIC_RELAXED,
											   __HIP_MEMORY_SCOPE_WORKGROUP);
				}
			}
		}
		hcnt = 0;
	}

	__device__ inline void flush(uint32_t lane) {
		unsigned long long base = 0;
		if (lane == 0) base = atomicAdd(counter, (unsigned long long)hcnt);
		base = __shfl(base, 0);
		for (uint32_t i = lane; i < hcnt; i += 64)
			if (base + i < cap) hits[base + i] = hb[i];
		hcnt = 0;
	}

	// h: lanes (filters of block b) matched by group query jq (index qi)
	__device__ inline void add(uint64_t h, uint32_t jq, uint32_t qi, uint32_t b, uint64_t kv, uint32_t lane) {
		const bool mine = (h >> lane) & 1;
		if (MODE == 0 && KLOG) {
			const uint32_t n = (uint32_t)__popcll(h);
			if (hcnt + n > 64u) drain(lane);
			if (mine) klog[hcnt + mbcnt64(h)] = make_uint4(jq, (uint32_t)kv, (uint32_t)(kv >> 32), 0u);
			hcnt += n;
		} else if (MODE == 0) {
			if (mine) or_keys<LDS_BM>(bm_base + (size_t)jq * kw, kv, ovf, ovf_n, n_keys);
		} else {
			const uint32_t n = (uint32_t)__popcll(h);
			if (hcnt + n > (uint32_t)HBUF) flush(lane);
			if (mine) hb[hcnt + mbcnt64(h)] = make_uint2(qi, b * 64 + lane);
			hcnt += n;
		}
	}
};

// MODE 0: key bitmaps (LDS_BM: in LDS, else in gbm with workgroup-scope
// atomics); MODE 1: hit pairs.  VAR 0: candidate query words from LDS, two
// tests per iteration; VAR 1: from pre-inverted query words in HBM through
// scalar loads, four independent tests per iteration.
// The default one-tile, 4-wave instance (large and compacted steps) is held
// to 8 waves per SIMD (64 VGPRs, a few spilled outside the test loop): at
// 10M filters its compacted launch took 2.25 instead of 2.32 ms and the
// every-match step 642 instead of 675 ms.  The 8- and 16-wave instances of
// small steps are left free: held the same way, 1024-query serving steps
// took 0.19 instead of 0.16 ms.
template <int MODE, bool LDS_BM, int TILES, int NW, int VAR>
__global__ void __launch_bounds__(NW * 64)
__attribute__((amdgpu_waves_per_eu(TILES == 1 && VAR == 2 && NW == 4 ? 8 : 1)))
k_match_q(mq_args a) {
	constexpr int NQ = TILES * 64;
	constexpr int NT = NW * 64;
	constexpr bool KLOG = MODE == 0 && (VAR & 2);
	__shared__ __align__(16) uint32_t s_q[NQ * 8];
	__shared__ uint32_t s_item[2];
	__shared__ uint32_t s_hits, s_keys;
	__shared__ uint2 s_hbuf[MODE == 1 ? NW * HBUF : 1];
	__shared__ uint4 s_klog[KLOG ? NW * 64 : 1];
	extern __shared__ uint32_t s_bm[];        // LDS_BM: NQ x kw
	const uint32_t tid = threadIdx.x, lane = tid & 63;
	const uint32_t wave = __builtin_amdgcn_readfirstlane(tid >> 6);
	const uint32_t kw = a.f.kw;
	const uint32_t home = blockIdx.x % NQUEUE;
	const uint32_t P = a.f.n_parts;
	// statistics of this wave over every item it works on, added to the
	// global counters once at the end (all waves adding to one line per
	// item serializes small launches)
	uint64_t w_hits = 0, w_visit = 0, w_test = 0, w_hev = 0;
	bool bm_dirty = true;   // LDS bitmap rows hold bits of an earlier item (uniform)
	// counts written by earlier kernels of the step (compacted submissions)
	const uint32_t nq_all = a.nq_dev ? sld1(a.nq_dev) : a.nq;
	const uint32_t ntiles_all = a.ntiles_dev ? sld1(a.ntiles_dev) : a.ntiles;
	const uint32_t nseg_all = a.ntiles_dev ? sld1(a.ntiles_dev + 1) : a.nseg;
	// the inserts the step sees: from memory when a captured graph replays it
	const uint32_t epoch = a.step_vals ? sld1(a.step_vals) : a.epoch;
	const uint32_t ovf_n = a.step_vals ? sld1(a.step_vals + 1) : a.f.ovf_n;
	const uint32_t seg_blocks_all = a.ntiles_dev ? sld1(a.ntiles_dev + 2) : a.seg_blocks;
	// bitmap row (and flag) of local query i
	auto row_of = [&](uint32_t i) -> uint32_t { return a.map_rows ? a.qmap[i] : i; };

	for (;;) {
		uint64_t t_start = 0;
		if (tid == 0) {
			unsigned long long item = ~0ULL;
			for (int k = 0; k < NQUEUE; ++k) {
				const uint32_t x = (home + k) % NQUEUE;
				if (a.qnext[x * QPAD] >= a.qbeg[x + 1]) continue;
				const unsigned long long i = atomicAdd(&a.qnext[x * QPAD], 1ULL);
				if (i < a.qbeg[x + 1]) { item = i; break; }
			}
			s_item[0] = (uint32_t)item;
			s_item[1] = item == ~0ULL;
			s_hits = 0;
			s_keys = 0;
			t_start = wall_clock64();
		}
		__syncthreads();
		if (s_item[1]) break;
		const uint32_t item0 = __builtin_amdgcn_readfirstlane(s_item[0]);
		// A work item is a query tile, or with nseg > 1 a tile restricted to
		// one contiguous range of blocks (small submissions spread over the
		// whole GPU); segment-major, so concurrent items share filter blocks.
		uint32_t item, seg_b0 = 0, seg_b1 = 0xFFFFFFFFu, r_begin = 0, r_end = P;
		uint32_t q0, q_end = nq_all;
		if (a.units) {
			// explicit (partition, query range) units: the item's unit by
			// binary search over the item prefix; segment-major in the unit
			uint32_t lo = 0, hi = a.n_units;
			while (hi - lo > 1) {
				const uint32_t mid = (lo + hi) >> 1;
				if (sld1(a.uprefix + mid) <= item0) lo = mid; else hi = mid;
			}
			const uint4 u0 = sld4(a.units + (size_t)lo * 8);       // part, first query, queries, tiles
			const uint4 u1 = sld4(a.units + (size_t)lo * 8 + 4);   // segments, seg blocks, first, end block
			const uint32_t local = item0 - sld1(a.uprefix + lo);
			const uint32_t seg = local / u0.w;
			item = local - seg * u0.w;
			q0 = u0.y + item * NQ;
			q_end = u0.y + u0.z;
			seg_b0 = u1.z + seg * u1.y;
			seg_b1 = min(seg_b0 + u1.y, u1.w);
			r_begin = u0.x;
			r_end = seg_b1 > seg_b0 ? u0.x + 1 : u0.x;
		} else if (nseg_all > 1 || a.blk_lo != 0 || a.blk_hi < a.f.n_blocks) {
			const uint32_t seg = item0 / ntiles_all;
			item = item0 - seg * ntiles_all;
			if (a.segb) {   // segments of equal estimated work
				seg_b0 = sld1(a.segb + seg);
				seg_b1 = sld1(a.segb + seg + 1);
			} else {
				seg_b0 = a.blk_lo + seg * seg_blocks_all;
				seg_b1 = min(seg_b0 + seg_blocks_all, a.blk_hi);
			}
			if (seg_b1 > seg_b0) {
				r_begin = sld1(a.f.blk_info + (size_t)seg_b0 * 8 + 7);
				r_end = sld1(a.f.blk_info + (size_t)(seg_b1 - 1) * 8 + 7) + 1;
			} else {
				r_begin = r_end = 0;
			}
		} else {
			item = a.iorder ? sld1(a.iorder + item0) : item0;
		}
		if (!a.units) q0 = item * NQ;

		for (uint32_t t = tid; t < (uint32_t)NQ; t += NT) {
			uint32_t * d = &s_q[t * 8];
			const uint32_t qi = q0 + t;
			if (qi < q_end) {
				const uint32_t gq = a.qmap ? a.qmap[qi] : qi;
				const uint4 x = *reinterpret_cast<const uint4 *>(a.q + (size_t)gq * 8);
				const uint2 y = *reinterpret_cast<const uint2 *>(a.q + (size_t)gq * 8 + 4);
				d[0] = ~x.x; d[1] = ~x.y; d[2] = ~x.z; d[3] = ~x.w; d[4] = ~y.x; d[5] = ~y.y;
				d[6] = qi;
			} else {
				d[0] = d[1] = d[2] = d[3] = d[4] = d[5] = 0;
				d[6] = NO_QUERY;
			}
		}
		hit_sink<MODE, LDS_BM, KLOG> sink;
		sink.klog = &s_klog[KLOG ? wave * 64 : 0];
		sink.kw = kw;
		sink.hcnt = 0;
		sink.ovf = a.f.ovf;
		sink.ovf_n = ovf_n;
		sink.n_keys = a.f.n_keys;
		sink.last_mask = a.last_mask;
		sink.counter = &a.counters[0];
		sink.hits = a.hits;
		sink.cap = a.hit_cap;
		sink.hb = &s_hbuf[MODE == 1 ? wave * HBUF : 0];
		sink.bm_base = nullptr;
		if (MODE == 0) {
			if (LDS_BM) {
				if (bm_dirty)   // an item without hits leaves the rows at zero
					for (uint32_t w = tid; w < (uint32_t)NQ * kw; w += NT) s_bm[w] = 0;
				bm_dirty = false;
				sink.bm_base = s_bm;
			} else {
				sink.bm_base = a.gbm + (size_t)q0 * kw;
				const uint32_t nw = (min(q_end, q0 + NQ) - q0) * kw;
				for (uint32_t w = tid; w < nw; w += NT) sink.bm_base[w] = 0;
			}
		}
		__syncthreads();

		uint32_t qv[TILES][6];
		uint64_t qok[TILES];
#pragma unroll
		for (int t = 0; t < TILES; ++t) {
			const uint32_t * d = &s_q[(t * 64 + lane) * 8];
#pragma unroll
			for (int j = 0; j < 6; ++j) qv[t][j] = d[j];
			qok[t] = __ballot(d[6] != NO_QUERY);
		}

		// full-key-set flags of the item's queries, reloaded at a partition
		// only while a selected query is not known to be full
		uint64_t qfullm[TILES];
		const bool use_flags = MODE == 0 && LDS_BM && a.full_exit && a.qfull;
		auto load_flags = [&](int t) -> uint64_t {
			const uint32_t qrow = q0 + t * 64 + lane;
			return __ballot(qrow < q_end && __hip_atomic_load(&a.qfull[row_of(qrow)], __ATOMIC_RELAXED,
															  __HIP_MEMORY_SCOPE_AGENT) != 0u);
		};
#pragma unroll
		for (int t = 0; t < TILES; ++t) qfullm[t] = use_flags ? load_flags(t) : 0ULL;
		uint32_t hit_acc = 0, n_visit = 0, n_test = 0, n_hev = 0, key_acc = 0;
		// a hit event: lanes h of block b matched local query jl of the item
		auto on_hit = [&](uint64_t h, uint32_t jl, uint32_t b, uint64_t kv) {
			++n_hev;
			hit_acc += (uint32_t)__popcll(h);
			const uint32_t ql = q0 + jl;
			const uint32_t gq = (MODE == 1 && a.qmap) ? sld1(a.qmap + ql) : ql;
			if (a.log) {   // uniform: a scalar branch around the per-lane count
				if ((h >> lane) & 1) {
					// the empty asm keeps the count's arithmetic inside the
					// branch: hoisted, it cost every block visit about 17
					// VALU instructions in steps that log nothing
					uint64_t kx = kv;
					asm volatile("" : "+v"(kx));
					key_acc += key_count(kx, a.f.ovf, ovf_n);
				}
			}
			sink.add(h, jl, gq, b, kv, lane);
		};
		for (uint32_t r = r_begin; r < r_end; ++r) {
			const uint32_t p = a.porder && nseg_all <= 1 ? sld1(a.porder + r) : r;
			const uint32_t * pi = a.f.part_info + (size_t)p * 8;
			const uint4 pa = sld4(pi);
			const uint4 pb = sld4(pi + 4);
			// this item's share of the partition's blocks
			const uint32_t fb = max(pb.z, seg_b0), fe = min(pb.z + pb.w, seg_b1);
			const uint32_t nb = fe > fb ? fe - fb : 0;
			const bool counts_here = pb.z >= seg_b0;   // count each partition once
			uint64_t qsel[TILES];
			uint64_t anyq = 0;
			uint32_t nsel = 0;
#pragma unroll
			for (int t = 0; t < TILES; ++t) {
				const uint32_t x = test6s(pa, pb, qv[t]);
				qsel[t] = __ballot(x == 0) & qok[t];
				anyq |= qsel[t];
				nsel += (uint32_t)__popcll(qsel[t]);
			}
			if (wave == 0 && lane == 0 && a.pcount && counts_here && nsel) atomicAdd(&a.pcount[p], nsel);
			if (anyq == 0 || nb == 0) continue;
			if (MODE == 0 && LDS_BM && a.full_exit) {
				// A query whose key bitmap already holds every key of the
				// dictionary cannot gain anything: drop it from the rest of
				// the item.  Bitmaps only grow, so a stale read only delays
				// the drop.
				anyq = 0;
#pragma unroll
				for (int t = 0; t < TILES; ++t) {
					if (qsel[t] == 0) continue;
					// full: words 0..kw-2 all ones, word kw-1 holds the last K % 32 keys.
					// With block segments the other segments of this tile merge
					// their rows into gbm and flag a query whose merged row is
					// full (one coalesced load per tile here); in one segment
					// the LDS row is the query's whole bitmap.
					// Flags also mark the queries whose rows the rich-filter pass
					// (k_rich_rows) already filled.
					if (use_flags && (qsel[t] & ~qfullm[t])) qfullm[t] |= load_flags(t);
					qsel[t] &= ~qfullm[t];
					if (nseg_all <= 1 && qsel[t])
						qsel[t] &= ~__ballot(row_full(&s_bm[(t * 64 + lane) * kw], kw, lane, a.last_mask));
					anyq |= qsel[t];
				}
				if (anyq == 0) continue;
			}
			// Subset tests of block b's filters (words cf, key slots kv, nv valid
			// lanes) against its candidate queries cand.
			auto test_block = [&](uint32_t b, const uint32_t (&cf)[6], uint64_t kv, const uint64_t (&cand)[TILES],
								  uint32_t nv) {
				const uint32_t f0 = cf[0], f1 = cf[1], f2 = cf[2], f3 = cf[3], f4 = cf[4], f5 = cf[5];
				uint64_t lmask = nv >= 64 ? ~0ULL : ((1ULL << nv) - 1ULL);
				if (a.slot_epoch) lmask &= __ballot(a.slot_epoch[(size_t)b * 64 + lane] <= epoch);
#pragma unroll
				for (int t = 0; t < TILES; ++t) {
					uint64_t c = cand[t];
					n_test += (uint32_t)__popcll(c);
					if ((VAR & 5) == 0) {
						// candidate query words broadcast from LDS, two per iteration:
						// the lowest and the highest candidate left (one and the same
						// for the last of an odd count).  Scalar work per pair kept
						// small (the scalar unit is as busy as the vector one here):
						// bits cleared by index, and one branch over both results for
						// the common pair without hits.
						while (c) {
							const uint32_t j0 = (uint32_t)__builtin_ctzll(c);
							const uint32_t j1 = 63u - (uint32_t)__builtin_clzll(c);
							c &= ~((1ULL << j0) | (1ULL << j1));
							const bool has1 = j1 != j0;
							const uint32_t * qq0 = &s_q[(t * 64 + j0) * 8];
							const uint32_t * qq1 = &s_q[(t * 64 + j1) * 8];
							const uint4 a0 = *reinterpret_cast<const uint4 *>(qq0);
							const uint2 c0 = *reinterpret_cast<const uint2 *>(qq0 + 4);
							const uint4 a1 = *reinterpret_cast<const uint4 *>(qq1);
							const uint2 c1 = *reinterpret_cast<const uint2 *>(qq1 + 4);
							const uint32_t acc0 = test6(f0, f1, f2, f3, f4, f5, a0.x, a0.y, a0.z, a0.w, c0.x, c0.y);
							const uint32_t acc1 = test6(f0, f1, f2, f3, f4, f5, a1.x, a1.y, a1.z, a1.w, c1.x, c1.y);
							const uint64_t r0 = __ballot(acc0 == 0), r1 = __ballot(acc1 == 0);
							if (((r0 | r1) & lmask) == 0) continue;   // (r1 = r0 when j1 = j0)
							const uint64_t h0 = r0 & lmask;
							const uint64_t h1 = has1 ? r1 & lmask : 0ULL;
							if (h0) on_hit(h0, t * 64 + j0, b, kv);
							if (h1) on_hit(h1, t * 64 + j1, b, kv);
						}
					} else if ((VAR & 1) == 0) {
						// early exit variant: words 0-1 first, two per iteration
						while (c) {
							const uint32_t j0 = (uint32_t)__builtin_ctzll(c);
							c &= c - 1;
							const uint64_t keep1 = c ? ~0ULL : 0ULL;
							const uint32_t j1 = c ? (uint32_t)__builtin_ctzll(c) : j0;
							c &= c - 1;
							const uint32_t * qq0 = &s_q[(t * 64 + j0) * 8];
							const uint32_t * qq1 = &s_q[(t * 64 + j1) * 8];
							uint32_t acc0, acc1;
							if (VAR & 4) {
								const uint2 a0 = *reinterpret_cast<const uint2 *>(qq0);
								const uint2 a1 = *reinterpret_cast<const uint2 *>(qq1);
								acc0 = (f0 & a0.x) | (f1 & a0.y);
								acc1 = (f0 & a1.x) | (f1 & a1.y);
								// early exit: words 0-1 already reject every filter
								// for both candidates (the common case)
								const uint64_t alive = (__ballot(acc0 == 0) | (__ballot(acc1 == 0) & keep1)) & lmask;
								if (alive == 0) continue;
								const uint2 b0 = *reinterpret_cast<const uint2 *>(qq0 + 2);
								const uint2 c0 = *reinterpret_cast<const uint2 *>(qq0 + 4);
								const uint2 b1 = *reinterpret_cast<const uint2 *>(qq1 + 2);
								const uint2 c1 = *reinterpret_cast<const uint2 *>(qq1 + 4);
								acc0 |= (f2 & b0.x) | (f3 & b0.y) | (f4 & c0.x) | (f5 & c0.y);
								acc1 |= (f2 & b1.x) | (f3 & b1.y) | (f4 & c1.x) | (f5 & c1.y);
							} else {
								const uint4 a0 = *reinterpret_cast<const uint4 *>(qq0);
								const uint2 c0 = *reinterpret_cast<const uint2 *>(qq0 + 4);
								const uint4 a1 = *reinterpret_cast<const uint4 *>(qq1);
								const uint2 c1 = *reinterpret_cast<const uint2 *>(qq1 + 4);
								acc0 = (f0 & a0.x) | (f1 & a0.y) | (f2 & a0.z) | (f3 & a0.w) | (f4 & c0.x) | (f5 & c0.y);
								acc1 = (f0 & a1.x) | (f1 & a1.y) | (f2 & a1.z) | (f3 & a1.w) | (f4 & c1.x) | (f5 & c1.y);
							}
							const uint64_t h0 = __ballot(acc0 == 0) & lmask;
							const uint64_t h1 = __ballot(acc1 == 0) & lmask & keep1;
							if (h0) on_hit(h0, t * 64 + j0, b, kv);
							if (h1) on_hit(h1, t * 64 + j1, b, kv);
						}
					} else {
						// candidate query words through scalar loads, two per iteration
						const uint32_t * qbase = a.qinv + (size_t)(q0 + t * 64) * 8;
						while (c) {
							const uint32_t j0 = (uint32_t)__builtin_ctzll(c);
							c &= c - 1;
							const uint64_t keep1 = c ? ~0ULL : 0ULL;
							const uint32_t j1 = c ? (uint32_t)__builtin_ctzll(c) : j0;
							c &= c - 1;
							const uint2 a0 = sld2(qbase + j0 * 8), a1 = sld2(qbase + j1 * 8);
							uint32_t acc0 = (f0 & a0.x) | (f1 & a0.y);
							uint32_t acc1 = (f0 & a1.x) | (f1 & a1.y);
							if (VAR & 4) {
								const uint64_t alive = (__ballot(acc0 == 0) | (__ballot(acc1 == 0) & keep1)) & lmask;
								if (alive == 0) continue;
							}
							const uint4 b0 = sld4(qbase + j0 * 8 + 2), b1 = sld4(qbase + j1 * 8 + 2);
							acc0 |= (f2 & b0.x) | (f3 & b0.y) | (f4 & b0.z) | (f5 & b0.w);
							acc1 |= (f2 & b1.x) | (f3 & b1.y) | (f4 & b1.z) | (f5 & b1.w);
							const uint64_t h0 = __ballot(acc0 == 0) & lmask;
							const uint64_t h1 = __ballot(acc1 == 0) & lmask & keep1;
							if (h0) on_hit(h0, t * 64 + j0, b, kv);
							if (h1) on_hit(h1, t * 64 + j1, b, kv);
						}
					}
				}
			};
			// candidate queries of a block from its AND mask (ma, mb: info words 0-7)
			auto block_cands = [&](const uint4 & ma, const uint4 & mb, uint64_t (&cand)[TILES]) -> uint64_t {
				uint64_t any = 0;
#pragma unroll
				for (int t = 0; t < TILES; ++t) {
					if (qsel[t] == 0) { cand[t] = 0; continue; }
					const uint32_t x = test6s(ma, mb, qv[t]);
					cand[t] = __ballot(x == 0) & qsel[t];
					any |= cand[t];
				}
				return any;
			};
			const uint32_t be = fb + nb;
			if (VAR & 16) {
				// GB blocks per round: their info with one scalar-load wait, then
				// the words of every block that has candidates with one vector
				// wait, then the tests; the loads of a round overlap.
				constexpr int GB = 4;
				for (uint32_t b0 = fb + wave; b0 < be; b0 += NW * GB) {
					uint64_t cand[GB][TILES];
					uint64_t any[GB];
					uint32_t nv[GB];
#pragma unroll
					for (int g = 0; g < GB; ++g) {
						// past the end: read the last block, use nothing
						const uint32_t b = min(b0 + g * NW, be - 1);
						const uint4 ma = sld4(a.f.blk_info + (size_t)b * 8);
						const uint4 mb = sld4(a.f.blk_info + (size_t)b * 8 + 4);
						nv[g] = mb.z;
						any[g] = block_cands(ma, mb, cand[g]);
						if (b0 + g * NW >= be) any[g] = 0;
					}
					uint64_t anyall = 0;
#pragma unroll
					for (int g = 0; g < GB; ++g) anyall |= any[g];
					if (anyall == 0) continue;
					uint32_t cf[GB][6];
					uint64_t kv[GB];
#pragma unroll
					for (int g = 0; g < GB; ++g) {
						kv[g] = 0;
#pragma unroll
						for (int j = 0; j < 6; ++j) cf[g][j] = 0;
						if (any[g]) {
							const uint32_t b = b0 + g * NW;
							const uint32_t * bw = a.f.blk_words + (size_t)b * BLOCK_WORDS + lane;
#pragma unroll
							for (int j = 0; j < 6; ++j) cf[g][j] = bw[j * 64];
							if (MODE == 0) kv[g] = a.f.slot_keys[(size_t)b * 64 + lane];
						}
					}
#pragma unroll
					for (int g = 0; g < GB; ++g) {
						if (!any[g]) continue;
						++n_visit;
						test_block(b0 + g * NW, cf[g], kv[g], cand[g], nv[g]);
					}
				}
			} else if (VAR & 8) {
				// software pipeline: the next block's mask, words and key slots
				// are loaded while this block is tested
				uint4 p_ma = make_uint4(0, 0, 0, 0), p_mb = make_uint4(0, 0, 0, 0);
				uint32_t p_f[6] = {0, 0, 0, 0, 0, 0};
				uint64_t p_kv = 0;
				if (fb + wave < be) {
					const uint32_t b = fb + wave;
					p_ma = sld4(a.f.blk_info + (size_t)b * 8);
					p_mb = sld4(a.f.blk_info + (size_t)b * 8 + 4);
					const uint32_t * bw = a.f.blk_words + (size_t)b * BLOCK_WORDS + lane;
#pragma unroll
					for (int j = 0; j < 6; ++j) p_f[j] = bw[j * 64];
					if (MODE == 0) p_kv = a.f.slot_keys[(size_t)b * 64 + lane];
				}
				for (uint32_t b = fb + wave; b < be; b += NW) {
					const uint4 ma = p_ma, mb = p_mb;
					uint32_t c_f[6];
#pragma unroll
					for (int j = 0; j < 6; ++j) c_f[j] = p_f[j];
					const uint64_t c_kv = p_kv;
					const uint32_t bn = b + NW;
					if (bn < be) {
						p_ma = sld4(a.f.blk_info + (size_t)bn * 8);
						p_mb = sld4(a.f.blk_info + (size_t)bn * 8 + 4);
						const uint32_t * bw = a.f.blk_words + (size_t)bn * BLOCK_WORDS + lane;
#pragma unroll
						for (int j = 0; j < 6; ++j) p_f[j] = bw[j * 64];
						if (MODE == 0) p_kv = a.f.slot_keys[(size_t)bn * 64 + lane];
					}
					uint64_t cand[TILES];
					if (block_cands(ma, mb, cand) == 0) continue;
					++n_visit;
					test_block(b, c_f, c_kv, cand, mb.z);
				}
			} else {
				// one block: its candidates from its mask, then its tests
				auto visit = [&](uint32_t b) {
					const uint32_t * bi = a.f.blk_info + (size_t)b * 8;
					const uint4 ma = sld4(bi), mb = sld4(bi + 4);
					uint64_t cand[TILES];
					if (block_cands(ma, mb, cand) == 0) return;
					++n_visit;
					uint32_t c_f[6];
					const uint32_t * bw = a.f.blk_words + (size_t)b * BLOCK_WORDS + lane;
#pragma unroll
					for (int j = 0; j < 6; ++j) c_f[j] = bw[j * 64];
					const uint64_t c_kv = MODE == 0 ? a.f.slot_keys[(size_t)b * 64 + lane] : 0;
					test_block(b, c_f, c_kv, cand, mb.z);
				};
				if (a.groups && a.f.sup_info) {
					// Compacted steps: groups of SUP_BLOCKS blocks (aligned in
					// the table), one group per wave at a time; a group whose
					// AND of block masks no selected query contains is skipped
					// whole (at 10M filters, 21M block scans per compacted
					// launch become 14M group and block scans in a CPU model of
					// the same layout, and the kernel 2.65 -> 2.42 ms).  Other
					// steps keep one block per wave in turn: their items are
					// short, and a group per wave leaves waves idle (a
					// 1024-query step took 0.27 ms instead of 0.20).
					const uint32_t g0 = fb / SUP_BLOCKS, g1 = (be + SUP_BLOCKS - 1) / SUP_BLOCKS;
					for (uint32_t gi = g0 + wave; gi < g1; gi += NW) {
						const uint32_t * si = a.f.sup_info + (size_t)gi * 8;
						const uint4 sa = sld4(si), sb = sld4(si + 4);
						uint64_t scand[TILES];
						if (block_cands(sa, sb, scand) == 0) continue;
						const uint32_t b1 = min(gi * SUP_BLOCKS + SUP_BLOCKS, be);
						for (uint32_t b = max(gi * SUP_BLOCKS, fb); b < b1; ++b) visit(b);
					}
				} else {
					for (uint32_t b = fb + wave; b < be; b += NW) visit(b);
				}
			}
		}
		if (MODE == 1 && sink.hcnt) sink.flush(lane);
		if (KLOG && sink.hcnt) sink.drain(lane);
		w_hits += hit_acc;
		w_visit += n_visit;
		w_test += n_test;
		w_hev += n_hev;
		if (lane == 0 && hit_acc) atomicAdd(&s_hits, hit_acc);
		if (a.log) {
			for (int d = 32; d > 0; d >>= 1) key_acc += __shfl_xor(key_acc, d, 64);
			if (lane == 0 && key_acc) atomicAdd(&s_keys, key_acc);
		}
		__syncthreads();
		// An item without hits set no bit: rows that are merged into gbm
		// (several segments, or rows filled before the launch) need no flush.
		const bool item_hits = s_hits != 0;
		if (LDS_BM) bm_dirty = item_hits;
		if (MODE == 0 && LDS_BM && (item_hits || (nseg_all <= 1 && !a.qfull))) {
			const uint32_t nrows = min(q_end, q0 + NQ) - q0;
			const uint32_t nw = nrows * kw;
			uint32_t * dst = a.gbm + (size_t)q0 * kw;
			if (nseg_all > 1) {
				// Other segments of the tile write the same rows: an atomic
				// only for bits a row does not hold yet (rows fill up, and most
				// words are then complete).  One wave per row; a row that is
				// now full flags its query for the full-key-set exit.
				for (uint32_t j = wave; j < nrows; j += NW) {
					const uint32_t * src = s_bm + (size_t)j * kw;
					const uint32_t grow = row_of(q0 + j);
					uint32_t * drow = a.gbm + (size_t)grow * kw;
					uint32_t any = 0;
					for (uint32_t w = lane; w < kw; w += 64) any |= src[w];
					if (__ballot(any != 0) == 0) continue;   // nothing from this segment
					uint32_t all = 0xFFFFFFFFu;
					for (uint32_t w = lane; w < kw; w += 64) {
						const uint32_t v = src[w], d = drow[w];
						if (v & ~d) atomicOr(&drow[w], v);
						uint32_t u = v | d;
						if (w + 1 == kw) u |= ~a.last_mask;
						all &= u;
					}
					if (a.qfull && __ballot(all != 0xFFFFFFFFu) == 0 && lane == 0)
						atomicMax(&a.qfull[grow], 1u);
				}
			} else {
				// with flags (rich-filter pass, explicit units) the rows hold
				// keys already: merge
				if (a.map_rows) {
					for (uint32_t w = tid; w < nw; w += NT) {
						const uint32_t j = w / kw;
						a.gbm[(size_t)row_of(q0 + j) * kw + (w - j * kw)] |= s_bm[w];
					}
				} else if (a.qfull) {
					for (uint32_t w = tid; w < nw; w += NT) dst[w] |= s_bm[w];
				} else {
					for (uint32_t w = tid; w < nw; w += NT) dst[w] = s_bm[w];
				}
			}
		}
		if (tid == 0) {
			if (a.log && a.log_unit >= 0) {   // plain: every item of the launch adds up
				atomicAdd(&a.log[(size_t)a.log_unit * 3], s_hits);
				atomicAdd(&a.log[(size_t)a.log_unit * 3 + 1], s_keys);
				atomicAdd(&a.log[(size_t)a.log_unit * 3 + 2], (uint32_t)(wall_clock64() - t_start));
			} else if (a.log && item0 < a.log_cap) {
				a.log[(size_t)item0 * 3] = s_hits;
				a.log[(size_t)item0 * 3 + 1] = 0;
				a.log[(size_t)item0 * 3 + 2] = (uint32_t)(wall_clock64() - t_start);
			}
		}
		__syncthreads();
	}
	if (lane == 0 && (w_hits || w_visit)) {
		atomicAdd(&a.counters[1], (unsigned long long)w_hits);
		atomicAdd(&a.counters[2], (unsigned long long)w_visit);
		atomicAdd(&a.counters[3], (unsigned long long)w_test);
		atomicAdd(&a.counters[4], (unsigned long long)w_hev);
	}
}

// Inverted query words (6 + 2 pad per query) for the scalar-load variant.
__global__ void __launch_bounds__(256)
k_invert_queries(const uint32_t * __restrict__ q, uint32_t nq, uint32_t * __restrict__ qinv) {
	const uint32_t i = blockIdx.x * 256 + threadIdx.x;
	if (i >= nq * 8) return;
	qinv[i] = (i & 7) < 6 ? ~q[i] : 0u;
}

// Queue bounds for `items` work items, NQUEUE contiguous ranges.
__global__ void k_init_queues(uint64_t items, uint64_t * __restrict__ qbeg, unsigned long long * __restrict__ qnext) {
	const uint32_t t = threadIdx.x;
	if (t <= NQUEUE) qbeg[t] = items * t / NQUEUE;
	if (t < NQUEUE) qnext[t * QPAD] = items * t / NQUEUE;
}

// ------------------------------------------------------------- finalize

// KEYS_UNIQUE: one wave per query counts its bitmap bits.
__global__ void __launch_bounds__(256)
k_bm_count(const uint32_t * __restrict__ gbm, uint32_t kw, uint32_t nq, uint32_t * __restrict__ qcount) {
	const uint32_t qi = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
	if (qi >= nq) return;
	uint32_t c = 0;
	for (uint32_t w = lane; w < kw; w += 64) c += __popc(gbm[(size_t)qi * kw + w]);
	for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
	if (lane == 0) qcount[qi] = c;
}

// KEY_BITMAP, packed delivery: flag each query whose row is not the whole
// dictionary (bits 0..K-1 set, last_mask in the last word).  A wave takes 64
// queries; a query whose full flag this step set (qfull, when valid) is full
// without reading its row, and the wave reads the rows of the others one at
// a time, coalesced.
__global__ void __launch_bounds__(256)
k_row_partial(const uint32_t * __restrict__ gbm, uint32_t kw, uint32_t nq, uint32_t last_mask,
			  const uint32_t * __restrict__ qfull, uint32_t * __restrict__ flag) {
	const uint32_t lane = threadIdx.x & 63;
	const uint32_t base = (blockIdx.x * 4 + (threadIdx.x >> 6)) * 64;
	if (base >= nq) return;
	const uint32_t qi = base + lane;
	const bool valid = qi < nq;
	const bool known = valid && qfull && qfull[qi] != 0u;
	uint64_t check = __ballot(valid && !known);
	bool partial = false;
	while (check) {
		const uint32_t j = (uint32_t)__builtin_ctzll(check);
		check &= check - 1;
		const uint32_t * row = gbm + (size_t)(base + j) * kw;
		bool diff = false;
		for (uint32_t w = lane; w < kw; w += 64) diff |= row[w] != (w + 1 == kw ? last_mask : 0xFFFFFFFFu);
		const bool p = __ballot(diff) != 0;
		if (lane == j) partial = p;
	}
	if (valid) flag[qi] = partial;
}

// KEY_BITMAP, packed delivery: each query's entry (TM_ROW_FULL or the index
// of its row among the flagged ones, in query order) and the flagged rows,
// packed.  A wave takes 64 queries and copies their flagged rows one at a
// time, coalesced.
__global__ void __launch_bounds__(256)
k_row_pack(const uint32_t * __restrict__ gbm, uint32_t kw, uint32_t nq, const uint32_t * __restrict__ flag,
		   const uint64_t * __restrict__ base_of, uint32_t * __restrict__ row_of, uint32_t * __restrict__ prow) {
	const uint32_t lane = threadIdx.x & 63;
	const uint32_t base = (blockIdx.x * 4 + (threadIdx.x >> 6)) * 64;
	if (base >= nq) return;
	const uint32_t qi = base + lane;
	const bool valid = qi < nq;
	const bool partial = valid && flag[qi] != 0u;
	const uint64_t b = valid ? base_of[qi] : 0;
	if (valid) row_of[qi] = partial ? (uint32_t)b : TM_ROW_FULL;
	uint64_t copy = __ballot(partial);
	while (copy) {
		const uint32_t j = (uint32_t)__builtin_ctzll(copy);
		copy &= copy - 1;
		const uint64_t dst = __shfl(b, (int)j, 64);
		const uint32_t * src = gbm + (size_t)(base + j) * kw;
		for (uint32_t w = lane; w < kw; w += 64) prow[dst * kw + w] = src[w];
	}
}

// KEYS_UNIQUE: one wave per query writes its keys in ascending id order,
// which is ascending key order (the dictionary is sorted).  Lanes take 64
// consecutive key ids at a time: a set bit's output slot is the count of set
// bits below it (mbcnt), so each store writes consecutive values (rows are
// mostly dense) and the dictionary reads are coalesced.  The dictionary
// reads of BM_GROUP such 64-id groups are issued before any of their stores,
// so a wave waits for one round trip per BM_GROUP groups, not per group.
// A workgroup takes BM_WAVES consecutive queries: their outputs are one
// contiguous stretch written from one CU (16 per group measured 0.6 ms
// faster per 1M-query step than 4, and 1 slower than 4).
constexpr uint32_t BM_GROUP = 8;
constexpr uint32_t BM_WAVES = 16;

__global__ void __launch_bounds__(BM_WAVES * 64)
k_bm_write(const uint32_t * __restrict__ gbm, uint32_t kw, uint32_t nq, const uint64_t * __restrict__ off,
		   const uint32_t * __restrict__ dict, uint32_t * __restrict__ vals, uint64_t cap) {
	const uint32_t qi = blockIdx.x * BM_WAVES + (threadIdx.x >> 6), lane = threadIdx.x & 63;
	if (qi >= nq) return;
	const uint32_t * row = gbm + (size_t)qi * kw;
	uint64_t pos = off[qi];
	for (uint32_t wb = 0; wb < kw; wb += 64) {   // 64 words (2048 ids) per pass
		const uint32_t v = wb + lane < kw ? row[wb + lane] : 0u;
		const uint32_t nw = min(64u, kw - wb);
		for (uint32_t c0 = 0; c0 < nw; c0 += 2 * BM_GROUP) {
			uint64_t m[BM_GROUP];
			uint32_t d[BM_GROUP];
#pragma unroll
			for (uint32_t k = 0; k < BM_GROUP; ++k) {
				const uint32_t c = c0 + 2 * k;   // words c, c + 1 (c + 1 < 64: shuffles stay in the wave)
				m[k] = (uint64_t)(uint32_t)__shfl(v, (int)c, 64) | ((uint64_t)(uint32_t)__shfl(v, (int)c + 1, 64) << 32);
				// a set bit names a key id below the dictionary's size (bits past
				// nw words are zero: v is zero past the row's kw words)
				d[k] = ((m[k] >> lane) & 1) ? dict[(wb + c) * 32 + lane] : 0u;
			}
#pragma unroll
			for (uint32_t k = 0; k < BM_GROUP; ++k) {
				if ((m[k] >> lane) & 1) {
					const uint64_t p = pos + mbcnt64(m[k]);
					if (p < cap) vals[p] = d[k];
				}
				pos += (uint64_t)__popcll(m[k]);
			}
		}
	}
}

__device__ inline uint32_t slot_key_count(uint64_t kv, const uint32_t * __restrict__ ovf, uint32_t ovf_n) {
	uint32_t off;
	if ((kv & 0xFFFFu) == KEY_NONE) return ovf_list(kv, ovf, ovf_n, off);
	uint32_t n = 1;
	if (((kv >> 16) & 0xFFFFu) != KEY_NONE) ++n; else return n;
	if (((kv >> 32) & 0xFFFFu) != KEY_NONE) ++n; else return n;
	if ((kv >> 48) != KEY_NONE) ++n;
	return n;
}

// Key i < slot_key_count(kv) of a slot (a dense id, clamped into the
// dictionary).
__device__ inline uint32_t slot_key_id(uint64_t kv, const uint32_t * __restrict__ ovf, uint32_t i, uint32_t n_keys) {
	const uint32_t id = (kv & 0xFFFFu) == KEY_NONE ? ovf[(uint32_t)(kv >> 16) + 1 + i]
												   : (uint32_t)((kv >> (16 * i)) & 0xFFFFu);
	return id < n_keys ? id : 0u;
}

// HITS / KEYS_ALL: values per query.
template <bool KEYS>
__global__ void __launch_bounds__(256)
k_hits_count(const uint2 * __restrict__ hits, const unsigned long long * __restrict__ counters,
			 uint64_t cap, dev_fib_view f, uint32_t * __restrict__ qcount) {
	uint64_t n = counters[0];
	if (n > cap) n = cap;
	const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
	for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
		uint2 h = hits[i];
		uint32_t c = KEYS ? slot_key_count(f.slot_keys[h.y], f.ovf, f.ovf_n) : 1u;
		atomicAdd(&qcount[h.x], c);
	}
}

template <bool KEYS>
__global__ void __launch_bounds__(256)
k_hits_write(const uint2 * __restrict__ hits, const unsigned long long * __restrict__ counters,
			 uint64_t cap, dev_fib_view f, unsigned long long * __restrict__ qcur,
			 uint32_t * __restrict__ vals, uint64_t vcap) {
	uint64_t n = counters[0];
	if (n > cap) n = cap;
	const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
	for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
		uint2 h = hits[i];
		if (KEYS) {
			uint64_t kv = f.slot_keys[h.y];
			uint32_t c = slot_key_count(kv, f.ovf, f.ovf_n);
			unsigned long long p = atomicAdd(&qcur[h.x], (unsigned long long)c);
			for (uint32_t k = 0; k < c; ++k)
				if (p + k < vcap) vals[p + k] = f.key_dict[slot_key_id(kv, f.ovf, k, f.n_keys)];
		} else {
			unsigned long long p = atomicAdd(&qcur[h.x], 1ULL);
			if (p < vcap) vals[p] = f.slot_fid[h.y];
		}
	}
}

// Collect the submission's counters in one array for a single D2H copy.
__global__ void k_gather_stats(const uint64_t * __restrict__ pbase, uint32_t P,
							   const unsigned long long * __restrict__ counters,
							   const uint64_t * __restrict__ off, uint32_t nq,
							   const uint64_t * __restrict__ item_prefix, uint64_t * __restrict__ out) {
	if (threadIdx.x == 0 && blockIdx.x == 0) {
		out[0] = pbase[P];        // (query, partition) pairs
		out[1] = counters[0];     // hit slots taken
		out[2] = counters[1];     // hits
		out[3] = off ? off[nq] : 0;   // result values (none: KEY_BITMAP rows)
		out[4] = item_prefix[P];  // work items
		out[5] = counters[2];     // block visits
		out[6] = counters[3];     // subset tests
		out[7] = counters[4];     // tests with hits
	}
}

// Work queues of every plain unit: unit u's NQUEUE queues split its items.
__global__ void __launch_bounds__(256)
k_init_unit_queues(const uint32_t * __restrict__ items, uint32_t n_units, uint64_t * __restrict__ qbeg,
				   unsigned long long * __restrict__ qnext) {
	const uint32_t u = blockIdx.x * 256 + threadIdx.x;
	if (u >= n_units) return;
	const uint64_t n = items[u];
	for (int x = 0; x <= NQUEUE; ++x) qbeg[(size_t)u * (NQUEUE + 1) + x] = n * x / NQUEUE;
	for (int x = 0; x < NQUEUE; ++x) qnext[((size_t)u * NQUEUE + x) * QPAD] = n * x / NQUEUE;
}

// OR one plain unit's staging rows into its queries' rows and clear them
// for the next unit.  The queries of a batch are distinct and launches on
// one stream run in order, so plain read-modify-writes suffice.
__global__ void __launch_bounds__(256)
k_or_rows(uint32_t * __restrict__ stage, uint32_t kw, const uint32_t * __restrict__ qmap, uint32_t nqv,
		  uint32_t * __restrict__ gbm) {
	const uint32_t w = blockIdx.x * 256 + threadIdx.x;
	if (w >= nqv * kw) return;
	const uint32_t v = stage[w];
	if (v == 0) return;
	const uint32_t t = w / kw;
	gbm[(size_t)qmap[t] * kw + (w - t * kw)] |= v;
	stage[w] = 0;
}

// Estimated work of each block for a 64-query tile whose queries have bit
// density rho (l2rho = log2 rho): a fixed part for the candidate pass and
// the subset tests, plus the expected hits, 64 x the sum over valid lanes of
// rho^popcount (a random query of density rho contains a filter of
// popcount k with probability rho^k).  One wave per block.
__global__ void __launch_bounds__(256)
k_block_weights(tmx::dev_fib_view f, uint32_t nb, float l2rho, uint32_t * __restrict__ w) {
	const uint32_t lane = threadIdx.x & 63;
	const uint32_t b = blockIdx.x * 4 + (threadIdx.x >> 6);
	if (b >= nb) return;
	const uint32_t nv = f.blk_info[(size_t)b * 8 + 6];
	float e = 0.f;
	if (lane < nv) {
		uint32_t pop = 0;
		for (int j = 0; j < 6; ++j) pop += __popc(f.blk_words[(size_t)b * BLOCK_WORDS + j * 64 + lane]);
		// a hit costs one step, plus one per 64 keys of a long overflow list
		const uint32_t nk = key_count(f.slot_keys[(size_t)b * 64 + lane], f.ovf, f.ovf_n);
		e = exp2f(l2rho * (float)pop) * (1.f + (nk > COOP_KEYS ? (float)nk / 64.f : 0.f));
	}
	for (int d = 32; d > 0; d >>= 1) e += __shfl_xor(e, d, 64);
	if (lane == 0) w[b] = WSEG_BASE + (uint32_t)fminf(64.f * e, 1048576.f);
}

// Bounds of nseg segments of about equal weight: segb[k] is the first block
// whose exclusive weight prefix reaches k/nseg of the total; segb[0] = 0,
// segb[nseg] = nb.  Non-decreasing in k; a segment may be empty.
__global__ void __launch_bounds__(256)
k_seg_bounds(const uint64_t * __restrict__ prefix, uint32_t nb, uint32_t nseg, uint32_t * __restrict__ segb) {
	const uint32_t k = blockIdx.x * 256 + threadIdx.x;
	if (k > nseg) return;
	if (k == 0 || k == nseg) {
		segb[k] = k == 0 ? 0u : nb;
		return;
	}
	const uint64_t target = (prefix[nb] * k + nseg - 1) / nseg;
	uint32_t lo = 0, hi = nb;
	while (lo < hi) {
		const uint32_t mid = (lo + hi) >> 1;
		if (prefix[mid] >= target) hi = mid; else lo = mid + 1;
	}
	segb[k] = lo;
}

// Rich-filter pass: the key bitmaps of the rich filters that a query
// contains, ORed into its row, which this writes whole; flag 2 when the row
// then holds every key (the match kernel's full-key-set exit drops the
// query), else 0.  Exact: a filter contained in the query matches it,
// whatever the partition lookup.  A wave takes qpw queries (at most 64):
// first one lane per query against the rich filters that hold every key
// (most queries contain one), whose rows the wave then writes full, row by
// row; then the other queries one at a time with the whole wave, 64 rich
// filters per test and eight bitmaps in flight.  Large steps use 64 queries
// per wave; small ones fewer, so that their few incomplete queries are
// spread over more waves.
__global__ void __launch_bounds__(256)
k_rich_rows(const uint32_t * __restrict__ q, uint32_t nq, uint32_t qpw, const uint32_t * __restrict__ rw,
			const uint32_t * __restrict__ rbm, uint32_t nr, uint32_t nfull, uint32_t kw, uint32_t last_mask,
			uint32_t * __restrict__ gbm, uint32_t * __restrict__ qfull) {
	const uint32_t lane = threadIdx.x & 63;
	const uint32_t base = (blockIdx.x * 4 + (threadIdx.x >> 6)) * qpw;
	if (base >= nq) return;
	const uint32_t qi = base + lane;
	const bool valid = lane < qpw && qi < nq;
	uint4 x = make_uint4(0, 0, 0, 0);
	uint2 y = make_uint2(0, 0);
	if (valid) {
		x = *reinterpret_cast<const uint4 *>(q + (size_t)qi * 8);
		y = *reinterpret_cast<const uint2 *>(q + (size_t)qi * 8 + 4);
	}
	// the first nfull rich filters hold every key: one contained is enough
	bool full = false;
	if (qpw == 1) {
		// one query (lane 0's): lanes are rich filters, 64 at a time
		const uint32_t qx = __shfl(x.x, 0, 64), qy = __shfl(x.y, 0, 64), qz = __shfl(x.z, 0, 64),
					   qw = __shfl(x.w, 0, 64), q4 = __shfl(y.x, 0, 64), q5 = __shfl(y.y, 0, 64);
		bool any_full = false;
		for (uint32_t fb = 0; fb < nfull && !any_full; fb += 64) {
			const uint32_t f = fb + lane;
			bool hit = false;
			if (f < nfull) {
				const uint4 a = *reinterpret_cast<const uint4 *>(rw + (size_t)f * 8);
				const uint2 b = *reinterpret_cast<const uint2 *>(rw + (size_t)f * 8 + 4);
				hit = ((a.x & ~qx) | (a.y & ~qy) | (a.z & ~qz) | (a.w & ~qw) | (b.x & ~q4) | (b.y & ~q5)) == 0;
			}
			any_full = __ballot(hit) != 0;
		}
		full = valid && any_full;
	} else {
		// one lane per query; the filters' words are the same for every lane
		for (uint32_t f = 0; f < nfull; ++f) {
			const uint4 a = sld4(rw + (size_t)f * 8);
			const uint2 b = sld2(rw + (size_t)f * 8 + 4);
			full |= ((a.x & ~x.x) | (a.y & ~x.y) | (a.z & ~x.z) | (a.w & ~x.w) | (b.x & ~y.x) | (b.y & ~y.y)) == 0;
			if ((f & 7) == 7 && __ballot(valid && !full) == 0) break;
		}
	}
	uint64_t m = __ballot(valid && full);
	while (m) {
		const uint32_t j = (uint32_t)__builtin_ctzll(m);
		m &= m - 1;
		uint32_t * row = gbm + (size_t)(base + j) * kw;
		for (uint32_t w = lane; w < kw; w += 64) row[w] = w + 1 == kw ? last_mask : 0xFFFFFFFFu;
	}
	bool done_full = full;
	uint64_t rest = __ballot(valid && !full);
	while (rest) {
		const uint32_t j = (uint32_t)__builtin_ctzll(rest);
		rest &= rest - 1;
		const uint32_t qx = __shfl(x.x, (int)j, 64), qy = __shfl(x.y, (int)j, 64), qz = __shfl(x.z, (int)j, 64),
					   qw = __shfl(x.w, (int)j, 64), q4 = __shfl(y.x, (int)j, 64), q5 = __shfl(y.y, (int)j, 64);
		uint32_t * row = gbm + (size_t)(base + j) * kw;
		uint32_t all = 0xFFFFFFFFu;
		for (uint32_t wb = 0; wb < kw; wb += 64) {   // 64 words (2048 keys) per pass
			const uint32_t w = wb + lane;
			uint32_t acc = 0;
			for (uint32_t fb = 0; fb < nr; fb += 64) {
				const uint32_t f = fb + lane;
				bool hit = false;
				if (f < nr) {
					const uint4 a = *reinterpret_cast<const uint4 *>(rw + (size_t)f * 8);
					const uint2 b = *reinterpret_cast<const uint2 *>(rw + (size_t)f * 8 + 4);
					hit = ((a.x & ~qx) | (a.y & ~qy) | (a.z & ~qz) | (a.w & ~qw) | (b.x & ~q4) | (b.y & ~q5)) == 0;
				}
				uint64_t hm = __ballot(hit);
				while (hm) {   // eight bitmaps in flight per round (a repeat ORs twice)
					uint32_t jj[8];
#pragma unroll
					for (int k = 0; k < 8; ++k) {
						jj[k] = hm ? (uint32_t)__builtin_ctzll(hm) : jj[0];
						hm &= hm - 1;
					}
					if (w < kw) {
						uint32_t v[8];
#pragma unroll
						for (int k = 0; k < 8; ++k) v[k] = rbm[(size_t)(fb + jj[k]) * kw + w];
#pragma unroll
						for (int k = 0; k < 8; ++k) acc |= v[k];
					}
				}
			}
			if (w < kw) {
				row[w] = acc;
				all &= w + 1 == kw ? (acc | ~last_mask) : acc;
			}
		}
		const bool f = __ballot(all != 0xFFFFFFFFu) == 0;
		if (lane == j) done_full = f;
	}
	if (valid) qfull[qi] = done_full ? 2u : 0u;
}

// Queries whose rows are not yet full after the rich-filter pass, listed
// for a match over them alone, in order of a key made of their ACT_KEY_BITS
// rarest bits (the table's significance order), those holding the rarest
// bits first.  A tile then holds queries that share their rare bits, and a
// block, whose AND mask holds the rare bits of its filters, has a candidate
// in fewer tiles: at 10M filters and about 10k active queries, 8.8M
// (tile, block) visits instead of 12.5M in arbitrary order (measured on the
// CPU over the same layout).  Order within a key is arbitrary; results do
// not depend on the order.
constexpr uint32_t ACT_KEY_BITS = 12;
constexpr uint32_t ACT_BUCKETS = 1u << ACT_KEY_BITS;

struct act_key_bits {
	uint8_t pos[ACT_KEY_BITS];   // bit positions, rarest first
};

__device__ inline uint32_t act_bucket(const uint32_t * __restrict__ qw, const act_key_bits & kb) {
	uint32_t k = 0;
#pragma unroll
	for (uint32_t i = 0; i < ACT_KEY_BITS; ++i) {
		const uint32_t b = kb.pos[i];
		k = (k << 1) | ((qw[b >> 5] >> (b & 31)) & 1u);
	}
	return (ACT_BUCKETS - 1u) - k;
}

// Active queries per key bucket (LDS counters, one global atomic per
// non-empty bucket and workgroup).
__global__ void __launch_bounds__(256)
k_act_hist(const uint32_t * __restrict__ q, uint32_t nq, const uint32_t * __restrict__ qfull, act_key_bits kb,
		   uint32_t * __restrict__ hist) {
	__shared__ uint32_t h[ACT_BUCKETS];
	for (uint32_t i = threadIdx.x; i < ACT_BUCKETS; i += 256) h[i] = 0;
	__syncthreads();
	for (uint32_t i = blockIdx.x * 256 + threadIdx.x; i < nq; i += gridDim.x * 256)
		if (qfull[i] == 0u) atomicAdd(&h[act_bucket(q + (size_t)i * 8, kb)], 1u);
	__syncthreads();
	for (uint32_t i = threadIdx.x; i < ACT_BUCKETS; i += 256)
		if (h[i]) atomicAdd(&hist[i], h[i]);
}

// Each active query to its bucket's next slot; the active count to
// n_active[0] (base[ACT_BUCKETS] is the total).
__global__ void __launch_bounds__(256)
k_act_scatter(const uint32_t * __restrict__ q, uint32_t nq, const uint32_t * __restrict__ qfull, act_key_bits kb,
			  const uint64_t * __restrict__ base, uint32_t * __restrict__ cursor, uint32_t * __restrict__ list,
			  uint32_t * __restrict__ n_active) {
	const uint32_t i = blockIdx.x * 256 + threadIdx.x;
	if (i == 0) n_active[0] = (uint32_t)base[ACT_BUCKETS];
	if (i >= nq || qfull[i] != 0u) return;
	const uint32_t b = act_bucket(q + (size_t)i * 8, kb);
	list[base[b] + atomicAdd(&cursor[b], 1u)] = i;
}

// Shape of a compacted step from its active query count: ceil(n / nqg)
// tiles, cut into segments as the host cuts other steps (about two items
// per resident workgroup, at most seg_max blocks each).  act[1..3] = tiles,
// segments, blocks per segment; queue bounds and the item count follow.
__global__ void k_init_queues_active(uint32_t * __restrict__ act, uint32_t nqg, uint32_t nblk, uint32_t resident,
									 uint32_t seg_max, uint64_t * __restrict__ qbeg,
									 unsigned long long * __restrict__ qnext, uint64_t * __restrict__ n_items) {
	const uint32_t t = threadIdx.x;
	const uint64_t tiles = ((uint64_t)act[0] + nqg - 1) / nqg;
	uint32_t nseg = 1;
	if (tiles > 0) {
		nseg = (uint32_t)min((2ull * resident + tiles - 1) / tiles, (uint64_t)max(nblk / 16, 1u));
		nseg = max(nseg, (nblk + seg_max - 1) / seg_max);
	}
	const uint32_t sb = nseg > 1 ? (nblk + nseg - 1) / nseg : max(nblk, 1u);
	if (nseg > 1) nseg = (nblk + sb - 1) / sb;
	const uint64_t items = tiles * nseg;
	if (t <= NQUEUE) qbeg[t] = items * t / NQUEUE;
	if (t < NQUEUE) qnext[t * QPAD] = items * t / NQUEUE;
	if (t == 0) {
		act[1] = (uint32_t)tiles;
		act[2] = nseg;
		act[3] = sb;
		*n_items = items;
	}
}

// Queries per partition (the lookup's counts).  Lanes are partitions, 64 at
// a time; the wave's share of the queries is read with scalar loads, and
// each lane counts the queries that select its partition in a register.
// Counts meet in LDS (P counters), then one global atomic per partition per
// workgroup.
__global__ void __launch_bounds__(256)
k_lookup_hist(const uint32_t * __restrict__ q, uint32_t nq, const uint32_t * __restrict__ part_info, uint32_t P,
			  uint32_t * __restrict__ pcount) {
	extern __shared__ uint32_t s_cnt[];
	for (uint32_t p = threadIdx.x; p < P; p += 256) s_cnt[p] = 0;
	__syncthreads();
	const uint32_t lane = threadIdx.x & 63;
	const uint32_t gw = __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
	const uint32_t nwaves = (gridDim.x * blockDim.x) >> 6;
	const uint32_t per = (nq + nwaves - 1) / nwaves;
	const uint32_t qa = min(gw * per, nq), qb = min(qa + per, nq);
	for (uint32_t p0 = 0; p0 < P; p0 += 64) {
		const uint32_t p = p0 + lane;
		uint32_t m[6] = {0, 0, 0, 0, 0, 0};
		if (p < P)
			for (int j = 0; j < 6; ++j) m[j] = part_info[(size_t)p * 8 + j];
		uint32_t cnt = 0;
#pragma unroll 4
		for (uint32_t i = qa; i < qb; ++i) {
			const uint4 x = sld4(q + (size_t)i * 8);
			const uint2 y = sld2(q + (size_t)i * 8 + 4);
			const uint32_t t = (m[0] & ~x.x) | (m[1] & ~x.y) | (m[2] & ~x.z) | (m[3] & ~x.w) | (m[4] & ~y.x) |
							   (m[5] & ~y.y);
			cnt += t == 0;
		}
		if (p < P && cnt) atomicAdd(&s_cnt[p], cnt);
	}
	__syncthreads();
	for (uint32_t p = threadIdx.x; p < P; p += 256)
		if (s_cnt[p]) atomicAdd(&pcount[p], s_cnt[p]);
}

// ------------------------------------------------------------- kernel table

using mq_fn = void (*)(mq_args);

template <int MODE, bool LDS_BM, int TILES, int VAR>
mq_fn pick_q_waves(int nw) {
	switch (nw) {
	case 4: return k_match_q<MODE, LDS_BM, TILES, 4, VAR>;
	case 16: return k_match_q<MODE, LDS_BM, TILES, 16, VAR>;
	default: return k_match_q<MODE, LDS_BM, TILES, 8, VAR>;
	}
}

template <int MODE, bool LDS_BM, int VAR>
mq_fn pick_q_tiles(int tiles, int nw) {
	switch (tiles) {
	case 1: return pick_q_waves<MODE, LDS_BM, 1, VAR>(nw);
	case 2: return pick_q_waves<MODE, LDS_BM, 2, VAR>(nw);
	default: return pick_q_waves<MODE, LDS_BM, 4, VAR>(nw);
	}
}

template <int VAR>
mq_fn pick_q_mode(int mode, bool lds_bm, int tiles, int nw) {
	if (mode == 0) return lds_bm ? pick_q_tiles<0, true, VAR>(tiles, nw) : pick_q_tiles<0, false, VAR>(tiles, nw);
	return pick_q_tiles<1, false, VAR>(tiles, nw);
}

mq_fn pick_match_q(int mode, bool lds_bm, int tiles, int nw, int var) {
	switch (var) {
	case 1: return pick_q_mode<1>(mode, lds_bm, tiles, nw);
	case 2: return pick_q_mode<2>(mode, lds_bm, tiles, nw);
	case 3: return pick_q_mode<3>(mode, lds_bm, tiles, nw);
	case 4: return pick_q_mode<4>(mode, lds_bm, tiles, nw);
	case 5: return pick_q_mode<5>(mode, lds_bm, tiles, nw);
	case 6: return pick_q_mode<6>(mode, lds_bm, tiles, nw);
	case 7: return pick_q_mode<7>(mode, lds_bm, tiles, nw);
	case 10: return pick_q_mode<10>(mode, lds_bm, tiles, nw);
	case 18: return pick_q_mode<18>(mode, lds_bm, tiles, nw);
	default: return pick_q_mode<2>(mode, lds_bm, tiles, nw);
	}
}

} // namespace

// ------------------------------------------------------------- ring

struct tm_ring {
	tm_fib * fib = nullptr;
	tm_match_config cfg;
	uint32_t max_q = 0;
	uint32_t P = 0, kw = 0;
	hipStream_t s = nullptr;
	hipEvent_t e0 = nullptr, em0 = nullptr, em1 = nullptr, e1 = nullptr;

	uint32_t * h_q = nullptr;      // pinned staging, max_q x 8
	uint32_t * d_q = nullptr;      // max_q x 8
	uint32_t * d_qinv = nullptr;   // max_q x 8, inverted (variant 1)
	uint32_t * d_pcount = nullptr; // P
	uint64_t * d_pbase = nullptr;  // P + 1
	unsigned long long * d_pcur = nullptr;
	uint32_t * d_list = nullptr;
	uint64_t list_cap = 0;
	uint4 * d_sched = nullptr;
	uint64_t * d_iprefix = nullptr;
	uint64_t * d_qbeg = nullptr;
	unsigned long long * d_qnext = nullptr;
	uint32_t * d_porder = nullptr;
	bool have_order = false;
	uint32_t * d_gbm = nullptr;
	uint32_t * d_qfull = nullptr;   // max_q: full-key-set flags of merged rows
	uint32_t * d_active = nullptr;  // max_q: queries left after the rich-filter pass
	uint32_t * d_nact = nullptr;    // active queries, their tiles, segments, blocks per segment
	uint32_t * d_ahist = nullptr;   // 2 x ACT_BUCKETS: active queries per key bucket, then cursors
	uint64_t * d_abase = nullptr;   // ACT_BUCKETS + 1: first slot of each key bucket
	uint32_t * d_bw = nullptr;      // per block: work estimate (weighted segments)
	uint64_t * d_bprefix = nullptr; // n_blocks + 1: exclusive prefix of d_bw
	uint64_t * d_bscan = nullptr;   // scan scratch for d_bprefix
	uint32_t * d_segb = nullptr;    // n_blocks + 1: segment bounds
	uint32_t bw_cap = 0, bw_nb = 0;
	uint64_t bw_version = ~0ull;
	float bw_l2 = 0.f;              // log2 density the weights were computed for
	float rho_l2 = -1.f;            // log2 mean bit density of the last host queries
	uint2 * d_hits = nullptr;
	uint64_t hit_cap = 0;
	uint32_t * d_qcount = nullptr;
	uint64_t * d_off = nullptr;
	unsigned long long * d_qcur = nullptr;
	uint32_t * d_vals = nullptr;
	uint64_t val_cap = 0;
	unsigned long long * d_cnt = nullptr;
	uint64_t * d_scan = nullptr;
	size_t scan_words = 0;
	uint64_t * d_stat = nullptr;
	uint64_t * h_stat = nullptr;   // pinned, 8
	uint32_t * d_log = nullptr;
	uint64_t log_cap = 0;
	uint32_t * d_stage = nullptr;     // plain: key bitmap rows of one unit's batch
	uint64_t stage_cap = 0;
	uint32_t * d_uitems = nullptr;    // plain: work items per unit
	uint64_t uitems_cap = 0;
	uint64_t * d_ubeg = nullptr;      // plain: per unit NQUEUE + 1 queue bounds
	unsigned long long * d_unext = nullptr;
	uint32_t uq_cap = 0;
	// plain unit log: per launch (partition, batch, queries) and features
	uint4 * d_units = nullptr;
	float * d_ufeat = nullptr;
	uint64_t units_cap = 0;
	uint32_t n_units = 0;

	// cost model (persistent schedule)
	tm_mlp * model = nullptr;
	uint32_t model_flags = 0;
	float * d_tfeat = nullptr;     // (tile, partition) feature rows
	float * d_tpred = nullptr;     // predictions, 2 per row
	float * d_tcost = nullptr;     // per tile: predicted ticks
	float * d_thits = nullptr;     // per tile: predicted hits
	float * d_tkeys = nullptr;     // per tile: predicted key instances
	uint32_t * d_iorder = nullptr; // per work item: query tile
	uint64_t trows_cap = 0, tiles_cap = 0;
	bool use_iorder = false;

	// effective launch parameters
	int tiles = 4, wgs_per_cu = 2, waves = 8, variant = 1;
	int wgs_for[3] = {1, 1, 1};     // resident workgroups per CU with 4, 8, 16 waves
	uint32_t chunk = 256;
	uint32_t plain_bq = 256;   // PLAIN: queries per (partition, batch) unit
	bool lds_bm = false;       // key bitmaps in LDS
	size_t stage_lds = 0;      // dynamic LDS bytes of the match launch

	// graph replay
	// captured steps, one per step size (graph replays run power-of-two sizes)
	struct graph_entry {
		hipGraphExec_t exec = nullptr;
		uint32_t nq = 0, epoch = 0;
		const uint32_t * slot_epoch = nullptr;
		bool weighted = false;       // reads the block weights, refreshed outside the graph
		bool qfull_valid = false;    // writes d_qfull
	};
	bool capturing = false;          // enqueue_persistent runs inside a stream capture
	bool qfull_valid = false;        // the last step wrote d_qfull: a set flag means a full row
	bool step_weighted = false;      // the last enqueue_persistent cut weighted segments
	std::vector<graph_entry> graphs;

	uint32_t last_nq = 0;
	uint32_t run_nq = 0;       // queries the step's kernels run over: last_nq, or more
	                           // (zero queries appended) when graphs replay fixed sizes
	uint64_t fib_version = 0;  // FIB version the device arrays are sized for
	// explicit units (tm_submit_units): the last submission's units
	bool units_mode = false;
	std::vector<uint32_t> u_part, u_off;
	uint32_t * d_xunits = nullptr;     // 8 words per unit
	uint32_t * d_xprefix = nullptr;    // items prefix, n + 1
	uint32_t * d_iota = nullptr;       // 0 .. max_q - 1 (plain units' query lists)
	uint64_t xunits_cap = 0, xprefix_cap = 0;
	uint32_t last_nseg = 1;    // block segments per tile of the last persistent step
	uint32_t epoch = 0;        // FIB insert epoch the submission sees
	bool no_weights = false;          // segments by block count only (no per-block weight arrays)
	uint32_t seg_bucket[TM_SIZE_BUCKETS] = {};   // explicit units: blocks per item by partition size (0: auto)
	uint32_t * d_step = nullptr;      // graph replays of key-row steps: epoch, valid overflow words
	uint32_t * h_step = nullptr;      // pinned source of d_step
	uint32_t * host_rows = nullptr;   // KEY_BITMAP: rows delivered here at the end of each step
	uint64_t host_rows_cap = 0;       // words
	// KEY_BITMAP, packed delivery (tm_ring_set_host_rows_packed)
	uint32_t * host_row_of = nullptr; // per query: TM_ROW_FULL or its row's index in host_prow
	uint32_t * host_prow = nullptr;   // the rows that are not the whole dictionary, packed
	uint64_t host_row_of_cap = 0, host_prow_cap = 0;   // entries, words
	uint32_t * d_rflag = nullptr;     // per query: row is not full
	uint64_t * d_rbase = nullptr;     // exclusive scan of d_rflag (nq + 1)
	uint64_t * d_rscan = nullptr;     // scan partials
	uint32_t * d_row_of = nullptr;
	uint32_t * d_prow = nullptr;
	uint64_t pack_q_cap = 0;          // queries the packing buffers hold
	uint64_t * h_npack = nullptr;     // pinned: packed rows of the last step
	uint64_t pack_sent = 0;           // rows copied with the step (the rest in tm_wait)
	uint64_t pack_guess = 0;          // rows to copy with the next step
	bool pending = false;
	bool match_events = false;
	int retries = 0;
	tm_status st;

	void release();
};

void tm_ring::release() {
	for (auto & g : graphs) (void)hipGraphExecDestroy(g.exec);
	graphs.clear();
	if (h_q) (void)hipHostFree(h_q);
	if (h_stat) (void)hipHostFree(h_stat);
	if (h_step) (void)hipHostFree(h_step);
	if (h_npack) (void)hipHostFree(h_npack);
	h_npack = nullptr;
	dfree(d_rflag); dfree(d_rbase); dfree(d_rscan); dfree(d_row_of); dfree(d_prow);
	pack_q_cap = 0;
	h_q = nullptr;
	h_stat = nullptr;
	h_step = nullptr;
	dfree(d_step);
	dfree(d_q); dfree(d_qinv); dfree(d_pcount); dfree(d_pbase); dfree(d_pcur); dfree(d_list);
	dfree(d_sched); dfree(d_iprefix); dfree(d_qbeg); dfree(d_qnext); dfree(d_porder);
	dfree(d_gbm); dfree(d_qfull); dfree(d_active); dfree(d_nact); dfree(d_ahist); dfree(d_abase); dfree(d_bw); dfree(d_bprefix); dfree(d_bscan); dfree(d_segb); dfree(d_hits); dfree(d_qcount); dfree(d_off); dfree(d_qcur);
	dfree(d_vals); dfree(d_cnt); dfree(d_scan); dfree(d_stat); dfree(d_log);
	dfree(d_units); dfree(d_ufeat); dfree(d_stage); dfree(d_uitems); dfree(d_ubeg); dfree(d_unext);
	dfree(d_xunits); dfree(d_xprefix); dfree(d_iota);
	xunits_cap = xprefix_cap = 0;
	stage_cap = uitems_cap = 0;
	uq_cap = 0;
	dfree(d_tfeat); dfree(d_tpred); dfree(d_tcost); dfree(d_thits); dfree(d_tkeys); dfree(d_iorder);
	units_cap = trows_cap = tiles_cap = 0;
	if (e0) (void)hipEventDestroy(e0);
	if (em0) (void)hipEventDestroy(em0);
	if (em1) (void)hipEventDestroy(em1);
	if (e1) (void)hipEventDestroy(e1);
	e0 = em0 = em1 = e1 = nullptr;
	if (s) (void)hipStreamDestroy(s);
	s = nullptr;
}

namespace {

// KEYS_UNIQUE and KEY_BITMAP both build per-query key bitmap rows; only
// KEYS_UNIQUE expands them into sorted key lists.
bool keyrows(const tm_match_config & c) { return c.mode == TM_MODE_KEYS_UNIQUE || c.mode == TM_MODE_KEY_BITMAP; }
int mode_of(const tm_match_config & c) { return keyrows(c) ? 0 : 1; }

// Partition-size bucket of a partition of nb blocks (tm_ring_set_bucket_segments).
uint32_t size_bucket(uint32_t nb) {
	uint32_t b = 0;
	for (uint32_t lim = 16; b + 1 < TM_SIZE_BUCKETS && nb >= lim; lim *= 4) ++b;
	return b;
}

// Effective launch parameters from the config and the FIB.
int resolve_params(tm_ring * r) {
	const tm_match_config & c = r->cfg;
	if (c.mode > TM_MODE_KEY_BITMAP) return set_error(TM_ERR_INVALID, "unknown mode %u", c.mode);
	if (c.schedule > TM_SCHED_PERSISTENT) return set_error(TM_ERR_INVALID, "unknown schedule %u", c.schedule);
	if (keyrows(c) && r->fib->n_keys > 65536)
		return set_error(TM_ERR_INVALID,
						 "KEYS_UNIQUE and KEY_BITMAP support up to 65536 distinct keys (FIB has %u); use KEYS_ALL",
						 r->fib->n_keys);
	// default batch: the reference's 256 for the plain form (parameters.hh:22),
	// 128-query tiles for the persistent kernel (measured best at 10M filters)
	uint32_t bs = c.batch_size ? c.batch_size : (c.schedule == TM_SCHED_PLAIN ? 256 : 64);
	if (bs != 64 && bs != 128 && bs != 256)
		return set_error(TM_ERR_INVALID, "batch_size must be 64, 128 or 256 (got %u)", bs);
	// plain: bs queries per (partition, batch) unit, run as 128-query tiles
	r->plain_bq = bs;
	r->tiles = c.schedule == TM_SCHED_PLAIN ? (bs == 64 ? 1 : 2) : (int)(bs / 64);
	r->chunk = 1u << 30;   // work items of the lookup: one per (partition, batch)
	const size_t max_lds = (size_t)dev().max_lds;
	const size_t nq_grp = (size_t)r->tiles * 64;
	r->variant = ((c.variant >= 1 && c.variant <= 7) || c.variant == 10 || c.variant == 18) ? (int)c.variant : 2;
	int nw = c.waves_per_group ? (int)c.waves_per_group : (c.schedule == TM_SCHED_PLAIN ? 8 : 4);
	if (nw != 4 && nw != 8 && nw != 16)
		return set_error(TM_ERR_INVALID, "waves_per_group must be 4, 8 or 16 (got %d)", nw);
	r->waves = nw;
	size_t static_lds = nq_grp * 32 + 64 + (mode_of(c) == 1 ? (size_t)nw * HBUF * 8 : 8) +
						((mode_of(c) == 0 && (r->variant & 2)) ? (size_t)nw * 64 * 16 : 16);
	size_t bm = nq_grp * r->kw * 4;
	r->lds_bm = (mode_of(c) == 0) && bm + static_lds <= max_lds;
	r->stage_lds = r->lds_bm ? bm : 0;
	const int want = c.groups_per_cu ? (int)c.groups_per_cu : 64;
	for (int i = 0; i < 3; ++i) {
		// resident workgroups per CU from the kernel's registers and LDS
		const int w = 4 << i;
		const mq_fn fn = pick_match_q(mode_of(c), r->lds_bm, r->tiles, w, r->variant);
		int n = 0;
		if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, reinterpret_cast<const void *>(fn), w * 64,
														  r->stage_lds) != hipSuccess || n < 1) {
			const size_t st = nq_grp * 32 + 64 + (mode_of(c) == 1 ? (size_t)w * HBUF * 8 : 8) +
							  ((mode_of(c) == 0 && (r->variant & 2)) ? (size_t)w * 64 * 16 : 16);
			n = std::max(1, std::min((int)(std::min<size_t>(160 * 1024, max_lds) / (r->stage_lds + st)), 28 / w));
		}
		r->wgs_for[i] = std::max(1, std::min(want, n));
	}
	r->wgs_per_cu = r->wgs_for[nw == 4 ? 0 : nw == 8 ? 1 : 2];
	return TM_OK;
}

// Waves per workgroup of a persistent step of `tiles` query tiles: the
// configured value, or by default more waves for fewer tiles (a small step
// has few items in flight per tile, and more waves share each item's blocks).
int step_waves(const tm_ring * r, uint64_t tiles, bool compacted) {
	if (r->cfg.waves_per_group) return r->waves;
	// a compacted step has few tiles but many segments: 4 waves (measured at
	// 10M filters, 1M queries: 2.8 ms against 3.0 ms with 8)
	if (compacted) return 4;
	return tiles <= 16 ? 16 : tiles <= 256 ? 8 : 4;
}
int wgs_for_waves(const tm_ring * r, int nw) { return r->wgs_for[nw == 4 ? 0 : nw == 8 ? 1 : 2]; }

// Block weights for equal-work segments (k_block_weights + prefix sum), kept
// while the FIB and the submissions' bit density, rounded to a quarter of a
// binary order, stay the same: they only balance the segments.
void ensure_weights(tm_ring * r, hipStream_t s) {
	const uint32_t nblk = r->fib->n_blocks;
	const float l2 = std::round(r->rho_l2 * 4.f) / 4.f;
	if (r->bw_version == r->fib->version && r->bw_nb == nblk && r->bw_l2 == l2) return;
	hipLaunchKernelGGL(k_block_weights, dim3((nblk + 3) / 4), dim3(256), 0, s, r->fib->view(), nblk, l2, r->d_bw);
	scan_u32_to_u64(r->d_bw, r->d_bprefix, nblk, r->d_bscan, s);
	r->bw_version = r->fib->version;
	r->bw_nb = nblk;
	r->bw_l2 = l2;
}

int grow_u32(uint32_t ** p, uint64_t & cap, uint64_t need) {
	if (need <= cap && *p) return TM_OK;
	dfree(*p);
	uint64_t n = std::max<uint64_t>(need, 1);
	TM_CHECK(dalloc(p, n));
	cap = n;
	return TM_OK;
}

// Enqueue lookup + item building (shared by both schedules).
void enqueue_front(tm_ring * r, uint32_t nq) {
	hipStream_t s = r->s;
	const tm_fib * f = r->fib;
	const uint32_t P = r->P;
	fill_u32(r->d_pcount, 0, P, s);
	fill_u64(reinterpret_cast<uint64_t *>(r->d_cnt), 0, 8, s);
	uint32_t g = (nq + 255) / 256;
	if (g == 0) g = 1;
	hipLaunchKernelGGL(k_lookup_count, dim3(g), dim3(256), 0, s, r->d_q, nq, f->part_info, P, r->d_pcount);
	scan_u32_to_u64(r->d_pcount, r->d_pbase, P, r->d_scan, s);
	copy_u64(reinterpret_cast<uint64_t *>(r->d_pcur), r->d_pbase, P, s);
	hipLaunchKernelGGL(k_lookup_fill, dim3(g), dim3(256), 0, s, r->d_q, nq, f->part_info, P, r->d_pcur,
					   r->d_list, r->list_cap);
	hipLaunchKernelGGL(k_build_items, dim3(1), dim3(1024), 0, s, r->d_pcount, r->d_pbase, r->list_cap,
					   f->part_info, P, r->have_order ? r->d_porder : (const uint32_t *)nullptr,
					   r->plain_bq, r->chunk, r->d_sched, r->d_iprefix, r->d_qbeg, r->d_qnext);
}

void enqueue_finalize(tm_ring * r, uint32_t nq) {
	hipStream_t s = r->s;
	const int mode = r->cfg.mode;
	uint32_t gq = (nq + 3) / 4;
	if (gq == 0) gq = 1;
	if (mode == TM_MODE_KEY_BITMAP) {
		// the rows are the results: no expansion
	} else if (mode == TM_MODE_KEYS_UNIQUE) {
		hipLaunchKernelGGL(k_bm_count, dim3(gq), dim3(256), 0, s, r->d_gbm, r->kw, nq, r->d_qcount);
		scan_u32_to_u64(r->d_qcount, r->d_off, nq, r->d_scan, s);
		hipLaunchKernelGGL(k_bm_write, dim3(std::max(1u, (nq + BM_WAVES - 1) / BM_WAVES)), dim3(BM_WAVES * 64), 0, s,
						   r->d_gbm, r->kw, nq, r->d_off, r->fib->key_dict, r->d_vals, r->val_cap);
	} else {
		const bool keys = mode == TM_MODE_KEYS_ALL;
		dev_fib_view v = r->fib->view();
		fill_u32(r->d_qcount, 0, nq, s);
		uint32_t gh = 2048;
		if (keys)
			hipLaunchKernelGGL(k_hits_count<true>, dim3(gh), dim3(256), 0, s, r->d_hits, r->d_cnt, r->hit_cap, v, r->d_qcount);
		else
			hipLaunchKernelGGL(k_hits_count<false>, dim3(gh), dim3(256), 0, s, r->d_hits, r->d_cnt, r->hit_cap, v, r->d_qcount);
		scan_u32_to_u64(r->d_qcount, r->d_off, nq, r->d_scan, s);
		copy_u64(reinterpret_cast<uint64_t *>(r->d_qcur), r->d_off, nq, s);
		if (keys)
			hipLaunchKernelGGL(k_hits_write<true>, dim3(gh), dim3(256), 0, s, r->d_hits, r->d_cnt, r->hit_cap, v,
							   r->d_qcur, r->d_vals, r->val_cap);
		else
			hipLaunchKernelGGL(k_hits_write<false>, dim3(gh), dim3(256), 0, s, r->d_hits, r->d_cnt, r->hit_cap, v,
							   r->d_qcur, r->d_vals, r->val_cap);
	}
	hipLaunchKernelGGL(k_gather_stats, dim3(1), dim3(64), 0, s, r->d_pbase, r->P, r->d_cnt,
					   mode == TM_MODE_KEY_BITMAP ? (const uint64_t *)nullptr : r->d_off, nq, r->d_iprefix, r->d_stat);
}

// The persistent step: everything between the query upload and the stats.
int enqueue_persistent(tm_ring * r, uint32_t nq, bool events) {
	hipStream_t s = r->s;
	const uint32_t P = r->P;
	fill_u32(r->d_pcount, 0, P, s);
	fill_u64(reinterpret_cast<uint64_t *>(r->d_cnt), 0, 8, s);
	const uint64_t nqg = (uint64_t)r->tiles * 64;
	const uint64_t tiles = (nq + nqg - 1) / nqg;
	const bool exit_on = keyrows(r->cfg) && r->lds_bm && r->kw > 0 && r->cfg.full_exit != 2;
	const bool rich_pass = exit_on && r->fib->n_rich > 0 && nq > 0;
	// Compacted step: only the queries whose rows the pass leaves incomplete
	// are matched (large steps only: the extra kernels cost about 0.1 ms,
	// more than a step of up to ~8k queries gains)
	const uint64_t compact_min = r->cfg.compact_tiles ? r->cfg.compact_tiles : COMPACT_MIN_TILES;
	const bool compact = rich_pass && tiles > compact_min && !r->use_iorder && !r->cfg.log_units &&
						 (r->variant & 1) == 0;
	const int nw = step_waves(r, tiles, compact);
	// Small submissions: cut each tile's work into block segments so that
	// the items fill the GPU (about two per resident workgroup).  Needs the
	// key bitmaps in LDS (flushed with atomicOr) or the hit-list modes.
	const uint64_t resident = (uint64_t)dev().num_cu * wgs_for_waves(r, nw);
	// Segments also bound the blocks a work item streams: with segment-major
	// items, the workgroups on an XCD share one segment's blocks (at most
	// SEG_MAX_BLOCKS x 2 KiB) in its L2 instead of each streaming the whole
	// table from HBM.
	uint32_t nseg = 1;
	const uint32_t nblk = r->fib->n_blocks;
	const bool can_cut = (r->lds_bm || !keyrows(r->cfg)) && !r->use_iorder && !r->cfg.log_units;
	if (can_cut && tiles > 0) {
		nseg = (uint32_t)std::min<uint64_t>((2 * resident + tiles - 1) / tiles, std::max<uint32_t>(nblk / 16, 1));
		const uint32_t seg_max = keyrows(r->cfg) && r->lds_bm && r->cfg.full_exit != 2
									 ? SEG_MAX_BLOCKS_EXIT : SEG_MAX_BLOCKS;
		nseg = std::max<uint32_t>(nseg, (nblk + seg_max - 1) / seg_max);
	}
	if (r->cfg.chunk_blocks && (r->lds_bm || !keyrows(r->cfg)))   // fixed segment size
		nseg = std::max<uint32_t>(1, (nblk + r->cfg.chunk_blocks - 1) / r->cfg.chunk_blocks);
	const uint32_t seg_blocks = nseg > 1 ? (nblk + nseg - 1) / nseg : std::max<uint32_t>(nblk, 1);
	if (nseg > 1) nseg = (nblk + seg_blocks - 1) / seg_blocks;
	const uint64_t items = tiles * nseg;
	// Few tiles: segments of equal estimated work rather than equal block
	// counts.  Blocks of low-popcount filters are matched by most dense
	// queries; uniform segments put them in a few long work items, which
	// then set the step's time.
	const bool weighted = nseg > 1 && tiles <= WSEG_MAX_TILES && !r->cfg.chunk_blocks && !r->no_weights;
	r->step_weighted = weighted;
	if (weighted) {
		// a captured step leaves the weights out: the caller refreshes them
		// before each replay
		if (!r->capturing) ensure_weights(r, s);
		hipLaunchKernelGGL(k_seg_bounds, dim3((nseg + 256) / 256), dim3(256), 0, s, r->d_bprefix, nblk, nseg,
						   r->d_segb);
	}
	r->last_nseg = nseg;
	const uint32_t last_mask = (r->fib->n_keys & 31) ? (1u << (r->fib->n_keys & 31)) - 1u : 0xFFFFFFFFu;
	if (rich_pass) {
		// rows start with the keys of the rich filters each query contains
		// (written whole, with the flags); full rows need no test at all
		// queries per wave: 64 for large steps, fewer (down to 1) below
		// 4096 waves' worth
		uint32_t qpw = 64;
		while (qpw > 1 && (uint64_t)nq < (uint64_t)qpw * 4096) qpw >>= 1;
		const uint32_t waves = (nq + qpw - 1) / qpw;
		hipLaunchKernelGGL(k_rich_rows, dim3((waves + 3) / 4), dim3(256), 0, s, r->d_q, nq, qpw, r->fib->d_rich_words,
						   r->fib->d_rich_bm, r->fib->n_rich, r->fib->n_rich_full, r->kw, last_mask, r->d_gbm,
						   r->d_qfull);
	} else if (nseg > 1 && keyrows(r->cfg)) {
		fill_u32(r->d_gbm, 0, (uint64_t)nq * r->kw, s);
		fill_u32(r->d_qfull, 0, nq, s);
	}
	// Compacted step: the active queries in tiles of their own (rows and
	// flags through the list); their count stays on the device.  The lookup
	// counts of every query come from k_lookup_hist, as the match sees only
	// the active ones.
	if (compact) {
		fill_u32(r->d_nact, 0, 4, s);
		if (P <= 16384) {
			const uint32_t gh = (uint32_t)std::min<uint64_t>((nq + 255) / 256, 2048);
			hipLaunchKernelGGL(k_lookup_hist, dim3(gh), dim3(256), (size_t)P * 4, s, r->d_q, nq, r->fib->part_info, P,
							   r->d_pcount);
		} else {
			hipLaunchKernelGGL(k_lookup_count, dim3((nq + 255) / 256), dim3(256), 0, s, r->d_q, nq, r->fib->part_info,
							   P, r->d_pcount);
		}
		act_key_bits kb;
		for (uint32_t i = 0; i < ACT_KEY_BITS; ++i) kb.pos[i] = (uint8_t)r->fib->h_order[i];
		fill_u32(r->d_ahist, 0, 2 * ACT_BUCKETS, s);   // counts, then cursors
		hipLaunchKernelGGL(k_act_hist, dim3((uint32_t)std::min<uint64_t>((nq + 255) / 256, 1024)), dim3(256), 0, s,
						   r->d_q, nq, r->d_qfull, kb, r->d_ahist);
		scan_u32_to_u64(r->d_ahist, r->d_abase, ACT_BUCKETS, r->d_scan, s);
		hipLaunchKernelGGL(k_act_scatter, dim3((nq + 255) / 256), dim3(256), 0, s, r->d_q, nq, r->d_qfull, kb,
						   r->d_abase, r->d_ahist + ACT_BUCKETS, r->d_active, r->d_nact);
		const uint32_t seg_max = r->cfg.chunk_blocks ? r->cfg.chunk_blocks : SEG_MAX_BLOCKS_COMPACT;
		hipLaunchKernelGGL(k_init_queues_active, dim3(1), dim3(64), 0, s, r->d_nact, (uint32_t)nqg, nblk,
						   (uint32_t)resident, seg_max, r->d_qbeg, r->d_qnext, r->d_iprefix + P);
	} else {
		hipLaunchKernelGGL(k_init_queues, dim3(1), dim3(64), 0, s, items, r->d_qbeg, r->d_qnext);
	}
	mq_args a;
	std::memset(&a, 0, sizeof(a));
	a.f = r->fib->view();
	a.q = r->d_q;
	a.nq = nq;
	a.nq_dev = compact ? r->d_nact : nullptr;
	a.ntiles_dev = compact ? r->d_nact + 1 : nullptr;
	a.map_rows = compact;
	a.porder = r->have_order ? r->d_porder : nullptr;
	a.iorder = r->use_iorder ? r->d_iorder : nullptr;
	a.full_exit = exit_on;
	a.groups = compact;
	a.last_mask = last_mask;
	a.slot_epoch = r->fib->slot_epoch;
	a.epoch = r->epoch;
	a.step_vals = (r->cfg.use_graph && keyrows(r->cfg) && r->d_step) ? r->d_step : nullptr;
	a.ntiles = (uint32_t)tiles;
	a.nseg = nseg;
	a.seg_blocks = seg_blocks;
	a.segb = weighted && !compact ? r->d_segb : nullptr;
	a.blk_lo = 0;
	a.blk_hi = nblk;
	a.log_unit = -1;
	a.qmap = compact ? r->d_active : nullptr;
	a.qnext = r->d_qnext;
	a.qbeg = r->d_qbeg;
	a.pcount = compact ? nullptr : r->d_pcount;
	a.gbm = r->d_gbm;
	// flags hold this submission's state: cleared or set by the pass above
	a.qfull = (nseg > 1 || rich_pass) ? r->d_qfull : nullptr;
	r->qfull_valid = a.qfull != nullptr;   // set flags mean full rows this step
	a.hits = r->d_hits;
	a.hit_cap = r->hit_cap;
	a.counters = r->d_cnt;
	a.log = r->cfg.log_units ? r->d_log : nullptr;
	a.log_cap = r->log_cap;
	if (r->variant & 1) {
		a.qinv = r->d_qinv;
		const uint64_t nw = (uint64_t)nq * 8;
		hipLaunchKernelGGL(k_invert_queries, dim3((uint32_t)std::max<uint64_t>(1, (nw + 255) / 256)), dim3(256), 0, s,
						   r->d_q, nq, r->d_qinv);
	}
	mq_fn fn = pick_match_q(mode_of(r->cfg), r->lds_bm, r->tiles, nw, r->variant);
	uint32_t grid = (uint32_t)std::min<uint64_t>(resident, std::max<uint64_t>(items, 1));
	if (events) TM_HIP(hipEventRecord(r->em0, s));
	hipLaunchKernelGGL(fn, dim3(grid), dim3(nw * 64), r->stage_lds, s, a);
	TM_HIP(hipGetLastError());
	if (events) TM_HIP(hipEventRecord(r->em1, s));
	scan_u32_to_u64(r->d_pcount, r->d_pbase, P, r->d_scan, s);
	if (!compact) fill_u64(r->d_iprefix + P, items, 1, s);
	enqueue_finalize(r, nq);
	return TM_OK;
}

// A (partition, query batch) unit of the plain schedule: its queries are
// list[qbase .. qbase + nqv) (the lookup list, or the identity for explicit
// units), its blocks [blo, bhi), cut into nseg segments of seg_blocks.
struct punit {
	uint32_t p, b, nqv, blo, bhi, ntiles, nseg, seg_blocks;
	uint64_t qbase;
};

void plain_segments(const tm_ring * r, punit & u) {
	const uint32_t nqg = 64 * r->tiles;
	const uint64_t resident = (uint64_t)dev().num_cu * r->wgs_per_cu;
	const uint32_t nb = u.bhi - u.blo;
	u.ntiles = (u.nqv + nqg - 1) / nqg;
	// segments: about two items per resident workgroup, at least
	// r->cfg.chunk_blocks (or 64) blocks each: an item's fixed cost (its
	// bitmap rows cleared and merged) is about that of 30 block visits
	const uint32_t min_seg = r->cfg.chunk_blocks ? r->cfg.chunk_blocks : 64;
	const uint64_t want = (2 * resident + std::max<uint32_t>(u.ntiles, 1) - 1) / std::max<uint32_t>(u.ntiles, 1);
	uint32_t nseg = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(want, (nb + min_seg - 1) / min_seg));
	if (keyrows(r->cfg) && !r->lds_bm) nseg = 1;   // rows in HBM: one workgroup per tile
	u.seg_blocks = std::max<uint32_t>(1, (nb + nseg - 1) / nseg);
	u.nseg = nb ? (nb + u.seg_blocks - 1) / u.seg_blocks : 1;
}

int run_plain(tm_ring * r, uint32_t nq, const uint32_t * list, std::vector<punit> & units);

// The plain step: one launch per (partition, batch) unit, as the reference
// launches one kernel per (partition, batch) (back_end.cc:243-286).  Each
// launch runs the persistent kernel's inner loop over the unit only: the
// batch's queries (through the lookup list) against the partition's blocks,
// cut into block segments so that the launch spreads over the GPU; key
// bitmaps go to staging rows that k_or_rows merges into the queries' rows.
int enqueue_plain(tm_ring * r, uint32_t nq) {
	hipStream_t s = r->s;
	if (keyrows(r->cfg)) fill_u32(r->d_gbm, 0, (uint64_t)nq * r->kw, s);
	enqueue_front(r, nq);
	std::vector<uint4> sched(r->P);
	std::vector<uint64_t> pbase(r->P + 1, 0);
	if (r->P) {
		TM_HIP(hipMemcpyAsync(sched.data(), r->d_sched, sizeof(uint4) * r->P, hipMemcpyDeviceToHost, s));
		TM_HIP(hipMemcpyAsync(pbase.data(), r->d_pbase, sizeof(uint64_t) * (r->P + 1), hipMemcpyDeviceToHost, s));
		TM_HIP(hipStreamSynchronize(s));
	}
	const uint32_t bq = r->plain_bq;
	const tm_fib * f = r->fib;
	std::vector<punit> units;
	for (uint32_t k = 0; k < r->P; ++k) {
		const uint4 sc = sched[k];   // partition, batches, 1, queries
		if (sc.z == 0) continue;
		const uint32_t p = sc.x;
		const uint32_t blo = f->h_part_info[(size_t)p * 8 + 6], nb = f->h_part_info[(size_t)p * 8 + 7];
		for (uint32_t b = 0; b < sc.y; ++b) {
			punit u;
			u.p = p;
			u.b = b;
			u.nqv = std::min(bq, sc.w - b * bq);
			u.blo = blo;
			u.bhi = blo + nb;
			u.qbase = pbase[p] + (uint64_t)b * bq;
			plain_segments(r, u);
			units.push_back(u);
		}
	}
	return run_plain(r, nq, r->d_list, units);
}

int run_plain(tm_ring * r, uint32_t nq, const uint32_t * list, std::vector<punit> & units) {
	hipStream_t s = r->s;
	const uint32_t bq = r->plain_bq;
	const uint64_t resident = (uint64_t)dev().num_cu * r->wgs_per_cu;
	const tm_fib * f = r->fib;
	const uint32_t nu = (uint32_t)units.size();
	// per-unit work queues, initialised in one launch
	TM_CHECK(grow_u32(&r->d_uitems, r->uitems_cap, std::max<uint32_t>(nu, 1)));
	if (r->uq_cap < nu) {
		dfree(r->d_ubeg);
		dfree(r->d_unext);
		r->uq_cap = 0;
		TM_CHECK(dalloc(&r->d_ubeg, (size_t)std::max<uint32_t>(nu, 1) * (NQUEUE + 1)));
		TM_CHECK(dalloc(&r->d_unext, (size_t)std::max<uint32_t>(nu, 1) * NQUEUE * QPAD));
		r->uq_cap = nu;
	}
	if (keyrows(r->cfg) && r->stage_cap < (uint64_t)bq * r->kw) {
		TM_CHECK(grow_u32(&r->d_stage, r->stage_cap, (uint64_t)bq * r->kw));
		fill_u32(r->d_stage, 0, r->stage_cap, s);
	}
	r->n_units = 0;
	if (r->cfg.log_units) {
		// one log row per launch: (partition, batch, queries in the batch)
		std::vector<uint4> ud(nu);
		for (uint32_t i = 0; i < nu; ++i)
			ud[i] = make_uint4(units[i].p, units[i].b, units[i].nqv, (uint32_t)units[i].qbase);
		const uint32_t nl = (uint32_t)std::min<uint64_t>(nu, r->log_cap);
		if (nl > r->units_cap) {
			dfree(r->d_units);
			dfree(r->d_ufeat);
			r->units_cap = 0;
			TM_CHECK(dalloc(&r->d_units, nl));
			TM_CHECK(dalloc(&r->d_ufeat, (size_t)nl * N_FEAT));
			r->units_cap = nl;
		}
		if (nl) {
			TM_HIP(hipMemcpyAsync(r->d_units, ud.data(), sizeof(uint4) * nl, hipMemcpyHostToDevice, s));
			fill_u32(r->d_log, 0, (uint64_t)nl * 3, s);
			launch_unit_features_plain(f->view(), r->d_q, list, r->d_units, nl, r->d_ufeat, s);
			TM_HIP(hipStreamSynchronize(s));   // ud is pageable host memory
		}
		r->n_units = nl;
	}
	std::vector<uint32_t> items(std::max<uint32_t>(nu, 1), 0);
	for (uint32_t i = 0; i < nu; ++i) items[i] = units[i].ntiles * units[i].nseg;
	TM_HIP(hipMemcpyAsync(r->d_uitems, items.data(), 4ull * nu, hipMemcpyHostToDevice, s));
	if (nu)
		hipLaunchKernelGGL(k_init_unit_queues, dim3((nu + 255) / 256), dim3(256), 0, s, r->d_uitems, nu, r->d_ubeg,
						   r->d_unext);
	TM_HIP(hipStreamSynchronize(s));   // items is pageable host memory

	mq_args a;
	std::memset(&a, 0, sizeof(a));
	a.f = f->view();
	a.q = r->d_q;
	a.hits = r->d_hits;
	a.hit_cap = r->hit_cap;
	a.counters = r->d_cnt;
	a.log_cap = r->log_cap;
	a.full_exit = 0;   // the staging rows start empty for every unit: nothing to learn
	a.slot_epoch = f->slot_epoch;
	a.epoch = r->epoch;
	a.last_mask = 0xFFFFFFFFu;
	mq_fn fn = pick_match_q(mode_of(r->cfg), r->lds_bm, r->tiles, r->waves, r->variant);
	TM_HIP(hipEventRecord(r->em0, s));
	for (uint32_t i = 0; i < nu; ++i) {
		const punit & u = units[i];
		a.qmap = list + u.qbase;
		a.nq = u.nqv;
		a.gbm = r->d_stage;
		a.blk_lo = u.blo;
		a.blk_hi = u.bhi;
		a.ntiles = u.ntiles;
		a.nseg = u.nseg;
		a.seg_blocks = u.seg_blocks;
		a.qbeg = r->d_ubeg + (size_t)i * (NQUEUE + 1);
		a.qnext = r->d_unext + (size_t)i * NQUEUE * QPAD;
		a.log = (r->cfg.log_units && i < r->n_units) ? r->d_log : nullptr;
		a.log_unit = a.log ? (int32_t)i : -1;
		const uint64_t nit = (uint64_t)u.ntiles * u.nseg;
		const uint32_t grid = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(resident, nit));
		hipLaunchKernelGGL(fn, dim3(grid), dim3(r->waves * 64), r->stage_lds, s, a);
		if (keyrows(r->cfg))
			hipLaunchKernelGGL(k_or_rows, dim3((u.nqv * r->kw + 255) / 256), dim3(256), 0, s, r->d_stage, r->kw,
							   a.qmap, u.nqv, r->d_gbm);
	}
	TM_HIP(hipGetLastError());
	TM_HIP(hipEventRecord(r->em1, s));
	r->st.n_items = nu;
	enqueue_finalize(r, nq);
	return TM_OK;
}

// Copies of a ring's buffers go through its own non-blocking stream, which is
// then waited for: while another thread of the process captures a graph, the
// runtime refuses operations on the legacy stream ("would make the legacy
// stream depend on a capturing blocking stream").
hipError_t ring_copy(tm_ring * r, void * dst, const void * src, size_t bytes, hipMemcpyKind kind) {
	hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, r->s);
	return e == hipSuccess ? hipStreamSynchronize(r->s) : e;
}

int ensure_capacity(tm_ring * r) {
	const uint32_t P = r->P;
	const uint32_t nblk = r->fib->n_blocks;
	if (r->bw_cap < nblk || !r->d_bw) {   // weighted segments: sized by the table
		dfree(r->d_bw); dfree(r->d_bprefix); dfree(r->d_bscan); dfree(r->d_segb);
		const uint32_t cap = std::max<uint32_t>(nblk, 1);
		TM_CHECK(dalloc(&r->d_bw, cap));
		TM_CHECK(dalloc(&r->d_bprefix, (size_t)cap + 1));
		TM_CHECK(dalloc(&r->d_bscan, scan_tmp_words(cap)));
		TM_CHECK(dalloc(&r->d_segb, (size_t)cap + 1));
		r->bw_cap = cap;
		r->bw_version = ~0ull;
	}
	if (!r->d_pcount) {
		TM_CHECK(dalloc(&r->d_pcount, P + 1));
		TM_CHECK(dalloc(&r->d_pbase, P + 1));
		TM_CHECK(dalloc(&r->d_pcur, P + 1));
		TM_CHECK(dalloc(&r->d_sched, P + 1));
		TM_CHECK(dalloc(&r->d_iprefix, P + 1));
		TM_CHECK(dalloc(&r->d_qbeg, NQUEUE + 1));
		TM_CHECK(dalloc(&r->d_qnext, NQUEUE * QPAD));
		TM_CHECK(dalloc(&r->d_porder, P + 1));
		TM_CHECK(dalloc(&r->d_cnt, 8));
		TM_CHECK(dalloc(&r->d_stat, 8));
		TM_CHECK(dalloc(&r->d_qcount, (size_t)r->max_q + 1));
		TM_CHECK(dalloc(&r->d_off, (size_t)r->max_q + 1));
		TM_CHECK(dalloc(&r->d_qcur, (size_t)r->max_q + 1));
		r->scan_words = std::max(scan_tmp_words(r->max_q), scan_tmp_words(P));
		TM_CHECK(dalloc(&r->d_scan, r->scan_words));
	}
	if (keyrows(r->cfg) && !r->d_gbm) {
		size_t n = (size_t)r->max_q * r->kw;
		TM_CHECK(dalloc(&r->d_gbm, n));
		TM_HIP(hipMemsetAsync(r->d_gbm, 0, std::max<size_t>(n, 1) * 4, r->s));
		TM_HIP(hipStreamSynchronize(r->s));
	}
	if (keyrows(r->cfg) && !r->d_qfull) {
		TM_CHECK(dalloc(&r->d_qfull, std::max<uint32_t>(r->max_q, 1)));
		TM_CHECK(dalloc(&r->d_active, std::max<uint32_t>(r->max_q, 1)));
		TM_CHECK(dalloc(&r->d_nact, 4));
		TM_CHECK(dalloc(&r->d_ahist, 2 * ACT_BUCKETS));
		TM_CHECK(dalloc(&r->d_abase, ACT_BUCKETS + 1));
	}
	if (r->scan_words < scan_tmp_words(ACT_BUCKETS)) {   // the key buckets' scan
		dfree(r->d_scan);
		r->scan_words = scan_tmp_words(ACT_BUCKETS);
		TM_CHECK(dalloc(&r->d_scan, r->scan_words));
	}
	if (!keyrows(r->cfg) && !r->d_hits) {
		uint64_t cap = r->cfg.hit_capacity ? r->cfg.hit_capacity : std::max<uint64_t>(1 << 20, (uint64_t)r->max_q * 64);
		TM_CHECK(dalloc(&r->d_hits, cap));
		r->hit_cap = cap;
	}
	if (!r->d_list && r->cfg.schedule == TM_SCHED_PLAIN) {
		uint64_t cap = (uint64_t)r->max_q * std::min<uint32_t>(std::max<uint32_t>(P, 1), 64);
		TM_CHECK(grow_u32(&r->d_list, r->list_cap, cap));
	}
	if (!r->d_vals) {
		// KEYS_UNIQUE: every key for every query when that is at most 64M
		// values (no overflow retry then), else a first guess that tm_wait grows
		uint64_t per = r->cfg.mode == TM_MODE_KEYS_UNIQUE ? std::min<uint32_t>(r->fib->n_keys, 256) : 64;
		if (r->cfg.mode == TM_MODE_KEYS_UNIQUE && (uint64_t)r->max_q * r->fib->n_keys <= (1ull << 26))
			per = std::max<uint32_t>(r->fib->n_keys, 1);
		if (r->cfg.mode == TM_MODE_KEY_BITMAP) per = 0;   // the rows are the results
		TM_CHECK(grow_u32(&r->d_vals, r->val_cap, (uint64_t)r->max_q * std::max<uint64_t>(per, 1)));
	}
	if (r->cfg.log_units && !r->d_log) {
		r->log_cap = 1 << 22;
		TM_CHECK(dalloc(&r->d_log, r->log_cap * 3));
	}
	return TM_OK;
}

__global__ void __launch_bounds__(256) k_iota(uint32_t * __restrict__ p, uint32_t n) {
	const uint32_t i = blockIdx.x * 256 + threadIdx.x;
	if (i < n) p[i] = i;
}

// Explicit units (tm_submit_units): queries [u_off[u], u_off[u+1]) against
// partition u_part[u] only, without the lookup.  PLAIN: one launch per
// (unit, batch); PERSISTENT: one work-stealing launch over every unit's
// (segment, tile) items.
int enqueue_units(tm_ring * r, uint32_t nq) {
	hipStream_t s = r->s;
	const tm_fib * f = r->fib;
	const uint32_t n = (uint32_t)r->u_part.size();
	if (keyrows(r->cfg)) {
		fill_u32(r->d_gbm, 0, (uint64_t)nq * r->kw, s);
		fill_u32(r->d_qfull, 0, nq, s);
	}
	fill_u32(r->d_pcount, 0, r->P, s);
	fill_u64(reinterpret_cast<uint64_t *>(r->d_cnt), 0, 8, s);
	scan_u32_to_u64(r->d_pcount, r->d_pbase, r->P, r->d_scan, s);
	if (r->cfg.schedule == TM_SCHED_PLAIN) {
		if (!r->d_iota) {
			TM_CHECK(dalloc(&r->d_iota, std::max<uint32_t>(r->max_q, 1)));
			hipLaunchKernelGGL(k_iota, dim3((r->max_q + 255) / 256), dim3(256), 0, s, r->d_iota, r->max_q);
		}
		const uint32_t bq = r->plain_bq;
		std::vector<punit> units;
		for (uint32_t i = 0; i < n; ++i) {
			const uint32_t p = r->u_part[i], q0 = r->u_off[i], qn = r->u_off[i + 1] - q0;
			const uint32_t blo = f->h_part_info[(size_t)p * 8 + 6], nb = f->h_part_info[(size_t)p * 8 + 7];
			if (nb == 0) continue;
			for (uint32_t b = 0; b * bq < qn; ++b) {
				punit u;
				u.p = p;
				u.b = b;
				u.nqv = std::min(bq, qn - b * bq);
				u.blo = blo;
				u.bhi = blo + nb;
				u.qbase = (uint64_t)q0 + (uint64_t)b * bq;
				plain_segments(r, u);
				units.push_back(u);
			}
		}
		return run_plain(r, nq, r->d_iota, units);
	}
	// persistent: segment size from the whole submission's work
	const uint32_t nqg = 64 * r->tiles;
	const uint64_t resident = (uint64_t)dev().num_cu * r->wgs_per_cu;
	uint64_t work = 0;   // tile x blocks
	for (uint32_t i = 0; i < n; ++i) {
		const uint32_t p = r->u_part[i], qn = r->u_off[i + 1] - r->u_off[i];
		work += (uint64_t)((qn + nqg - 1) / nqg) * f->h_part_info[(size_t)p * 8 + 7];
	}
	uint32_t seg = r->cfg.chunk_blocks ? r->cfg.chunk_blocks
									   : (uint32_t)std::max<uint64_t>(4, (work + 4 * resident - 1) / (4 * resident));
	if (keyrows(r->cfg) && !r->lds_bm) seg = 0xFFFFFFFFu;   // rows in HBM: one workgroup per tile
	std::vector<uint32_t> tab((size_t)std::max<uint32_t>(n, 1) * 8, 0), pre(n + 1, 0);
	uint32_t max_nseg = 1;
	uint64_t items = 0;
	const bool tuned = !(keyrows(r->cfg) && !r->lds_bm) && !r->cfg.chunk_blocks;
	for (uint32_t i = 0; i < n; ++i) {
		const uint32_t p = r->u_part[i], q0 = r->u_off[i], qn = r->u_off[i + 1] - q0;
		const uint32_t blo = f->h_part_info[(size_t)p * 8 + 6], nb = f->h_part_info[(size_t)p * 8 + 7];
		const uint32_t ntiles = (qn + nqg - 1) / nqg;
		const uint32_t useg = tuned && r->seg_bucket[size_bucket(nb)] ? r->seg_bucket[size_bucket(nb)] : seg;
		const uint32_t sb = std::max<uint32_t>(1, std::min<uint32_t>(useg, std::max<uint32_t>(nb, 1)));
		const uint32_t nseg = nb ? (nb + sb - 1) / sb : 1;
		uint32_t * t = &tab[(size_t)i * 8];
		t[0] = p; t[1] = q0; t[2] = qn; t[3] = std::max<uint32_t>(ntiles, 1);
		t[4] = nseg; t[5] = sb; t[6] = blo; t[7] = blo + nb;
		pre[i] = (uint32_t)items;
		items += ntiles ? (uint64_t)ntiles * nseg : 0;
		if (ntiles == 0) t[3] = 1;   // no items: the value only avoids a division
		max_nseg = std::max(max_nseg, nseg);
	}
	if (items > 0xFFFFFFF0ull) return set_error(TM_ERR_INVALID, "too many work items");
	pre[n] = (uint32_t)items;
	TM_CHECK(grow_u32(&r->d_xunits, r->xunits_cap, tab.size()));
	TM_CHECK(grow_u32(&r->d_xprefix, r->xprefix_cap, pre.size()));
	TM_HIP(hipMemcpyAsync(r->d_xunits, tab.data(), tab.size() * 4, hipMemcpyHostToDevice, s));
	TM_HIP(hipMemcpyAsync(r->d_xprefix, pre.data(), pre.size() * 4, hipMemcpyHostToDevice, s));
	hipLaunchKernelGGL(k_init_queues, dim3(1), dim3(64), 0, s, items, r->d_qbeg, r->d_qnext);
	mq_args a;
	std::memset(&a, 0, sizeof(a));
	a.f = f->view();
	a.q = r->d_q;
	a.nq = nq;
	a.units = r->d_xunits;
	a.uprefix = r->d_xprefix;
	a.slot_epoch = f->slot_epoch;
	a.epoch = r->epoch;
	a.n_units = n;
	a.ntiles = 1;
	a.nseg = max_nseg;
	a.seg_blocks = 1;
	a.blk_lo = 0;
	a.blk_hi = f->n_blocks;
	a.log_unit = -1;
	a.qnext = r->d_qnext;
	a.qbeg = r->d_qbeg;
	a.gbm = r->d_gbm;
	a.qfull = r->d_qfull;
	a.hits = r->d_hits;
	a.hit_cap = r->hit_cap;
	a.counters = r->d_cnt;
	a.full_exit = keyrows(r->cfg) && r->lds_bm && r->kw > 0 && r->cfg.full_exit != 2;
	a.last_mask = (f->n_keys & 31) ? (1u << (f->n_keys & 31)) - 1u : 0xFFFFFFFFu;
	mq_fn fn = pick_match_q(mode_of(r->cfg), r->lds_bm, r->tiles, r->waves, r->variant);
	const uint32_t grid = (uint32_t)std::min<uint64_t>(resident, std::max<uint64_t>(items, 1));
	TM_HIP(hipEventRecord(r->em0, s));
	hipLaunchKernelGGL(fn, dim3(grid), dim3(r->waves * 64), r->stage_lds, s, a);
	TM_HIP(hipGetLastError());
	TM_HIP(hipEventRecord(r->em1, s));
	TM_HIP(hipStreamSynchronize(s));   // tab and pre are pageable host memory
	fill_u64(r->d_iprefix + r->P, items, 1, s);
	enqueue_finalize(r, nq);
	return TM_OK;
}

// KEY_BITMAP with host delivery: the step's rows into the caller's host
// buffer, on the ring's stream after the step (outside any captured graph).
int enqueue_delivery(tm_ring * r, uint32_t nq) {
	r->pack_sent = 0;
	if (r->cfg.mode != TM_MODE_KEY_BITMAP || nq == 0 || r->kw == 0) return TM_OK;
	if (r->host_rows) {
		TM_HIP(hipMemcpyAsync(r->host_rows, r->d_gbm, (size_t)nq * r->kw * 4, hipMemcpyDeviceToHost, r->s));
		return TM_OK;
	}
	if (!r->host_row_of) return TM_OK;
	// packed: flag the rows that are not the whole dictionary, number them,
	// pack them; the entries and as many rows as the last step had go to the
	// host now, any further rows in tm_wait once the count is known
	hipStream_t s = r->s;
	const uint32_t kw = r->kw, bits = r->fib->n_keys % 32;
	const uint32_t last_mask = bits ? (1u << bits) - 1u : 0xFFFFFFFFu;
	const uint32_t g = (nq + 255) / 256;   // a wave per 64 queries
	hipLaunchKernelGGL(k_row_partial, dim3(g), dim3(256), 0, s, r->d_gbm, kw, nq, last_mask,
					   r->qfull_valid ? r->d_qfull : nullptr, r->d_rflag);
	scan_u32_to_u64(r->d_rflag, r->d_rbase, nq, r->d_rscan, s);
	hipLaunchKernelGGL(k_row_pack, dim3(g), dim3(256), 0, s, r->d_gbm, kw, nq, r->d_rflag, r->d_rbase, r->d_row_of,
					   r->d_prow);
	TM_HIP(hipGetLastError());
	TM_HIP(hipMemcpyAsync(r->host_row_of, r->d_row_of, (size_t)nq * 4, hipMemcpyDeviceToHost, s));
	TM_HIP(hipMemcpyAsync(r->h_npack, r->d_rbase + nq, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
	r->pack_sent = std::min<uint64_t>(r->pack_guess, nq);
	if (r->pack_sent)
		TM_HIP(hipMemcpyAsync(r->host_prow, r->d_prow, r->pack_sent * kw * 4, hipMemcpyDeviceToHost, s));
	return TM_OK;
}

// The packing buffers for nq queries.
int ensure_pack(tm_ring * r, uint64_t nq) {
	if (!r->h_npack) TM_HIP(hipHostMalloc((void **)&r->h_npack, sizeof(uint64_t)));
	if (nq <= r->pack_q_cap) return TM_OK;
	dfree(r->d_rflag); dfree(r->d_rbase); dfree(r->d_rscan); dfree(r->d_row_of); dfree(r->d_prow);
	r->pack_q_cap = 0;
	TM_CHECK(dalloc(&r->d_rflag, nq));
	TM_CHECK(dalloc(&r->d_rbase, nq + 1));
	TM_CHECK(dalloc(&r->d_rscan, scan_tmp_words(nq)));
	TM_CHECK(dalloc(&r->d_row_of, nq));
	TM_CHECK(dalloc(&r->d_prow, std::max<uint64_t>(nq * r->kw, 1)));
	r->pack_q_cap = nq;
	return TM_OK;
}

int enqueue_step(tm_ring * r, bool allow_graph) {
	const uint32_t nq = r->run_nq;
	r->qfull_valid = false;   // set by the persistent path that writes the flags
	TM_HIP(hipEventRecord(r->e0, r->s));
	if (r->units_mode) {
		TM_CHECK(enqueue_units(r, nq));
		r->match_events = true;
		TM_CHECK(enqueue_delivery(r, r->last_nq));
		TM_HIP(hipMemcpyAsync(r->h_stat, r->d_stat, 8 * sizeof(uint64_t), hipMemcpyDeviceToHost, r->s));
		TM_HIP(hipEventRecord(r->e1, r->s));
		return TM_OK;
	}
	const bool persistent = r->cfg.schedule == TM_SCHED_PERSISTENT;
	// Key-row steps of a graph ring read the inserts they see from memory
	// (d_step), so a live insert needs no new capture; the hit-list modes'
	// finalize reads the overflow words' count from its arguments, so their
	// graphs are captured again after an insert.
	const bool vals = persistent && r->cfg.use_graph && keyrows(r->cfg);
	if (vals) {
		r->h_step[0] = r->epoch;
		r->h_step[1] = r->fib->ovf_valid;
		TM_HIP(hipMemcpyAsync(r->d_step, r->h_step, 2 * sizeof(uint32_t), hipMemcpyHostToDevice, r->s));
	}
	if (persistent && allow_graph && r->cfg.use_graph) {
		tm_ring::graph_entry * ge = nullptr;
		for (auto & g : r->graphs)
			if (g.nq == nq) ge = &g;
		if (ge && ((!vals && ge->epoch != r->epoch) || ge->slot_epoch != r->fib->slot_epoch)) {
			(void)hipGraphExecDestroy(ge->exec);
			r->graphs.erase(r->graphs.begin() + (ge - r->graphs.data()));
			ge = nullptr;
		}
		if (!ge) {
			if (r->graphs.size() >= MAX_GRAPHS) {   // the oldest size goes
				(void)hipGraphExecDestroy(r->graphs.front().exec);
				r->graphs.erase(r->graphs.begin());
			}
			tm_ring::graph_entry e;
			hipGraph_t g = nullptr;
			TM_HIP(hipStreamBeginCapture(r->s, hipStreamCaptureModeThreadLocal));
			r->capturing = true;
			int rc = enqueue_persistent(r, nq, false);
			r->capturing = false;
			hipError_t ce = hipStreamEndCapture(r->s, &g);
			if (rc != TM_OK) {
				if (g) (void)hipGraphDestroy(g);
				return rc;
			}
			TM_HIP(ce);
			hipError_t ie = hipGraphInstantiate(&e.exec, g, nullptr, nullptr, 0);
			(void)hipGraphDestroy(g);
			TM_HIP(ie);
			e.nq = nq;
			e.epoch = r->epoch;
			e.slot_epoch = r->fib->slot_epoch;
			e.weighted = r->step_weighted;
			e.qfull_valid = r->qfull_valid;
			r->graphs.push_back(e);
			ge = &r->graphs.back();
		}
		if (ge->weighted) ensure_weights(r, r->s);
		r->qfull_valid = ge->qfull_valid;
		TM_HIP(hipGraphLaunch(ge->exec, r->s));
		r->match_events = false;
	} else if (persistent) {
		TM_CHECK(enqueue_persistent(r, nq, true));
		r->match_events = true;
	} else {
		TM_CHECK(enqueue_plain(r, nq));
		r->match_events = true;
	}
	TM_CHECK(enqueue_delivery(r, r->last_nq));
	TM_HIP(hipMemcpyAsync(r->h_stat, r->d_stat, 8 * sizeof(uint64_t), hipMemcpyDeviceToHost, r->s));
	TM_HIP(hipEventRecord(r->e1, r->s));
	return TM_OK;
}

void drop_graph(tm_ring * r) {
	for (auto & g : r->graphs) (void)hipGraphExecDestroy(g.exec);
	r->graphs.clear();
}

// Price the submission's (query tile, partition) units with the cost model:
// order tiles longest-first (dealt round-robin over the NQUEUE queues so
// every queue starts with heavy tiles) and grow result buffers to the
// predicted volume.  Predictions never remove work: an under-estimate is
// caught by the overflow retry in tm_wait.
int price_submission(tm_ring * r, uint32_t nq) {
	const bool want = r->model && r->cfg.schedule == TM_SCHED_PERSISTENT && nq > 0 && r->P > 0;
	const bool had_order = r->use_iorder;
	r->use_iorder = false;
	if (!want) {
		if (had_order) drop_graph(r);
		return TM_OK;
	}
	const auto t0 = std::chrono::steady_clock::now();
	hipStream_t s = r->s;
	const uint32_t nqt = 64 * r->tiles;
	const uint32_t n_all = (nq + nqt - 1) / nqt;
	// ordering pays when tiles outnumber the resident workgroups and the
	// step is not cut into block segments (small submissions and tables
	// above SEG_MAX_BLOCKS are, in enqueue_persistent)
	const uint64_t resident = (uint64_t)dev().num_cu * r->wgs_per_cu;
	const uint32_t nblk = r->fib->n_blocks;
	const bool one_segment = nblk <= SEG_MAX_BLOCKS || r->cfg.chunk_blocks >= nblk;
	const bool order = (r->model_flags & TM_COST_ORDER) && n_all >= 2 * resident && one_segment;
	// Without an order to compute, only totals are needed: price a prefix of
	// PRICE_SAMPLE_TILES tiles (queries arrive in no particular order) and
	// scale.  Pricing every tile of a large step costs more than the step.
	const uint32_t n_tiles = order ? n_all : std::min<uint32_t>(n_all, PRICE_SAMPLE_TILES);
	const uint32_t nq_priced = std::min<uint32_t>(nq, n_tiles * nqt);
	const double scale = (double)nq / std::max<uint32_t>(nq_priced, 1);
	const uint64_t rows = (uint64_t)n_tiles * r->P;
	if (rows > 0xFFFFFFFFull) return set_error(TM_ERR_INVALID, "too many cost model rows");
	if (rows > r->trows_cap) {
		dfree(r->d_tfeat);
		dfree(r->d_tpred);
		r->trows_cap = 0;
		TM_CHECK(dalloc(&r->d_tfeat, rows * N_FEAT));
		TM_CHECK(dalloc(&r->d_tpred, rows * N_TARGET));
		r->trows_cap = rows;
	}
	if (n_tiles > r->tiles_cap) {
		const bool graph_uses = r->d_iorder != nullptr;
		dfree(r->d_tcost);
		dfree(r->d_thits);
		dfree(r->d_tkeys);
		dfree(r->d_iorder);
		r->tiles_cap = 0;
		TM_CHECK(dalloc(&r->d_tcost, n_tiles));
		TM_CHECK(dalloc(&r->d_thits, n_tiles));
		TM_CHECK(dalloc(&r->d_tkeys, n_tiles));
		TM_CHECK(dalloc(&r->d_iorder, n_tiles));
		r->tiles_cap = n_tiles;
		if (graph_uses) drop_graph(r);
	}
	launch_unit_features_tiles(r->fib->view(), r->d_q, nq_priced, nqt, n_tiles, r->d_tfeat, s);
	TM_CHECK(mlp_infer(r->model, r->d_tfeat, (uint32_t)rows, r->d_tpred, s));
	launch_tile_costs(r->d_tfeat, r->d_tpred, n_tiles, r->P, r->d_tcost, r->d_thits, r->d_tkeys, s);
	TM_HIP(hipGetLastError());
	std::vector<float> cost(n_tiles), hits(n_tiles), keys(n_tiles);
	TM_HIP(hipMemcpyAsync(cost.data(), r->d_tcost, 4ull * n_tiles, hipMemcpyDeviceToHost, s));
	TM_HIP(hipMemcpyAsync(hits.data(), r->d_thits, 4ull * n_tiles, hipMemcpyDeviceToHost, s));
	TM_HIP(hipMemcpyAsync(keys.data(), r->d_tkeys, 4ull * n_tiles, hipMemcpyDeviceToHost, s));
	TM_HIP(hipStreamSynchronize(s));
	double total = 0.0, total_keys = 0.0;
	for (uint32_t t = 0; t < n_tiles; ++t) {
		total += hits[t];
		total_keys += keys[t];
	}
	total *= scale;
	total_keys *= scale;
	r->st.pred_hits = total;

	if (order) {
		std::vector<uint32_t> idx(n_tiles);
		std::iota(idx.begin(), idx.end(), 0u);
		std::stable_sort(idx.begin(), idx.end(), [&](uint32_t x, uint32_t y) { return cost[x] > cost[y]; });
		// queue x serves items [n*x/NQUEUE, n*(x+1)/NQUEUE) (k_init_queues)
		std::vector<uint32_t> order(n_tiles);
		std::vector<uint64_t> fill(NQUEUE), end(NQUEUE);
		for (int x = 0; x < NQUEUE; ++x) {
			fill[x] = (uint64_t)n_tiles * x / NQUEUE;
			end[x] = (uint64_t)n_tiles * (x + 1) / NQUEUE;
		}
		int x = 0;
		for (uint32_t k = 0; k < n_tiles; ++k) {
			while (fill[x] >= end[x]) x = (x + 1) % NQUEUE;
			order[fill[x]++] = idx[k];
			x = (x + 1) % NQUEUE;
		}
		TM_HIP(hipMemcpyAsync(r->d_iorder, order.data(), 4ull * n_tiles, hipMemcpyHostToDevice, s));
		TM_HIP(hipStreamSynchronize(s));   // `order` is pageable and local
		r->use_iorder = true;
	}
	if (r->use_iorder != had_order) drop_graph(r);

	if (r->model_flags & TM_COST_PRESIZE) {
		const tm_fib * f = r->fib;
		const double margin = 1.25;   // predictions are fitted in log space: allow for spread
		uint64_t slots = 0, results = 0;
		if (r->cfg.mode == TM_MODE_KEY_BITMAP) {
			results = 0;   // rows, sized for every query already
		} else if (r->cfg.mode == TM_MODE_KEYS_UNIQUE) {
			// expected distinct keys among k key instances drawn from K keys,
			// from the tile's mean instances per query (concave in k, so the
			// tile mean over-estimates: a safe direction)
			const double K = std::max<double>(f->n_keys, 1.0);
			double est = 0.0;
			for (uint32_t t = 0; t < n_tiles; ++t) {
				const uint32_t qn = std::min<uint32_t>(nqt, nq_priced - t * nqt);
				const double k = keys[t] / std::max<uint32_t>(qn, 1);
				est += qn * K * (1.0 - std::exp(-k / K));
			}
			results = (uint64_t)(est * scale * margin) + 4096;
		} else {
			slots = (uint64_t)(total * margin) + 4096;
			results = r->cfg.mode == TM_MODE_KEYS_ALL ? (uint64_t)(total_keys * margin) + 4096 : slots;
			if (slots > r->hit_cap) {
				dfree(r->d_hits);
				r->hit_cap = 0;
				TM_CHECK(dalloc(&r->d_hits, slots));
				r->hit_cap = slots;
				drop_graph(r);
			}
		}
		if (results > r->val_cap) {
			TM_CHECK(grow_u32(&r->d_vals, r->val_cap, results));
			drop_graph(r);
		}
	}
	r->st.model_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
	return TM_OK;
}

// After tm_fib_insert changed the partition count, the key count or the
// FIB's device arrays: re-size what depends on them and drop any graph
// captured with the old pointers.
int sync_fib(tm_ring * r) {
	const tm_fib * f = r->fib;
	if (r->fib_version == f->version) return TM_OK;
	drop_graph(r);
	if (r->P != f->n_parts) {
		dfree(r->d_pcount); dfree(r->d_pbase); dfree(r->d_pcur); dfree(r->d_sched); dfree(r->d_iprefix);
		dfree(r->d_qbeg); dfree(r->d_qnext); dfree(r->d_porder); dfree(r->d_cnt); dfree(r->d_stat);
		dfree(r->d_qcount); dfree(r->d_off); dfree(r->d_qcur); dfree(r->d_scan);
		r->have_order = false;
		r->P = f->n_parts;
	}
	const uint32_t kw = (f->n_keys + 31) / 32;
	if (kw != r->kw) {
		dfree(r->d_gbm);
		r->kw = kw;
	}
	TM_CHECK(resolve_params(r));
	r->fib_version = f->version;
	return ensure_capacity(r);
}

void end_pending(tm_ring * r) {
	if (!r->pending) return;
	r->pending = false;
	r->fib->in_flight.fetch_sub(1);
}

int submit_common(tm_ring * r, uint32_t nq) {
	r->st = tm_status();
	r->last_nq = nq;
	r->retries = 0;
	r->epoch = r->fib->epoch;   // the inserts this submission sees
	TM_CHECK(sync_fib(r));
	TM_CHECK(ensure_capacity(r));
	if (r->cfg.mode == TM_MODE_KEY_BITMAP && r->host_rows && (uint64_t)nq * r->kw > r->host_rows_cap)
		return set_error(TM_ERR_OVERFLOW, "host rows hold %llu words, the submission needs %llu",
						 (unsigned long long)r->host_rows_cap, (unsigned long long)nq * r->kw);
	if (r->cfg.mode == TM_MODE_KEY_BITMAP && r->host_row_of) {
		if (nq > r->host_row_of_cap || (uint64_t)nq * r->kw > r->host_prow_cap)
			return set_error(TM_ERR_OVERFLOW, "packed host rows hold %llu entries and %llu words, the submission "
							 "needs %u and up to %llu", (unsigned long long)r->host_row_of_cap,
							 (unsigned long long)r->host_prow_cap, nq, (unsigned long long)nq * r->kw);
		TM_CHECK(ensure_pack(r, std::max<uint64_t>(nq, 1)));
	}
	if (!r->units_mode) TM_CHECK(price_submission(r, r->run_nq));
	int rc = enqueue_step(r, true);
	if (rc != TM_OK) return rc;
	r->pending = true;
	r->fib->in_flight.fetch_add(1);
	return TM_OK;
}

} // namespace

extern "C" {

void tm_match_config_default(tm_match_config * c) {
	std::memset(c, 0, sizeof(*c));
	c->mode = TM_MODE_KEYS_UNIQUE;
	c->schedule = TM_SCHED_PERSISTENT;
}

int tm_ring_create(tm_fib * fib, uint32_t max_queries, const tm_match_config * cfg, tm_ring ** out) {
	if (!fib || !out) return set_error(TM_ERR_INVALID, "NULL argument");
	if (max_queries == 0) return set_error(TM_ERR_INVALID, "max_queries must be > 0");
	TM_CHECK(require_device());
	tm_ring * r = new tm_ring();
	r->fib = fib;
	if (cfg) r->cfg = *cfg; else tm_match_config_default(&r->cfg);
	r->max_q = max_queries;
	r->P = fib->n_parts;
	r->kw = (fib->n_keys + 31) / 32;
	r->fib_version = fib->version;
	int rc = resolve_params(r);
	if (rc != TM_OK) { delete r; return rc; }
	hipError_t e = hipStreamCreateWithFlags(&r->s, hipStreamNonBlocking);
	if (e == hipSuccess) e = hipEventCreate(&r->e0);
	if (e == hipSuccess) e = hipEventCreate(&r->em0);
	if (e == hipSuccess) e = hipEventCreate(&r->em1);
	if (e == hipSuccess) e = hipEventCreate(&r->e1);
	if (e == hipSuccess) e = hipHostMalloc((void **)&r->h_q, (size_t)max_queries * 32);
	if (e == hipSuccess) e = hipHostMalloc((void **)&r->h_stat, 8 * sizeof(uint64_t));
	if (e == hipSuccess) e = hipHostMalloc((void **)&r->h_step, 2 * sizeof(uint32_t));
	if (e != hipSuccess) {
		int code = hip_error(e, "ring setup", __FILE__, __LINE__);
		r->release();
		delete r;
		return code;
	}
	rc = dalloc(&r->d_q, (size_t)max_queries * 8);
	if (rc == TM_OK) rc = dalloc(&r->d_step, 2);
	if (rc == TM_OK) rc = dalloc(&r->d_qinv, (size_t)max_queries * 8);
	if (rc == TM_OK) rc = ensure_capacity(r);
	if (rc != TM_OK) { r->release(); delete r; return rc; }
	*out = r;
	return TM_OK;
}

int tm_ring_destroy(tm_ring * r) {
	if (!r) return TM_OK;
	if (dev().device >= 0) (void)hipSetDevice(dev().device);
	if (r->s) (void)hipStreamSynchronize(r->s);
	end_pending(r);
	r->release();
	delete r;
	return TM_OK;
}

int tm_ring_config(tm_ring * r, const tm_match_config * cfg) {
	if (!r || !cfg) return set_error(TM_ERR_INVALID, "NULL argument");
	if (r->pending) return set_error(TM_ERR_BUSY, "ring has a submission in flight");
	TM_CHECK(require_device());
	tm_match_config old = r->cfg;
	r->cfg = *cfg;
	int rc = resolve_params(r);
	if (rc != TM_OK) { r->cfg = old; (void)resolve_params(r); return rc; }
	for (auto & g : r->graphs) (void)hipGraphExecDestroy(g.exec);
	r->graphs.clear();
	return ensure_capacity(r);
}

int tm_ring_query_buffer(tm_ring * r, uint32_t ** dev_words) {
	if (!r || !dev_words) return set_error(TM_ERR_INVALID, "NULL argument");
	*dev_words = r->d_q;
	return TM_OK;
}

int tm_submit(tm_ring * r, const uint32_t * words, uint32_t nq) {
	if (!r) return set_error(TM_ERR_INVALID, "ring is NULL");
	if (nq > r->max_q) return set_error(TM_ERR_INVALID, "nq %u exceeds ring capacity %u", nq, r->max_q);
	if (nq && !words) return set_error(TM_ERR_INVALID, "query words are NULL");
	if (r->pending) return set_error(TM_ERR_BUSY, "ring has a submission in flight");
	TM_CHECK(require_device());
	uint64_t bits = 0;
	for (uint32_t i = 0; i < nq; ++i) {
		std::memcpy(r->h_q + (size_t)i * 8, words + (size_t)i * 6, 24);
		r->h_q[(size_t)i * 8 + 6] = 0;
		r->h_q[(size_t)i * 8 + 7] = 0;
		if (i < 4096)
			for (int j = 0; j < 6; ++j) bits += (uint64_t)__builtin_popcount(words[(size_t)i * 6 + j]);
	}
	if (nq) {   // mean bit density of the queries, for weighted segments
		const double rho = (double)bits / (192.0 * std::min<uint32_t>(nq, 4096));
		r->rho_l2 = (float)std::max(-16.0, std::log2(std::max(rho, 1e-6)));
	}
	// Graph replays run fixed sizes: the submission is padded with zero
	// queries up to the next power of two (at least 64).  A zero query
	// selects only partitions with an empty mask and contains no rich
	// filter, so the padding costs next to nothing; its rows and keys lie
	// past the submission's results and are never read.
	uint32_t nrun = nq;
	if (r->cfg.use_graph && r->cfg.schedule == TM_SCHED_PERSISTENT && nq) {
		nrun = 64;
		while (nrun < nq) nrun <<= 1;
		nrun = std::min(nrun, r->max_q);
		std::memset(r->h_q + (size_t)nq * 8, 0, (size_t)(nrun - nq) * 32);
	}
	if (nrun) TM_HIP(hipMemcpyAsync(r->d_q, r->h_q, (size_t)nrun * 32, hipMemcpyHostToDevice, r->s));
	r->units_mode = false;
	r->run_nq = nrun;
	return submit_common(r, nq);
}

int tm_submit_units(tm_ring * r, const uint32_t * words, uint32_t nq, const uint32_t * unit_part,
					const uint32_t * unit_off, uint32_t n_units) {
	if (!r) return set_error(TM_ERR_INVALID, "ring is NULL");
	if (nq > r->max_q) return set_error(TM_ERR_INVALID, "nq %u exceeds ring capacity %u", nq, r->max_q);
	if (n_units && (!unit_part || !unit_off)) return set_error(TM_ERR_INVALID, "unit arrays are NULL");
	if (r->pending) return set_error(TM_ERR_BUSY, "ring has a submission in flight");
	if (n_units == 0 && nq) return set_error(TM_ERR_INVALID, "queries without units");
	if (n_units) {
		if (unit_off[0] != 0 || unit_off[n_units] != nq)
			return set_error(TM_ERR_INVALID, "unit offsets must run from 0 to nq (%u)", nq);
		for (uint32_t i = 0; i < n_units; ++i) {
			if (unit_off[i + 1] < unit_off[i]) return set_error(TM_ERR_INVALID, "unit offsets not monotone at %u", i);
			if (unit_part[i] >= r->fib->n_parts)
				return set_error(TM_ERR_INVALID, "unit %u: partition %u of %u", i, unit_part[i], r->fib->n_parts);
		}
	}
	TM_CHECK(require_device());
	if (words && nq) {
		for (uint32_t i = 0; i < nq; ++i) {
			std::memcpy(r->h_q + (size_t)i * 8, words + (size_t)i * 6, 24);
			r->h_q[(size_t)i * 8 + 6] = 0;
			r->h_q[(size_t)i * 8 + 7] = 0;
		}
		TM_HIP(hipMemcpyAsync(r->d_q, r->h_q, (size_t)nq * 32, hipMemcpyHostToDevice, r->s));
	}
	r->u_part.assign(unit_part, unit_part + n_units);
	r->u_off.assign(unit_off, unit_off + n_units + 1);
	if (n_units == 0) r->u_off.assign(1, 0);
	r->units_mode = true;
	r->have_order = false;
	r->run_nq = nq;
	int rc = submit_common(r, nq);
	r->units_mode = rc == TM_OK;   // kept for the overflow retry in tm_wait
	return rc;
}

int tm_submit_resident(tm_ring * r, uint32_t nq) {
	if (!r) return set_error(TM_ERR_INVALID, "ring is NULL");
	if (nq > r->max_q) return set_error(TM_ERR_INVALID, "nq %u exceeds ring capacity %u", nq, r->max_q);
	if (r->pending) return set_error(TM_ERR_BUSY, "ring has a submission in flight");
	TM_CHECK(require_device());
	r->units_mode = false;
	r->run_nq = nq;   // resident queries past nq are the caller's: no padding
	return submit_common(r, nq);
}

int tm_poll(tm_ring * r) {
	if (!r) return set_error(TM_ERR_INVALID, "ring is NULL");
	if (!r->pending) return TM_OK;
	hipError_t e = hipEventQuery(r->e1);
	if (e == hipErrorNotReady) return TM_ERR_BUSY;
	TM_HIP(e);
	return TM_OK;
}

int tm_wait(tm_ring * r, tm_status * st) {
	if (!r) return set_error(TM_ERR_INVALID, "ring is NULL");
	TM_CHECK(require_device());
	if (r->pending) {
		for (;;) {
			TM_HIP(hipEventSynchronize(r->e1));
			const uint64_t pairs = r->h_stat[0], slots = r->h_stat[1], results = r->h_stat[3];
			bool again = false;
			if (r->cfg.schedule == TM_SCHED_PLAIN && pairs > r->list_cap) {
				TM_CHECK(grow_u32(&r->d_list, r->list_cap, pairs + pairs / 4 + 1024));
				again = true;
			}
			if (!keyrows(r->cfg) && slots > r->hit_cap) {
				dfree(r->d_hits);
				uint64_t cap = slots + slots / 4 + 1024;
				TM_CHECK(dalloc(&r->d_hits, cap));
				r->hit_cap = cap;
				again = true;
			}
			if (results > r->val_cap) {
				TM_CHECK(grow_u32(&r->d_vals, r->val_cap, results + results / 4 + 1024));
				again = true;
			}
			if (!again) break;
			if (++r->retries > 4) {
				end_pending(r);
				return set_error(TM_ERR_OVERFLOW, "capacity retry limit reached");
			}
			// overflow-retry: capacities changed, so drop any captured graph
			for (auto & g : r->graphs) (void)hipGraphExecDestroy(g.exec);
			r->graphs.clear();
			TM_CHECK(enqueue_step(r, false));
		}
		if (r->cfg.mode == TM_MODE_KEY_BITMAP && r->host_row_of && r->last_nq && r->kw) {
			// packed rows beyond those copied with the step
			const uint64_t n = *r->h_npack;
			if (n > r->pack_sent) {
				const hipError_t e = ring_copy(r, r->host_prow + r->pack_sent * r->kw, r->d_prow + r->pack_sent * r->kw,
											   (n - r->pack_sent) * r->kw * 4, hipMemcpyDeviceToHost);
				if (e != hipSuccess) {
					end_pending(r);
					return hip_error(e, "packed rows", __FILE__, __LINE__);
				}
			}
			// the next step copies this many rows with itself (with room to grow)
			r->pack_guess = n + n / 8 + 64;
		}
		end_pending(r);
		float dms = 0.f, mms = 0.f;
		(void)hipEventElapsedTime(&dms, r->e0, r->e1);
		if (r->match_events) (void)hipEventElapsedTime(&mms, r->em0, r->em1);
		r->st.n_queries = r->last_nq;
		r->st.n_pairs = r->h_stat[0];
		r->st.n_hits = r->h_stat[2];
		r->st.n_results = r->h_stat[3];
		if (r->cfg.schedule == TM_SCHED_PERSISTENT) r->st.n_items = (uint32_t)r->h_stat[4];
		r->st.hit_capacity = r->hit_cap;
		r->st.overflow = r->retries > 0;
		r->st.error = TM_OK;
		r->st.device_ms = dms;
		r->st.match_ms = mms;
		r->st.n_block_visits = r->h_stat[5];
		r->st.n_cand_tests = r->h_stat[6];
		r->st.n_hit_events = r->h_stat[7];
		if (r->cfg.mode == TM_MODE_KEY_BITMAP && r->host_row_of) r->st.n_results = r->last_nq ? *r->h_npack : 0;
	}
	if (st) *st = r->st;
	return TM_OK;
}

int tm_read_results(tm_ring * r, uint64_t * offsets, uint32_t * values, uint64_t capacity) {
	if (!r || !offsets) return set_error(TM_ERR_INVALID, "NULL argument");
	if (r->pending) return set_error(TM_ERR_BUSY, "call tm_wait() first");
	if (r->cfg.mode == TM_MODE_KEY_BITMAP)
		return set_error(TM_ERR_STATE, "KEY_BITMAP results are rows: tm_read_key_bitmaps");
	TM_CHECK(require_device());
	const uint32_t nq = r->last_nq;
	if (nq == 0) {
		offsets[0] = 0;
		return TM_OK;
	}
	TM_HIP(ring_copy(r, offsets, r->d_off, ((size_t)nq + 1) * sizeof(uint64_t), hipMemcpyDeviceToHost));
	uint64_t total = offsets[nq];
	if (!values) return TM_OK;
	if (total > capacity)
		return set_error(TM_ERR_OVERFLOW, "%llu result values, capacity %llu", (unsigned long long)total,
						 (unsigned long long)capacity);
	if (total) TM_HIP(ring_copy(r, values, r->d_vals, total * sizeof(uint32_t), hipMemcpyDeviceToHost));
	return TM_OK;
}

int tm_read_key_bitmaps(tm_ring * r, uint32_t * rows, uint64_t capacity_words, uint32_t * kw) {
	if (!r) return set_error(TM_ERR_INVALID, "ring is NULL");
	if (r->pending) return set_error(TM_ERR_BUSY, "call tm_wait() first");
	if (!keyrows(r->cfg)) return set_error(TM_ERR_STATE, "the ring's mode has no key bitmap rows");
	if (kw) *kw = r->kw;
	const uint64_t n = (uint64_t)r->last_nq * r->kw;
	if (n > capacity_words)
		return set_error(TM_ERR_OVERFLOW, "%llu row words, capacity %llu", (unsigned long long)n,
						 (unsigned long long)capacity_words);
	if (n == 0) return TM_OK;
	if (!rows) return set_error(TM_ERR_INVALID, "rows is NULL");
	TM_CHECK(require_device());
	TM_HIP(ring_copy(r, rows, r->d_gbm, n * 4, hipMemcpyDeviceToHost));
	return TM_OK;
}

int tm_ring_set_host_rows(tm_ring * r, uint32_t * host_rows, uint64_t capacity_words) {
	if (!r) return set_error(TM_ERR_INVALID, "ring is NULL");
	if (r->pending) return set_error(TM_ERR_BUSY, "ring has a submission in flight");
	if (host_rows && r->cfg.mode != TM_MODE_KEY_BITMAP)
		return set_error(TM_ERR_STATE, "host rows are delivered in KEY_BITMAP mode only");
	r->host_rows = host_rows;
	r->host_rows_cap = host_rows ? capacity_words : 0;
	if (host_rows) {   // one delivery form at a time
		r->host_row_of = r->host_prow = nullptr;
		r->host_row_of_cap = r->host_prow_cap = 0;
	}
	return TM_OK;
}

int tm_ring_set_host_rows_packed(tm_ring * r, uint32_t * row_of, uint64_t row_of_capacity, uint32_t * rows,
								 uint64_t rows_capacity_words) {
	if (!r) return set_error(TM_ERR_INVALID, "ring is NULL");
	if (r->pending) return set_error(TM_ERR_BUSY, "ring has a submission in flight");
	if (row_of && r->cfg.mode != TM_MODE_KEY_BITMAP)
		return set_error(TM_ERR_STATE, "host rows are delivered in KEY_BITMAP mode only");
	if (row_of && !rows) return set_error(TM_ERR_INVALID, "rows is NULL");
	r->host_row_of = row_of;
	r->host_prow = row_of ? rows : nullptr;
	r->host_row_of_cap = row_of ? row_of_capacity : 0;
	r->host_prow_cap = row_of ? rows_capacity_words : 0;
	r->pack_guess = 0;
	if (row_of) {   // one delivery form at a time
		r->host_rows = nullptr;
		r->host_rows_cap = 0;
	}
	return TM_OK;
}

int tm_ring_results_device(tm_ring * r, const uint64_t ** off, const uint32_t ** vals) {
	if (!r || !off || !vals) return set_error(TM_ERR_INVALID, "NULL argument");
	if (r->pending) return set_error(TM_ERR_BUSY, "call tm_wait() first");
	*off = r->d_off;
	*vals = r->d_vals;
	return TM_OK;
}

int tm_ring_autotune(tm_ring * r, uint32_t nq, const tm_match_config * cands, uint32_t n_cands, uint32_t reps,
					 uint32_t * best, double * times_ms) {
	if (!r || !cands || n_cands == 0 || !best) return set_error(TM_ERR_INVALID, "NULL argument");
	if (nq == 0 || nq > r->max_q) return set_error(TM_ERR_INVALID, "nq %u not in 1..%u", nq, r->max_q);
	if (r->pending) return set_error(TM_ERR_BUSY, "ring has a submission in flight");
	if (reps == 0) reps = 1;
	TM_CHECK(require_device());
	for (uint32_t c = 0; c < n_cands; ++c)
		if (cands[c].mode != r->cfg.mode) return set_error(TM_ERR_INVALID, "candidate %u changes the mode", c);
	const tm_match_config saved = r->cfg;
	auto fail = [&](int rc) {   // a failed run leaves the ring as it was configured
		const std::string err = tm_last_error();
		(void)tm_ring_config(r, &saved);
		return set_error(rc, "%s", err.c_str());
	};
	double best_ms = 1e300;
	uint32_t bi = 0;
	for (uint32_t c = 0; c < n_cands; ++c) {
		int rc = tm_ring_config(r, &cands[c]);
		if (rc != TM_OK) return fail(rc);
		double t = 1e300;
		for (uint32_t k = 0; k <= reps; ++k) {   // the first run warms up (capacities, graph)
			rc = tm_submit_resident(r, nq);
			tm_status st;
			if (rc == TM_OK) rc = tm_wait(r, &st);
			if (rc != TM_OK) return fail(rc);
			if (k > 0) t = std::min(t, st.device_ms);
		}
		if (times_ms) times_ms[c] = t;
		if (t < best_ms) {
			best_ms = t;
			bi = c;
		}
	}
	*best = bi;
	return tm_ring_config(r, &cands[bi]);
}

int tm_ring_set_bucket_segments(tm_ring * r, const uint32_t * seg, uint32_t n) {
	if (!r || (n && !seg)) return set_error(TM_ERR_INVALID, "NULL argument");
	if (n > TM_SIZE_BUCKETS) return set_error(TM_ERR_INVALID, "%u buckets, at most %u", n, (unsigned)TM_SIZE_BUCKETS);
	if (r->pending) return set_error(TM_ERR_BUSY, "ring has a submission in flight");
	for (uint32_t i = 0; i < TM_SIZE_BUCKETS; ++i) r->seg_bucket[i] = i < n ? seg[i] : 0u;
	return TM_OK;
}

int tm_ring_get_bucket_segments(const tm_ring * r, uint32_t * seg, uint32_t n) {
	if (!r || (n && !seg)) return set_error(TM_ERR_INVALID, "NULL argument");
	for (uint32_t i = 0; i < n && i < TM_SIZE_BUCKETS; ++i) seg[i] = r->seg_bucket[i];
	return TM_OK;
}

int tm_ring_autotune_buckets(tm_ring * r, const uint32_t * unit_part, const uint32_t * unit_off, uint32_t n_units,
							 const uint32_t * cands, uint32_t n_cands, uint32_t reps, double * base_ms,
							 double * best_ms) {
	if (!r || !unit_part || !unit_off || n_units == 0 || !cands || n_cands == 0)
		return set_error(TM_ERR_INVALID, "NULL argument");
	if (r->pending) return set_error(TM_ERR_BUSY, "ring has a submission in flight");
	if (reps == 0) reps = 1;
	const uint32_t nq = unit_off[n_units];
	auto run = [&](double * out) -> int {
		double t = 1e300;
		for (uint32_t k = 0; k <= reps; ++k) {   // the first run warms up
			TM_CHECK(tm_submit_units(r, nullptr, nq, unit_part, unit_off, n_units));
			tm_status st;
			TM_CHECK(tm_wait(r, &st));
			if (k > 0) t = std::min(t, st.device_ms);
		}
		*out = t;
		return TM_OK;
	};
	// buckets the units reach
	bool present[TM_SIZE_BUCKETS] = {};
	for (uint32_t i = 0; i < n_units; ++i) {
		if (unit_part[i] >= r->fib->n_parts) return set_error(TM_ERR_INVALID, "unit %u: partition out of range", i);
		present[size_bucket(r->fib->h_part_info[(size_t)unit_part[i] * 8 + 7])] = true;
	}
	uint32_t saved[TM_SIZE_BUCKETS];
	std::copy(r->seg_bucket, r->seg_bucket + TM_SIZE_BUCKETS, saved);
	auto fail = [&](int rc) {   // a failed run leaves the table as it was
		std::copy(saved, saved + TM_SIZE_BUCKETS, r->seg_bucket);
		return rc;
	};
	double cur;
	int rc = run(&cur);
	if (rc != TM_OK) return fail(rc);
	if (base_ms) *base_ms = cur;
	for (uint32_t b = 0; b < TM_SIZE_BUCKETS; ++b) {
		if (!present[b]) continue;
		const uint32_t keep = r->seg_bucket[b];
		uint32_t best = keep;
		for (uint32_t c = 0; c < n_cands; ++c) {
			if (cands[c] == keep) continue;
			r->seg_bucket[b] = cands[c];
			double t;
			rc = run(&t);
			if (rc != TM_OK) return fail(rc);
			if (t < cur) {
				cur = t;
				best = cands[c];
			}
		}
		r->seg_bucket[b] = best;
	}
	if (best_ms) *best_ms = cur;
	return TM_OK;
}

int tm_ring_unit_log(tm_ring * r, float * features, uint32_t * hits, uint32_t * keys, uint32_t * ticks,
					 uint32_t * units, uint32_t capacity, uint32_t * n_units) {
	if (!r) return set_error(TM_ERR_INVALID, "ring is NULL");
	if (r->pending) return set_error(TM_ERR_BUSY, "call tm_wait() first");
	if (n_units) *n_units = r->n_units;
	const uint32_t n = std::min(capacity, r->n_units);
	if (n == 0) return TM_OK;
	TM_CHECK(require_device());
	if (features)
		TM_HIP(ring_copy(r, features, r->d_ufeat, (size_t)n * N_FEAT * sizeof(float), hipMemcpyDeviceToHost));
	if (units) {
		std::vector<uint4> u(n);
		TM_HIP(ring_copy(r, u.data(), r->d_units, (size_t)n * sizeof(uint4), hipMemcpyDeviceToHost));
		for (uint32_t i = 0; i < n; ++i) {
			units[(size_t)i * 3] = u[i].x;
			units[(size_t)i * 3 + 1] = u[i].y;
			units[(size_t)i * 3 + 2] = u[i].z;
		}
	}
	if (hits || keys || ticks) {
		std::vector<uint32_t> log((size_t)n * 3);
		TM_HIP(ring_copy(r, log.data(), r->d_log, (size_t)n * 3 * sizeof(uint32_t), hipMemcpyDeviceToHost));
		for (uint32_t i = 0; i < n; ++i) {
			if (hits) hits[i] = log[(size_t)i * 3];
			if (keys) keys[i] = log[(size_t)i * 3 + 1];
			if (ticks) ticks[i] = log[(size_t)i * 3 + 2];
		}
	}
	return TM_OK;
}

int tm_ring_set_cost_model(tm_ring * r, tm_mlp * m, uint32_t flags) {
	if (!r) return set_error(TM_ERR_INVALID, "ring is NULL");
	if (r->pending) return set_error(TM_ERR_BUSY, "ring has a submission in flight");
	if (m && !mlp_fused_shape(m))
		return set_error(TM_ERR_INVALID, "cost model must be %d-64-64-%d", N_FEAT, N_TARGET);
	if (flags & ~(TM_COST_PRESIZE | TM_COST_ORDER)) return set_error(TM_ERR_INVALID, "unknown cost model flags");
	r->model = m;
	r->model_flags = m ? flags : 0;
	return TM_OK;
}

int tm_ring_lookup_counts(tm_ring * r, uint32_t * counts, uint32_t n_parts) {
	if (!r || !counts) return set_error(TM_ERR_INVALID, "NULL argument");
	if (n_parts != r->P) return set_error(TM_ERR_INVALID, "n_parts %u != %u", n_parts, r->P);
	if (r->pending) return set_error(TM_ERR_BUSY, "call tm_wait() first");
	TM_CHECK(require_device());
	if (n_parts) TM_HIP(ring_copy(r, counts, r->d_pcount, (size_t)n_parts * 4, hipMemcpyDeviceToHost));
	return TM_OK;
}

int tm_match(tm_fib * fib, const uint32_t * words, uint32_t nq, const tm_match_config * cfg,
			 uint64_t * offsets, uint32_t * values, uint64_t capacity, tm_status * st) {
	tm_ring * r = nullptr;
	TM_CHECK(tm_ring_create(fib, nq ? nq : 1, cfg, &r));
	int rc = tm_submit(r, words, nq);
	if (rc == TM_OK) rc = tm_wait(r, st);
	if (rc == TM_OK) rc = tm_read_results(r, offsets, values, capacity);
	tm_ring_destroy(r);
	return rc;
}

} // extern "C"

// ------------------------------------------------------------------------
// No-allocation entry point (tm_match_rows_static).
//
// The persistent KEY_BITMAP step of a ring, for callers that cannot
// allocate device memory: every buffer the step's kernels need beyond the
// caller's queries and rows (lookup counts and their prefix sums, the work
// queues, counters, full-key-set flags and the active-query list) is carved
// out of `scratch`, a file-scope device array of fixed size.  Each GPU the
// library runs on holds its own copy of it until the process ends.  Host
// code must not use the array's name as a device address: its address on
// each device comes from hipGetSymbolAddress, taken once per device on the
// entry point's first call there.  Calls on one GPU share that GPU's array;
// the call clears what must start at zero, and calls that overlap on one GPU
// race (the call is synchronous, so calls from one thread never overlap).

namespace {

constexpr size_t SC_ALIGN = 256;
constexpr size_t sc_round(size_t b) { return (b + SC_ALIGN - 1) / SC_ALIGN * SC_ALIGN; }
constexpr size_t SC_P = (size_t)TM_STATIC_MAX_PARTS + 1;
constexpr size_t SC_Q = (size_t)TM_STATIC_MAX_QUERIES;
// byte offsets of the sub-arrays
constexpr size_t SC_PCOUNT = 0;                                          // u32 [P + 1]
constexpr size_t SC_PBASE = SC_PCOUNT + sc_round(SC_P * 4);             // u64 [P + 1]
constexpr size_t SC_IPREFIX = SC_PBASE + sc_round(SC_P * 8);            // u64 [P + 1]
constexpr size_t SC_QBEG = SC_IPREFIX + sc_round(SC_P * 8);             // u64 [NQUEUE + 1]
constexpr size_t SC_QNEXT = SC_QBEG + sc_round((NQUEUE + 1) * 8);       // u64 [NQUEUE * QPAD]
constexpr size_t SC_CNT = SC_QNEXT + sc_round((size_t)NQUEUE * QPAD * 8);   // u64 [8]
constexpr size_t SC_STAT = SC_CNT + sc_round(8 * 8);                    // u64 [8]
constexpr size_t SC_NACT = SC_STAT + sc_round(8 * 8);                   // u32 [4]
constexpr size_t SC_SCAN = SC_NACT + sc_round(4 * 4);                   // u64 [scan words of P]
constexpr size_t SC_SCAN_WORDS = (SC_P + 255) / 256 + 2;                // >= scan_tmp_words(P)
constexpr size_t SC_AHIST = SC_SCAN + sc_round(SC_SCAN_WORDS * 8);      // u32 [2 x ACT_BUCKETS]
constexpr size_t SC_ABASE = SC_AHIST + sc_round(2 * ACT_BUCKETS * 4);    // u64 [ACT_BUCKETS + 1]
constexpr size_t SC_QFULL = SC_ABASE + sc_round((ACT_BUCKETS + 1) * 8);  // u32 [Q]
constexpr size_t SC_ACTIVE = SC_QFULL + sc_round(SC_Q * 4);             // u32 [Q]
constexpr size_t SC_BYTES = SC_ACTIVE + sc_round(SC_Q * 4);

static __device__ __align__(256) unsigned char scratch[SC_BYTES];

constexpr int SC_MAX_DEVICES = 64;
std::mutex sc_mu;
unsigned char * sc_addr[SC_MAX_DEVICES] = {};

// The device address of `scratch` on the current device.
int scratch_on_device(unsigned char ** out) {
	const int d = dev().device;
	if (d < 0 || d >= SC_MAX_DEVICES) return set_error(TM_ERR_STATE, "no device");
	std::lock_guard<std::mutex> lk(sc_mu);
	if (!sc_addr[d]) {
		void * p = nullptr;
		TM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
		if (!p) return set_error(TM_ERR_HIP, "no device address for the scratch array");
		sc_addr[d] = static_cast<unsigned char *>(p);
	}
	*out = sc_addr[d];
	return TM_OK;
}

} // namespace

extern "C" {

int tm_match_rows_static(tm_fib * fib, const uint32_t * dev_queries, uint32_t nq, uint32_t * dev_rows,
						 void * stream) {
	if (!fib) return set_error(TM_ERR_INVALID, "fib is NULL");
	if (nq > TM_STATIC_MAX_QUERIES)
		return set_error(TM_ERR_INVALID, "%u queries: this entry point takes at most %u", nq,
						 (unsigned)TM_STATIC_MAX_QUERIES);
	if (nq && (!dev_queries || !dev_rows)) return set_error(TM_ERR_INVALID, "queries or rows are NULL");
	if (fib->h_order.size() != 192 || !fib->blk_words) return set_error(TM_ERR_STATE, "no table uploaded");
	if (fib->n_parts > TM_STATIC_MAX_PARTS)
		return set_error(TM_ERR_INVALID, "the FIB has %u partitions: this entry point takes at most %u",
						 fib->n_parts, (unsigned)TM_STATIC_MAX_PARTS);
	TM_CHECK(require_device());
	if (nq == 0) return TM_OK;
	unsigned char * sc = nullptr;
	TM_CHECK(scratch_on_device(&sc));
	tm_ring r;   // host-side state only; no buffer of its own
	r.fib = fib;
	tm_match_config_default(&r.cfg);
	r.cfg.mode = TM_MODE_KEY_BITMAP;
	r.max_q = nq;
	r.P = fib->n_parts;
	r.kw = (fib->n_keys + 31) / 32;
	TM_CHECK(resolve_params(&r));
	if (scan_tmp_words(r.P) > SC_SCAN_WORDS || scan_tmp_words(ACT_BUCKETS) > SC_SCAN_WORDS)
		return set_error(TM_ERR_STATE, "scan scratch too small");
	r.no_weights = true;
	r.s = static_cast<hipStream_t>(stream);
	r.d_q = const_cast<uint32_t *>(dev_queries);
	r.d_gbm = dev_rows;
	r.d_pcount = reinterpret_cast<uint32_t *>(sc + SC_PCOUNT);
	r.d_pbase = reinterpret_cast<uint64_t *>(sc + SC_PBASE);
	r.d_iprefix = reinterpret_cast<uint64_t *>(sc + SC_IPREFIX);
	r.d_qbeg = reinterpret_cast<uint64_t *>(sc + SC_QBEG);
	r.d_qnext = reinterpret_cast<unsigned long long *>(sc + SC_QNEXT);
	r.d_cnt = reinterpret_cast<unsigned long long *>(sc + SC_CNT);
	r.d_stat = reinterpret_cast<uint64_t *>(sc + SC_STAT);
	r.d_nact = reinterpret_cast<uint32_t *>(sc + SC_NACT);
	r.d_scan = reinterpret_cast<uint64_t *>(sc + SC_SCAN);
	r.d_ahist = reinterpret_cast<uint32_t *>(sc + SC_AHIST);
	r.d_abase = reinterpret_cast<uint64_t *>(sc + SC_ABASE);
	r.d_qfull = reinterpret_cast<uint32_t *>(sc + SC_QFULL);
	r.d_active = reinterpret_cast<uint32_t *>(sc + SC_ACTIVE);
	r.epoch = fib->epoch;   // the inserts this call sees
	fib->in_flight.fetch_add(1);
	int rc = enqueue_persistent(&r, nq, false);
	const hipError_t e = hipStreamSynchronize(r.s);
	fib->in_flight.fetch_sub(1);
	if (rc != TM_OK) return rc;
	TM_HIP(e);
	return TM_OK;
}

} // extern "C"
