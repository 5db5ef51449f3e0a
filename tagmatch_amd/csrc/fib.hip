// FIB compilation and upload: the host-side layout pass that replaces the
// reference's back_end::compile_fibs (reference back_end.cc:363-453).
//
// Differences from the reference, by design:
//   * one structure-of-arrays table for all partitions, 64 filters per
//     block (one per lane of a wave64), word planes contiguous per block;
//   * per-block AND masks, used by the match kernel to discard whole blocks
//     for a tile of queries before testing single filters;
//   * keys resident in HBM as dense ids (reference keeps them on the host,
//     back_end.cc:69-88, and expands hits there, back_end.cc:288-301);
//   * within a partition filters are re-ordered so that the rarest bits are
//     the most significant sort positions, which makes block masks sparser
//     in common bits and more selective (the partition's membership, and so
//     the results, are unchanged).
#include "fib.h"

#include <algorithm>
#include <array>
#include <cstring>
#include <numeric>
#include <thread>

using namespace tmx;

void tm_fib::release() {
	dfree(blk_words);
	dfree(blk_info);
	dfree(slot_keys);
	dfree(slot_fid);
	dfree(ovf);
	dfree(key_dict);
	dfree(part_info);
	dfree(part_stat);
	dfree(sup_info);
	sup_groups = 0;
	dfree(slot_epoch);
	dfree(d_rich_words);
	dfree(d_rich_bm);
	n_rich = n_rich_full = 0;
	epoch = 0;
	n_parts = n_blocks = n_keys = 0;
	n_filters = n_key_instances = device_bytes = max_part = 0;
	ovf_cap = 0;
	h_part_info.clear();
	h_part_size.clear();
	h_part_stat.clear();
	h_part_cap.clear();
	h_ovf.clear();
	h_dict.clear();
	h_order.clear();
}

namespace {

template <typename F>
void parallel_ranges(size_t n, F && fn) {
	unsigned T = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
	if (n < 2 || T == 1) {
		fn(0, n);
		return;
	}
	std::vector<std::thread> th;
	size_t chunk = (n + T - 1) / T;
	for (unsigned t = 0; t < T; ++t) {
		size_t b = t * chunk, e = std::min(n, b + chunk);
		if (b >= e) break;
		th.emplace_back([&fn, b, e] { fn(b, e); });
	}
	for (auto & t : th) t.join();
}

template <typename T>
int upload(T ** dst, const std::vector<T> & src, uint64_t & bytes) {
	TM_CHECK(dalloc(dst, src.size()));
	if (!src.empty()) TM_HIP(hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
	bytes += std::max<size_t>(1, src.size()) * sizeof(T);
	return TM_OK;
}

// Host side of a block range under construction: partitions are laid out
// one after another from block `base`, each followed by its reserve.
struct block_build {
	std::vector<uint32_t> words, info, fid, part, stat, cap;
	std::vector<uint64_t> keys;
	uint32_t nb = 0;
};

// Dense key slot of filter i: up to four inline 16-bit ids, else an
// overflow list (count, ids...) appended to ovf.
int key_slot(const uint32_t * keys, uint32_t kb, uint32_t ke, const std::vector<uint32_t> & dict, bool inline_ok,
			 std::vector<uint32_t> & ovf, uint64_t & out) {
	auto key_id = [&](uint32_t k) -> uint32_t {
		return (uint32_t)(std::lower_bound(dict.begin(), dict.end(), k) - dict.begin());
	};
	const uint32_t kn = ke - kb;
	if (inline_ok && kn <= 4) {
		uint64_t v = ~0ULL;
		for (uint32_t x = 0; x < kn; ++x) {
			const uint64_t id = key_id(keys[kb + x]);
			v &= ~(0xFFFFULL << (16 * x));
			v |= id << (16 * x);
		}
		out = v;
		return TM_OK;
	}
	const uint64_t off = ovf.size();
	if (off > 0xFFFFFFFFULL) return set_error(TM_ERR_INVALID, "overflow key table too large");
	ovf.push_back(kn);
	for (uint32_t x = 0; x < kn; ++x) ovf.push_back(key_id(keys[kb + x]));
	out = 0xFFFFULL | (off << 16);
	return TM_OK;
}

uint32_t reserve_blocks(uint32_t nb, uint32_t pct) { return pct ? (nb * pct + 99) / 100 + 1 : 0; }

// Lay out partitions [0, n_parts) of `words` (members of p are
// part_offsets[p] .. part_offsets[p+1] in partitioned order) as blocks,
// partition ids starting at p_base, blocks at b_base.  Members are sorted
// within each partition by the significance order `order` (rarest bit
// first) so that block AND masks are selective.
int build_blocks(const uint32_t * words, const uint32_t * part_offsets, const uint32_t * part_masks, uint32_t n_parts,
				 const uint32_t * key_offsets, const uint32_t * keys, const uint32_t * filter_ids, uint64_t fid_base,
				 const std::vector<int> & order, const std::vector<uint32_t> & dict, bool inline_ok,
				 std::vector<uint32_t> & ovf, uint32_t p_base, uint32_t b_base, uint32_t reserve_pct,
				 block_build & bb, uint32_t last_cap = 0) {
	const uint32_t n = part_offsets[n_parts];
	std::vector<uint32_t> pblock(n_parts + 1, 0);
	for (uint32_t p = 0; p < n_parts; ++p) {
		const uint32_t sz = part_offsets[p + 1] - part_offsets[p];
		const uint32_t used = (sz + 63) / 64;
		const uint32_t cap = (last_cap && p + 1 == n_parts) ? std::max(last_cap, used)
															: used + reserve_blocks(used, reserve_pct);
		pblock[p + 1] = pblock[p] + cap;
	}
	const uint32_t nb = pblock[n_parts];
	std::vector<uint32_t> member(n);  // member[k] = input index at partitioned position k
	std::iota(member.begin(), member.end(), 0u);
	// sort keys: the filter's bits in significance order, most significant
	// first; a table per input byte gives that byte's share of the key
	std::vector<std::array<uint64_t, 3>> tab((size_t)24 * 256, std::array<uint64_t, 3>{0, 0, 0});
	{
		std::vector<int> rank(192, 0);
		for (int r = 0; r < 192 && r < (int)order.size(); ++r) rank[order[r]] = r;
		for (int bp = 0; bp < 24; ++bp)
			for (int v = 0; v < 256; ++v)
				for (int i = 0; i < 8; ++i)
					if ((v >> i) & 1) {
						const int r = rank[bp * 8 + i];
						tab[(size_t)bp * 256 + v][r >> 6] |= 1ULL << (63 - (r & 63));
					}
	}
	parallel_ranges(n_parts, [&](size_t pb, size_t pe) {
		std::vector<std::array<uint64_t, 3>> key;
		for (size_t p = pb; p < pe; ++p) {
			uint32_t b = part_offsets[p], e = part_offsets[p + 1];
			if (e - b < 2) continue;
			key.resize(e - b);
			for (uint32_t i = b; i < e; ++i) {
				std::array<uint64_t, 3> k{0, 0, 0};
				const uint8_t * wb = reinterpret_cast<const uint8_t *>(words + (size_t)i * 6);
				for (int bp = 0; bp < 24; ++bp) {
					const auto & t = tab[(size_t)bp * 256 + wb[bp]];
					k[0] |= t[0];
					k[1] |= t[1];
					k[2] |= t[2];
				}
				key[i - b] = k;
			}
			std::stable_sort(member.begin() + b, member.begin() + e, [&](uint32_t x, uint32_t y) {
				return key[x - b] < key[y - b];
			});
		}
	});

	bb.nb = nb;
	bb.words.assign((size_t)nb * BLOCK_WORDS, 0);
	bb.info.assign((size_t)nb * 8, 0xFFFFFFFFu);
	bb.keys.assign((size_t)nb * 64, ~0ULL);
	bb.fid.assign((size_t)nb * 64, 0xFFFFFFFFu);
	bb.part.assign((size_t)n_parts * 8, 0);
	bb.stat.assign((size_t)n_parts * 8, 0);
	bb.cap.assign(n_parts, 0);
	for (uint32_t p = 0; p < n_parts; ++p) {
		const uint32_t b0 = part_offsets[p], sz = part_offsets[p + 1] - b0;
		for (int j = 0; j < 6; ++j) bb.part[(size_t)p * 8 + j] = part_masks[(size_t)p * 6 + j];
		bb.part[(size_t)p * 8 + 6] = b_base + pblock[p];
		bb.part[(size_t)p * 8 + 7] = (sz + 63) / 64;
		bb.cap[p] = pblock[p + 1] - pblock[p];
		for (uint32_t blk = pblock[p]; blk < pblock[p + 1]; ++blk) {   // empty: nv 0, mask all ones
			bb.info[(size_t)blk * 8 + 6] = 0;
			bb.info[(size_t)blk * 8 + 7] = p_base + p;
		}
		uint32_t * ps = &bb.stat[(size_t)p * 8];
		for (uint32_t k = 0; k < sz; ++k) {
			const uint32_t i = member[b0 + k];
			const uint32_t blk = pblock[p] + k / 64, lane = k % 64;
			const uint32_t * w = words + (size_t)i * 6;
			uint32_t * info = &bb.info[(size_t)blk * 8];
			for (int j = 0; j < 6; ++j) {
				bb.words[(size_t)blk * BLOCK_WORDS + j * 64 + lane] = w[j];
				info[j] &= w[j];
				ps[j] |= w[j];
				ps[7] += (uint32_t)__builtin_popcount(w[j]);
			}
			info[6] = lane + 1;
			ps[6] += 1;
			bb.fid[(size_t)blk * 64 + lane] = filter_ids ? filter_ids[i] : (uint32_t)(fid_base + i);
			TM_CHECK(key_slot(keys, key_offsets[i], key_offsets[i + 1], dict, inline_ok, ovf,
							  bb.keys[(size_t)blk * 64 + lane]));
		}
	}
	return TM_OK;
}

int check_table(const uint32_t * words, uint32_t n, const uint32_t * key_offsets, const uint32_t * keys) {
	if (n && (!words || !key_offsets || !keys)) return set_error(TM_ERR_INVALID, "NULL input array");
	if (n == 0) return TM_OK;
	if (key_offsets[0] != 0) return set_error(TM_ERR_INVALID, "key_offsets[0] must be 0");
	for (uint32_t i = 0; i < n; ++i)
		if (key_offsets[i + 1] <= key_offsets[i])
			return set_error(TM_ERR_INVALID, "filter %u has no keys (key_offsets must increase)", i);
	return TM_OK;
}

// Copy `old_n` elements of *p into a new allocation of new_n elements.
template <typename T>
int regrow(T ** p, size_t old_n, size_t new_n, hipStream_t s) {
	T * q = nullptr;
	TM_CHECK(dalloc(&q, new_n));
	if (old_n && *p) {
		hipError_t e = hipMemcpyAsync(q, *p, old_n * sizeof(T), hipMemcpyDeviceToDevice, s);
		if (e == hipSuccess) e = hipStreamSynchronize(s);
		if (e != hipSuccess) {
			dfree(q);
			return hip_error(e, "regrow copy", __FILE__, __LINE__);
		}
	}
	dfree(*p);
	*p = q;
	return TM_OK;
}

template <typename T>
int put(T * dst, const std::vector<T> & src, size_t n = SIZE_MAX) {
	const size_t k = std::min(n, src.size());
	if (k) TM_HIP(hipMemcpy(dst, src.data(), k * sizeof(T), hipMemcpyHostToDevice));
	return TM_OK;
}

// One inserted filter: its lane and content.
struct ins_rec {
	uint32_t block, lane, fid, pad;
	uint32_t w[6];
	uint64_t kv;
};

// Writes each inserted filter into its lane and folds it into the block's
// AND mask and valid-lane count.
__global__ void __launch_bounds__(256)
k_insert_lanes(const ins_rec * __restrict__ recs, uint32_t n, uint32_t n_blocks, uint32_t * __restrict__ blk_words,
			   uint32_t * __restrict__ blk_info, uint32_t * __restrict__ sup_info, uint64_t * __restrict__ slot_keys,
			   uint32_t * __restrict__ slot_fid, uint32_t * __restrict__ slot_epoch, uint32_t epoch) {
	const uint32_t i = blockIdx.x * 256 + threadIdx.x;
	if (i >= n) return;
	const ins_rec r = recs[i];
	// records are made against the current table; one naming a block it
	// does not have is dropped rather than written out of bounds
	if (r.block >= n_blocks || r.lane >= 64) return;
	// a submission older than this insert sees the lane with its old epoch
	// (none yet: invisible) or the new one (newer than it: invisible)
	if (slot_epoch) slot_epoch[(size_t)r.block * 64 + r.lane] = epoch;
	for (int j = 0; j < 6; ++j) {
		blk_words[(size_t)r.block * BLOCK_WORDS + j * 64 + r.lane] = r.w[j];
		atomicAnd(&blk_info[(size_t)r.block * 8 + j], r.w[j]);
		if (sup_info) atomicAnd(&sup_info[(size_t)(r.block / SUP_BLOCKS) * 8 + j], r.w[j]);
	}
	atomicMax(&blk_info[(size_t)r.block * 8 + 6], r.lane + 1);
	slot_keys[(size_t)r.block * 64 + r.lane] = r.kv;
	slot_fid[(size_t)r.block * 64 + r.lane] = r.fid;
}

// Group masks: AND of the masks of blocks [8g, 8g + 8) (words 0-5; empty
// blocks hold all ones), one thread per group.
__global__ void __launch_bounds__(256)
k_sup_masks(const uint32_t * __restrict__ blk_info, uint32_t nb, uint32_t * __restrict__ sup_info) {
	const uint32_t gi = blockIdx.x * 256 + threadIdx.x;
	if ((uint64_t)gi * SUP_BLOCKS >= nb) return;
	uint32_t m[6] = {~0u, ~0u, ~0u, ~0u, ~0u, ~0u};
	const uint32_t b1 = min(gi * SUP_BLOCKS + SUP_BLOCKS, nb);
	for (uint32_t b = gi * SUP_BLOCKS; b < b1; ++b)
		for (int j = 0; j < 6; ++j) m[j] &= blk_info[(size_t)b * 8 + j];
	for (int j = 0; j < 6; ++j) sup_info[(size_t)gi * 8 + j] = m[j];
	sup_info[(size_t)gi * 8 + 6] = 0;
	sup_info[(size_t)gi * 8 + 7] = 0;
}

// Lane epochs of blocks [b0, nb): lanes already valid get `visible`, the
// others (spare) the never-visible maximum.
__global__ void __launch_bounds__(256)
k_init_epochs(const uint32_t * __restrict__ blk_info, uint32_t b0, uint32_t nb, uint32_t * __restrict__ slot_epoch,
			  uint32_t visible) {
	const uint64_t i = (uint64_t)b0 * 64 + (uint64_t)blockIdx.x * 256 + threadIdx.x;
	if (i >= (uint64_t)nb * 64) return;
	const uint32_t b = (uint32_t)(i >> 6), lane = (uint32_t)(i & 63);
	slot_epoch[i] = lane < blk_info[(size_t)b * 8 + 6] ? visible : 0xFFFFFFFFu;
}

// Dense ids of every slot through `map` (old id -> new id) after the key
// dictionary grew; overflow-list slots are rewritten on the host.
__global__ void __launch_bounds__(256)
k_remap_keys(uint64_t * __restrict__ slot_keys, uint64_t n, const uint32_t * __restrict__ map) {
	const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
	if (i >= n) return;
	uint64_t v = slot_keys[i];
	if ((v & 0xFFFFu) == KEY_NONE) return;   // empty slot or overflow list
	uint64_t o = 0;
	for (int x = 0; x < 4; ++x) {
		const uint64_t id = (v >> (16 * x)) & 0xFFFFu;
		o |= (id == KEY_NONE ? (uint64_t)KEY_NONE : (uint64_t)map[id]) << (16 * x);
	}
	slot_keys[i] = o;
}

// Group masks of every block, into a fresh array when the block count
// outgrew the current one.  `fresh` says whether sup_info moved (rings must
// then re-read the FIB, as for any other array that moved).
int refresh_sup(tm_fib * fib, bool & fresh) {
	const uint64_t ng = ((uint64_t)fib->n_blocks + SUP_BLOCKS - 1) / SUP_BLOCKS;
	fresh = false;
	if (!fib->sup_info || ng > fib->sup_groups) {
		dfree(fib->sup_info);
		fib->sup_groups = 0;
		TM_CHECK(dalloc(&fib->sup_info, std::max<uint64_t>(ng, 1) * 8));
		fib->sup_groups = std::max<uint64_t>(ng, 1);
		fresh = true;
	}
	if (ng) hipLaunchKernelGGL(k_sup_masks, dim3((uint32_t)((ng + 255) / 256)), dim3(256), 0, 0, fib->blk_info,
							   fib->n_blocks, fib->sup_info);
	TM_HIP(hipGetLastError());
	TM_HIP(hipDeviceSynchronize());
	return TM_OK;
}

} // namespace

extern "C" {

int tm_fib_create(tm_fib ** out) {
	if (!out) return set_error(TM_ERR_INVALID, "out is NULL");
	*out = new tm_fib();
	return TM_OK;
}

int tm_fib_destroy(tm_fib * fib) {
	if (!fib) return TM_OK;
	if (dev().device >= 0) (void)hipSetDevice(dev().device);
	if (fib->ins_stream) (void)hipStreamSynchronize(fib->ins_stream);
	fib->release();
	if (fib->d_recs) (void)hipFree(fib->d_recs);
	if (fib->ins_stream) (void)hipStreamDestroy(fib->ins_stream);
	delete fib;
	return TM_OK;
}

int tm_fib_key_dictionary(const tm_fib * fib, uint32_t * keys, uint64_t capacity, uint64_t * n) {
	if (!fib || !n) return set_error(TM_ERR_INVALID, "NULL argument");
	std::lock_guard<std::mutex> lk(const_cast<tm_fib *>(fib)->host_mu);   // inserts grow the dictionary
	*n = fib->h_dict.size();
	if (!keys) return TM_OK;
	if (capacity < fib->h_dict.size())
		return set_error(TM_ERR_OVERFLOW, "%llu keys, capacity %llu", (unsigned long long)fib->h_dict.size(),
						 (unsigned long long)capacity);
	if (!fib->h_dict.empty()) std::memcpy(keys, fib->h_dict.data(), fib->h_dict.size() * 4);
	return TM_OK;
}

int tm_fib_stats_get(const tm_fib * fib, tm_fib_stats * st) {
	if (!fib || !st) return set_error(TM_ERR_INVALID, "NULL argument");
	st->n_filters = fib->n_filters;
	st->n_blocks = fib->n_blocks;
	st->n_parts = fib->n_parts;
	st->n_keys = fib->n_key_instances;
	st->n_distinct_keys = fib->n_keys;
	st->device_bytes = fib->device_bytes;
	st->max_part_size = fib->max_part;
	return TM_OK;
}

int tm_fib_set_reserve(tm_fib * fib, uint32_t percent) {
	if (!fib) return set_error(TM_ERR_INVALID, "fib is NULL");
	if (percent > 400) return set_error(TM_ERR_INVALID, "reserve %u%% above 400%%", percent);
	fib->reserve_pct = percent;
	return TM_OK;
}

int tm_fib_upload(tm_fib * fib, const uint32_t * words, uint32_t n,
				  const uint32_t * part_offsets, const uint32_t * part_masks, uint32_t n_parts,
				  const uint32_t * key_offsets, const uint32_t * keys,
				  const uint32_t * filter_ids) {
	if (!fib) return set_error(TM_ERR_INVALID, "fib is NULL");
	std::lock_guard<std::mutex> lk(fib->host_mu);
	if (n && (!words || !key_offsets || !keys)) return set_error(TM_ERR_INVALID, "NULL input array");
	if (!part_offsets || (n_parts && !part_masks)) return set_error(TM_ERR_INVALID, "NULL partition array");
	if (part_offsets[0] != 0 || part_offsets[n_parts] != n)
		return set_error(TM_ERR_INVALID, "part_offsets must run from 0 to n (%u)", n);
	for (uint32_t p = 0; p < n_parts; ++p)
		if (part_offsets[p + 1] < part_offsets[p])
			return set_error(TM_ERR_INVALID, "part_offsets not monotone at %u", p);
	if (key_offsets[0] != 0) return set_error(TM_ERR_INVALID, "key_offsets[0] must be 0");
	for (uint32_t i = 0; i < n; ++i)
		if (key_offsets[i + 1] <= key_offsets[i])
			return set_error(TM_ERR_INVALID, "filter %u has no keys (key_offsets must increase)", i);
	if (fib->in_flight.load() > 0) return set_error(TM_ERR_BUSY, "a ring on this FIB has a submission in flight");

	// Every partition mask must be a subset of each member filter, or the
	// lookup would drop matches (reference front_end.cc:465-477 relies on the
	// same property, partitioning.cc:389-402).
	for (uint32_t p = 0; p < n_parts; ++p)
		for (uint32_t i = part_offsets[p]; i < part_offsets[p + 1]; ++i)
			for (int j = 0; j < 6; ++j)
				if (part_masks[p * 6 + j] & ~words[(size_t)i * 6 + j])
					return set_error(TM_ERR_INVALID, "mask of partition %u is not a subset of filter %u", p, i);

	TM_CHECK(require_device());

	// Key dictionary: sorted distinct keys, dense ids.
	uint64_t nkeys = key_offsets[n];
	std::vector<uint32_t> dict(keys, keys + nkeys);
	std::sort(dict.begin(), dict.end());
	dict.erase(std::unique(dict.begin(), dict.end()), dict.end());
	if (dict.size() > 0xFFFFFFF0u) return set_error(TM_ERR_INVALID, "too many distinct keys");
	const uint32_t K = (uint32_t)dict.size();
	const bool inline_ok = K < KEY_NONE;

	// Global bit frequencies -> significance order, rarest bit first.
	std::vector<uint64_t> freq(192, 0);
	for (uint32_t i = 0; i < n; ++i)
		for (int j = 0; j < 6; ++j) {
			uint32_t w = words[(size_t)i * 6 + j];
			while (w) {
				int b = __builtin_ctz(w);
				freq[j * 32 + b]++;
				w &= w - 1;
			}
		}
	std::vector<int> order(192);
	std::iota(order.begin(), order.end(), 0);
	std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return freq[a] < freq[b]; });

	std::vector<uint32_t> h_ovf(1, 0);   // overflow key lists (count, ids...)
	block_build bb;
	// With a reserve, one more partition: empty, mask 0 (every query selects
	// it), room for about 1% of the table.  Inserted filters that contain no
	// partition's mask go there instead of needing a new partition, so that
	// they can be inserted live.
	const bool catch_all = fib->reserve_pct > 0;
	std::vector<uint32_t> off2, masks2;
	const uint32_t * po = part_offsets;
	const uint32_t * pm = part_masks;
	uint32_t np = n_parts, last_cap = 0;
	if (catch_all) {
		off2.assign(part_offsets, part_offsets + n_parts + 1);
		off2.push_back(n);
		masks2.assign(part_masks, part_masks + (size_t)n_parts * 6);
		masks2.resize(masks2.size() + 6, 0u);
		po = off2.data();
		pm = masks2.data();
		np = n_parts + 1;
		last_cap = std::max<uint32_t>(16, (n / 64) / 100);
	}
	TM_CHECK(build_blocks(words, po, pm, np, key_offsets, keys, filter_ids, 0, order, dict, inline_ok, h_ovf, 0, 0,
						  fib->reserve_pct, bb, last_cap));
	uint64_t maxp = 0;
	std::vector<uint32_t> h_psize(np, 0);
	for (uint32_t p = 0; p < np; ++p) {
		h_psize[p] = po[p + 1] - po[p];
		maxp = std::max<uint64_t>(maxp, h_psize[p]);
	}

	// rich filters: the MAX_RICH with the most keys (at least RICH_MIN_KEYS),
	// their words and their key sets as dense bitmaps
	std::vector<uint32_t> rich_words, rich_bm;
	uint32_t n_rich_full = 0;
	{
		std::vector<std::pair<uint32_t, uint32_t>> cand;   // (keys, filter)
		for (uint32_t i = 0; i < n; ++i) {
			const uint32_t kn = key_offsets[i + 1] - key_offsets[i];
			if (kn >= RICH_MIN_KEYS) cand.push_back({kn, i});
		}
		const size_t nr = std::min<size_t>(cand.size(), MAX_RICH);
		std::partial_sort(cand.begin(), cand.begin() + nr, cand.end(),
						  [](const auto & x, const auto & y) { return x.first > y.first || (x.first == y.first && x.second < y.second); });
		const uint32_t kwr = (K + 31) / 32;
		std::vector<uint32_t> bw(nr * 8, 0), bm(nr * kwr, 0);
		std::vector<uint8_t> full(nr, 0);
		for (size_t r = 0; r < nr; ++r) {
			const uint32_t i = cand[r].second;
			for (int j = 0; j < 6; ++j) bw[r * 8 + j] = words[(size_t)i * 6 + j];
			uint32_t bits = 0;
			for (uint32_t x = key_offsets[i]; x < key_offsets[i + 1]; ++x) {
				const uint32_t id = (uint32_t)(std::lower_bound(dict.begin(), dict.end(), keys[x]) - dict.begin());
				bits += !((bm[r * kwr + id / 32] >> (id % 32)) & 1u);
				bm[r * kwr + id / 32] |= 1u << (id % 32);
			}
			full[r] = bits == K;
		}
		// filters holding every key first: a query that contains one needs
		// nothing else
		rich_words.reserve(nr * 8);
		rich_bm.reserve(nr * kwr);
		for (int pass = 0; pass < 2; ++pass)
			for (size_t r = 0; r < nr; ++r) {
				if (full[r] != (pass == 0)) continue;
				rich_words.insert(rich_words.end(), bw.begin() + r * 8, bw.begin() + r * 8 + 8);
				rich_bm.insert(rich_bm.end(), bm.begin() + r * kwr, bm.begin() + (r + 1) * kwr);
				n_rich_full += pass == 0;
			}
	}

	fib->release();
	uint64_t bytes = 0;
	TM_CHECK(upload(&fib->blk_words, bb.words, bytes));
	TM_CHECK(upload(&fib->blk_info, bb.info, bytes));
	TM_CHECK(upload(&fib->slot_keys, bb.keys, bytes));
	TM_CHECK(upload(&fib->slot_fid, bb.fid, bytes));
	fib->ovf_cap = std::max<size_t>(h_ovf.size() * 2, 1024);
	TM_CHECK(dalloc(&fib->ovf, fib->ovf_cap));
	TM_CHECK(put(fib->ovf, h_ovf));
	bytes += fib->ovf_cap * 4;
	TM_CHECK(upload(&fib->key_dict, dict, bytes));
	TM_CHECK(upload(&fib->part_info, bb.part, bytes));
	TM_CHECK(upload(&fib->part_stat, bb.stat, bytes));
	if (!rich_words.empty()) {
		TM_CHECK(upload(&fib->d_rich_words, rich_words, bytes));
		TM_CHECK(upload(&fib->d_rich_bm, rich_bm, bytes));
	}
	fib->n_rich = (uint32_t)(rich_words.size() / 8);
	fib->n_rich_full = n_rich_full;
	fib->n_parts = np;
	fib->n_blocks = bb.nb;
	{
		bool fresh;
		TM_CHECK(refresh_sup(fib, fresh));
		bytes += fib->sup_groups * 32;
	}
	fib->n_keys = K;
	fib->n_filters = n;
	fib->n_key_instances = nkeys;
	fib->device_bytes = bytes;
	fib->max_part = maxp;
	fib->h_part_info = std::move(bb.part);
	fib->h_part_stat = std::move(bb.stat);
	fib->h_part_cap = std::move(bb.cap);
	fib->h_part_size = std::move(h_psize);
	fib->h_ovf = std::move(h_ovf);
	fib->h_dict = std::move(dict);
	fib->h_order = std::move(order);
	fib->ovf_valid = (uint32_t)fib->h_ovf.size();
	fib->prep_epoch = fib->epoch;
	fib->version++;
	return TM_OK;
}

} // extern "C"

namespace tmx {

struct live_plan {
	std::vector<ins_rec> recs;
	size_t ovf_from = 0, ovf_to = 0;              // overflow words this insert appended
	std::vector<uint32_t> part_info, part_stat;   // P x 8, as after this insert
	uint32_t epoch = 0;
};

namespace {

int create_epochs(tm_fib * fib) {
	if (!fib->ins_stream) TM_HIP(hipStreamCreateWithFlags(&fib->ins_stream, hipStreamNonBlocking));
	TM_CHECK(dalloc(&fib->slot_epoch, (size_t)std::max<uint32_t>(fib->n_blocks, 1) * 64));
	hipLaunchKernelGGL(k_init_epochs, dim3((uint32_t)((fib->n_blocks * 64ull + 255) / 256)), dim3(256), 0,
					   fib->ins_stream, fib->blk_info, 0u, fib->n_blocks, fib->slot_epoch, 0u);
	TM_HIP(hipGetLastError());
	TM_HIP(hipStreamSynchronize(fib->ins_stream));
	return TM_OK;
}

} // namespace

// Host half of a live insert, with fib->host_mu held: checks that nothing
// but spare lanes, known keys and room in the overflow lists is needed
// (TM_ERR_BUSY and no change otherwise), chooses each filter's partition,
// builds the lane records and updates the host copies.
int live_prepare_locked(tm_fib * fib, const uint32_t * words, uint32_t n, const uint32_t * key_offsets,
						const uint32_t * keys, const uint32_t * filter_ids, live_plan ** out) {
	*out = nullptr;
	const uint64_t nkeys = key_offsets[n];
	for (uint64_t x = 0; x < nkeys; ++x)
		if (!std::binary_search(fib->h_dict.begin(), fib->h_dict.end(), keys[x]))
			return set_error(TM_ERR_BUSY, "key %u is new to the table: use tm_fib_insert", keys[x]);
	const bool inline_ok = fib->n_keys < KEY_NONE;
	size_t ovf_need = 0;
	for (uint32_t i = 0; i < n; ++i) {
		const uint32_t kn = key_offsets[i + 1] - key_offsets[i];
		if (!inline_ok || kn > 4) ovf_need += 1 + kn;
	}
	if (fib->h_ovf.size() + ovf_need > fib->ovf_cap)
		return set_error(TM_ERR_BUSY, "overflow key lists are full: use tm_fib_insert");
	const uint32_t P = fib->n_parts;
	std::vector<uint32_t> pmpop(P);
	for (uint32_t p = 0; p < P; ++p) {
		uint32_t c = 0;
		for (int j = 0; j < 6; ++j) c += __builtin_popcount(fib->h_part_info[(size_t)p * 8 + j]);
		pmpop[p] = c;
	}
	std::vector<uint32_t> size = fib->h_part_size, target(n);
	auto choose = [&](const uint32_t * w, const std::vector<uint32_t> & sz) -> int64_t {
		int64_t best = -1;
		for (uint32_t p = 0; p < P; ++p) {
			const uint32_t * m = &fib->h_part_info[(size_t)p * 8];
			if ((m[0] & ~w[0]) | (m[1] & ~w[1]) | (m[2] & ~w[2]) | (m[3] & ~w[3]) | (m[4] & ~w[4]) | (m[5] & ~w[5]))
				continue;
			if (sz[p] >= (uint64_t)fib->h_part_cap[p] * 64) continue;
			if (best < 0 || pmpop[p] > pmpop[best] || (pmpop[p] == pmpop[best] && sz[p] < sz[best])) best = p;
		}
		return best;
	};
	// first choices in parallel on the sizes before the batch, then in
	// order: a choice that the batch itself has filled is made again
	std::vector<int64_t> first(n);
	parallel_ranges(n, [&](size_t b0, size_t b1) {
		for (size_t i = b0; i < b1; ++i) first[i] = choose(words + i * 6, size);
	});
	for (uint32_t i = 0; i < n; ++i) {
		int64_t best = first[i];
		if (best >= 0 && size[best] >= (uint64_t)fib->h_part_cap[best] * 64) best = choose(words + (size_t)i * 6, size);
		if (best < 0) return set_error(TM_ERR_BUSY, "filter %u fits no spare lane: use tm_fib_insert", i);
		target[i] = (uint32_t)best;
		size[best]++;
	}
	// every check passed: lanes and host copies
	live_plan * pl = new live_plan();
	pl->epoch = fib->prep_epoch + 1;
	pl->ovf_from = fib->h_ovf.size();
	pl->recs.resize(n);
	for (uint32_t i = 0; i < n; ++i) {
		const uint32_t p = target[i];
		const uint32_t * w = words + (size_t)i * 6;
		const uint32_t k = fib->h_part_size[p]++;
		ins_rec & r = pl->recs[i];
		r.block = fib->h_part_info[(size_t)p * 8 + 6] + k / 64;
		r.lane = k % 64;
		r.fid = filter_ids ? filter_ids[i] : (uint32_t)(fib->n_filters + i);
		r.pad = 0;
		for (int j = 0; j < 6; ++j) r.w[j] = w[j];
		const int rc = key_slot(keys, key_offsets[i], key_offsets[i + 1], fib->h_dict, inline_ok, fib->h_ovf, r.kv);
		if (rc != TM_OK) {   // cannot happen after the capacity check above
			delete pl;
			return rc;
		}
		fib->h_part_info[(size_t)p * 8 + 7] = (fib->h_part_size[p] + 63) / 64;
		uint32_t * ps = &fib->h_part_stat[(size_t)p * 8];
		for (int j = 0; j < 6; ++j) {
			ps[j] |= w[j];
			ps[7] += (uint32_t)__builtin_popcount(w[j]);
		}
		ps[6] += 1;
		fib->max_part = std::max<uint64_t>(fib->max_part, fib->h_part_size[p]);
	}
	pl->ovf_to = fib->h_ovf.size();
	pl->part_info = fib->h_part_info;
	pl->part_stat = fib->h_part_stat;
	fib->prep_epoch = pl->epoch;
	fib->n_filters += n;
	fib->n_key_instances += nkeys;
	*out = pl;
	return TM_OK;
}

// Device half, with fib->host_mu held, on the insert stream: the new
// overflow words, the lanes (written with the plan's epoch), then the block
// counts and statistics.  A submission that sees a new block count sees
// lanes that its epoch hides.
int live_commit_locked(tm_fib * fib, live_plan * pl) {
	const uint32_t n = (uint32_t)pl->recs.size();
	if (!fib->ins_stream) TM_HIP(hipStreamCreateWithFlags(&fib->ins_stream, hipStreamNonBlocking));
	hipStream_t s = fib->ins_stream;
	if (fib->recs_cap < n) {
		if (fib->d_recs) {
			TM_HIP(hipStreamSynchronize(s));
			(void)hipFree(fib->d_recs);
			fib->d_recs = nullptr;
			fib->recs_cap = 0;
		}
		TM_HIP(hipMalloc(&fib->d_recs, sizeof(ins_rec) * (size_t)n * 2));
		fib->recs_cap = (size_t)n * 2;
	}
	if (pl->ovf_to > pl->ovf_from)
		TM_HIP(hipMemcpyAsync(fib->ovf + pl->ovf_from, fib->h_ovf.data() + pl->ovf_from,
							  (pl->ovf_to - pl->ovf_from) * 4, hipMemcpyHostToDevice, s));
	if (n) {
		TM_HIP(hipMemcpyAsync(fib->d_recs, pl->recs.data(), sizeof(ins_rec) * n, hipMemcpyHostToDevice, s));
		hipLaunchKernelGGL(k_insert_lanes, dim3((n + 255) / 256), dim3(256), 0, s, (const ins_rec *)fib->d_recs, n,
						   fib->n_blocks, fib->blk_words, fib->blk_info, fib->sup_info, fib->slot_keys, fib->slot_fid, fib->slot_epoch, pl->epoch);
		TM_HIP(hipGetLastError());
	}
	TM_HIP(hipMemcpyAsync(fib->part_info, pl->part_info.data(), pl->part_info.size() * 4, hipMemcpyHostToDevice, s));
	TM_HIP(hipMemcpyAsync(fib->part_stat, pl->part_stat.data(), pl->part_stat.size() * 4, hipMemcpyHostToDevice, s));
	TM_HIP(hipStreamSynchronize(s));
	fib->ovf_valid = (uint32_t)std::max<size_t>(fib->ovf_valid, pl->ovf_to);
	fib->epoch = pl->epoch;
	return TM_OK;
}

int live_prepare(tm_fib * fib, const uint32_t * words, uint32_t n, const uint32_t * key_offsets,
				 const uint32_t * keys, const uint32_t * filter_ids, live_plan ** out) {
	*out = nullptr;
	if (!fib) return set_error(TM_ERR_INVALID, "fib is NULL");
	TM_CHECK(check_table(words, n, key_offsets, keys));
	if (fib->h_order.size() != 192) return set_error(TM_ERR_STATE, "no table uploaded");
	std::lock_guard<std::mutex> lk(fib->host_mu);
	// lane epochs are created by a commit-side call only (tm_fib_insert or
	// tm_fib_insert_live), in the thread that makes the submissions
	if (!fib->slot_epoch) return set_error(TM_ERR_BUSY, "no lane epochs yet");
	if (n == 0) {
		*out = new live_plan();
		(*out)->epoch = ++fib->prep_epoch;
		(*out)->ovf_from = (*out)->ovf_to = fib->h_ovf.size();
		(*out)->part_info = fib->h_part_info;
		(*out)->part_stat = fib->h_part_stat;
		return TM_OK;
	}
	return live_prepare_locked(fib, words, n, key_offsets, keys, filter_ids, out);
}

int live_commit(tm_fib * fib, live_plan * plan) {
	int rc = TM_OK;
	{
		std::lock_guard<std::mutex> lk(fib->host_mu);
		rc = require_device();
		if (rc == TM_OK) rc = live_commit_locked(fib, plan);
	}
	delete plan;
	return rc;
}

void live_discard(live_plan * plan) { delete plan; }

} // namespace tmx

extern "C" {

int tm_fib_insert(tm_fib * fib, const uint32_t * words, uint32_t n, const uint32_t * key_offsets,
				  const uint32_t * keys, const uint32_t * filter_ids, uint32_t maxp) {
	if (!fib) return set_error(TM_ERR_INVALID, "fib is NULL");
	std::lock_guard<std::mutex> lk(fib->host_mu);
	TM_CHECK(check_table(words, n, key_offsets, keys));
	if (maxp == 0) return set_error(TM_ERR_INVALID, "maxp must be > 0");
	if (fib->in_flight.load() > 0) return set_error(TM_ERR_BUSY, "a ring on this FIB has a submission in flight");
	if (n == 0) return TM_OK;
	if (fib->h_order.size() != 192) return set_error(TM_ERR_STATE, "no table uploaded");
	TM_CHECK(require_device());
	bool bump = false;   // device arrays, P or K change: rings must re-read the FIB
	const uint32_t e = std::max(fib->epoch, fib->prep_epoch) + 1;
	if (!fib->slot_epoch) {
		// lane epochs from now on, so that later inserts can run live
		TM_CHECK(dalloc(&fib->slot_epoch, (size_t)std::max<uint32_t>(fib->n_blocks, 1) * 64));
		hipLaunchKernelGGL(k_init_epochs, dim3((uint32_t)((fib->n_blocks * 64ull + 255) / 256)), dim3(256), 0, 0,
						   fib->blk_info, 0u, fib->n_blocks, fib->slot_epoch, 0u);
		TM_HIP(hipGetLastError());
		TM_HIP(hipDeviceSynchronize());
	}

	// 1. keys new to the dictionary: grow it and remap dense ids
	const uint64_t nkeys = key_offsets[n];
	std::vector<uint32_t> fresh;
	for (uint64_t x = 0; x < nkeys; ++x)
		if (!std::binary_search(fib->h_dict.begin(), fib->h_dict.end(), keys[x])) fresh.push_back(keys[x]);
	if (!fresh.empty()) {
		std::sort(fresh.begin(), fresh.end());
		fresh.erase(std::unique(fresh.begin(), fresh.end()), fresh.end());
		std::vector<uint32_t> dict;
		std::merge(fib->h_dict.begin(), fib->h_dict.end(), fresh.begin(), fresh.end(), std::back_inserter(dict));
		const bool was_inline = fib->n_keys < KEY_NONE;
		if (was_inline && dict.size() >= KEY_NONE)
			return set_error(TM_ERR_STATE, "key dictionary outgrows inline ids: upload the table again");
		if (dict.size() > 0xFFFFFFF0u) return set_error(TM_ERR_INVALID, "too many distinct keys");
		std::vector<uint32_t> map(fib->h_dict.size());
		for (size_t i = 0; i < map.size(); ++i)
			map[i] = (uint32_t)(std::lower_bound(dict.begin(), dict.end(), fib->h_dict[i]) - dict.begin());
		if (was_inline && !map.empty()) {
			uint32_t * d_map = nullptr;
			TM_CHECK(dalloc(&d_map, map.size()));
			int rc = put(d_map, map);
			const uint64_t slots = (uint64_t)fib->n_blocks * 64;
			if (rc == TM_OK) {
				hipLaunchKernelGGL(k_remap_keys, dim3((uint32_t)((slots + 255) / 256)), dim3(256), 0, 0,
								   fib->slot_keys, slots, d_map);
				hipError_t e = hipGetLastError();
				if (e == hipSuccess) e = hipDeviceSynchronize();
				if (e != hipSuccess) rc = hip_error(e, "key remap", __FILE__, __LINE__);
			}
			dfree(d_map);
			TM_CHECK(rc);
		}
		// overflow lists: (count, ids...) from offset 1
		for (size_t o = 1; o < fib->h_ovf.size();) {
			const uint32_t c = fib->h_ovf[o];
			for (uint32_t x = 0; x < c; ++x) fib->h_ovf[o + 1 + x] = map[fib->h_ovf[o + 1 + x]];
			o += 1 + c;
		}
		TM_CHECK(put(fib->ovf, fib->h_ovf));
		dfree(fib->key_dict);
		uint64_t b = 0;
		TM_CHECK(upload(&fib->key_dict, dict, b));
		fib->h_dict = std::move(dict);
		fib->n_keys = (uint32_t)fib->h_dict.size();
		dfree(fib->d_rich_words);   // the bitmaps hold the old dense ids
		dfree(fib->d_rich_bm);
		fib->n_rich = fib->n_rich_full = 0;
		bump = true;
	}
	const bool inline_ok = fib->n_keys < KEY_NONE;
	const size_t ovf_before = fib->h_ovf.size();

	// 2. each filter joins the most selective partition whose mask it
	//    contains and that has a free reserved lane
	const uint32_t P = fib->n_parts;
	std::vector<uint32_t> pmpop(P);
	for (uint32_t p = 0; p < P; ++p) {
		uint32_t c = 0;
		for (int j = 0; j < 6; ++j) c += __builtin_popcount(fib->h_part_info[(size_t)p * 8 + j]);
		pmpop[p] = c;
	}
	std::vector<ins_rec> recs;
	std::vector<uint32_t> left;   // filters no partition takes
	recs.reserve(n);
	for (uint32_t i = 0; i < n; ++i) {
		const uint32_t * w = words + (size_t)i * 6;
		int64_t best = -1;
		for (uint32_t p = 0; p < P; ++p) {
			const uint32_t * m = &fib->h_part_info[(size_t)p * 8];
			if ((m[0] & ~w[0]) | (m[1] & ~w[1]) | (m[2] & ~w[2]) | (m[3] & ~w[3]) | (m[4] & ~w[4]) | (m[5] & ~w[5]))
				continue;
			if (fib->h_part_size[p] >= (uint64_t)fib->h_part_cap[p] * 64) continue;
			if (best < 0 || pmpop[p] > pmpop[best] ||
				(pmpop[p] == pmpop[best] && fib->h_part_size[p] < fib->h_part_size[best]))
				best = p;
		}
		if (best < 0) {
			left.push_back(i);
			continue;
		}
		const uint32_t p = (uint32_t)best;
		const uint32_t k = fib->h_part_size[p]++;
		ins_rec r;
		r.block = fib->h_part_info[(size_t)p * 8 + 6] + k / 64;
		r.lane = k % 64;
		r.fid = filter_ids ? filter_ids[i] : (uint32_t)(fib->n_filters + i);
		r.pad = 0;
		for (int j = 0; j < 6; ++j) r.w[j] = w[j];
		TM_CHECK(key_slot(keys, key_offsets[i], key_offsets[i + 1], fib->h_dict, inline_ok, fib->h_ovf, r.kv));
		recs.push_back(r);
		fib->h_part_info[(size_t)p * 8 + 7] = (fib->h_part_size[p] + 63) / 64;
		uint32_t * ps = &fib->h_part_stat[(size_t)p * 8];
		for (int j = 0; j < 6; ++j) {
			ps[j] |= w[j];
			ps[7] += (uint32_t)__builtin_popcount(w[j]);
		}
		ps[6] += 1;
		fib->max_part = std::max<uint64_t>(fib->max_part, fib->h_part_size[p]);
	}

	// 3. the rest: partitioned on its own, appended as new partitions
	block_build bb;
	uint32_t np = 0;
	if (!left.empty()) {
		const uint32_t m = (uint32_t)left.size();
		std::vector<uint32_t> lw((size_t)m * 6), lko(m + 1, 0), lk, lfid(m);
		for (uint32_t x = 0; x < m; ++x) {
			const uint32_t i = left[x];
			std::memcpy(&lw[(size_t)x * 6], words + (size_t)i * 6, 24);
			lfid[x] = filter_ids ? filter_ids[i] : (uint32_t)(fib->n_filters + i);
		}
		std::vector<uint32_t> perm(m), poff(m + 2), pmask((size_t)(m + 1) * 6);
		TM_CHECK(tm_partition(lw.data(), m, maxp, perm.data(), poff.data(), pmask.data(), m + 1, &np));
		std::vector<uint32_t> pw((size_t)m * 6), pfid(m);
		for (uint32_t x = 0; x < m; ++x) {
			const uint32_t i = left[perm[x]];
			std::memcpy(&pw[(size_t)x * 6], words + (size_t)i * 6, 24);
			pfid[x] = lfid[perm[x]];
			for (uint32_t y = key_offsets[i]; y < key_offsets[i + 1]; ++y) lk.push_back(keys[y]);
			lko[x + 1] = (uint32_t)lk.size();
		}
		TM_CHECK(build_blocks(pw.data(), poff.data(), pmask.data(), np, lko.data(), lk.data(), pfid.data(), 0,
							  fib->h_order, fib->h_dict, inline_ok, fib->h_ovf, P, fib->n_blocks, fib->reserve_pct,
							  bb));
	}

	// 4. device: grow arrays for new partitions and overflow lists
	hipStream_t s = nullptr;
	if (np) {
		const size_t ob = fib->n_blocks, nbt = ob + bb.nb;
		TM_CHECK(regrow(&fib->blk_words, ob * BLOCK_WORDS, nbt * BLOCK_WORDS, s));
		TM_CHECK(regrow(&fib->blk_info, ob * 8, nbt * 8, s));
		TM_CHECK(regrow(&fib->slot_keys, ob * 64, nbt * 64, s));
		TM_CHECK(regrow(&fib->slot_fid, ob * 64, nbt * 64, s));
		if (fib->slot_epoch) TM_CHECK(regrow(&fib->slot_epoch, ob * 64, nbt * 64, s));
		TM_HIP(hipMemcpy(fib->blk_words + ob * BLOCK_WORDS, bb.words.data(), bb.words.size() * 4, hipMemcpyHostToDevice));
		TM_HIP(hipMemcpy(fib->blk_info + ob * 8, bb.info.data(), bb.info.size() * 4, hipMemcpyHostToDevice));
		TM_HIP(hipMemcpy(fib->slot_keys + ob * 64, bb.keys.data(), bb.keys.size() * 8, hipMemcpyHostToDevice));
		TM_HIP(hipMemcpy(fib->slot_fid + ob * 64, bb.fid.data(), bb.fid.size() * 4, hipMemcpyHostToDevice));
		if (fib->slot_epoch) {
			hipLaunchKernelGGL(k_init_epochs, dim3((uint32_t)((bb.nb * 64ull + 255) / 256)), dim3(256), 0, 0,
							   fib->blk_info, (uint32_t)ob, (uint32_t)nbt, fib->slot_epoch, e);
			TM_HIP(hipGetLastError());
			TM_HIP(hipDeviceSynchronize());
		}
		fib->device_bytes += (uint64_t)bb.nb * (BLOCK_WORDS * 4 + 32 + 64 * 12);
		fib->n_blocks = (uint32_t)nbt;
		fib->h_part_info.insert(fib->h_part_info.end(), bb.part.begin(), bb.part.end());
		fib->h_part_stat.insert(fib->h_part_stat.end(), bb.stat.begin(), bb.stat.end());
		fib->h_part_cap.insert(fib->h_part_cap.end(), bb.cap.begin(), bb.cap.end());
		for (uint32_t p = 0; p < np; ++p) {
			const uint32_t sz = bb.stat[(size_t)p * 8 + 6];
			fib->h_part_size.push_back(sz);
			fib->max_part = std::max<uint64_t>(fib->max_part, sz);
		}
		dfree(fib->part_info);
		dfree(fib->part_stat);
		uint64_t b = 0;
		TM_CHECK(upload(&fib->part_info, fib->h_part_info, b));
		TM_CHECK(upload(&fib->part_stat, fib->h_part_stat, b));
		fib->n_parts = P + np;
		bump = true;
	}
	if (fib->h_ovf.size() > fib->ovf_cap) {
		const size_t cap = fib->h_ovf.size() * 2;
		dfree(fib->ovf);
		TM_CHECK(dalloc(&fib->ovf, cap));
		fib->device_bytes += (cap - fib->ovf_cap) * 4;
		fib->ovf_cap = cap;
		TM_CHECK(put(fib->ovf, fib->h_ovf));
		bump = true;
	} else if (fib->h_ovf.size() > ovf_before) {
		TM_HIP(hipMemcpy(fib->ovf + ovf_before, fib->h_ovf.data() + ovf_before, (fib->h_ovf.size() - ovf_before) * 4,
						 hipMemcpyHostToDevice));
	}

	// 5. device: lanes into reserved blocks, then the partitions' new block
	//    counts and statistics
	if (!recs.empty()) {
		ins_rec * d_recs = nullptr;
		TM_CHECK(dalloc(&d_recs, recs.size()));
		int rc = put(d_recs, recs);
		if (rc == TM_OK) {
			hipLaunchKernelGGL(k_insert_lanes, dim3((uint32_t)((recs.size() + 255) / 256)), dim3(256), 0, 0, d_recs,
							   (uint32_t)recs.size(), fib->n_blocks, fib->blk_words, fib->blk_info, fib->sup_info, fib->slot_keys, fib->slot_fid,
							   fib->slot_epoch, e);
			hipError_t e = hipGetLastError();
			if (e == hipSuccess) e = hipDeviceSynchronize();
			if (e != hipSuccess) rc = hip_error(e, "insert", __FILE__, __LINE__);
		}
		dfree(d_recs);
		TM_CHECK(rc);
		TM_CHECK(put(fib->part_info, fib->h_part_info, (size_t)P * 8));
		TM_CHECK(put(fib->part_stat, fib->h_part_stat, (size_t)P * 8));
	}
	// group masks of the new blocks (and of the blocks the lanes above joined)
	{
		bool fresh;
		TM_CHECK(refresh_sup(fib, fresh));
		if (fresh) bump = true;
	}
	fib->n_filters += n;
	fib->n_key_instances += nkeys;
	fib->epoch = e;
	fib->prep_epoch = e;
	fib->ovf_valid = (uint32_t)fib->h_ovf.size();
	if (bump) fib->version++;
	return TM_OK;
}

int tm_fib_insert_live(tm_fib * fib, const uint32_t * words, uint32_t n, const uint32_t * key_offsets,
					   const uint32_t * keys, const uint32_t * filter_ids) {
	if (!fib) return set_error(TM_ERR_INVALID, "fib is NULL");
	TM_CHECK(check_table(words, n, key_offsets, keys));
	if (n == 0) return TM_OK;
	if (fib->h_order.size() != 192) return set_error(TM_ERR_STATE, "no table uploaded");
	TM_CHECK(require_device());
	std::lock_guard<std::mutex> lk(fib->host_mu);
	// Submissions made before lane epochs existed do not filter lanes by
	// epoch: while one runs, a lane could become valid before its key slot
	// is written.  Epochs are created only with nothing in flight.
	if (!fib->slot_epoch) {
		if (fib->in_flight.load() > 0)
			return set_error(TM_ERR_BUSY, "the first live insert needs no submission in flight: use tm_fib_insert");
		TM_CHECK(create_epochs(fib));
	}
	tmx::live_plan * plan = nullptr;
	TM_CHECK(tmx::live_prepare_locked(fib, words, n, key_offsets, keys, filter_ids, &plan));
	const int rc = tmx::live_commit_locked(fib, plan);
	delete plan;
	return rc;
}

} // extern "C"

// ------------------------------------------------------------- FIB image

namespace {

constexpr char IMAGE_MAGIC[8] = {'T', 'M', 'F', 'I', 'B', 'I', 'M', 'G'};
constexpr uint32_t IMAGE_VERSION = 1;

struct image_header {
	char magic[8];
	uint32_t version, reserved;
	uint64_t n_parts, n_blocks, n_keys, n_filters, n_key_instances, max_part, reserve_pct, n_rich, n_rich_full;
};

template <typename T>
bool write_array(std::FILE * f, const T * p, uint64_t n) {
	return std::fwrite(&n, 8, 1, f) == 1 && (n == 0 || std::fwrite(p, sizeof(T), n, f) == n);
}

template <typename T>
int device_array_out(std::FILE * f, const T * d, uint64_t n) {
	std::vector<T> h(n);
	if (n) TM_HIP(hipMemcpy(h.data(), d, n * sizeof(T), hipMemcpyDeviceToHost));
	return write_array(f, h.data(), n) ? TM_OK : set_error(TM_ERR_INVALID, "image write failed");
}

// An array of exactly `want` elements.
template <typename T>
int read_array(std::FILE * f, std::vector<T> & v, uint64_t want, const char * what) {
	uint64_t n = 0;
	if (std::fread(&n, 8, 1, f) != 1) return set_error(TM_ERR_INVALID, "image truncated before %s", what);
	if (n != want)
		return set_error(TM_ERR_INVALID, "image %s holds %llu elements, header says %llu", what, (unsigned long long)n,
						 (unsigned long long)want);
	v.resize(n);
	if (n && std::fread(v.data(), sizeof(T), n, f) != n) return set_error(TM_ERR_INVALID, "image truncated in %s", what);
	return TM_OK;
}

// Consistency of a loaded image before it replaces the FIB: every index a
// kernel or a later insert follows must stay inside the arrays it indexes.
int check_image(uint64_t nb, uint64_t P, uint64_t K, const std::vector<uint32_t> & info,
				const std::vector<uint64_t> & skeys, const std::vector<uint32_t> & ovf,
				const std::vector<uint32_t> & dict, const std::vector<uint32_t> & pinfo,
				const std::vector<uint32_t> & psize, const std::vector<uint32_t> & pcap, const std::vector<int> & order) {
	std::vector<char> seen(192, 0);
	for (int b : order) {
		if (b < 0 || b >= 192 || seen[b]) return set_error(TM_ERR_INVALID, "image: bit order is not a permutation of 0..191");
		seen[b] = 1;
	}
	for (uint64_t i = 1; i < dict.size(); ++i)
		if (dict[i - 1] >= dict[i]) return set_error(TM_ERR_INVALID, "image: key dictionary is not strictly increasing");
	for (uint64_t p = 0; p < P; ++p) {
		const uint64_t first = pinfo[p * 8 + 6], used = pinfo[p * 8 + 7];
		if (used > pcap[p] || first + pcap[p] > nb || psize[p] > used * 64)
			return set_error(TM_ERR_INVALID, "image: partition %llu's blocks lie outside the table",
							 (unsigned long long)p);
	}
	for (uint64_t b = 0; b < nb; ++b) {
		const uint32_t valid = info[b * 8 + 6], part = info[b * 8 + 7];
		if (valid > 64 || (valid && part >= P))
			return set_error(TM_ERR_INVALID, "image: block %llu has %u lanes or partition %u", (unsigned long long)b,
							 valid, part);
		for (uint32_t l = 0; l < valid; ++l) {
			const uint64_t kv = skeys[b * 64 + l];
			if ((kv & 0xFFFF) == KEY_NONE) {
				const uint64_t at = (kv >> 16) & 0xFFFFFFFFull;
				if (at >= ovf.size() || at + 1 + ovf[at] > ovf.size())
					return set_error(TM_ERR_INVALID, "image: block %llu lane %u: overflow list out of range",
									 (unsigned long long)b, l);
				for (uint64_t j = 0; j < ovf[at]; ++j)
					if (ovf[at + 1 + j] >= K) return set_error(TM_ERR_INVALID, "image: key id out of range");
			} else {
				for (int j = 0; j < 4; ++j) {
					const uint32_t id = (uint32_t)(kv >> (16 * j)) & 0xFFFF;
					if (id != KEY_NONE && id >= K) return set_error(TM_ERR_INVALID, "image: key id out of range");
				}
			}
		}
	}
	return TM_OK;
}

} // namespace

extern "C" {

int tm_fib_save(tm_fib * fib, const char * path) {
	if (!fib || !path) return set_error(TM_ERR_INVALID, "NULL argument");
	if (fib->h_order.size() != 192) return set_error(TM_ERR_STATE, "no table uploaded");
	TM_CHECK(require_device());
	std::lock_guard<std::mutex> lk(fib->host_mu);
	if (fib->prep_epoch != fib->epoch) return set_error(TM_ERR_BUSY, "a prepared live insert is not committed yet");
	const size_t P0 = fib->n_parts;
	if (fib->h_part_info.size() != P0 * 8 || fib->h_part_stat.size() != P0 * 8 || fib->h_part_size.size() != P0 ||
		fib->h_part_cap.size() != P0 || fib->h_dict.size() != fib->n_keys || fib->h_ovf.size() < fib->ovf_valid)
		return set_error(TM_ERR_STATE, "host copies of the table are inconsistent");
	std::FILE * f = std::fopen(path, "wb");
	if (!f) return set_error(TM_ERR_INVALID, "cannot create %s", path);
	image_header h;
	std::memset(&h, 0, sizeof(h));
	std::memcpy(h.magic, IMAGE_MAGIC, 8);
	h.version = IMAGE_VERSION;
	h.n_parts = fib->n_parts;
	h.n_blocks = fib->n_blocks;
	h.n_keys = fib->n_keys;
	h.n_filters = fib->n_filters;
	h.n_key_instances = fib->n_key_instances;
	h.max_part = fib->max_part;
	h.reserve_pct = fib->reserve_pct;
	h.n_rich = fib->n_rich;
	h.n_rich_full = fib->n_rich_full;
	const uint64_t nb = fib->n_blocks, P = fib->n_parts, kw = (fib->n_keys + 31) / 32;
	int rc = std::fwrite(&h, sizeof(h), 1, f) == 1 ? TM_OK : set_error(TM_ERR_INVALID, "image write failed");
	if (rc == TM_OK) rc = device_array_out(f, fib->blk_words, nb * BLOCK_WORDS);
	if (rc == TM_OK) rc = device_array_out(f, fib->blk_info, nb * 8);
	if (rc == TM_OK) rc = device_array_out(f, fib->slot_keys, nb * 64);
	if (rc == TM_OK) rc = device_array_out(f, fib->slot_fid, nb * 64);
	if (rc == TM_OK) rc = device_array_out(f, fib->d_rich_words, (uint64_t)fib->n_rich * 8);
	if (rc == TM_OK) rc = device_array_out(f, fib->d_rich_bm, (uint64_t)fib->n_rich * kw);
	bool ok = rc == TM_OK;
	ok = ok && write_array(f, fib->h_ovf.data(), fib->h_ovf.size());
	ok = ok && write_array(f, fib->h_dict.data(), fib->h_dict.size());
	ok = ok && write_array(f, fib->h_part_info.data(), P * 8);
	ok = ok && write_array(f, fib->h_part_stat.data(), P * 8);
	ok = ok && write_array(f, fib->h_part_size.data(), P);
	ok = ok && write_array(f, fib->h_part_cap.data(), P);
	ok = ok && write_array(f, fib->h_order.data(), (uint64_t)fib->h_order.size());
	if (std::fclose(f) != 0) ok = false;
	if (rc == TM_OK && !ok) rc = set_error(TM_ERR_INVALID, "image write failed: %s", path);
	return rc;
}

int tm_fib_load(tm_fib * fib, const char * path) {
	if (!fib || !path) return set_error(TM_ERR_INVALID, "NULL argument");
	std::lock_guard<std::mutex> lk(fib->host_mu);
	if (fib->in_flight.load() > 0) return set_error(TM_ERR_BUSY, "a ring on this FIB has a submission in flight");
	if (fib->prep_epoch != fib->epoch) return set_error(TM_ERR_BUSY, "a prepared live insert is not committed yet");
	TM_CHECK(require_device());
	std::FILE * f = std::fopen(path, "rb");
	if (!f) return set_error(TM_ERR_INVALID, "cannot open %s", path);
	struct closer {
		std::FILE * f;
		~closer() { std::fclose(f); }
	} c{f};
	image_header h;
	if (std::fread(&h, sizeof(h), 1, f) != 1 || std::memcmp(h.magic, IMAGE_MAGIC, 8) != 0)
		return set_error(TM_ERR_INVALID, "%s is not a FIB image", path);
	if (h.version != IMAGE_VERSION)
		return set_error(TM_ERR_INVALID, "%s: image version %u, this library reads %u", path, h.version, IMAGE_VERSION);
	if (h.n_parts > 0xFFFFFFFFull || h.n_blocks > 0xFFFFFFFFull || h.n_keys > 0xFFFFFFFFull || h.n_rich > MAX_RICH ||
		h.n_rich_full > h.n_rich)
		return set_error(TM_ERR_INVALID, "%s: sizes out of range", path);
	const uint64_t nb = h.n_blocks, P = h.n_parts, kw = (h.n_keys + 31) / 32;
	std::vector<uint32_t> words, info, fid, rich_w, rich_bm, ovf, dict, pinfo, pstat, psize, pcap;
	std::vector<uint64_t> skeys;
	std::vector<int> order;
	TM_CHECK(read_array(f, words, nb * BLOCK_WORDS, "block words"));
	TM_CHECK(read_array(f, info, nb * 8, "block info"));
	TM_CHECK(read_array(f, skeys, nb * 64, "key slots"));
	TM_CHECK(read_array(f, fid, nb * 64, "filter ids"));
	TM_CHECK(read_array(f, rich_w, h.n_rich * 8, "rich filters"));
	TM_CHECK(read_array(f, rich_bm, h.n_rich * kw, "rich key bitmaps"));
	uint64_t n_ovf = 0;
	{   // overflow lists: any length
		const long at = std::ftell(f);
		if (std::fread(&n_ovf, 8, 1, f) != 1 || std::fseek(f, at, SEEK_SET) != 0)
			return set_error(TM_ERR_INVALID, "image truncated before overflow lists");
		if (n_ovf == 0 || n_ovf > 0xFFFFFFFFull) return set_error(TM_ERR_INVALID, "%s: overflow lists out of range", path);
	}
	TM_CHECK(read_array(f, ovf, n_ovf, "overflow lists"));
	TM_CHECK(read_array(f, dict, h.n_keys, "key dictionary"));
	TM_CHECK(read_array(f, pinfo, P * 8, "partition info"));
	TM_CHECK(read_array(f, pstat, P * 8, "partition statistics"));
	TM_CHECK(read_array(f, psize, P, "partition sizes"));
	TM_CHECK(read_array(f, pcap, P, "partition capacities"));
	TM_CHECK(read_array(f, order, 192, "bit order"));
	TM_CHECK(check_image(nb, P, h.n_keys, info, skeys, ovf, dict, pinfo, psize, pcap, order));

	fib->release();
	uint64_t bytes = 0;
	TM_CHECK(upload(&fib->blk_words, words, bytes));
	TM_CHECK(upload(&fib->blk_info, info, bytes));
	TM_CHECK(upload(&fib->slot_keys, skeys, bytes));
	TM_CHECK(upload(&fib->slot_fid, fid, bytes));
	fib->ovf_cap = std::max<size_t>(ovf.size() * 2, 1024);
	TM_CHECK(dalloc(&fib->ovf, fib->ovf_cap));
	TM_CHECK(put(fib->ovf, ovf));
	bytes += fib->ovf_cap * 4;
	TM_CHECK(upload(&fib->key_dict, dict, bytes));
	TM_CHECK(upload(&fib->part_info, pinfo, bytes));
	TM_CHECK(upload(&fib->part_stat, pstat, bytes));
	if (h.n_rich) {
		TM_CHECK(upload(&fib->d_rich_words, rich_w, bytes));
		TM_CHECK(upload(&fib->d_rich_bm, rich_bm, bytes));
	}
	fib->n_rich = (uint32_t)h.n_rich;
	fib->n_rich_full = (uint32_t)h.n_rich_full;
	fib->n_parts = (uint32_t)P;
	fib->n_blocks = (uint32_t)nb;
	{
		bool fresh;
		TM_CHECK(refresh_sup(fib, fresh));
		bytes += fib->sup_groups * 32;
	}
	fib->n_keys = (uint32_t)h.n_keys;
	fib->n_filters = h.n_filters;
	fib->n_key_instances = h.n_key_instances;
	fib->max_part = h.max_part;
	fib->reserve_pct = (uint32_t)h.reserve_pct;
	fib->device_bytes = bytes;
	fib->h_part_info = std::move(pinfo);
	fib->h_part_stat = std::move(pstat);
	fib->h_part_size = std::move(psize);
	fib->h_part_cap = std::move(pcap);
	fib->h_ovf = std::move(ovf);
	fib->h_dict = std::move(dict);
	fib->h_order = std::move(order);
	fib->ovf_valid = (uint32_t)fib->h_ovf.size();
	fib->prep_epoch = fib->epoch;
	fib->version++;
	return TM_OK;
}

} // extern "C"
