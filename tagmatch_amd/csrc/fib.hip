// FIB compilation and upload: the host-side layout pass that replaces the
// reference's back_end::compile_fibs (reference back_end.cc:404-494).
//
// Differences from the reference, by design:
//   * one structure-of-arrays table for all partitions, 64 filters per
//     block (one per lane of a wave64), word planes contiguous per block;
//   * per-block AND masks, used by the match kernel to discard whole blocks
//     for a tile of queries before testing single filters;
//   * keys resident in HBM as dense ids (reference keeps them on the host,
//     back_end.cc:109-129, and expands hits there, back_end.cc:329-343);
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
	n_parts = n_blocks = n_keys = 0;
	n_filters = n_key_instances = device_bytes = max_part = 0;
	h_part_info.clear();
	h_part_size.clear();
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
	fib->release();
	delete fib;
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

int tm_fib_upload(tm_fib * fib, const uint32_t * words, uint32_t n,
				  const uint32_t * part_offsets, const uint32_t * part_masks, uint32_t n_parts,
				  const uint32_t * key_offsets, const uint32_t * keys,
				  const uint32_t * filter_ids) {
	if (!fib) return set_error(TM_ERR_INVALID, "fib is NULL");
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

	// Every partition mask must be a subset of each member filter, or the
	// lookup would drop matches (reference front_end.cc:465-477 relies on the
	// same property, partitioning.cc:412-423).
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

	// Per partition: member order and block count.
	std::vector<uint32_t> pblock(n_parts + 1, 0);
	uint64_t maxp = 0;
	for (uint32_t p = 0; p < n_parts; ++p) {
		uint32_t sz = part_offsets[p + 1] - part_offsets[p];
		maxp = std::max<uint64_t>(maxp, sz);
		pblock[p + 1] = pblock[p] + (sz + 63) / 64;
	}
	const uint32_t nb = pblock[n_parts];
	std::vector<uint32_t> member(n);  // member[k] = input index at partitioned position k
	std::iota(member.begin(), member.end(), 0u);
	parallel_ranges(n_parts, [&](size_t pb, size_t pe) {
		std::vector<std::array<uint64_t, 3>> key;
		for (size_t p = pb; p < pe; ++p) {
			uint32_t b = part_offsets[p], e = part_offsets[p + 1];
			if (e - b < 2) continue;
			key.resize(e - b);
			for (uint32_t i = b; i < e; ++i) {
				std::array<uint64_t, 3> k{0, 0, 0};
				const uint32_t * w = words + (size_t)i * 6;
				for (int r = 0; r < 192; ++r) {
					int bit = order[r];
					if ((w[bit >> 5] >> (bit & 31)) & 1)
						k[r >> 6] |= 1ULL << (63 - (r & 63));
				}
				key[i - b] = k;
			}
			std::stable_sort(member.begin() + b, member.begin() + e, [&](uint32_t x, uint32_t y) {
				return key[x - b] < key[y - b];
			});
		}
	});

	std::vector<uint32_t> h_words((size_t)nb * BLOCK_WORDS, 0);
	std::vector<uint32_t> h_info((size_t)nb * 8, 0);
	std::vector<uint64_t> h_keys((size_t)nb * 64, ~0ULL);
	std::vector<uint32_t> h_fid((size_t)nb * 64, 0xFFFFFFFFu);
	std::vector<uint32_t> h_part((size_t)n_parts * 8, 0);
	std::vector<uint32_t> h_psize(n_parts, 0);
	std::vector<uint32_t> h_pstat((size_t)n_parts * 8, 0);
	// overflow key lists (count, ids...) for filters with more than 4 keys
	std::vector<uint32_t> h_ovf;
	h_ovf.push_back(0);

	auto key_id = [&](uint32_t k) -> uint32_t {
		return (uint32_t)(std::lower_bound(dict.begin(), dict.end(), k) - dict.begin());
	};

	for (uint32_t p = 0; p < n_parts; ++p) {
		uint32_t b0 = part_offsets[p], sz = part_offsets[p + 1] - b0;
		for (int j = 0; j < 6; ++j) h_part[(size_t)p * 8 + j] = part_masks[(size_t)p * 6 + j];
		h_part[(size_t)p * 8 + 6] = pblock[p];
		h_part[(size_t)p * 8 + 7] = pblock[p + 1] - pblock[p];
		h_psize[p] = sz;
		for (uint32_t k = 0; k < sz; ++k) {
			uint32_t i = member[b0 + k];
			uint32_t blk = pblock[p] + k / 64, lane = k % 64;
			const uint32_t * w = words + (size_t)i * 6;
			uint32_t * info = &h_info[(size_t)blk * 8];
			if (lane == 0)
				for (int j = 0; j < 6; ++j) info[j] = 0xFFFFFFFFu;
			for (int j = 0; j < 6; ++j) {
				h_words[(size_t)blk * BLOCK_WORDS + j * 64 + lane] = w[j];
				info[j] &= w[j];
			}
			info[6] = lane + 1;
			info[7] = p;
			uint32_t * ps = &h_pstat[(size_t)p * 8];
			for (int j = 0; j < 6; ++j) {
				ps[j] |= w[j];
				ps[7] += (uint32_t)__builtin_popcount(w[j]);
			}
			ps[6] += 1;
			h_fid[(size_t)blk * 64 + lane] = filter_ids ? filter_ids[i] : i;
			uint32_t kb = key_offsets[i], ke = key_offsets[i + 1], kn = ke - kb;
			if (inline_ok && kn <= 4) {
				uint64_t v = ~0ULL;
				for (uint32_t x = 0; x < kn; ++x) {
					uint64_t id = key_id(keys[kb + x]);
					v &= ~(0xFFFFULL << (16 * x));
					v |= id << (16 * x);
				}
				h_keys[(size_t)blk * 64 + lane] = v;
			} else {
				uint64_t off = h_ovf.size();
				if (off > 0xFFFFFFFFULL) return set_error(TM_ERR_INVALID, "overflow key table too large");
				h_ovf.push_back(kn);
				for (uint32_t x = 0; x < kn; ++x) h_ovf.push_back(key_id(keys[kb + x]));
				h_keys[(size_t)blk * 64 + lane] = 0xFFFFULL | (off << 16);
			}
		}
	}

	fib->release();
	uint64_t bytes = 0;
	TM_CHECK(upload(&fib->blk_words, h_words, bytes));
	TM_CHECK(upload(&fib->blk_info, h_info, bytes));
	TM_CHECK(upload(&fib->slot_keys, h_keys, bytes));
	TM_CHECK(upload(&fib->slot_fid, h_fid, bytes));
	TM_CHECK(upload(&fib->ovf, h_ovf, bytes));
	TM_CHECK(upload(&fib->key_dict, dict, bytes));
	TM_CHECK(upload(&fib->part_info, h_part, bytes));
	TM_CHECK(upload(&fib->part_stat, h_pstat, bytes));
	fib->n_parts = n_parts;
	fib->n_blocks = nb;
	fib->n_keys = K;
	fib->n_filters = n;
	fib->n_key_instances = nkeys;
	fib->device_bytes = bytes;
	fib->max_part = maxp;
	fib->h_part_info = std::move(h_part);
	fib->h_part_size = std::move(h_psize);
	return TM_OK;
}

} // extern "C"
