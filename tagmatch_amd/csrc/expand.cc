// Host expansion of KEY_BITMAP rows into KEYS_UNIQUE lists.
//
// A row holds one bit per key of the FIB's sorted dictionary; the query's
// match_unique result (reference tagmatch_query.hh:99-119: keys collected by
// add_output, sorted, duplicates removed) is the dictionary keys of its set
// bits, in increasing bit order.  Queries are cut into contiguous ranges,
// one per thread: a pass of popcounts gives the offsets, a second pass
// writes each range's keys.
#include "../../include/tagmatch_amd.h"

#include <algorithm>
#include <cstring>
#include <thread>
#include <vector>

namespace tmx {
int set_error(int code, const char * fmt, ...);   // text for tm_last_error()
}

namespace {

uint32_t row_count(const uint32_t * row, uint32_t kw_used, uint32_t last_mask) {
	uint32_t c = 0;
	for (uint32_t w = 0; w + 1 < kw_used; ++w) c += (uint32_t)__builtin_popcount(row[w]);
	if (kw_used) c += (uint32_t)__builtin_popcount(row[kw_used - 1] & last_mask);
	return c;
}

void row_write(const uint32_t * row, uint32_t kw_used, uint32_t last_mask, const uint32_t * dict, uint32_t * out) {
	for (uint32_t w = 0; w < kw_used; ++w) {
		uint32_t bits = row[w] & (w + 1 == kw_used ? last_mask : 0xFFFFFFFFu);
		if (bits == 0xFFFFFFFFu) {   // a full word: 32 consecutive dictionary keys
			std::memcpy(out, dict + (size_t)w * 32, 32 * 4);
			out += 32;
			continue;
		}
		while (bits) {
			*out++ = dict[(size_t)w * 32 + (uint32_t)__builtin_ctz(bits)];
			bits &= bits - 1;
		}
	}
}

template <typename F>
void parallel_ranges(uint64_t n, uint32_t threads, F fn) {
	const uint32_t T = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(threads, (n + 1023) / 1024));
	if (T == 1) {
		fn(0, n);
		return;
	}
	std::vector<std::thread> th;
	th.reserve(T);
	for (uint32_t t = 0; t < T; ++t) th.emplace_back(fn, n * t / T, n * (t + 1) / T);
	for (auto & x : th) x.join();
}

// Offsets, then values, of nq queries whose rows row_at(q) gives (NULL: the
// whole dictionary).
template <typename R>
int expand(R row_at, uint64_t nq, uint32_t kw, const uint32_t * dict, uint32_t n_keys, uint64_t * offsets,
		   uint32_t * values, uint64_t capacity, uint32_t threads) {
	const uint32_t kw_used = std::min<uint32_t>(kw, (n_keys + 31) / 32);
	const uint32_t last_mask = (n_keys & 31) ? (1u << (n_keys & 31)) - 1u : 0xFFFFFFFFu;
	if (threads == 0) threads = std::max(1u, std::thread::hardware_concurrency());
	offsets[0] = 0;
	parallel_ranges(nq, threads, [&](uint64_t a, uint64_t b) {
		for (uint64_t q = a; q < b; ++q) {
			const uint32_t * row = row_at(q);
			offsets[q + 1] = row ? row_count(row, kw_used, last_mask) : n_keys;
		}
	});
	for (uint64_t q = 0; q < nq; ++q) offsets[q + 1] += offsets[q];
	if (!values) return TM_OK;
	if (offsets[nq] > capacity)
		return tmx::set_error(TM_ERR_OVERFLOW, "%llu keys, capacity %llu", (unsigned long long)offsets[nq],
							  (unsigned long long)capacity);
	parallel_ranges(nq, threads, [&](uint64_t a, uint64_t b) {
		for (uint64_t q = a; q < b; ++q) {
			const uint32_t * row = row_at(q);
			if (row)
				row_write(row, kw_used, last_mask, dict, values + offsets[q]);
			else if (n_keys)
				std::memcpy(values + offsets[q], dict, (size_t)n_keys * 4);
		}
	});
	return TM_OK;
}

} // namespace

extern "C" int tm_expand_key_rows_packed(const uint32_t * row_of, const uint32_t * rows, uint64_t n_rows, uint64_t nq,
										 uint32_t kw, const uint32_t * dict, uint32_t n_keys, uint64_t * offsets,
										 uint32_t * values, uint64_t capacity, uint32_t threads) {
	if (!offsets) return tmx::set_error(TM_ERR_INVALID, "offsets is NULL");
	if (nq && (!row_of || (n_keys && !dict))) return tmx::set_error(TM_ERR_INVALID, "row_of or dictionary is NULL");
	if (n_rows && !rows) return tmx::set_error(TM_ERR_INVALID, "rows is NULL");
	for (uint64_t q = 0; q < nq; ++q)
		if (row_of[q] != TM_ROW_FULL && row_of[q] >= n_rows)
			return tmx::set_error(TM_ERR_INVALID, "query %llu: row %u of %llu", (unsigned long long)q, row_of[q],
								  (unsigned long long)n_rows);
	return expand([&](uint64_t q) { return row_of[q] == TM_ROW_FULL ? nullptr : rows + (uint64_t)row_of[q] * kw; }, nq,
				  kw, dict, n_keys, offsets, values, capacity, threads);
}

extern "C" int tm_expand_key_bitmaps(const uint32_t * rows, uint64_t nq, uint32_t kw, const uint32_t * dict,
									 uint32_t n_keys, uint64_t * offsets, uint32_t * values, uint64_t capacity,
									 uint32_t threads) {
	if (!offsets) return tmx::set_error(TM_ERR_INVALID, "offsets is NULL");
	if (nq && (!rows || (n_keys && !dict))) return tmx::set_error(TM_ERR_INVALID, "rows or dictionary is NULL");
	return expand([&](uint64_t q) { return rows + q * kw; }, nq, kw, dict, n_keys, offsets, values, capacity,
				  threads);
}
