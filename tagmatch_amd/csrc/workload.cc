// Seeded synthetic workload generator for TagMatch-style subset matching.
//
// Filters and queries are 192-bit Bloom filters of tag sets, stored as six
// little-endian 32-bit words: bit p of the filter is bit (p % 32) of word
// (p / 32).  This is the in-memory layout of the reference's bitvector<192>
// read as uint32 words (reference bitvector.hh:104-139, :210-212).
//
// Every random draw is a pure function of (seed, stream, item index, draw
// index) built from splitmix64, so any slice of the workload can be produced
// independently, by any number of threads, with identical results.
//
//   vocabulary : n_tags tags named "t%07u"; tag rank r has Zipf weight
//                1/(r+1)^s; sampled by inverse CDF with a guide table.
//   tag -> bits: h = splitmix64(fnv1a64(name) ^ seed); bit j of the tag is
//                splitmix64(h + j * 0x9E3779B97F4A7C15) mod 192, j < k.
//   filter i   : U{ftag_min..ftag_max} distinct tags, U{key_min..key_max}
//                distinct keys, key = (tree << 16) | iface with tree
//                U{0..trees-1} and iface U{0..ifaces-1}.
//   query j    : U{qtag_min..qtag_max} distinct tags; with probability
//                superset_frac it starts from the tags of a random filter
//                and adds Zipf tags up to its count.
#include <cstdint>
#include <cstring>
#include <cmath>
#include <cstdio>
#include <vector>
#include <thread>
#include <algorithm>
#include <functional>

#include "../../include/tagmatch_amd.h"

namespace {

inline uint64_t splitmix64(uint64_t x) {
	uint64_t z = x + 0x9E3779B97F4A7C15ULL;
	z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
	z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
	return z ^ (z >> 31);
}

inline uint64_t fnv1a64(const char * s, size_t n) {
	uint64_t h = 0xcbf29ce484222325ULL;
	for (size_t i = 0; i < n; ++i) {
		h ^= (unsigned char)s[i];
		h *= 0x100000001b3ULL;
	}
	return h;
}

// Sub-streams of one item's draws.
enum : uint64_t { SUB_COUNT = 0, SUB_TAG = 1, SUB_KEY = 2, SUB_PICK = 3, SUB_EXTRA = 4 };
enum : uint64_t { STREAM_FILTER = 0x66696c74ULL, STREAM_QUERY = 0x71756572ULL };

inline uint64_t draw(uint64_t base, uint64_t item, uint64_t sub, uint64_t k) {
	return splitmix64(splitmix64(base + item) ^ ((sub << 40) + k));
}

inline double unit(uint64_t r) { return (double)(r >> 11) * (1.0 / 9007199254740992.0); }

inline uint32_t uniform_range(uint64_t r, uint32_t lo, uint32_t hi) {
	return lo + (uint32_t)(r % (uint64_t)(hi - lo + 1));
}

} // namespace

struct tm_gen {
	tm_workload_params p;
	uint64_t filter_base, query_base;
	std::vector<double> cdf;        // cumulative Zipf weights
	std::vector<uint32_t> guide;    // guide[g] = first rank with cdf > g/G
	std::vector<uint8_t> tag_bits;  // n_tags * k bit positions

	uint32_t sample_tag(uint64_t r) const {
		double total = cdf.back();
		double u = unit(r) * total;
		size_t G = guide.size() - 1;
		size_t g = (size_t)(unit(r) * (double)G);
		if (g >= G) g = G - 1;
		uint32_t lo = guide[g], hi = guide[g + 1];
		// first rank in [lo, hi] with cdf[rank] > u
		while (lo < hi) {
			uint32_t mid = lo + (hi - lo) / 2;
			if (cdf[mid] > u) hi = mid; else lo = mid + 1;
		}
		if (lo >= p.n_tags) lo = p.n_tags - 1;
		return lo;
	}

	void add_tag(uint32_t tag, uint32_t * w) const {
		const uint8_t * b = &tag_bits[(size_t)tag * p.bits_per_tag];
		for (uint32_t j = 0; j < p.bits_per_tag; ++j)
			w[b[j] >> 5] |= 1u << (b[j] & 31);
	}

	// Tags of filter i, in draw order; returns the count.
	uint32_t filter_tags(uint64_t i, uint32_t * tags) const {
		uint32_t n = uniform_range(draw(filter_base, i, SUB_COUNT, 0), p.ftag_min, p.ftag_max);
		uint32_t have = 0;
		for (uint64_t d = 0; have < n; ++d) {
			uint32_t t = sample_tag(draw(filter_base, i, SUB_TAG, d));
			bool dup = false;
			for (uint32_t x = 0; x < have; ++x) dup |= (tags[x] == t);
			if (!dup) tags[have++] = t;
			if (d > 100000) break;  // vocabulary smaller than the tag count
		}
		return have;
	}
};

extern "C" {

void tm_workload_default_params(tm_workload_params * p) {
	std::memset(p, 0, sizeof(*p));
	p->seed = 0x7461676D61746301ULL;
	p->n_tags = 1000000;
	p->zipf_s = 1.0;
	p->bits_per_tag = 7;
	p->ftag_min = 1;  p->ftag_max = 10;
	p->key_min = 1;   p->key_max = 4;
	p->trees = 8;     p->ifaces = 256;
	p->qtag_min = 10; p->qtag_max = 30;
	p->superset_frac = 0.3;
}

tm_gen * tm_gen_create(const tm_workload_params * params) {
	if (!params || params->n_tags == 0 || params->bits_per_tag == 0 || params->bits_per_tag > 64
		|| params->ftag_min > params->ftag_max || params->key_min > params->key_max
		|| params->qtag_min > params->qtag_max || params->trees == 0 || params->ifaces == 0
		|| params->trees > 65536 || params->ifaces > 65536 || params->key_max > TM_GEN_MAX_KEYS
		|| params->qtag_max > TM_GEN_MAX_TAGS || params->ftag_max > TM_GEN_MAX_TAGS)
		return nullptr;
	tm_gen * g = new tm_gen();
	g->p = *params;
	g->filter_base = splitmix64(params->seed ^ STREAM_FILTER);
	g->query_base = splitmix64(params->seed ^ STREAM_QUERY);
	uint32_t n = params->n_tags;
	g->cdf.resize(n);
	double acc = 0.0;
	for (uint32_t r = 0; r < n; ++r) {
		acc += 1.0 / std::pow((double)(r + 1), params->zipf_s);
		g->cdf[r] = acc;
	}
	size_t G = std::max<size_t>(1, std::min<size_t>(n, 1u << 16));
	g->guide.resize(G + 1);
	uint32_t r = 0;
	for (size_t k = 0; k <= G; ++k) {
		double thr = acc * (double)k / (double)G;
		while (r < n - 1 && g->cdf[r] <= thr) ++r;
		g->guide[k] = r;
	}
	g->guide[G] = n - 1;
	g->tag_bits.resize((size_t)n * params->bits_per_tag);
	char name[32];
	for (uint32_t t = 0; t < n; ++t) {
		int len = std::snprintf(name, sizeof(name), "t%07u", t);
		uint64_t h = splitmix64(fnv1a64(name, (size_t)len) ^ params->seed);
		for (uint32_t j = 0; j < params->bits_per_tag; ++j)
			g->tag_bits[(size_t)t * params->bits_per_tag + j]
				= (uint8_t)(splitmix64(h + (uint64_t)j * 0x9E3779B97F4A7C15ULL) % 192);
	}
	return g;
}

void tm_gen_destroy(tm_gen * g) { delete g; }

int tm_gen_tag_bits(const tm_gen * g, uint32_t tag, uint8_t * bits) {
	if (!g || tag >= g->p.n_tags) return TM_ERR_INVALID;
	std::memcpy(bits, &g->tag_bits[(size_t)tag * g->p.bits_per_tag], g->p.bits_per_tag);
	return TM_OK;
}

static void parallel_for(uint64_t n, int threads, const std::function<void(uint64_t, uint64_t)> & fn);

int tm_gen_filters(const tm_gen * g, uint64_t first, uint64_t n, uint32_t * words,
				   uint32_t * key_counts, uint32_t * keys, int threads) {
	if (!g || (n && (!words || !key_counts || !keys))) return TM_ERR_INVALID;
	const tm_workload_params & p = g->p;
	parallel_for(n, threads, [&](uint64_t b, uint64_t e) {
		uint32_t tags[TM_GEN_MAX_TAGS];
		for (uint64_t x = b; x < e; ++x) {
			uint64_t i = first + x;
			uint32_t * w = words + x * 6;
			std::memset(w, 0, 24);
			uint32_t nt = g->filter_tags(i, tags);
			for (uint32_t t = 0; t < nt; ++t) g->add_tag(tags[t], w);
			uint32_t nk = uniform_range(draw(g->filter_base, i, SUB_COUNT, 1), p.key_min, p.key_max);
			uint32_t * k = keys + x * p.key_max;
			uint32_t have = 0;
			for (uint64_t d = 0; have < nk && d < 100000; ++d) {
				uint64_t r = draw(g->filter_base, i, SUB_KEY, d);
				uint32_t key = ((uint32_t)(r % p.trees) << 16) | (uint32_t)((r >> 32) % p.ifaces);
				bool dup = false;
				for (uint32_t y = 0; y < have; ++y) dup |= (k[y] == key);
				if (!dup) k[have++] = key;
			}
			for (uint32_t y = have; y < p.key_max; ++y) k[y] = 0;
			key_counts[x] = have;
		}
	});
	return TM_OK;
}

int tm_gen_queries(const tm_gen * g, uint64_t n_filters, uint64_t first, uint64_t n,
				   uint32_t * words, uint8_t * is_superset, int threads) {
	if (!g || (n && !words)) return TM_ERR_INVALID;
	const tm_workload_params & p = g->p;
	parallel_for(n, threads, [&](uint64_t b, uint64_t e) {
		uint32_t tags[TM_GEN_MAX_TAGS];
		for (uint64_t x = b; x < e; ++x) {
			uint64_t j = first + x;
			uint32_t * w = words + x * 6;
			std::memset(w, 0, 24);
			uint32_t nt = uniform_range(draw(g->query_base, j, SUB_COUNT, 0), p.qtag_min, p.qtag_max);
			uint32_t have = 0;
			bool sup = n_filters > 0 && unit(draw(g->query_base, j, SUB_PICK, 0)) < p.superset_frac;
			if (sup) {
				uint64_t f = draw(g->query_base, j, SUB_PICK, 1) % n_filters;
				have = g->filter_tags(f, tags);
			}
			for (uint64_t d = 0; have < nt && d < 100000; ++d) {
				uint32_t t = g->sample_tag(draw(g->query_base, j, SUB_TAG, d));
				bool dup = false;
				for (uint32_t y = 0; y < have; ++y) dup |= (tags[y] == t);
				if (!dup) tags[have++] = t;
			}
			for (uint32_t t = 0; t < have; ++t) g->add_tag(tags[t], w);
			if (is_superset) is_superset[x] = sup ? 1 : 0;
		}
	});
	return TM_OK;
}

} // extern "C"

static void parallel_for(uint64_t n, int threads, const std::function<void(uint64_t, uint64_t)> & fn) {
	if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
	if (n < 4096 || threads == 1) { fn(0, n); return; }
	std::vector<std::thread> T;
	uint64_t chunk = (n + threads - 1) / threads;
	for (int t = 0; t < threads; ++t) {
		uint64_t b = t * chunk, e = std::min<uint64_t>(n, b + chunk);
		if (b >= e) break;
		T.emplace_back(fn, b, e);
	}
	for (auto & t : T) t.join();
}
