// Test infrastructure: drives the reference's compact Patricia-trie matcher
// (compact_patricia_predicate.hh, key_array.cc) as a CPU baseline and as a
// second checker of C2 results.  Compiled by `make ref` into
// _ref/libref_patricia.so from the reference's own sources; not part of the
// product.
//
// The trie is built as compact_patricia_matcher.cc:34-57 builds it (one
// add() per filter, its keys appended to the mapped key_array, then
// consolidate()), except that filters are handed over in the trie's final
// order through use_pre_sorted_filters (compact_patricia_predicate.hh:149-155)
// so that no std::map of the whole table is built: the order is the
// reference's own filter_t::operator< (bitvector.hh:389), checked here, and
// identical filters merge their keys as the map would.
//
// A query's result is match_unique's (tagmatch_query.hh:99-119): the keys of
// every filter that is a subset of it, sorted, without duplicates.  The
// reference driver's result collector (compact_patricia_matcher.cc:165-183)
// indexes an empty vector, so its own result path is not used.
#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "filter.hh"
#include "key_array.hh"
#include "compact_patricia_predicate.hh"

namespace {

typedef compact_patricia_predicate<key_array> predicate;

filter_t to_filter(const uint32_t * w) {
	block_t b[filter_t::WIDTH / 64];
	static_assert(sizeof(b) == 24, "192-bit filters");
	std::memcpy(b, w, sizeof(b));   // little-endian words, as bitvector::read_binary
	return filter_t(b);
}

class collect : public predicate::match_handler {
public:
	std::vector<tagmatch_key_t> keys;
	bool match(const key_array & k) override {
		keys.insert(keys.end(), k.begin(), k.end());
		return false;
	}
	void finish() {
		std::sort(keys.begin(), keys.end());
		keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
	}
};

struct trie {
	predicate p;
	uint64_t n_nodes = 0;
};

template <typename F>
void parallel(uint64_t n, int threads, F && fn) {
	threads = std::max(1, threads);
	std::vector<std::thread> ts;
	for (int t = 0; t < threads; ++t)
		ts.emplace_back([&, t] { fn(n * t / threads, n * (t + 1) / threads); });
	for (auto & t : ts) t.join();
}

} // namespace

extern "C" {

// Builds the trie over n filters (6 LE u32 words each) and their keys
// (key_off[n + 1] offsets into keys).  Returns NULL on an empty table or a
// filter with more keys than a key_array holds.
void * rp_build(const uint32_t * words, uint64_t n, const uint64_t * key_off, const uint32_t * keys) {
	if (n == 0 || n > 0xFFFFFFFFull) return nullptr;
	std::vector<filter_t> f(n);
	for (uint64_t i = 0; i < n; ++i) f[i] = to_filter(words + 6 * i);
	std::vector<uint32_t> order(n);
	for (uint64_t i = 0; i < n; ++i) order[i] = (uint32_t)i;
	std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return f[a] < f[b]; });
	uint64_t distinct = 0;
	for (uint64_t i = 0; i < n; ++i)
		if (i == 0 || f[order[i - 1]] < f[order[i]]) ++distinct;
	trie * t = new trie;
	t->p.use_pre_sorted_filters((unsigned)distinct, false);
	key_array * cur = nullptr;
	for (uint64_t i = 0; i < n; ++i) {
		const uint32_t j = order[i];
		if (i == 0 || f[order[i - 1]] < f[j]) cur = &t->p.add(f[j]);
		const uint64_t cnt = key_off[j + 1] - key_off[j];
		if (cnt > 0xFFFF) {
			delete t;
			return nullptr;
		}
		for (uint64_t k = key_off[j]; k < key_off[j + 1]; ++k) cur->add(keys[k]);
	}
	t->p.consolidate();
	t->n_nodes = distinct;
	return t;
}

uint64_t rp_nodes(void * h) { return h ? static_cast<trie *>(h)->n_nodes : 0; }

// Matches nq queries on `threads` threads.  Per query, the number of
// distinct keys goes to counts[i]; when `out` is non-NULL the keys, sorted,
// go to out + off[i] where off is the exclusive prefix sum of the counts of
// an earlier call (two-pass use).  Returns the total number of keys.
uint64_t rp_match_unique(void * h, const uint32_t * q, uint64_t nq, int threads, uint32_t * counts,
						 const uint64_t * off, uint32_t * out) {
	trie * t = static_cast<trie *>(h);
	std::atomic<uint64_t> total{0};
	parallel(nq, threads, [&](uint64_t b, uint64_t e) {
		collect c;
		uint64_t s = 0;
		for (uint64_t i = b; i < e; ++i) {
			c.keys.clear();
			t->p.find_all_subsets(to_filter(q + 6 * i), c);
			c.finish();
			counts[i] = (uint32_t)c.keys.size();
			s += c.keys.size();
			if (out) std::copy(c.keys.begin(), c.keys.end(), out + off[i]);
		}
		total += s;
	});
	return total.load();
}

void rp_free(void * h) { delete static_cast<trie *>(h); }

} // extern "C"
