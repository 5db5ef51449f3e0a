// Reference-style tests of the C++ facade `class tagmatch`.
//
// Mirrors the cases of the reference's test/test_tagmatch.cc:191-638
// (lifecycle, find_one_filter, clear, add_same_filter, add_and_find_multi,
// add_and_clear, add_many_and_find, misc_subsets, misc_subsets_Q1,
// corner_case_filters, single_bit_filters, deepest_trie), plus checks of
// match() (keys with duplicates), remove(), all-zero filters and a random
// table against a brute-force scan.
//
// usage: tm_api_tests <vectors.txt>   (lines "F <bits>" and "Q <bits>")
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <random>
#include <set>
#include <sstream>
#include <string>
#include <vector>

#include "tagmatch.hh"
#include "synchronous_match_handler.hh"

using std::set;
using std::vector;

static int failures = 0;
#define CHECK(cond)                                                                \
	do {                                                                           \
		if (!(cond)) {                                                             \
			std::fprintf(stderr, "  FAIL %s:%d: %s\n", __FILE__, __LINE__, #cond); \
			++failures;                                                            \
		}                                                                          \
	} while (0)

static vector<filter_t> F, Q;
static filter_t ALL_ONES, ALL_ZEROS;

class general_matcher : public synchronous_match_handler {
public:
	set<tagmatch_key_t> keys;
	vector<tagmatch_key_t> raw;
	virtual void process_results(query * q) {
		for (tagmatch_key_t k : q->output_keys) keys.insert(k);
		raw = q->output_keys;
	}
};

class finder_matcher : public synchronous_match_handler {
public:
	tagmatch_key_t target;
	bool found = false;
	explicit finder_matcher(tagmatch_key_t t) : target(t) {}
	virtual void process_results(query * q) {
		for (tagmatch_key_t k : q->output_keys)
			if (k == target) found = true;
	}
};

// Submit queries (before or after start), check after stop.
class deferred_checks {
	struct check {
		tagmatch_query * q;
		general_matcher * h;
		set<tagmatch_key_t> expected;
	};
	vector<check> checks;

public:
	void run_query(const filter_t & f, const set<tagmatch_key_t> & s) {
		check c{new tagmatch_query(f), new general_matcher(), s};
		checks.push_back(c);
		tagmatch::match_unique(c.q, c.h);
	}
	void check_results() {
		for (size_t i = 0; i < checks.size(); ++i) {
			checks[i].h->synchronize_and_process();
			if (checks[i].h->keys != checks[i].expected) {
				std::fprintf(stderr, "  query %zu: unexpected results (%zu keys, expected %zu)\n", i,
							 checks[i].h->keys.size(), checks[i].expected.size());
				++failures;
			}
		}
	}
	~deferred_checks() {
		for (check & c : checks) {
			delete c.q;
			delete c.h;
		}
	}
};

static void init_test() {
	tagmatch::set_latency_limit_ms(200);
	tagmatch::clear();
}

static filter_t bits(std::initializer_list<unsigned> b) {
	filter_t f;
	for (unsigned x : b) f.set_bit(x);
	return f;
}

// ------------------------------------------------------------ cases

static void consolidate_empty() {
	tagmatch::set_latency_limit_ms(200);
	tagmatch::consolidate();
	tagmatch::clear();
}

static void start_and_stop() {
	tagmatch::consolidate();
	tagmatch::start();
	tagmatch::stop();
	tagmatch::start(4, 1);
	tagmatch::stop();
	tagmatch::clear();
	tagmatch::consolidate();
	tagmatch::start(4, 1);
	tagmatch::stop();
	tagmatch::start();
	tagmatch::stop();
	tagmatch::clear();
}

static void consolidate_no_match_action() {
	for (int r = 0; r < 2; ++r) {
		for (int i = 0; i < 5; ++i) tagmatch::add(F[i], i);
		tagmatch::consolidate();
		tagmatch::start();
		tagmatch::stop();
		tagmatch::start(4, 1);
		tagmatch::stop();
		tagmatch::clear();
	}
}

static void repeated_stop() {
	tagmatch::start();
	tagmatch::stop();
	tagmatch::stop();
	tagmatch::clear();
	tagmatch::clear();
	tagmatch::start();
	for (int i = 0; i < 8; ++i) tagmatch::stop();
	tagmatch::clear();
}

static void find_one_filter() {
	tagmatch::clear();
	tagmatch::add(F[1], 1);
	tagmatch::consolidate();
	tagmatch::start();
	finder_matcher f1(1), f2(1), f3(1), f4(1);
	tagmatch_query q1(F[1]), q2(F[2]), q3(ALL_ZEROS), q4(ALL_ONES);
	tagmatch::match_unique(&q1, &f1);
	tagmatch::match_unique(&q2, &f2);
	tagmatch::match_unique(&q3, &f3);
	tagmatch::match_unique(&q4, &f4);
	tagmatch::stop();
	f1.synchronize_and_process();
	f2.synchronize_and_process();
	f3.synchronize_and_process();
	f4.synchronize_and_process();
	CHECK(f1.found);
	CHECK(!f2.found);
	CHECK(!f3.found);
	CHECK(f4.found);
	tagmatch::clear();
}

static void clear_case() {
	tagmatch::clear();
	tagmatch::add(F[1], 1);
	tagmatch::consolidate();
	tagmatch::start();
	finder_matcher f1(1);
	tagmatch_query q1(F[1]);
	tagmatch::match_unique(&q1, &f1);
	tagmatch::stop();
	f1.synchronize_and_process();
	CHECK(f1.found);
	tagmatch::clear();
	tagmatch::consolidate();
	tagmatch::start();
	finder_matcher f2(1);
	tagmatch_query q2(F[1]);
	tagmatch::match_unique(&q2, &f2);
	tagmatch::stop();
	f2.synchronize_and_process();
	CHECK(!f2.found);
}

static void add_same_filter() {
	tagmatch::clear();
	tagmatch::add(F[1], 2);
	tagmatch::add(F[1], 3);
	tagmatch::add(F[1], 4);
	tagmatch::consolidate();
	tagmatch::start();
	general_matcher m1, m2;
	tagmatch_query q1(F[1]), q2(F[2]);
	tagmatch::match_unique(&q1, &m1);
	tagmatch::match_unique(&q2, &m2);
	tagmatch::stop();
	m1.synchronize_and_process();
	m2.synchronize_and_process();
	CHECK((m1.keys == set<tagmatch_key_t>{2, 3, 4}));
	CHECK(m2.keys.empty());
	tagmatch::clear();
}

static void add_and_find_multi() {
	tagmatch::set_latency_limit_ms(400);
	tagmatch::clear();
	tagmatch::add(F[1], 1);
	tagmatch::add(F[1], 2);
	tagmatch::add(F[2], 3);
	tagmatch::add(F[3], 4);
	tagmatch::consolidate();
	deferred_checks C;
	C.run_query(F[1], {1, 2});   // queued before start
	C.run_query(F[2], {3});
	C.run_query(F[3], {4});
	tagmatch::start();
	tagmatch::stop();
	C.check_results();
	tagmatch::clear();
}

static void add_and_find_multi2() {
	init_test();
	tagmatch::add(F[1], 1);
	tagmatch::add(F[1], 2);
	tagmatch::add(F[2], 3);
	tagmatch::add(F[3], 4);
	tagmatch::consolidate();
	tagmatch::start();
	deferred_checks C;
	C.run_query(F[1], {1, 2});
	C.run_query(F[2], {3});
	C.run_query(F[3], {4});
	tagmatch::stop();
	C.check_results();
	tagmatch::clear();
}

static void add_and_clear() {
	deferred_checks C;
	init_test();
	tagmatch::consolidate();
	tagmatch::start(4, 1);
	C.run_query(F[1], {});
	tagmatch::stop();
	tagmatch::add(F[1], 3);
	tagmatch::consolidate();
	tagmatch::start(4, 1);
	C.run_query(F[1], {3});
	tagmatch::stop();
	tagmatch::clear();
	tagmatch::consolidate();
	tagmatch::start(4, 1);
	C.run_query(F[1], {});
	tagmatch::stop();
	C.check_results();
	tagmatch::clear();
}

static void add_many_and_find() {
	init_test();
	deferred_checks C;
	for (int i = 1; i <= 10; ++i) tagmatch::add(F[i], i);
	tagmatch::consolidate();
	tagmatch::start(4, 1);
	for (int i = 1; i <= 10; ++i) C.run_query(F[i], {(tagmatch_key_t)i});
	tagmatch::stop();
	C.check_results();
	tagmatch::clear();
}

static void misc_subsets() {
	init_test();
	deferred_checks C;
	for (size_t i = 0; i < F.size(); ++i) tagmatch::add(F[i], i);
	tagmatch::consolidate();
	tagmatch::start(4, 1);
	for (size_t i = 0; i < Q.size(); ++i) {
		set<tagmatch_key_t> exp;
		for (size_t j = 0; j < F.size(); ++j)
			if (F[j].subset_of(Q[i])) exp.insert(j);
		C.run_query(Q[i], exp);
	}
	tagmatch::stop();
	C.check_results();
	tagmatch::clear();
}

static void misc_subsets_Q1() {
	init_test();
	for (size_t i = 0; i < F.size(); ++i) tagmatch::add(F[i], i);
	tagmatch::consolidate();
	tagmatch::start(4, 1);
	deferred_checks C;
	C.run_query(Q[1], {2});
	tagmatch::stop();
	C.check_results();
	tagmatch::clear();
}

static void corner_case_filters() {
	init_test();
	tagmatch::add(bits({0}), 1);
	tagmatch::add(bits({191}), 2);
	tagmatch::consolidate();
	tagmatch::start(4, 1);
	deferred_checks C;
	C.run_query(bits({1, 10, 100, 190}), {});
	C.run_query(bits({0, 10, 100, 190}), {1});
	C.run_query(bits({1, 10, 100, 191}), {2});
	C.run_query(bits({0, 10, 100, 191}), {1, 2});
	tagmatch::stop();
	C.check_results();
	tagmatch::clear();
}

static void single_bit_filters() {
	init_test();
	deferred_checks C;
	set<tagmatch_key_t> all;
	for (unsigned i = 0; i < 192; ++i) {
		tagmatch::add(bits({i}), i);
		all.insert(i);
	}
	tagmatch::consolidate();
	tagmatch::start(4, 1);
	C.run_query(ALL_ONES, all);
	tagmatch::stop();
	C.check_results();
	tagmatch::clear();
}

static void deepest_trie() {
	init_test();
	deferred_checks C;
	filter_t f;
	set<tagmatch_key_t> all;
	for (unsigned i = 0; i < 192; ++i) {
		f.set_bit(i);
		tagmatch::add(f, i);
		all.insert(i);
	}
	tagmatch::consolidate();
	tagmatch::start(4, 1);
	C.run_query(ALL_ONES, all);
	tagmatch::stop();
	C.check_results();
	tagmatch::clear();
}

// ------------------------------------------------------------ extensions

static void match_keeps_duplicates() {
	init_test();
	tagmatch::add(F[1], 7);
	tagmatch::add(F[1], 9);
	tagmatch::add(F[2], 7);
	tagmatch::consolidate();
	tagmatch::start();
	general_matcher m;
	tagmatch_query q(ALL_ONES);
	tagmatch::match(&q, &m);             // not unique: 7 appears twice
	tagmatch::stop();
	m.synchronize_and_process();
	vector<tagmatch_key_t> r = m.raw;
	std::sort(r.begin(), r.end());
	CHECK((r == vector<tagmatch_key_t>{7, 7, 9}));
	tagmatch::clear();
}

static void remove_keys() {
	init_test();
	tagmatch::add(F[1], {1, 2, 3});
	tagmatch::add(F[2], 5);
	tagmatch::consolidate();
	tagmatch::remove(F[1], 2);
	tagmatch::remove(F[2], 5);            // F[2] loses its only key: gone
	tagmatch::add(F[3], 8);
	tagmatch::consolidate();
	tagmatch::start();
	deferred_checks C;
	C.run_query(ALL_ONES, {1, 3, 8});
	C.run_query(F[2], {});
	tagmatch::stop();
	C.check_results();
	tagmatch::clear();
}

static void all_zero_filter_matches_everything() {
	init_test();
	tagmatch::add(ALL_ZEROS, 42);
	tagmatch::add(F[4], 4);
	tagmatch::consolidate();
	tagmatch::start();
	deferred_checks C;
	C.run_query(ALL_ZEROS, {42});
	C.run_query(F[4], {4, 42});
	C.run_query(ALL_ONES, {4, 42});
	tagmatch::stop();
	C.check_results();
	tagmatch::clear();
}

static void random_table_vs_scan() {
	init_test();
	std::mt19937_64 rng(2718);
	vector<filter_t> fs;
	vector<vector<tagmatch_key_t>> ks;
	for (int i = 0; i < 20000; ++i) {
		filter_t f;
		int nb = 1 + rng() % 30;
		for (int b = 0; b < nb; ++b) f.set_bit(rng() % 192);
		vector<tagmatch_key_t> k;
		int nk = 1 + rng() % 6;
		for (int j = 0; j < nk; ++j) k.push_back((tagmatch_key_t)(rng() % 5000));
		tagmatch::add(f, k);
		fs.push_back(f);
		ks.push_back(k);
	}
	tagmatch::consolidate(2000, 4);
	tagmatch::start();
	deferred_checks C;
	for (int i = 0; i < 600; ++i) {
		filter_t q;
		if (i % 3 == 0) q = fs[rng() % fs.size()];
		int nb = 60 + rng() % 80;
		for (int b = 0; b < nb; ++b) q.set_bit(rng() % 192);
		set<tagmatch_key_t> exp;
		for (size_t j = 0; j < fs.size(); ++j)
			if (fs[j].subset_of(q)) exp.insert(ks[j].begin(), ks[j].end());
		C.run_query(q, exp);
	}
	tagmatch::stop();
	C.check_results();
	CHECK(tagmatch::partition_count() > 1);
	tagmatch::clear();
}

static bool load_vectors(const char * path) {
	std::ifstream in(path);
	std::string tag, bitsstr;
	while (in >> tag >> bitsstr) {
		if (tag == "F") F.emplace_back(bitsstr);
		else if (tag == "Q") Q.emplace_back(bitsstr);
	}
	return F.size() == 11 && Q.size() == 13;
}

int main(int argc, char ** argv) {
	if (argc < 2 || !load_vectors(argv[1])) {
		std::fprintf(stderr, "usage: %s vectors.txt (11 F lines, 13 Q lines)\n", argv[0]);
		return 2;
	}
	ALL_ONES.fill();
	ALL_ZEROS.clear();
	struct {
		const char * name;
		void (*fn)();
	} cases[] = {
		{"consolidate_empty", consolidate_empty}, {"start_and_stop", start_and_stop},
		{"consolidate_no_match_action", consolidate_no_match_action}, {"repeated_stop", repeated_stop},
		{"find_one_filter", find_one_filter}, {"clear", clear_case}, {"add_same_filter", add_same_filter},
		{"add_and_find_multi", add_and_find_multi}, {"add_and_find_multi2", add_and_find_multi2},
		{"add_and_clear", add_and_clear}, {"add_many_and_find", add_many_and_find},
		{"misc_subsets", misc_subsets}, {"misc_subsets_Q1", misc_subsets_Q1},
		{"corner_case_filters", corner_case_filters}, {"single_bit_filters", single_bit_filters},
		{"deepest_trie", deepest_trie}, {"match_keeps_duplicates", match_keeps_duplicates},
		{"remove_keys", remove_keys}, {"all_zero_filter_matches_everything", all_zero_filter_matches_everything},
		{"random_table_vs_scan", random_table_vs_scan},
	};
	const char * only = argc > 2 ? argv[2] : nullptr;
	for (auto & c : cases) {
		if (only && std::string(only) != c.name) continue;
		int before = failures;
		std::printf("RUN  %s\n", c.name);
		std::fflush(stdout);
		c.fn();
		std::printf("%s %s\n", failures == before ? "PASS" : "FAIL", c.name);
		std::fflush(stdout);
	}
	tagmatch::clear();
	std::printf("%s: %d failure(s)\n", failures ? "FAILED" : "OK", failures);
	return failures ? 1 : 0;
}
