// Host side of the engine: the canonical filter table (merge of identical
// filters, key-set edits, lexicographic order) and the TagMatch C++ facade
// (class tagmatch) over the C ABI.
//
// Reference counterparts:
//   tagmatch::add/remove/consolidate   tagmatch.cc:22-77, :98-165
//   front_end::match / worker threads  front_end.cc:542-551, :563-691
//   front_end::stop (drain)            front_end.cc:693-771
//   tagmatch_query::finalize_matching  tagmatch_query.hh:99-119
// The query -> partition lookup, batching per partition, matching, key
// expansion and dedup all run on the device (tm_submit); the host thread
// only gathers queries into submissions and hands results back.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/tagmatch_amd.h"
#include "../../include/tagmatch/tagmatch.hh"

namespace {

// ------------------------------------------------------------ table

// Sort key of a filter: the bit string read position 0 first, as three
// big-endian 64-bit numbers (bit-reversed blocks), so that numeric order is
// the reference's lexicographic order (bitvector.hh:389-397).
struct sort_rec {
	uint64_t k0, k1, k2;
	uint64_t idx;
	bool operator<(const sort_rec & o) const {
		if (k0 != o.k0) return k0 < o.k0;
		if (k1 != o.k1) return k1 < o.k1;
		if (k2 != o.k2) return k2 < o.k2;
		return idx < o.idx;
	}
	bool same_filter(const sort_rec & o) const { return k0 == o.k0 && k1 == o.k1 && k2 == o.k2; }
};

inline uint64_t blk(const uint32_t * w, int i) { return (uint64_t)w[2 * i] | ((uint64_t)w[2 * i + 1] << 32); }

unsigned hw_threads() { return std::max(1u, std::min(16u, std::thread::hardware_concurrency())); }

// Parallel sort: sorted chunks, then pairwise merge rounds.
void psort(std::vector<sort_rec> & v) {
	const size_t n = v.size();
	unsigned T = hw_threads();
	if (n < (1u << 16) || T == 1) {
		std::sort(v.begin(), v.end());
		return;
	}
	std::vector<size_t> bounds;
	for (unsigned t = 0; t <= T; ++t) bounds.push_back(n * t / T);
	{
		std::vector<std::thread> th;
		for (unsigned t = 0; t < T; ++t)
			th.emplace_back([&, t] { std::sort(v.begin() + bounds[t], v.begin() + bounds[t + 1]); });
		for (auto & x : th) x.join();
	}
	std::vector<sort_rec> tmp(n);
	std::vector<sort_rec> * src = &v, * dst = &tmp;
	while (bounds.size() > 2) {
		std::vector<size_t> nb;
		std::vector<std::thread> th;
		for (size_t i = 0; i + 1 < bounds.size(); i += 2) {
			size_t a = bounds[i], m = bounds[i + 1], e = (i + 2 < bounds.size()) ? bounds[i + 2] : m;
			nb.push_back(a);
			th.emplace_back([=] {
				std::merge(src->begin() + a, src->begin() + m, src->begin() + m, src->begin() + e, dst->begin() + a);
			});
			if (i + 2 >= bounds.size()) break;
		}
		for (auto & x : th) x.join();
		nb.push_back(n);
		bounds.swap(nb);
		std::swap(src, dst);
	}
	if (src != &v) v.swap(*src);
}

struct table {
	std::vector<uint32_t> words;     // n x 6
	std::vector<uint64_t> key_off;   // n + 1
	std::vector<uint32_t> keys;
	uint64_t size() const { return words.size() / 6; }
	void clear() {
		words.clear();
		key_off.assign(1, 0);
		keys.clear();
	}
};

// One edit of a filter's key set.
struct edit {
	uint32_t w[6];
	uint32_t key;
	uint32_t remove;
};

// New table = old table with the edits applied: per filter,
// new keys = (old - removed) + (added - old), ascending; filters whose key
// set becomes empty are dropped; filters in lexicographic order
// (keyset_change::apply, reference tagmatch.cc:36-60, and :98-165).
table apply_edits(const table & old, const std::vector<edit> & ed) {
	const uint64_t n_old = old.size(), n_ed = ed.size();
	std::vector<sort_rec> rec(n_old + n_ed);
	for (uint64_t i = 0; i < n_old + n_ed; ++i) {
		const uint32_t * w = i < n_old ? &old.words[i * 6] : ed[i - n_old].w;
		rec[i] = sort_rec{__builtin_bitreverse64(blk(w, 0)), __builtin_bitreverse64(blk(w, 1)),
						  __builtin_bitreverse64(blk(w, 2)), i};
	}
	psort(rec);
	table t;
	t.key_off.assign(1, 0);
	std::vector<uint32_t> olds, add, rem, out;
	for (size_t i = 0; i < rec.size();) {
		size_t j = i;
		olds.clear(); add.clear(); rem.clear();
		while (j < rec.size() && rec[j].same_filter(rec[i])) {
			uint64_t x = rec[j].idx;
			if (x < n_old) olds.insert(olds.end(), old.keys.begin() + old.key_off[x], old.keys.begin() + old.key_off[x + 1]);
			else (ed[x - n_old].remove ? rem : add).push_back(ed[x - n_old].key);
			++j;
		}
		std::sort(olds.begin(), olds.end());
		olds.erase(std::unique(olds.begin(), olds.end()), olds.end());
		std::sort(add.begin(), add.end());
		std::sort(rem.begin(), rem.end());
		out.clear();
		for (uint32_t k : olds)
			if (!std::binary_search(rem.begin(), rem.end(), k)) out.push_back(k);
		for (uint32_t k : add)
			if (!std::binary_search(olds.begin(), olds.end(), k)) out.push_back(k);
		std::sort(out.begin(), out.end());
		out.erase(std::unique(out.begin(), out.end()), out.end());
		if (!out.empty()) {
			uint64_t x = rec[i].idx;
			const uint32_t * w = x < n_old ? &old.words[x * 6] : ed[x - n_old].w;
			t.words.insert(t.words.end(), w, w + 6);
			t.keys.insert(t.keys.end(), out.begin(), out.end());
			t.key_off.push_back(t.keys.size());
		}
		i = j;
	}
	return t;
}

// ------------------------------------------------------------ facade state

std::mutex g_api_mtx;                      // serialises add/consolidate/start/stop/clear
std::vector<edit> g_updates;
table g_table;
bool g_consolidated = false;
unsigned g_psize = 1000, g_pthreads = 1;   // reference tagmatch.cc:89-90
unsigned g_threads = 4, g_gpus = 1;        // reference tagmatch.cc:171-172
std::string g_dbname;                      // empty: no database file
std::atomic<int> g_status(TM_OK);
std::atomic<unsigned> g_latency_ms(0);

tm_fib * g_fib = nullptr;
uint32_t g_nparts = 0;
bool g_device = false;

// Single-producer / single-consumer queue of queries (the reference's
// matching job queue, front_end.cc:516-551).
constexpr size_t QCAP = 1u << 20;
std::vector<tagmatch_query *> g_queue(QCAP);
std::atomic<size_t> g_qhead(0), g_qtail(0);

std::thread g_engine;
std::atomic<bool> g_running(false), g_stop(false);

constexpr uint32_t SUBMIT_MAX = 8192;       // queries per device submission

void fail(int rc, const char * what) {
	g_status = rc;
	std::fprintf(stderr, "tagmatch: %s failed (%d): %s\n", what, rc, tm_last_error());
}

void ensure_device() {
	if (g_device) return;
	int rc = tm_dev_init(0);
	if (rc != TM_OK) {
		fail(rc, "tm_dev_init");
		throw std::runtime_error(std::string("tagmatch: no usable GPU: ") + tm_last_error());
	}
	g_device = true;
}

void write_database() {
	if (g_dbname.empty()) return;
	std::ofstream out(g_dbname, std::ios::binary);
	if (!out) {
		std::fprintf(stderr, "tagmatch: could not open database file: %s\n", g_dbname.c_str());
		return;
	}
	// fib_entry records (reference fib.cc:14-35): filter, u32 n, n keys
	for (uint64_t i = 0; i < g_table.size(); ++i) {
		filter_t f = filter_t::from_words(&g_table.words[i * 6]);
		f.write_binary(out);
		uint32_t n = (uint32_t)(g_table.key_off[i + 1] - g_table.key_off[i]);
		out.write(reinterpret_cast<const char *>(&n), 4);
		out.write(reinterpret_cast<const char *>(&g_table.keys[g_table.key_off[i]]), 4 * (size_t)n);
	}
}

// Partition the table on the GPU and make it resident.
void build_fib() {
	const uint64_t n = g_table.size();
	if (n > 0xFFFFFFF0ULL) throw std::runtime_error("tagmatch: table too large");
	if (g_fib) {
		tm_fib_destroy(g_fib);
		g_fib = nullptr;
	}
	g_nparts = 0;
	if (n == 0) return;
	ensure_device();
	const uint32_t N = (uint32_t)n;
	std::vector<uint32_t> perm(N), off(N + 2), masks((size_t)(N + 1) * 6);
	uint32_t np = 0;
	int rc = tm_partition(g_table.words.data(), N, g_psize ? g_psize : 1, perm.data(), off.data(), masks.data(),
						  N + 1, &np);
	if (rc != TM_OK) {
		fail(rc, "tm_partition");
		throw std::runtime_error(std::string("tagmatch: partitioning failed: ") + tm_last_error());
	}
	std::vector<uint32_t> w((size_t)N * 6), koff(N + 1), keys;
	keys.reserve(g_table.keys.size());
	koff[0] = 0;
	for (uint32_t i = 0; i < N; ++i) {
		uint32_t s = perm[i];
		std::memcpy(&w[(size_t)i * 6], &g_table.words[(size_t)s * 6], 24);
		keys.insert(keys.end(), g_table.keys.begin() + g_table.key_off[s], g_table.keys.begin() + g_table.key_off[s + 1]);
		koff[i + 1] = (uint32_t)keys.size();
	}
	rc = tm_fib_create(&g_fib);
	if (rc == TM_OK)
		rc = tm_fib_upload(g_fib, w.data(), N, off.data(), masks.data(), np, koff.data(), keys.data(), perm.data());
	if (rc != TM_OK) {
		fail(rc, "tm_fib_upload");
		throw std::runtime_error(std::string("tagmatch: FIB upload failed: ") + tm_last_error());
	}
	g_nparts = np;
}

// ------------------------------------------------------------ engine thread

struct batch_runner {
	tm_ring * ring = nullptr;
	uint32_t mode;
	std::vector<uint32_t> words;
	std::vector<uint64_t> off;
	std::vector<uint32_t> vals;

	explicit batch_runner(uint32_t m) : mode(m) {}
	~batch_runner() {
		if (ring) tm_ring_destroy(ring);
	}

	int ensure_ring() {
		if (ring || !g_fib) return TM_OK;
		tm_match_config c;
		tm_match_config_default(&c);
		c.mode = mode;
		int rc = tm_ring_create(g_fib, SUBMIT_MAX, &c, &ring);
		if (rc != TM_OK && mode == TM_MODE_KEYS_UNIQUE) {
			// more than 65536 distinct keys: all keys, dedup in finalize_matching
			mode = TM_MODE_KEYS_ALL;
			c.mode = mode;
			rc = tm_ring_create(g_fib, SUBMIT_MAX, &c, &ring);
		}
		return rc;
	}

	void run(std::vector<tagmatch_query *> & b) {
		if (b.empty()) return;
		const uint32_t nq = (uint32_t)b.size();
		int rc = TM_OK;
		if (g_fib) {
			rc = ensure_ring();
			words.resize((size_t)nq * 6);
			for (uint32_t i = 0; i < nq; ++i) b[i]->filter.copy_into_uint32_array(&words[(size_t)i * 6]);
			if (rc == TM_OK) rc = tm_submit(ring, words.data(), nq);
			if (rc == TM_OK) rc = tm_wait(ring, nullptr);
			off.resize((size_t)nq + 1);
			if (rc == TM_OK) rc = tm_read_results(ring, off.data(), nullptr, 0);
			if (rc == TM_OK) {
				vals.resize(off[nq]);
				rc = tm_read_results(ring, off.data(), vals.data(), vals.size());
			}
			if (rc != TM_OK) fail(rc, "match submission");
		}
		for (uint32_t i = 0; i < nq; ++i) {
			tagmatch_query * q = b[i];
			if (g_fib && rc == TM_OK) q->add_output(vals.data() + off[i], vals.data() + off[i + 1]);
			q->frontend_done();
			q->finalize_matching();
		}
		b.clear();
	}
};

void engine_loop() {
	batch_runner unique_runner(TM_MODE_KEYS_UNIQUE), all_runner(TM_MODE_KEYS_ALL);
	std::vector<tagmatch_query *> bu, ba;
	using clock = std::chrono::steady_clock;
	clock::time_point tu, ta;
	for (;;) {
		size_t head = g_qhead.load(std::memory_order_relaxed);
		size_t tail = g_qtail.load(std::memory_order_acquire);
		bool popped = false;
		while (head != tail && bu.size() < SUBMIT_MAX && ba.size() < SUBMIT_MAX) {
			tagmatch_query * q = g_queue[head];
			head = (head + 1) % QCAP;
			popped = true;
			auto & b = q->match_unique ? bu : ba;
			if (b.empty()) (q->match_unique ? tu : ta) = clock::now();
			b.push_back(q);
		}
		g_qhead.store(head, std::memory_order_release);
		const bool drained = head == g_qtail.load(std::memory_order_acquire);
		const bool stopping = g_stop.load(std::memory_order_acquire);
		const unsigned lim = g_latency_ms.load();
		auto due = [&](const std::vector<tagmatch_query *> & b, clock::time_point t0) {
			if (b.empty()) return false;
			if (b.size() >= SUBMIT_MAX) return true;
			if (stopping && drained) return true;
			return lim > 0 && clock::now() - t0 >= std::chrono::milliseconds(lim);
		};
		if (due(bu, tu)) unique_runner.run(bu);
		if (due(ba, ta)) all_runner.run(ba);
		if (stopping && drained && bu.empty() && ba.empty()) break;
		if (!popped) std::this_thread::sleep_for(std::chrono::microseconds(50));
	}
}

} // namespace

// ------------------------------------------------------------ C ABI helper

extern "C" int tm_canonicalize(const uint32_t * words, uint64_t n, const uint64_t * key_off, const uint32_t * keys,
							   uint32_t ** out_words, uint64_t * out_n, uint64_t ** out_key_off, uint32_t ** out_keys) {
	if (!out_words || !out_n || !out_key_off || !out_keys) return TM_ERR_INVALID;
	if (n && (!words || !key_off || !keys)) return TM_ERR_INVALID;
	std::vector<edit> ed;
	ed.reserve(key_off ? key_off[n] : 0);
	for (uint64_t i = 0; i < n; ++i)
		for (uint64_t k = key_off[i]; k < key_off[i + 1]; ++k) {
			edit e;
			std::memcpy(e.w, words + i * 6, 24);
			e.key = keys[k];
			e.remove = 0;
			ed.push_back(e);
		}
	table empty;
	empty.clear();
	table t = apply_edits(empty, ed);
	const uint64_t m = t.size();
	*out_words = (uint32_t *)std::malloc(std::max<size_t>(1, t.words.size()) * 4);
	*out_key_off = (uint64_t *)std::malloc((m + 1) * 8);
	*out_keys = (uint32_t *)std::malloc(std::max<size_t>(1, t.keys.size()) * 4);
	if (!*out_words || !*out_key_off || !*out_keys) return TM_ERR_NOMEM;
	std::memcpy(*out_words, t.words.data(), t.words.size() * 4);
	std::memcpy(*out_key_off, t.key_off.data(), (m + 1) * 8);
	std::memcpy(*out_keys, t.keys.data(), t.keys.size() * 4);
	*out_n = m;
	return TM_OK;
}

extern "C" void tm_free(void * p) { std::free(p); }

// ------------------------------------------------------------ class tagmatch

void tagmatch::add(const filter_t & s, const std::vector<tagmatch_key_t> & keys) {
	std::lock_guard<std::mutex> l(g_api_mtx);
	for (tagmatch_key_t k : keys) {
		edit e;
		s.copy_into_uint32_array(e.w);
		e.key = k;
		e.remove = 0;
		g_updates.push_back(e);
	}
}

void tagmatch::add(const filter_t & s, tagmatch_key_t key) {
	std::lock_guard<std::mutex> l(g_api_mtx);
	edit e;
	s.copy_into_uint32_array(e.w);
	e.key = key;
	e.remove = 0;
	g_updates.push_back(e);
}

void tagmatch::remove(const filter_t & s, tagmatch_key_t key) {
	std::lock_guard<std::mutex> l(g_api_mtx);
	edit e;
	s.copy_into_uint32_array(e.w);
	e.key = key;
	e.remove = 1;
	g_updates.push_back(e);
}

void tagmatch::consolidate(unsigned int psize, unsigned int threads) {
	{
		std::lock_guard<std::mutex> l(g_api_mtx);
		g_psize = psize;
		g_pthreads = threads;
	}
	consolidate();
}

void tagmatch::consolidate() {
	std::lock_guard<std::mutex> l(g_api_mtx);
	if (g_running) throw std::runtime_error("tagmatch: consolidate() while started; call stop() first");
	if (!g_consolidated) g_table.clear();
	g_table = apply_edits(g_table, g_updates);
	g_updates.clear();
	write_database();
	build_fib();
	g_consolidated = true;
	g_status = TM_OK;
}

const char * tagmatch::get_database_filename() { return g_dbname.c_str(); }

void tagmatch::set_database_filename(const char * name) { g_dbname = name ? name : ""; }

void tagmatch::start() { start(g_threads, g_gpus); }

void tagmatch::start(unsigned int threads, unsigned int gpu_count) {
	std::lock_guard<std::mutex> l(g_api_mtx);
	g_threads = threads;
	g_gpus = gpu_count;   // one GPU is used whatever the count
	if (g_running) return;
	if (g_fib) ensure_device();
	g_stop = false;
	g_running = true;
	g_engine = std::thread(engine_loop);
}

void tagmatch::stop() {
	std::lock_guard<std::mutex> l(g_api_mtx);
	if (!g_running) return;
	g_stop = true;
	g_engine.join();
	g_running = false;
	g_stop = false;
}

void tagmatch::clear() {
	stop();
	std::lock_guard<std::mutex> l(g_api_mtx);
	g_consolidated = false;
	g_updates.clear();
	g_table.clear();
	if (g_fib) {
		tm_fib_destroy(g_fib);
		g_fib = nullptr;
	}
	g_nparts = 0;
}

unsigned int tagmatch::get_latency_limit_ms() { return g_latency_ms.load(); }

void tagmatch::set_latency_limit_ms(unsigned int l) { g_latency_ms = l; }

void tagmatch::match(tagmatch_query * q, match_handler * h) noexcept {
	q->match_unique = false;
	q->set_match_handler(h);
	size_t tail = g_qtail.load(std::memory_order_relaxed);
	size_t next = (tail + 1) % QCAP;
	while (next == g_qhead.load(std::memory_order_acquire)) std::this_thread::yield();  // full
	g_queue[tail] = q;
	g_qtail.store(next, std::memory_order_release);
}

void tagmatch::match_unique(tagmatch_query * q, match_handler * h) noexcept {
	q->match_unique = true;
	q->set_match_handler(h);
	size_t tail = g_qtail.load(std::memory_order_relaxed);
	size_t next = (tail + 1) % QCAP;
	while (next == g_qhead.load(std::memory_order_acquire)) std::this_thread::yield();
	g_queue[tail] = q;
	g_qtail.store(next, std::memory_order_release);
}

int tagmatch::last_status() { return g_status.load(); }

size_t tagmatch::bytesize() {
	if (!g_fib) return 0;
	tm_fib_stats st;
	if (tm_fib_stats_get(g_fib, &st) != TM_OK) return 0;
	return (size_t)st.device_bytes;
}

unsigned int tagmatch::partition_count() { return g_nparts; }
