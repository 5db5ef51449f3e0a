// Serving loop: continuous batching of match requests over several rings
// in flight, admission by an HBM budget for result buffers, and filter
// inserts at version barriers.  It is a client of the C ABI only.
//
// The reference's serving path is the front end's per-partition queues
// flushed when full or when the oldest query passes the latency limit
// (front_end.cc:249-267, :350-376, :563-598), feeding two double-buffered
// streams per GPU (back_end.cc:131-311).  Here a request is a batch of
// queries; an engine thread admits as many waiting requests as fit one free
// ring (slot) and the budget, submits them as one device step, and returns
// each request's own results when the step completes.  Steps of different
// slots run on different HIP streams and overlap on the device.
#include "../../include/tagmatch_amd.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

namespace tmx {   // hip_util.hip, fib.hip
int set_error(int code, const char * fmt, ...);   // text for tm_last_error()
void * pinned_alloc(size_t bytes);                // pinned host memory, NULL on failure
void pinned_free(void * p);
struct live_plan;                                 // a live insert prepared on the host
int live_prepare(tm_fib * fib, const uint32_t * words, uint32_t n, const uint32_t * key_offsets,
				 const uint32_t * keys, const uint32_t * filter_ids, live_plan ** out);
int live_commit(tm_fib * fib, live_plan * plan);  // frees the plan
void live_discard(live_plan * plan);
}

namespace {

using clock_t_ = std::chrono::steady_clock;

double ms_since(clock_t_::time_point t0, clock_t_::time_point t1) {
	return std::chrono::duration<double, std::milli>(t1 - t0).count();
}

struct request {
	uint64_t id = 0;
	bool is_insert = false;
	uint32_t nq = 0;
	std::vector<uint32_t> q;          // nq x 6
	// insert payload
	std::vector<uint32_t> ins_words, ins_koff, ins_keys, ins_fid;
	uint32_t ins_n = 0, ins_maxp = 0;
	bool ins_tried_live = false;      // the live insert declined: apply at a barrier
	tmx::live_plan * plan = nullptr;  // prepared when submitted: the engine only commits it
	// results: offsets of its own, values a slice of its step's buffer; or,
	// for KEYS_UNIQUE, rows [row0, row0 + nq) of its step's key bitmap rows
	// over the dictionary the step ran with, expanded when they are read
	std::vector<uint64_t> off;
	std::shared_ptr<struct result_buf> rbuf;
	uint64_t vbase = 0, vn = 0;
	bool rows = false, counted = false;
	uint64_t row0 = 0;
	uint32_t kw = 0;
	std::shared_ptr<const std::vector<uint32_t>> dict;
	int status = TM_OK;
	std::string error;
	uint64_t reserved = 0;            // bytes of the HBM budget held
	clock_t_::time_point t_submit, t_admit, t_done;
	uint32_t step_queries = 0;        // queries of the device step it ran in
	bool done = false;
};

// Result values of one step in pinned host memory: filled by one DMA copy
// and shared by the step's requests until the last is released, then kept
// in the pool for later steps.
struct result_buf {
	uint32_t * p = nullptr;
	uint64_t cap = 0;   // values
};

struct buf_pool {
	std::mutex mu;
	std::vector<result_buf *> free;
	size_t keep = 8;
	~buf_pool() {
		for (result_buf * b : free) {
			tmx::pinned_free(b->p);
			delete b;
		}
	}
};

struct slot {
	tm_ring * ring = nullptr;
	std::vector<request *> reqs;
	uint64_t reserved = 0;
	bool busy = false;
	std::vector<uint32_t> words;
	std::vector<uint64_t> off;
	// KEYS_UNIQUE: the step's rows, delivered by its ring into this pinned
	// buffer, and the dictionary they index
	std::shared_ptr<struct result_buf> rows;
	std::shared_ptr<const std::vector<uint32_t>> dict;
	uint32_t kw = 0;
};

} // namespace

struct tm_server {
	tm_fib * fib = nullptr;
	tm_server_config cfg;
	tm_match_config mcfg;
	std::vector<slot> slots;
	std::mutex mu;
	std::condition_variable cv_engine, cv_done;
	std::deque<request *> waiting;
	std::unordered_map<uint64_t, request *> reqs;
	std::deque<uint64_t> finished;    // ids not yet collected by tm_server_collect
	uint64_t next_id = 1;
	uint64_t reserved = 0;            // bytes held by admitted requests
	double bytes_per_query = 0;       // running estimate of result bytes per query
	uint64_t fixed_per_query = 0;     // result bytes per query that do not depend on hits
	tm_server_stats st{};
	std::thread th;
	std::atomic<bool> stop{false};
	int fatal = TM_OK;
	std::string fatal_msg;
	std::shared_ptr<buf_pool> pool = std::make_shared<buf_pool>();
	std::mutex ins_mu;                // orders the preparation of inserts
	std::atomic<uint64_t> arrivals{0};   // requests and inserts submitted so far
	uint32_t unprepared = 0;          // queued inserts without a plan (they change the layout)
	// KEYS_UNIQUE requests are served as key bitmap rows (the slots' rings run
	// TM_MODE_KEY_BITMAP and copy each step's rows to the host as part of the
	// step); the dictionary is replaced when an insert changes it
	bool rows_mode = false;
	std::shared_ptr<const std::vector<uint32_t>> dict;
};

namespace {

// Result bytes a request of nq queries is charged against the budget.
uint64_t estimate(const tm_server * s, uint32_t nq) {
	return (uint64_t)((double)nq * (s->bytes_per_query * 1.5 + (double)s->fixed_per_query)) + 4096;
}

// The FIB's current key dictionary (rows mode).
int refresh_dict(tm_server * s) {
	uint64_t n = 0;
	int rc = tm_fib_key_dictionary(s->fib, nullptr, 0, &n);
	if (rc != TM_OK) return rc;
	auto d = std::make_shared<std::vector<uint32_t>>(n);
	rc = tm_fib_key_dictionary(s->fib, d->data(), n, &n);
	if (rc != TM_OK) return rc;
	d->resize(n);
	s->dict = std::move(d);
	return TM_OK;
}

void finish(tm_server * s, request * r) {
	r->t_done = clock_t_::now();
	r->done = true;
	s->finished.push_back(r->id);
	const double lat = ms_since(r->t_submit, r->t_done);
	if (!r->is_insert) {
		s->st.n_requests_done++;
		s->st.n_queries_done += r->nq;
		s->st.latency_sum_ms += lat;
		s->st.latency_max_ms = std::max(s->st.latency_max_ms, lat);
	}
}

// A pinned buffer of at least n values, from the pool or new; it returns to
// the pool when the last request holding it is released.
std::shared_ptr<result_buf> take_buf(tm_server * s, uint64_t n) {
	std::shared_ptr<buf_pool> pool = s->pool;
	result_buf * b = nullptr;
	{
		std::lock_guard<std::mutex> lk(pool->mu);
		for (size_t i = 0; i < pool->free.size(); ++i)
			if (pool->free[i]->cap >= n) {
				b = pool->free[i];
				pool->free.erase(pool->free.begin() + i);
				break;
			}
	}
	if (!b) {
		const uint64_t cap = std::max<uint64_t>(n + n / 4, 1u << 18);
		void * p = tmx::pinned_alloc(cap * 4);
		if (!p) return nullptr;
		b = new result_buf();
		b->p = static_cast<uint32_t *>(p);
		b->cap = cap;
	}
	return std::shared_ptr<result_buf>(b, [pool](result_buf * x) {
		std::lock_guard<std::mutex> lk(pool->mu);
		if (pool->free.size() < pool->keep) {
			pool->free.push_back(x);
		} else {
			tmx::pinned_free(x->p);
			delete x;
		}
	});
}

// Collect a finished step: split results per request, release the budget.
void complete(tm_server * s, slot & sl) {
	tm_status st;
	int rc = tm_wait(sl.ring, &st);
	const uint32_t nq = (uint32_t)(sl.words.size() / 6);
	if (s->rows_mode) {
		// the rows are in host memory already: hand each request its slice
		std::string err = rc == TM_OK ? std::string() : std::string(tm_last_error());
		std::lock_guard<std::mutex> lk(s->mu);
		uint32_t q0 = 0;
		for (request * r : sl.reqs) {
			r->status = rc;
			r->step_queries = nq;
			r->rows = true;
			r->rbuf = sl.rows;
			r->row0 = q0;
			r->kw = sl.kw;
			r->dict = sl.dict;
			if (rc != TM_OK) r->error = err;
			q0 += r->nq;
			finish(s, r);
		}
		if (rc == TM_OK) {
			s->st.n_steps++;
			s->st.device_ms_sum += st.device_ms;
		}
		s->reserved -= sl.reserved;
		sl.reserved = 0;
		sl.reqs.clear();
		sl.rows.reset();
		sl.dict.reset();
		sl.busy = false;
		s->cv_done.notify_all();
		return;
	}
	std::shared_ptr<result_buf> buf;
	if (rc == TM_OK) {
		sl.off.resize((size_t)nq + 1);
		rc = tm_read_results(sl.ring, sl.off.data(), nullptr, 0);
	}
	if (rc == TM_OK) {
		buf = take_buf(s, sl.off[nq]);
		rc = buf ? tm_read_results(sl.ring, sl.off.data(), buf->p, buf->cap)
				 : tmx::set_error(TM_ERR_NOMEM, "pinned result buffer of %llu values", (unsigned long long)sl.off[nq]);
	}
	std::string err = rc == TM_OK ? std::string() : std::string(tm_last_error());
	std::lock_guard<std::mutex> lk(s->mu);
	uint32_t q0 = 0;
	for (request * r : sl.reqs) {
		r->status = rc;
		r->step_queries = nq;
		if (rc == TM_OK) {
			const uint64_t base = sl.off[q0];
			r->off.resize((size_t)r->nq + 1);
			for (uint32_t i = 0; i <= r->nq; ++i) r->off[i] = sl.off[q0 + i] - base;
			r->rbuf = buf;
			r->vbase = base;
			r->vn = sl.off[q0 + r->nq] - base;
		} else {
			r->error = err;
		}
		q0 += r->nq;
		finish(s, r);
	}
	if (rc == TM_OK && nq) {
		// running estimate of result bytes per query for admission
		const double b = 4.0 * (double)sl.off[nq] / (double)nq;
		s->bytes_per_query = s->bytes_per_query == 0 ? b : 0.8 * s->bytes_per_query + 0.2 * b;
		s->st.n_steps++;
		s->st.device_ms_sum += st.device_ms;
		if (st.overflow) s->st.n_retries++;
	}
	s->reserved -= sl.reserved;
	sl.reserved = 0;
	sl.reqs.clear();
	sl.busy = false;
	s->cv_done.notify_all();
}

// Latency flush: whether free slots should wait for more queries.  True
// while fewer than min(flush_queries, slot_queries) queries wait at the head
// of the queue, the oldest of them has waited less than flush_us, no insert
// waits behind them and the server is not stopping.  Called under s->mu.
bool holding(tm_server * s, clock_t_::time_point now) {
	if (s->cfg.flush_queries == 0 || s->waiting.empty() || s->stop.load()) return false;
	const request * first = s->waiting.front();
	if (first->is_insert) return false;
	if (now - first->t_submit >= std::chrono::microseconds(s->cfg.flush_us)) return false;
	const uint64_t want = std::min(s->cfg.flush_queries, s->cfg.slot_queries);
	uint64_t nq = 0;
	for (const request * r : s->waiting) {
		if (r->is_insert) return false;
		nq += r->nq;
		if (nq >= want) return false;
	}
	return true;
}

void engine(tm_server * s) {
	for (;;) {
		bool progressed = false;
		// 1. completions
		for (auto & sl : s->slots) {
			if (!sl.busy) continue;
			if (tm_poll(sl.ring) == TM_ERR_BUSY) continue;
			complete(s, sl);
			progressed = true;
		}
		// 2. admission (and insert barriers), under the lock
		std::unique_lock<std::mutex> lk(s->mu);
		for (auto & sl : s->slots) {
			// the full-or-latency rule holds for every slot: what a slot
			// left behind waits for its own flush
			if (sl.busy || s->waiting.empty() || holding(s, clock_t_::now())) continue;
			if (s->waiting.front()->is_insert) break;
			uint32_t nq = 0;
			uint64_t bytes = 0;
			while (!s->waiting.empty() && !s->waiting.front()->is_insert) {
				request * r = s->waiting.front();
				const uint64_t need = estimate(s, r->nq);
				if (nq + r->nq > s->cfg.slot_queries && nq > 0) break;
				// the budget admits a request when it fits, or when nothing at
				// all is in flight (a request larger than the budget alone)
				const bool idle = s->reserved + bytes == 0;
				if (!idle && s->reserved + bytes + need > s->cfg.hbm_budget) {
					if (nq == 0) s->st.n_budget_waits++;
					break;
				}
				s->waiting.pop_front();
				r->reserved = need;
				r->t_admit = clock_t_::now();
				sl.reqs.push_back(r);
				nq += r->nq;
				bytes += need;
			}
			if (sl.reqs.empty()) break;   // budget full: later slots cannot do better
			s->reserved += bytes;
			sl.reserved = bytes;
			s->st.reserved_peak = std::max<uint64_t>(s->st.reserved_peak, s->reserved);
			lk.unlock();
			sl.words.resize((size_t)nq * 6);
			uint32_t q0 = 0;
			for (request * r : sl.reqs) {
				if (r->nq) std::memcpy(&sl.words[(size_t)q0 * 6], r->q.data(), (size_t)r->nq * 24);
				q0 += r->nq;
			}
			int rc = TM_OK;
			if (s->rows_mode) {
				// this step's rows go straight into a pinned buffer its
				// requests will share
				sl.dict = s->dict;
				sl.kw = (uint32_t)((sl.dict->size() + 31) / 32);
				sl.rows = take_buf(s, std::max<uint64_t>((uint64_t)nq * sl.kw, 1));
				rc = sl.rows ? tm_ring_set_host_rows(sl.ring, sl.rows->p, sl.rows->cap)
							 : tmx::set_error(TM_ERR_NOMEM, "pinned rows of %u queries", nq);
			}
			if (rc == TM_OK) rc = tm_submit(sl.ring, sl.words.data(), nq);
			lk.lock();
			if (rc != TM_OK) {
				const std::string err = tm_last_error();
				for (request * r : sl.reqs) {
					r->status = rc;
					r->error = err;
					finish(s, r);
				}
				s->reserved -= sl.reserved;
				sl.reserved = 0;
				sl.reqs.clear();
				sl.rows.reset();
				sl.dict.reset();
				s->cv_done.notify_all();
				continue;
			}
			sl.busy = true;
			s->st.n_batches++;
			s->st.requests_per_batch_sum += sl.reqs.size();
			progressed = true;
		}
		// inserts: live (beside the steps in flight, which do not see it) when
		// the batch fits spare lanes, else at a barrier once every earlier
		// request has completed
		if (!s->waiting.empty() && s->waiting.front()->is_insert && s->waiting.front()->plan) {
			// prepared when submitted: only the device half is left
			request * r = s->waiting.front();
			s->waiting.pop_front();
			tmx::live_plan * plan = r->plan;
			r->plan = nullptr;
			lk.unlock();
			int rc = tmx::live_commit(s->fib, plan);
			const std::string err = rc == TM_OK ? std::string() : std::string(tm_last_error());
			lk.lock();
			r->status = rc;
			r->error = err;
			s->st.n_inserts++;
			s->st.n_live_inserts += rc == TM_OK;
			finish(s, r);
			s->cv_done.notify_all();
			progressed = true;
		}
		if (!s->waiting.empty() && s->waiting.front()->is_insert && !s->waiting.front()->plan &&
			!s->waiting.front()->ins_tried_live) {
			request * r = s->waiting.front();
			lk.unlock();
			int rc = tm_fib_insert_live(s->fib, r->ins_words.data(), r->ins_n, r->ins_koff.data(),
										r->ins_keys.data(), r->ins_fid.empty() ? nullptr : r->ins_fid.data());
			const std::string err = (rc == TM_OK || rc == TM_ERR_BUSY) ? std::string() : std::string(tm_last_error());
			lk.lock();
			if (rc == TM_ERR_BUSY) {
				r->ins_tried_live = true;   // needs the barrier
			} else {
				s->waiting.pop_front();
				s->unprepared--;
				r->status = rc;
				r->error = err;
				s->st.n_inserts++;
				s->st.n_live_inserts += rc == TM_OK;
				finish(s, r);
				s->cv_done.notify_all();
				progressed = true;
			}
		}
		if (!s->waiting.empty() && s->waiting.front()->is_insert && s->waiting.front()->ins_tried_live) {
			bool idle = true;
			for (auto & sl : s->slots) idle = idle && !sl.busy;
			if (idle) {
				request * r = s->waiting.front();
				s->waiting.pop_front();
				lk.unlock();
				int rc = tm_fib_insert(s->fib, r->ins_words.data(), r->ins_n, r->ins_koff.data(), r->ins_keys.data(),
									   r->ins_fid.empty() ? nullptr : r->ins_fid.data(), r->ins_maxp);
				if (rc == TM_OK && s->rows_mode) rc = refresh_dict(s);   // new keys extend the dictionary
				const std::string err = rc == TM_OK ? std::string() : std::string(tm_last_error());
				lk.lock();
				s->unprepared--;
				r->status = rc;
				r->error = err;
				s->st.n_inserts++;
				finish(s, r);
				s->cv_done.notify_all();
				progressed = true;
			}
		}
		if (s->stop.load() && s->waiting.empty()) {
			bool idle = true;
			for (auto & sl : s->slots) idle = idle && !sl.busy;
			if (idle) return;
		}
		if (!progressed) {
			bool any_busy = false;
			for (auto & sl : s->slots) any_busy = any_busy || sl.busy;
			if (any_busy) {
				// steps in flight: poll their completions for a short while
				// without sleeping (a timed wait oversleeps by tens of
				// microseconds), then yield the lock briefly
				lk.unlock();
				bool any_free = false;
				for (auto & sl : s->slots) any_free = any_free || !sl.busy;
				const uint64_t seen = s->arrivals.load();
				const auto until = clock_t_::now() + std::chrono::microseconds(200);
				bool ready = false;
				while (!ready && clock_t_::now() < until) {
					for (auto & sl : s->slots)
						if (sl.busy && tm_poll(sl.ring) != TM_ERR_BUSY) ready = true;
					if (any_free && s->arrivals.load() != seen) ready = true;   // a request for a free slot
					if (!ready) std::this_thread::yield();
				}
				lk.lock();
				if (!ready && s->waiting.empty()) s->cv_engine.wait_for(lk, std::chrono::microseconds(20));
			} else if (holding(s, clock_t_::now())) {
				// a latency flush is pending: sleep until its deadline or an arrival
				// (requests are queued under the lock, so none is missed)
				const auto due = s->waiting.front()->t_submit + std::chrono::microseconds(s->cfg.flush_us);
				s->cv_engine.wait_until(lk, std::min(due, clock_t_::now() + std::chrono::milliseconds(5)),
										[&] { return !holding(s, clock_t_::now()); });
			} else {
				s->cv_engine.wait_for(lk, std::chrono::milliseconds(5));   // idle: sleep on the queue
			}
		}
	}
}

} // namespace

extern "C" {

void tm_server_config_default(tm_server_config * c) {
	std::memset(c, 0, sizeof(*c));
	c->n_slots = 4;
	c->slot_queries = 16384;
	c->hbm_budget = 0;
}

int tm_server_create(tm_fib * fib, const tm_server_config * cfg, const tm_match_config * mcfg, tm_server ** out) {
	if (!fib || !out) return tmx::set_error(TM_ERR_INVALID, "NULL argument");
	tm_server * s = new tm_server();
	s->fib = fib;
	if (cfg) s->cfg = *cfg; else tm_server_config_default(&s->cfg);
	if (mcfg) s->mcfg = *mcfg; else tm_match_config_default(&s->mcfg);
	if (s->cfg.n_slots == 0) s->cfg.n_slots = 4;
	if (s->cfg.slot_queries == 0) s->cfg.slot_queries = 16384;
	if (s->cfg.n_slots > 64) {
		delete s;
		return tmx::set_error(TM_ERR_INVALID, "n_slots above 64");
	}
	if (s->mcfg.mode == TM_MODE_KEY_BITMAP) {
		delete s;
		return tmx::set_error(TM_ERR_INVALID, "the server returns key lists: KEY_BITMAP rings are driven directly");
	}
	if (s->cfg.hbm_budget == 0) {
		size_t fr = 0, tot = 0;
		int rc = tm_mem_info(&fr, &tot);
		if (rc != TM_OK) {
			delete s;
			return rc;
		}
		s->cfg.hbm_budget = fr / 4;
	}
	tm_fib_stats fs;
	tm_fib_stats_get(fib, &fs);
	// key bitmap rows are per query and independent of hits
	s->fixed_per_query = s->mcfg.mode == TM_MODE_KEYS_UNIQUE ? 4 * ((fs.n_distinct_keys + 31) / 32) : 8;
	s->rows_mode = s->mcfg.mode == TM_MODE_KEYS_UNIQUE;
	if (s->rows_mode) {
		int rc = refresh_dict(s);
		if (rc != TM_OK) {
			delete s;
			return rc;
		}
	}
	tm_match_config ring_cfg = s->mcfg;
	if (s->rows_mode) ring_cfg.mode = TM_MODE_KEY_BITMAP;
	// small steps spend much of their time launching kernels: replay them
	// from graphs (a step of nq queries runs as the next power of two)
	if (s->cfg.step_graphs != 2 && ring_cfg.schedule == TM_SCHED_PERSISTENT) ring_cfg.use_graph = 1;
	s->slots.resize(s->cfg.n_slots);
	s->pool->keep = 2 * (size_t)s->cfg.n_slots + 4;
	for (auto & sl : s->slots) {
		int rc = tm_ring_create(fib, s->cfg.slot_queries, &ring_cfg, &sl.ring);
		if (rc != TM_OK) {
			for (auto & x : s->slots)
				if (x.ring) tm_ring_destroy(x.ring);
			delete s;
			return rc;
		}
	}
	s->th = std::thread(engine, s);
	*out = s;
	return TM_OK;
}

int tm_server_destroy(tm_server * s) {
	if (!s) return TM_OK;
	s->stop.store(true);
	s->cv_engine.notify_all();
	if (s->th.joinable()) s->th.join();
	for (auto & sl : s->slots)
		if (sl.ring) tm_ring_destroy(sl.ring);
	for (auto & kv : s->reqs) {
		if (kv.second->plan) tmx::live_discard(kv.second->plan);
		delete kv.second;
	}
	delete s;
	return TM_OK;
}

int tm_server_set_cost_model(tm_server * s, tm_mlp * model, uint32_t flags) {
	if (!s) return tmx::set_error(TM_ERR_INVALID, "server is NULL");
	// only between steps: wait until every slot is idle
	std::unique_lock<std::mutex> lk(s->mu);
	for (;;) {
		bool idle = s->waiting.empty();
		for (auto & sl : s->slots) idle = idle && !sl.busy;
		if (idle) break;
		s->cv_done.wait_for(lk, std::chrono::milliseconds(1));
	}
	for (auto & sl : s->slots) {
		int rc = tm_ring_set_cost_model(sl.ring, model, flags);
		if (rc != TM_OK) return rc;
	}
	return TM_OK;
}

int tm_server_submit(tm_server * s, const uint32_t * queries, uint32_t nq, uint64_t * id) {
	if (!s || !id || (nq && !queries)) return tmx::set_error(TM_ERR_INVALID, "NULL argument");
	if (nq > s->cfg.slot_queries)
		return tmx::set_error(TM_ERR_INVALID, "request of %u queries exceeds slot_queries %u", nq, s->cfg.slot_queries);
	request * r = new request();
	r->nq = nq;
	r->q.assign(queries, queries + (size_t)nq * 6);
	r->t_submit = clock_t_::now();
	{
		std::lock_guard<std::mutex> lk(s->mu);
		if (s->stop.load()) {
			delete r;
			return tmx::set_error(TM_ERR_STATE, "server is stopping");
		}
		r->id = s->next_id++;
		s->reqs[r->id] = r;
		s->waiting.push_back(r);
		s->st.n_requests++;
	}
	s->arrivals.fetch_add(1);
	s->cv_engine.notify_one();
	*id = r->id;
	return TM_OK;
}

int tm_server_insert(tm_server * s, const uint32_t * words, uint32_t n, const uint32_t * key_offsets,
					 const uint32_t * keys, const uint32_t * filter_ids, uint32_t maxp, uint64_t * id) {
	if (!s || !id || (n && (!words || !key_offsets || !keys))) return tmx::set_error(TM_ERR_INVALID, "NULL argument");
	request * r = new request();
	r->is_insert = true;
	r->ins_n = n;
	r->ins_maxp = maxp;
	r->ins_words.assign(words, words + (size_t)n * 6);
	r->ins_koff.assign(key_offsets, key_offsets + n + 1);
	if (n) r->ins_keys.assign(keys, keys + key_offsets[n]);
	if (filter_ids) r->ins_fid.assign(filter_ids, filter_ids + n);
	if (r->ins_keys.empty()) r->ins_keys.push_back(0);
	r->t_submit = clock_t_::now();
	// Prepare the insert now, in the caller's thread (partition choice and
	// lane records on the host), unless an insert that may change the
	// table's layout is queued ahead of it; the engine then only commits it
	// to the device when it reaches the front of the queue.
	std::lock_guard<std::mutex> og(s->ins_mu);
	bool can_prepare;
	{
		std::lock_guard<std::mutex> lk(s->mu);
		can_prepare = s->unprepared == 0;
	}
	if (can_prepare) {
		int rc = tmx::live_prepare(s->fib, r->ins_words.data(), n, r->ins_koff.data(), r->ins_keys.data(),
								   r->ins_fid.empty() ? nullptr : r->ins_fid.data(), &r->plan);
		if (rc != TM_OK) r->plan = nullptr;   // declined (TM_ERR_BUSY) or invalid: the engine path decides
	}
	{
		std::lock_guard<std::mutex> lk(s->mu);
		r->id = s->next_id++;
		s->reqs[r->id] = r;
		s->waiting.push_back(r);
		if (!r->plan) s->unprepared++;
	}
	s->arrivals.fetch_add(1);
	s->cv_engine.notify_one();
	*id = r->id;
	return TM_OK;
}

int tm_server_wait(tm_server * s, uint64_t id, double timeout_ms) {
	if (!s) return tmx::set_error(TM_ERR_INVALID, "server is NULL");
	std::unique_lock<std::mutex> lk(s->mu);
	auto it = s->reqs.find(id);
	if (it == s->reqs.end()) return tmx::set_error(TM_ERR_INVALID, "unknown request %llu", (unsigned long long)id);
	request * r = it->second;
	const auto deadline = clock_t_::now() + std::chrono::microseconds((int64_t)(timeout_ms * 1000));
	while (!r->done) {
		if (timeout_ms >= 0 && clock_t_::now() >= deadline) return TM_ERR_BUSY;
		s->cv_done.wait_for(lk, std::chrono::milliseconds(1));
	}
	return r->status;
}

int tm_server_collect(tm_server * s, uint64_t * ids, uint32_t cap, uint32_t * n) {
	if (!s || !n || (cap && !ids)) return tmx::set_error(TM_ERR_INVALID, "NULL argument");
	std::lock_guard<std::mutex> lk(s->mu);
	uint32_t k = 0;
	while (k < cap && !s->finished.empty()) {
		ids[k++] = s->finished.front();
		s->finished.pop_front();
	}
	*n = k;
	return TM_OK;
}

int tm_server_result(tm_server * s, uint64_t id, tm_request_info * info, uint64_t * offsets, uint32_t * values,
					 uint64_t capacity) {
	if (!s) return tmx::set_error(TM_ERR_INVALID, "server is NULL");
	std::unique_lock<std::mutex> lk(s->mu);
	auto it = s->reqs.find(id);
	if (it == s->reqs.end()) return tmx::set_error(TM_ERR_INVALID, "unknown request %llu", (unsigned long long)id);
	request * r = it->second;
	if (!r->done) return tmx::set_error(TM_ERR_BUSY, "request %llu still running", (unsigned long long)id);
	if (r->rows && r->status == TM_OK) {
		// key lists from the step's rows, expanded outside the server lock
		// (the caller does not release this request concurrently)
		const std::shared_ptr<result_buf> rb = r->rbuf;
		const std::shared_ptr<const std::vector<uint32_t>> dict = r->dict;
		const uint32_t * rows = rb->p + r->row0 * r->kw;
		const uint32_t nq = r->nq, kw = r->kw, K = (uint32_t)dict->size();
		if (!r->counted) {
			lk.unlock();
			std::vector<uint64_t> off((size_t)nq + 1);
			int rc = tm_expand_key_bitmaps(rows, nq, kw, dict->data(), K, off.data(), nullptr, 0, 1);
			if (rc != TM_OK) return rc;
			lk.lock();
			r->off = std::move(off);
			r->vn = r->off[nq];
			r->counted = true;
		}
		if (info) {
			info->status = r->status;
			info->n_queries = r->nq;
			info->n_values = r->vn;
			info->step_queries = r->step_queries;
			info->latency_ms = ms_since(r->t_submit, r->t_done);
			info->queue_ms = ms_since(r->t_submit, r->t_admit);
		}
		if (offsets) std::memcpy(offsets, r->off.data(), r->off.size() * sizeof(uint64_t));
		const uint64_t vn = r->vn;
		lk.unlock();
		if (values) {
			if (vn > capacity)
				return tmx::set_error(TM_ERR_OVERFLOW, "%llu values, capacity %llu", (unsigned long long)vn,
									  (unsigned long long)capacity);
			std::vector<uint64_t> off((size_t)nq + 1);
			return tm_expand_key_bitmaps(rows, nq, kw, dict->data(), K, off.data(), values, capacity, 1);
		}
		return TM_OK;
	}
	if (info) {
		info->status = r->status;
		info->n_queries = r->nq;
		info->n_values = r->vn;
		info->step_queries = r->step_queries;
		info->latency_ms = ms_since(r->t_submit, r->t_done);
		info->queue_ms = r->is_insert ? 0.0 : ms_since(r->t_submit, r->t_admit);
	}
	if (offsets && !r->off.empty()) std::memcpy(offsets, r->off.data(), r->off.size() * sizeof(uint64_t));
	if (values) {
		if (r->vn > capacity)
			return tmx::set_error(TM_ERR_OVERFLOW, "%llu values, capacity %llu", (unsigned long long)r->vn,
								  (unsigned long long)capacity);
		if (r->vn) std::memcpy(values, r->rbuf->p + r->vbase, r->vn * 4);
	}
	return TM_OK;
}

int tm_server_release(tm_server * s, uint64_t id) {
	if (!s) return tmx::set_error(TM_ERR_INVALID, "server is NULL");
	std::lock_guard<std::mutex> lk(s->mu);
	auto it = s->reqs.find(id);
	if (it == s->reqs.end()) return tmx::set_error(TM_ERR_INVALID, "unknown request %llu", (unsigned long long)id);
	if (!it->second->done) return tmx::set_error(TM_ERR_BUSY, "request %llu still running", (unsigned long long)id);
	delete it->second;
	s->reqs.erase(it);
	return TM_OK;
}

int tm_server_stats_get(tm_server * s, tm_server_stats * st) {
	if (!s || !st) return tmx::set_error(TM_ERR_INVALID, "NULL argument");
	std::lock_guard<std::mutex> lk(s->mu);
	*st = s->st;
	st->reserved_now = s->reserved;
	st->hbm_budget = s->cfg.hbm_budget;
	st->n_waiting = s->waiting.size();
	return TM_OK;
}

} // extern "C"
