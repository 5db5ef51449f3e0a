// Query with matching state (reference tagmatch_query.hh:25-120).
//
// The engine resolves a query in one device submission, so the per-partition
// bookkeeping of the reference (partition_enqueue / partition_done) is kept
// only for interface compatibility; finalize_matching() sorts and removes
// duplicate keys for match_unique queries and calls the handler once.
#ifndef TAGMATCH_AMD_TAGMATCH_QUERY_HH_INCLUDED
#define TAGMATCH_AMD_TAGMATCH_QUERY_HH_INCLUDED

#include <algorithm>
#include <atomic>
#include <mutex>
#include <vector>

#include "match_handler.hh"
#include "query.hh"

class tagmatch_query : public query {
	enum : unsigned char { FRONT_END = 0, BACK_END = 1, FINALIZED = 2 };
	std::atomic<unsigned char> state_;
	std::atomic<unsigned int> pending_partitions_;
	match_handler * handler_;
	std::mutex output_mtx_;

public:
	tagmatch_query() : query(), state_(FRONT_END), pending_partitions_(0), handler_(nullptr) {}
	tagmatch_query(const filter_t & f) : query(f), state_(FRONT_END), pending_partitions_(0), handler_(nullptr) {}
	tagmatch_query(const std::string & f) : query(f), state_(FRONT_END), pending_partitions_(0), handler_(nullptr) {}
	tagmatch_query(const tagmatch_query & x)
		: query(x), state_(x.state_.load()), pending_partitions_(x.pending_partitions_.load()), handler_(x.handler_) {}
	tagmatch_query(tagmatch_query && x)
		: query(std::move(x)), state_(x.state_.load()), pending_partitions_(x.pending_partitions_.load()),
		  handler_(x.handler_) {
		x.handler_ = nullptr;
	}

	void partition_enqueue() { ++pending_partitions_; }
	void partition_done() { --pending_partitions_; }
	void reset() {
		state_ = FRONT_END;
		pending_partitions_ = 0;
	}
	void frontend_done() {
		unsigned char s = FRONT_END;
		state_.compare_exchange_strong(s, BACK_END);
	}

	void add_output(tagmatch_key_t key) {
		std::lock_guard<std::mutex> lock(output_mtx_);
		output_keys.push_back(key);
	}
	void add_output(std::vector<tagmatch_key_t>::const_iterator k, std::vector<tagmatch_key_t>::const_iterator end) {
		std::lock_guard<std::mutex> lock(output_mtx_);
		output_keys.insert(output_keys.end(), k, end);
	}
	void add_output(const tagmatch_key_t * k, const tagmatch_key_t * end) {
		std::lock_guard<std::mutex> lock(output_mtx_);
		output_keys.insert(output_keys.end(), k, end);
	}

	void set_match_handler(match_handler * h) { handler_ = h; }
	match_handler * get_match_handler() const { return handler_; }

	// Called once all the query's results are in (reference
	// tagmatch_query.hh:99-119): dedup for match_unique, then the handler.
	bool finalize_matching() {
		if (pending_partitions_ > 0) return false;
		unsigned char s = BACK_END;
		if (!state_.compare_exchange_strong(s, FINALIZED)) return false;
		if (match_unique) {
			if (!std::is_sorted(output_keys.begin(), output_keys.end()))
				std::sort(output_keys.begin(), output_keys.end());
			output_keys.erase(std::unique(output_keys.begin(), output_keys.end()), output_keys.end());
		}
		if (handler_) handler_->match_done(this);
		return true;
	}
};

#endif
