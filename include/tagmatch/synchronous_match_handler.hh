// Synchronous result handler (reference synchronous_match_handler.hh:20-50):
// match_done() hands the query to a thread blocked in
// synchronize_and_process(), which then runs process_results().
#ifndef TAGMATCH_AMD_SYNCHRONOUS_MATCH_HANDLER_HH_INCLUDED
#define TAGMATCH_AMD_SYNCHRONOUS_MATCH_HANDLER_HH_INCLUDED

#include <condition_variable>
#include <mutex>

#include "match_handler.hh"

class synchronous_match_handler : public match_handler {
	std::mutex mtx_;
	std::condition_variable cv_;
	query * ready_ = nullptr;

public:
	virtual void process_results(query * q) = 0;

	virtual void match_done(query * q) {
		std::unique_lock<std::mutex> lock(mtx_);
		ready_ = q;
		cv_.notify_all();
	}

	virtual void synchronize_and_process() {
		query * q;
		{
			std::unique_lock<std::mutex> lock(mtx_);
			cv_.wait(lock, [this] { return ready_ != nullptr; });
			q = ready_;
			ready_ = nullptr;
		}
		process_results(q);
	}
};

#endif
