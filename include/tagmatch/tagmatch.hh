// The TagMatch API (reference tagmatch.hh:11-37), backed by the MI355X
// engine: consolidate() partitions on the GPU and makes the FIB resident in
// HBM; match() / match_unique() queue queries that the engine batches into
// device submissions; results arrive through the query's match_handler.
//
// Threading, as in the reference: add/remove/consolidate are not
// thread-safe; match/match_unique must be called from a single producer
// thread; handlers run on an engine thread; stop() drains all queued
// queries and may be called repeatedly; start()/stop() may repeat.
#ifndef TAGMATCH_AMD_TAGMATCH_HH_INCLUDED
#define TAGMATCH_AMD_TAGMATCH_HH_INCLUDED

#include <vector>

#include "filter.hh"
#include "key.hh"
#include "query.hh"
#include "tagmatch_query.hh"
#include "match_handler.hh"

class tagmatch {
public:
	static void add(const filter_t & set, const std::vector<tagmatch_key_t> & keys);
	static void add(const filter_t & set, tagmatch_key_t key);
	static void remove(const filter_t & set, tagmatch_key_t key);

	static void consolidate();
	static void consolidate(unsigned int partition_size, unsigned int threads);

	// Empty name (the default): no database file is written.  Otherwise
	// consolidate() writes the whole table there as fib_entry records
	// (reference fib.cc:14-35).
	static const char * get_database_filename();
	static void set_database_filename(const char *);

	static void start();
	static void start(unsigned int threads, unsigned int gpu_count);
	static void stop();
	static void clear();

	static unsigned int get_latency_limit_ms();
	static void set_latency_limit_ms(unsigned int latency_limit);

	static void match(tagmatch_query * q, match_handler * h) noexcept;
	static void match_unique(tagmatch_query * q, match_handler * h) noexcept;

	// Extensions (not in the reference interface).
	static int last_status();            // TM_* status of the last engine call
	static size_t bytesize();            // HBM held by the resident FIB (reference back_end::bytesize)
	static unsigned int partition_count();
};

#endif
