// Query types of the TagMatch API (reference query.hh:31-58).
#ifndef TAGMATCH_AMD_QUERY_HH_INCLUDED
#define TAGMATCH_AMD_QUERY_HH_INCLUDED

#include <string>
#include <utility>
#include <vector>

#include "filter.hh"
#include "key.hh"

// A "naked" query: a filter plus the match_unique flag.
class basic_query {
public:
	filter_t filter;
	bool match_unique;

	basic_query() : filter(), match_unique(false) {}
	basic_query(const filter_t & f) : filter(f), match_unique(false) {}
	basic_query(const std::string & f) : filter(f), match_unique(false) {}
	basic_query(const basic_query & q) : filter(q.filter), match_unique(q.match_unique) {}

	std::ostream & write_binary(std::ostream & o) const { return filter.write_binary(o); }
	std::istream & read_binary(std::istream & i) { return filter.read_binary(i); }
};

// A query plus its matching keys.
class query : public basic_query {
public:
	std::vector<tagmatch_key_t> output_keys;

	query() : basic_query() {}
	query(const filter_t & f) : basic_query(f) {}
	query(const std::string & f) : basic_query(f) {}
	query(const query & q) : basic_query(q), output_keys(q.output_keys) {}
	query(query && q) : basic_query(q), output_keys(std::move(q.output_keys)) {}
};

#endif
