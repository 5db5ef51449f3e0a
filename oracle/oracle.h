/*
 * oracle.h -- CPU restatement of the reference TagMatch algorithms.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker of the HIP engine:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * it.  The engine (libtagmatch_amd.so) never links or calls it.
 *
 * Filters are 6 little-endian uint32 words per 192-bit filter, bit p in word
 * p/32 (reference bitvector.hh:104-139, :210-212).
 */
#ifndef TM_ORACLE_H_INCLUDED
#define TM_ORACLE_H_INCLUDED

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* f subset of q: (f & ~q) == 0 on every word (bitvector.hh:282-296). */
int orc_subset(const uint32_t * f, const uint32_t * q);

/* Lexicographic order of the bit strings, bit 0 first
 * (bitvector.hh:389-397): -1, 0, +1. */
int orc_compare(const uint32_t * a, const uint32_t * b);

/* Position of the leftmost differing bit, 192 if equal (bitvector.hh:412-418). */
int orc_leftmost_diff(const uint32_t * a, const uint32_t * b);

/* Parse a string of '0'/'1' (bit 0 first) as the reference's string
 * constructor does (bitvector.hh:122-139). */
void orc_from_string(const char * s, uint32_t * out);

/* Consolidate a list of (filter, key) additions and (filter, key) removals
 * into the canonical table: identical filters merged, key sets union of
 * additions minus removals, empty key sets dropped, filters in lexicographic
 * order, keys ascending (tagmatch.cc:22-77, :98-165).
 *   ops: n_ops records of {6 words, key, op} with op 0 = add, 1 = remove.
 * Outputs are malloc'd; free with orc_free. */
int orc_consolidate(const uint32_t * ops, uint64_t n_ops,
					uint32_t ** words, uint64_t * n_filters,
					uint64_t ** key_off, uint32_t ** keys);

/* Brute-force match (test_matcher.cc:74-81) of nq queries against n filters.
 *   mode 0: sorted unique keys (match_unique, tagmatch_query.hh:99-119)
 *   mode 1: keys of all matched filters, sorted, duplicates kept (match)
 *   mode 2: indices of matched filters, ascending
 * off: nq + 1 offsets (caller-allocated), *vals malloc'd (orc_free). */
int orc_match(const uint32_t * words, uint64_t n, const uint64_t * key_off, const uint32_t * keys,
			  const uint32_t * queries, uint64_t nq, int mode, int threads,
			  uint64_t * off, uint32_t ** vals);

/* Recursive balanced partitioning (partitioning.cc:230-411), single thread,
 * on filters in the given order.  perm[i] = input index at partitioned
 * position i; part_off[np+1]; masks np x 6 (AND of members, all-ones when
 * empty).  Returns -1 on an all-zero filter (the reference does not
 * terminate on one), -2 if max_parts is exceeded. */
int orc_partition(const uint32_t * words, uint32_t n, uint32_t maxp, uint32_t * perm,
				  uint32_t * part_off, uint32_t * masks, uint32_t max_parts, uint32_t * n_parts);

/* Front-end lookup (front_end.cc:421-477): partitions registered in the
 * bucket of their least crowded mask bit; a query scans the buckets of its
 * set bits.  sel receives selected partition ids, ascending; returns count. */
typedef struct orc_lookup orc_lookup;
orc_lookup * orc_lookup_build(const uint32_t * masks, uint32_t np);
uint32_t orc_lookup_query(const orc_lookup * lk, const uint32_t * q, uint32_t * sel);
void orc_lookup_free(orc_lookup * lk);
uint32_t orc_lookup_size(const orc_lookup * lk);   /* partitions */

/* The reference schedule with key output: lookup (front_end.cc:465-477),
 * per-partition brute force (back_end.cc:490-516), then match_unique's sort
 * and dedup (tagmatch_query.hh:105-109) for mode 0; modes as orc_match.
 * words / key_off / keys in partitioned order (part_off[np] filters). */
int orc_schedule_match(const uint32_t * words, const uint64_t * key_off, const uint32_t * keys,
					   const uint32_t * part_off, const uint32_t * masks, uint32_t np, const uint32_t * queries,
					   uint64_t nq, int mode, int threads, uint64_t * off, uint32_t ** vals);

/* Version of the compiler that built this library. */
const char * orc_compiler(void);

/* CPU schedule of the reference back end (front_end.cc:465-477 lookup, then
 * back_end.cc:490-516 per-partition brute force); returns (query, filter)
 * hit count.  Used as a CPU baseline. */
uint64_t orc_schedule_count(const uint32_t * words, const uint32_t * part_off, const uint32_t * masks,
							uint32_t np, const uint32_t * queries, uint64_t nq, int threads);

/* Brute-force (query, filter) hit count, for timing. */
uint64_t orc_bruteforce_count(const uint32_t * words, uint64_t n, const uint32_t * queries,
							  uint64_t nq, int threads);

void orc_free(void * p);

#ifdef __cplusplus
}
#endif

#endif
