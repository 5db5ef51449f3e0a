/*
 * tagmatch_amd.h -- C ABI of the MI355X (gfx950) TagMatch matching engine.
 *
 * Plain C linkage, plain pointers and sizes, int status codes, no exceptions.
 * This is the device boundary that replaces the reference's CUDA wrapper
 * class `gpu` (reference gpu.hh:32-77) and the device-facing half of its
 * back end (reference back_end.hh:19-39).  The C++ facade `class tagmatch`
 * (include/tagmatch/tagmatch.hh) is built on top of it; other languages bind
 * these symbols directly (ctypes, cgo, JNI, N-API).
 *
 * Data conventions
 *   filter / query : 6 x uint32 little-endian words; bit p of the 192-bit
 *                    filter is bit (p % 32) of word (p / 32).  This is the
 *                    reference's bitvector<192> memory image read as uint32
 *                    (reference bitvector.hh:104-139, :210-212).
 *   key            : uint32 (reference key.hh:6).
 *   subset test    : f is matched by q iff (f[j] & ~q[j]) == 0 for j < 6
 *                    (reference bitvector.hh:290-296, gpu.cu:48-50).
 *
 * Ownership: the caller owns every host buffer passed in and keeps it valid
 * until the call returns (synchronous calls) or until tm_wait() reports the
 * submission finished (asynchronous calls).  The library owns all device
 * memory it allocates.  A tm_fib or tm_ring is used by one host thread at a
 * time; distinct rings may be driven from distinct threads.
 */
#ifndef TAGMATCH_AMD_H_INCLUDED
#define TAGMATCH_AMD_H_INCLUDED

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Status codes.  The reference prints and calls exit(0) on any CUDA error
 * (reference gpu.cu:14-24); every entry point here returns one of these. */
#define TM_OK             0
#define TM_ERR_INVALID   -1   /* bad argument or size                        */
#define TM_ERR_HIP       -2   /* HIP runtime error; see tm_last_error()      */
#define TM_ERR_NOMEM     -3   /* device or pinned allocation failed          */
#define TM_ERR_OVERFLOW  -4   /* result capacity exceeded; retry larger      */
#define TM_ERR_STATE     -5   /* call out of sequence (no device, no FIB)    */
#define TM_ERR_NODEV     -6   /* no HIP device / library built without one   */
#define TM_ERR_BUSY      -7   /* submission still running (tm_poll)          */

#define TM_FILTER_WORDS   6
#define TM_BLOCK_FILTERS  64  /* filters per device block (one per lane)    */

/* Thread-local text of the last error. */
const char * tm_last_error(void);
const char * tm_version(void);

/* ------------------------------------------------------------------------ */
/* Device lifecycle.
 * Replaces gpu::initialize / gpu::set_device / gpu::shutdown
 * (reference gpu.cu:778-794, :914-922) and gpu::mem_info (gpu.cu:796-807). */
int tm_dev_init(int device);
int tm_dev_shutdown(void);
int tm_dev_count(int * count);
int tm_mem_info(size_t * free_bytes, size_t * total_bytes);
int tm_dev_synchronize(void);
/* Name of the active device's architecture (e.g. "gfx950"). */
int tm_dev_arch(char * buf, size_t len);
/* Device memory for callers without the HIP headers (bindings, tests):
 * allocate, free, and synchronous copies to and from it.  The reference
 * allocates through its gpu class (gpu.cu:825-830). */
int tm_dev_alloc(void ** ptr, uint64_t bytes);
int tm_dev_free(void * ptr);
int tm_dev_upload(void * dev_dst, const void * host_src, uint64_t bytes);
int tm_dev_download(void * host_dst, const void * dev_src, uint64_t bytes);

/* ------------------------------------------------------------------------ */
/* Partitioner: the balanced partitioning of a filter table, run on the GPU.
 * Replaces partitioning::balanced_partitioning + partitioner_gpu
 * (reference partitioning.cc:361-411, partitioner_gpu.cu:15-158) with a
 * level-synchronous restatement of the same split rules: bit histograms,
 * pivot choice and stable pivot splits all run as HIP kernels.
 *
 *   words      : n filters (n x 6 words), usually lexicographically sorted.
 *   maxp       : maximum partition size (reference partitioning.cc:342-347).
 *   perm       : out, n entries; filter perm[i] of the input is at position i
 *                of the partitioned order.  Each partition is a contiguous
 *                run and keeps the input's relative order.
 *   part_offsets: out, capacity max_parts + 1; partition p is
 *                [part_offsets[p], part_offsets[p+1]).
 *   part_masks : out, capacity max_parts x 6; AND of the partition's filters
 *                (all-ones for an empty partition, reference
 *                partitioning.cc:389-402).  A partition of all-zero filters
 *                has an all-zero mask; the reference cannot terminate on it
 *                (partitioning.cc:230-250), this engine keeps it as a
 *                partition that every query selects.
 *   n_parts    : out, number of partitions.
 * Returns TM_ERR_OVERFLOW if more than max_parts partitions result. */
int tm_partition(const uint32_t * words, uint32_t n, uint32_t maxp,
				 uint32_t * perm, uint32_t * part_offsets, uint32_t * part_masks,
				 uint32_t max_parts, uint32_t * n_parts);

/* ------------------------------------------------------------------------ */
/* Device FIB: filters, partition masks and keys resident in HBM.
 * Replaces back_end::compile_fibs and its per-partition cudaMalloc/copy
 * (reference back_end.cc:363-453, gpu.cu:825-830); the keys, which the
 * reference keeps on the host (back_end.cc:69-88), are resident too. */
typedef struct tm_fib tm_fib;

typedef struct tm_fib_stats {
	uint64_t n_filters;      /* filters uploaded                           */
	uint64_t n_blocks;       /* 64-filter device blocks (partitions padded) */
	uint64_t n_parts;        /* partitions                                 */
	uint64_t n_keys;         /* key instances (sum of key-set sizes)       */
	uint64_t n_distinct_keys;/* size of the key dictionary                 */
	uint64_t device_bytes;   /* HBM held by this FIB                       */
	uint64_t max_part_size;  /* largest partition, filters                 */
} tm_fib_stats;

int tm_fib_create(tm_fib ** out);
int tm_fib_destroy(tm_fib * fib);

/* Upload a partitioned table (replaces any previous content).
 *   words        : n x 6 words, grouped by partition.
 *   part_offsets : n_parts + 1 entries, partition p = [off[p], off[p+1]).
 *   part_masks   : n_parts x 6 words; a partition is sent a query q iff
 *                  mask & ~q == 0 (reference front_end.cc:465-477).  Must be
 *                  a subset of every member filter.
 *   key_offsets  : n + 1 entries; keys of filter i = keys[ko[i] .. ko[i+1]).
 *   keys         : key values (any uint32).
 *   filter_ids   : n ids reported by TM_MODE_HITS (NULL: position i).
 * Within each partition the engine keeps the given order; sorted order makes
 * the per-block masks (AND of 64 neighbours) most selective. */
int tm_fib_upload(tm_fib * fib, const uint32_t * words, uint32_t n,
				  const uint32_t * part_offsets, const uint32_t * part_masks, uint32_t n_parts,
				  const uint32_t * key_offsets, const uint32_t * keys,
				  const uint32_t * filter_ids);
int tm_fib_stats_get(const tm_fib * fib, tm_fib_stats * st);

/* Versioned image of a resident table: blocks, key slots, overflow lists,
 * dictionary, partitions and rich filters, as uploaded and inserted so far.
 * Loading replaces the FIB's content without partitioning or laying out
 * blocks again.  Replaces the reference's tagmatch.db round trip
 * (tagmatch.cc:79-87, :98-134, written as partition_fib_entry, read back as
 * fib_entry), which re-partitions on every start.  Little-endian hosts;
 * a file of another image version is refused. */
int tm_fib_save(tm_fib * fib, const char * path);
int tm_fib_load(tm_fib * fib, const char * path);

/* Spare blocks reserved per partition at upload for tm_fib_insert, in
 * percent of its used blocks (plus one block); default 13, 0 = none.  With a
 * reserve, the upload also adds one empty partition with mask 0 (selected
 * by every query, room for about 1% of the table) for inserted filters that
 * contain no partition's mask; tm_fib_stats_get counts it in n_parts.
 * Applies to the next tm_fib_upload. */
int tm_fib_set_reserve(tm_fib * fib, uint32_t percent);

/* Insert n filters into a resident table without rebuilding it.  The
 * reference adds filters on the host and rebuilds the whole FIB
 * (tagmatch::add + consolidate, tagmatch.cc:22-77, :98-165, then
 * back_end::compile_fibs, back_end.cc:363-453).  Here each filter joins the
 * existing partition with the most mask bits among those whose mask it
 * contains and that have a spare lane; filters that fit none are
 * partitioned on their own with maxp and appended as new partitions.  Keys
 * new to the table extend its dictionary.  Filter ids default to the
 * running count of filters (the first upload's n, then n + each insert).
 * A filter identical to one already present is kept as its own entry: key
 * results are the same as after merging, TM_MODE_HITS reports both ids.
 * Returns TM_ERR_BUSY while a ring on this FIB has a submission that
 * tm_wait has not collected; rings pick the change up at their next
 * submission. */
int tm_fib_insert(tm_fib * fib, const uint32_t * words, uint32_t n, const uint32_t * key_offsets,
				  const uint32_t * keys, const uint32_t * filter_ids, uint32_t maxp);

/* Insert while submissions are in flight.  Only the fast case: every
 * filter takes a spare lane of an existing partition and every key is
 * already in the table (else TM_ERR_BUSY, nothing changed: collect the
 * submissions and call tm_fib_insert).  Each insert starts a new epoch; a
 * submission sees exactly the inserts that completed before it was made,
 * in flight or not: lanes carry their epoch and the match skips later ones.
 * Not to be called concurrently with tm_fib_insert / tm_fib_upload on the
 * same FIB, nor with a submission being made on it. */
int tm_fib_insert_live(tm_fib * fib, const uint32_t * words, uint32_t n, const uint32_t * key_offsets,
					   const uint32_t * keys, const uint32_t * filter_ids);

/* ------------------------------------------------------------------------ */
/* Matching.
 *
 * A ring is one in-flight submission context on its own HIP stream: a query
 * buffer, the device-side lookup lists and work queue, and output buffers.
 * Rings replace the reference's stream_handle double buffers
 * (reference back_end.cc:131-311) and its 256-query __constant__ upload
 * (gpu.cu:39-40, :809-811).  Several rings may be in flight at once. */
typedef struct tm_ring tm_ring;

/* Output modes. */
#define TM_MODE_KEYS_UNIQUE 0  /* per query: sorted, deduplicated keys
                                  (match_unique, reference
                                  tagmatch_query.hh:99-119)                  */
#define TM_MODE_KEYS_ALL    1  /* per query: keys of every matched filter,
                                  duplicates kept, order unspecified (match,
                                  reference tagmatch.cc:200-204)             */
#define TM_MODE_HITS        2  /* per query: ids of matched filters, order
                                  unspecified                                */
#define TM_MODE_KEY_BITMAP  3  /* per query: a row of ceil(K/32) uint32 words
                                  over the FIB's sorted key dictionary
                                  (tm_fib_key_dictionary, K keys): bit i
                                  (word i/32, bit i%32) is set iff dictionary
                                  key i is one of the query's keys.  The set
                                  bits in increasing order are exactly the
                                  KEYS_UNIQUE list (match_unique, reference
                                  tagmatch_query.hh:99-119) in a fixed-size
                                  form; read with tm_read_key_bitmaps or
                                  delivered by tm_ring_set_host_rows, expanded
                                  with tm_expand_key_bitmaps                 */

/* Schedules. */
#define TM_SCHED_PLAIN      0  /* one kernel launch per (partition, batch),
                                  the reference's dispatch shape
                                  (back_end.cc:243-286, gpu.cu:878-912)      */
#define TM_SCHED_PERSISTENT 1  /* one persistent, work-stealing launch over
                                  every (partition, query tile, block chunk) */

typedef struct tm_match_config {
	uint32_t mode;           /* TM_MODE_*                                   */
	uint32_t schedule;       /* TM_SCHED_*                                  */
	uint32_t batch_size;     /* PLAIN: queries per (partition, batch) launch,
	                            0 = 256 (reference parameters.hh:22);
	                            PERSISTENT: queries per work item (64, 128
	                            or 256), 0 = 64                              */
	uint32_t chunk_blocks;   /* PLAIN: least blocks per segment of a unit
	                            (0 = 64); PERSISTENT: blocks per segment of
	                            the table (0 = automatic: at most 1024, and
	                            smaller for submissions too small to fill
	                            the GPU)                                     */
	uint32_t waves_per_group;/* waves per workgroup (4, 8 or 16); 0 = by
	                            step size for PERSISTENT (16 up to 16 query
	                            tiles, 8 up to 256, else 4), 8 for PLAIN     */
	uint32_t groups_per_cu;  /* PERSISTENT: resident workgroups per CU; 0 = auto */
	uint64_t hit_capacity;   /* HITS / KEYS_ALL: hit slots; 0 = default      */
	uint32_t use_graph;      /* PERSISTENT: replay steady-state steps from a
	                            captured hipGraph                            */
	uint32_t log_units;      /* record per-work-item hit counts and cycles   */
	uint32_t order_by_cost;  /* PERSISTENT: visit partitions in the order set
	                            by tm_ring_set_order                          */
	uint32_t variant;        /* PERSISTENT inner loop, 1..7, 10 or 18, as
	                            bits; bit 0: candidate query words from
	                            scalar loads (else LDS broadcasts); bit 1:
	                            KEYS_UNIQUE hits go through a per-wave key
	                            log drained 64 lanes at a time; bit 2: early
	                            exit after words 0-1; bit 3: the next
	                            block's data is loaded while one is tested;
	                            bit 4: blocks in rounds of 4, whose loads
	                            overlap; 0 or other values: default (2)      */
	uint32_t full_exit;      /* PERSISTENT KEYS_UNIQUE: stop testing a query
	                            once its key set holds every key of the
	                            table (results unchanged); 0 = default (on),
	                            1 = on, 2 = off                              */
	uint32_t compact_tiles;  /* PERSISTENT KEYS_UNIQUE / KEY_BITMAP: steps of
	                            more query tiles than this match only the
	                            queries the rich-filter pass leaves
	                            incomplete, packed into tiles of their own;
	                            0 = default (128)                            */
	uint32_t reserved[4];
} tm_match_config;

void tm_match_config_default(tm_match_config * cfg);

typedef struct tm_status {
	uint32_t n_queries;      /* queries in the submission                   */
	uint32_t n_items;        /* work items (PERSISTENT) or launches (PLAIN)  */
	uint64_t n_pairs;        /* (query, selected partition) pairs           */
	uint64_t n_hits;         /* (query, filter) matches                     */
	uint64_t n_results;      /* values written to the result array          */
	uint64_t hit_capacity;   /* hit slots available in HITS / KEYS_ALL mode */
	int32_t  overflow;       /* 1 if hits exceeded hit_capacity             */
	int32_t  error;          /* TM_* status of the submission               */
	double   device_ms;      /* stream time from first to last kernel       */
	double   match_ms;       /* time of the match kernel(s) alone           */
	uint64_t n_block_visits; /* PERSISTENT: (query group, block) visits with
	                            at least one candidate query                */
	uint64_t n_cand_tests;   /* PERSISTENT: 64-filter subset tests run
	                            (one per candidate query per block)          */
	uint64_t n_hit_events;   /* PERSISTENT: subset tests with >= 1 hit       */
	double   pred_hits;      /* cost model: predicted hits of the submission
	                            (0 without a model)                          */
	double   model_ms;       /* cost model: host time spent pricing and
	                            ordering the submission                      */
} tm_status;

/* Create a ring for up to max_queries queries per submission. */
int tm_ring_create(tm_fib * fib, uint32_t max_queries, const tm_match_config * cfg,
				   tm_ring ** out);
int tm_ring_destroy(tm_ring * ring);
int tm_ring_config(tm_ring * ring, const tm_match_config * cfg);

/* Autotuner: run the nq queries already in the ring (a previous submission
 * or tm_ring_query_buffer) under each candidate configuration (same mode),
 * reps timed steps each after one warm-up, and keep the fastest by device
 * time; *best = its index, times_ms (may be NULL) = each candidate's best
 * step time.  Results do not depend on the configuration.  No counterpart
 * in the reference, whose launch shape is fixed (gpu.cu:878-912). */
int tm_ring_autotune(tm_ring * ring, uint32_t nq, const tm_match_config * cands, uint32_t n_cands, uint32_t reps,
					 uint32_t * best, double * times_ms);

/* Work-item size by partition-size bucket, for explicit units
 * (tm_submit_units, persistent schedule): a unit whose partition has nb
 * blocks (of 64 filters) is cut into work items of seg[bucket(nb)] blocks,
 * bucket(nb) = 0 below 16 blocks, 1 below 64, 2 below 256, 3 below 1024,
 * 4 below 4096, 5 from 4096; 0 = the automatic size (the submission's
 * tile-blocks spread over about four items per resident workgroup).  The
 * reference runs one fixed launch shape per (partition, batch)
 * (gpu.cu:878-912). */
#define TM_SIZE_BUCKETS 6
int tm_ring_set_bucket_segments(tm_ring * ring, const uint32_t * seg, uint32_t n);
int tm_ring_get_bucket_segments(const tm_ring * ring, uint32_t * seg, uint32_t n);
/* Tune that table on the caller's units (queries already in the ring):
 * bucket by bucket, each of the n_cands candidate sizes is run (reps timed
 * runs each after one warm-up) with the other buckets at their best so far,
 * and the fastest by device time is kept.  *base_ms: the step with the
 * table as it was; *best_ms: with the tuned table, which stays set.
 * Results do not depend on the table. */
int tm_ring_autotune_buckets(tm_ring * ring, const uint32_t * unit_part, const uint32_t * unit_off, uint32_t n_units,
							 const uint32_t * cands, uint32_t n_cands, uint32_t reps, double * base_ms,
							 double * best_ms);

/* Device pointer of the ring's query buffer (max_queries x 8 words: 6 filter
 * words + 2 pad), for callers that keep queries resident in HBM. */
int tm_ring_query_buffer(tm_ring * ring, uint32_t ** dev_words);

/* Submit nq queries (host words, nq x 6): asynchronous H2D copy + lookup +
 * match + key expansion/dedup on the ring's stream.  Returns immediately. */
int tm_submit(tm_ring * ring, const uint32_t * query_words, uint32_t nq);
/* Same, with the queries already in tm_ring_query_buffer(). */
int tm_submit_resident(tm_ring * ring, uint32_t nq);
/* Submit explicit (partition, query range) units instead of the lookup:
 * unit u matches queries [unit_off[u], unit_off[u+1]) against partition
 * unit_part[u] only (its mask still prunes, which is exact).  unit_off has
 * n_units + 1 entries from 0 to nq.  query_words: nq x 6 host words, or
 * NULL for the queries already in tm_ring_query_buffer().  This is the
 * reference's back-end entry, a batch of queries for one partition
 * (back_end::process_batch, back_end.cc:243-311), without the front end;
 * results per query as for tm_submit. */
int tm_submit_units(tm_ring * ring, const uint32_t * query_words, uint32_t nq, const uint32_t * unit_part,
					const uint32_t * unit_off, uint32_t n_units);
/* TM_OK when the last submission finished, TM_ERR_BUSY while running. */
int tm_poll(tm_ring * ring);
/* Block until the last submission finished; fill *st (may be NULL). */
int tm_wait(tm_ring * ring, tm_status * st);
/* Copy results of the finished submission to the host:
 *   offsets : nq + 1 entries; results of query i = values[off[i] .. off[i+1])
 *   values  : capacity entries (may be NULL to read offsets only).
 * TM_ERR_OVERFLOW if capacity is too small (offsets are still written). */
int tm_read_results(tm_ring * ring, uint64_t * offsets, uint32_t * values, uint64_t capacity);
/* Device pointers of the results of the finished submission. */
int tm_ring_results_device(tm_ring * ring, const uint64_t ** dev_offsets,
						   const uint32_t ** dev_values);

/* Key bitmap rows of the finished submission (KEYS_UNIQUE or KEY_BITMAP
 * rings): nq rows of *kw words each (kw = ceil(K/32)) into rows, which
 * holds capacity_words.  TM_ERR_OVERFLOW if nq * kw > capacity_words (*kw
 * is still set); TM_ERR_STATE for the other modes.  With the reference's
 * result delivery (back_end.cc:288-301, tagmatch_query.hh:99-119) the keys
 * reach the host one hit at a time; a row is a fixed 4 * kw bytes. */
int tm_read_key_bitmaps(tm_ring * ring, uint32_t * rows, uint64_t capacity_words, uint32_t * kw);

/* Host delivery for KEY_BITMAP rings: every later submission ends, on the
 * ring's stream, with its rows copied into host_rows (nq * kw words), so
 * that tm_poll / tm_wait report them already in host memory.  host_rows
 * should be pinned (tm_host_alloc) for the copy to overlap other rings'
 * work; it must stay valid until tm_wait of each submission returns.  A
 * submission with nq * kw > capacity_words fails with TM_ERR_OVERFLOW.
 * host_rows NULL stops the delivery. */
int tm_ring_set_host_rows(tm_ring * ring, uint32_t * host_rows, uint64_t capacity_words);

/* Packed host delivery for KEY_BITMAP rings: instead of every row, each
 * later submission delivers row_of (one entry per query: TM_ROW_FULL when
 * the query's key set is the whole dictionary, else the index of its row in
 * `rows`) and, packed in query order, the rows of the queries whose key set
 * is not the whole dictionary.  Lossless: a TM_ROW_FULL query's row is all
 * K bits.  row_of and rows as for tm_ring_set_host_rows (pinned, valid until
 * tm_wait returns; nq entries and nq * kw words of capacity are required).
 * The entries and part of the rows are copied on the ring's stream; rows
 * beyond that are copied by tm_wait, so they are complete when tm_wait
 * returns (tm_poll reporting completion is not enough).  tm_status.n_results
 * is the number of packed rows.  Setting either host delivery form replaces
 * the other; row_of NULL stops it.  The reference delivers each query's keys
 * through its match handler (back_end.cc:288-301, tagmatch_query.hh:99-119);
 * this form moves 4 bytes per query whose keys are every key. */
#define TM_ROW_FULL 0xFFFFFFFFu
int tm_ring_set_host_rows_packed(tm_ring * ring, uint32_t * row_of, uint64_t row_of_capacity, uint32_t * rows,
								 uint64_t rows_capacity_words);

/* Pinned (page-locked) host memory for result delivery, and its release. */
int tm_host_alloc(void ** ptr, uint64_t bytes);
int tm_host_free(void * ptr);

/* The FIB's sorted distinct keys, the dictionary KEY_BITMAP rows index:
 * *n = K; keys (capacity entries, may be NULL to read K only). */
int tm_fib_key_dictionary(const tm_fib * fib, uint32_t * keys, uint64_t capacity, uint64_t * n);

/* Expand nq KEY_BITMAP rows (kw words each) over dict (n_keys sorted keys)
 * into KEYS_UNIQUE lists on the host: offsets (nq + 1 entries) and values
 * (capacity entries; NULL to compute offsets only).  Runs on `threads` host
 * threads (0: one per core).  TM_ERR_OVERFLOW if the keys exceed capacity
 * (offsets are still written).  Bits at or above n_keys are ignored.  This
 * is the host half of the reference's add_output + finalize_matching
 * (tagmatch_query.hh:80-119) for a whole submission at once. */
int tm_expand_key_bitmaps(const uint32_t * rows, uint64_t nq, uint32_t kw, const uint32_t * dict, uint32_t n_keys,
						  uint64_t * offsets, uint32_t * values, uint64_t capacity, uint32_t threads);
/* The same from packed delivery (tm_ring_set_host_rows_packed): row_of[q]
 * TM_ROW_FULL gives the whole dictionary, else rows[row_of[q]] (n_rows rows
 * of kw words); an index at or above n_rows is TM_ERR_INVALID. */
int tm_expand_key_rows_packed(const uint32_t * row_of, const uint32_t * rows, uint64_t n_rows, uint64_t nq,
							  uint32_t kw, const uint32_t * dict, uint32_t n_keys, uint64_t * offsets,
							  uint32_t * values, uint64_t capacity, uint32_t threads);

/* Per-partition query counts of the finished submission (the on-device
 * lookup's result; the reference's per-partition queue fill,
 * front_end.cc:350-376).  counts: n_parts entries. */
int tm_ring_lookup_counts(tm_ring * ring, uint32_t * counts, uint32_t n_parts);

/* No-allocation match.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  It runs the persistent
 * KEY_BITMAP step of a ring -- lookup, rich-filter pass, compacted match --
 * over nq queries already in device memory (dev_queries: nq x 8 words, the
 * 6 filter words and 2 pad words, the layout of tm_ring_query_buffer) and
 * writes their key bitmap rows to device memory (dev_rows: nq x
 * ceil(K/32) words), on `stream` (a hipStream_t; NULL: the default stream),
 * and returns when they are written.  It allocates nothing: its working
 * buffers live in a fixed static device array of 0.88 MiB per GPU,
 * which bounds nq by TM_STATIC_MAX_QUERIES and the FIB's partitions by
 * TM_STATIC_MAX_PARTS (larger sizes: TM_ERR_INVALID).  Calls on one GPU share
 * that array: calls that overlap on one GPU (from several threads) race.
 * The FIB must not be changed by tm_fib_insert / upload / load during a call
 * (live inserts are safe: the call sees the inserts made before it). */
#define TM_STATIC_MAX_QUERIES 65536
#define TM_STATIC_MAX_PARTS   16384
int tm_match_rows_static(tm_fib * fib, const uint32_t * dev_queries, uint32_t nq, uint32_t * dev_rows,
						 void * stream);

/* One-shot synchronous match through a temporary ring. */
int tm_match(tm_fib * fib, const uint32_t * query_words, uint32_t nq,
			 const tm_match_config * cfg, uint64_t * offsets, uint32_t * values,
			 uint64_t capacity, tm_status * st);

/* ------------------------------------------------------------------------ */
/* Learned cost model.  No counterpart in the reference, whose scheduler has
 * no cost estimate (front_end.cc:350-376 flushes batches in arrival order and
 * back_end.cc:243-286 sizes every result buffer for the worst case).
 *
 * Work units are (partition, query batch) pairs.  The plain schedule logs one
 * row per launch when cfg.log_units is set: TM_UNIT_FEATURES float features,
 * the unit's hits, the key instances of its hit filters and its summed
 * workgroup time in 10 ns ticks.  A TM_UNIT_FEATURES-64-64-3 multilayer
 * perceptron (FP32, ReLU) trained on those rows with targets
 * log2(1 + x) / 16 of (hits, keys, ticks) prices the persistent schedule's
 * (query tile, partition) units.  Its
 * predictions pre-size result buffers and order work items; they never
 * skip work, and overflow-retry still catches an under-estimate.
 *
 * Feature order: log2(1+filters)/20, popcount(mask)/64, log2(1+B)/12, query
 * popcount mean, std, min, max (/192), fractions of queries with popcount in
 * [0,80) [80,110) [110,140) [140,192], mean member popcount/64,
 * popcount(OR of members)/192, mean and max popcount(query & OR)/192,
 * log2(1 + B * filters)/32. */
#define TM_UNIT_FEATURES 16

typedef struct tm_mlp tm_mlp;

/* n_in -> n_hidden -> n_hidden -> n_out, He-initialised from seed.  The
 * parameters are one FP32 array: W1[n_in][h], b1[h], W2[h][h], b2[h],
 * W3[h][n_out], b3[n_out]. */
int tm_mlp_create(uint32_t n_in, uint32_t n_hidden, uint32_t n_out, uint64_t seed, tm_mlp ** out);
int tm_mlp_destroy(tm_mlp * m);
uint64_t tm_mlp_param_count(const tm_mlp * m);
int tm_mlp_get_params(const tm_mlp * m, float * params, uint64_t n);
int tm_mlp_set_params(tm_mlp * m, const float * params, uint64_t n);
/* y (n x n_out) = network(x (n x n_in)), host buffers. */
int tm_mlp_forward(tm_mlp * m, const float * x, uint32_t n, float * y);
/* Mean squared error over all n x n_out outputs and its gradient with
 * respect to every parameter (grad: tm_mlp_param_count floats, may be NULL). */
int tm_mlp_loss_grad(tm_mlp * m, const float * x, const float * t, uint32_t n, float * loss, float * grad);
/* Adam (0.9, 0.999, 1e-8) on shuffled minibatches, all on the GPU; *loss =
 * mean minibatch loss of the last epoch. */
int tm_mlp_train(tm_mlp * m, const float * x, const float * t, uint32_t n, uint32_t epochs, uint32_t batch,
				 float lr, uint64_t seed, float * loss);

/* Unit log of the ring's last plain submission (cfg.log_units): up to
 * capacity rows of features (TM_UNIT_FEATURES floats), hits, key instances
 * of the hit filters, ticks and the unit itself (3 words: partition, batch
 * index, queries in the batch); *n_units = rows available.  Any output may
 * be NULL. */
int tm_ring_unit_log(tm_ring * ring, float * features, uint32_t * hits, uint32_t * keys, uint32_t * ticks,
					 uint32_t * units, uint32_t capacity, uint32_t * n_units);

/* Use a trained TM_UNIT_FEATURES-64-64-3 model in the ring's PERSISTENT
 * submissions (NULL: none).  The model must outlive its use by the ring. */
#define TM_COST_PRESIZE 1u  /* grow result buffers to the predicted size first */
#define TM_COST_ORDER   2u  /* visit query tiles longest-predicted first        */
int tm_ring_set_cost_model(tm_ring * ring, tm_mlp * model, uint32_t flags);

/* ------------------------------------------------------------------------ */
/* Serving loop.  Replaces the reference front end's batching and latency
 * flush (front_end.cc:249-267, :350-376, :563-598) and its per-GPU double
 * buffering (back_end.cc:131-311) with continuous batching: an engine thread
 * admits waiting requests (batches of queries) into free slots (rings, each
 * on its own HIP stream) as soon as a slot is free and the requests' result
 * bytes fit the HBM budget, so steps of several slots overlap on the
 * device.  A request is never split: its queries run in one device step,
 * possibly together with other requests.  Inserts queue in the same order
 * as requests.  One that fits spare lanes of existing partitions applies at
 * once (tm_fib_insert_live): the steps in flight do not see it, the
 * requests behind it do.  Any other waits until every earlier request has
 * completed (tm_fib_insert).  Either way each request sees exactly the
 * inserts submitted before it.  No counterpart in the reference for the budget or inserts:
 * its add/consolidate run only while the engine is stopped. */
typedef struct tm_server tm_server;

typedef struct tm_server_config {
	uint32_t n_slots;        /* rings in flight (0 = 4)                       */
	uint32_t slot_queries;   /* queries per device step and largest request
	                            (0 = 16384)                                   */
	uint64_t hbm_budget;     /* result bytes admitted at once (0 = a quarter
	                            of the free HBM at creation)                  */
	/* Latency flush (front_end.cc:249-267, :563-598: a batch is flushed when
	 * full or when its oldest query passes the latency limit).  With
	 * flush_queries > 0 a free slot waits until min(flush_queries,
	 * slot_queries) queries are waiting, or until the oldest waiting request
	 * has waited flush_us microseconds, or an insert is queued behind them,
	 * or the server stops.  0 (the default) starts a step at once. */
	uint32_t flush_queries;
	uint32_t flush_us;
	/* Steps replayed from captured HIP graphs, one per power-of-two step
	 * size (persistent schedule; requests padded with zero queries): 0 =
	 * default (on), 1 = on, 2 = off. */
	uint32_t step_graphs;
	uint32_t reserved[5];
} tm_server_config;

typedef struct tm_server_stats {
	uint64_t n_requests;          /* submitted                               */
	uint64_t n_requests_done;
	uint64_t n_queries_done;
	uint64_t n_batches;           /* device steps started                    */
	uint64_t n_steps;             /* device steps completed without error    */
	uint64_t n_retries;           /* steps that needed an overflow retry     */
	uint64_t n_inserts;
	uint64_t n_live_inserts;      /* inserts applied beside the steps in flight */
	uint64_t n_budget_waits;      /* admissions deferred by the HBM budget   */
	uint64_t requests_per_batch_sum;
	uint64_t reserved_peak;       /* largest budget use, bytes               */
	uint64_t reserved_now;
	uint64_t hbm_budget;
	uint64_t n_waiting;           /* requests and inserts not yet admitted   */
	double   latency_sum_ms;      /* submit -> results on the host           */
	double   latency_max_ms;
	double   device_ms_sum;       /* device time of the completed steps      */
} tm_server_stats;

typedef struct tm_request_info {
	int32_t  status;              /* TM_* of the request                     */
	uint32_t n_queries;
	uint64_t n_values;            /* result values (offsets has n_queries+1) */
	uint32_t step_queries;        /* queries of the device step it ran in    */
	uint32_t pad;
	double   latency_ms;          /* submit -> results on the host           */
	double   queue_ms;            /* submit -> admitted                      */
} tm_request_info;

void tm_server_config_default(tm_server_config * cfg);
/* mcfg: match configuration of every slot (NULL: tm_match_config_default). */
int tm_server_create(tm_fib * fib, const tm_server_config * cfg, const tm_match_config * mcfg, tm_server ** out);
/* Drains every queued request and insert, then stops. */
int tm_server_destroy(tm_server * srv);
/* Price the slots' steps with a cost model (tm_ring_set_cost_model); waits
 * until the server is idle. */
int tm_server_set_cost_model(tm_server * srv, tm_mlp * model, uint32_t flags);
/* Queue a request of nq queries (copied); returns at once with its id. */
int tm_server_submit(tm_server * srv, const uint32_t * queries, uint32_t nq, uint64_t * id);
/* Queue an insert (arguments as tm_fib_insert, copied); completes like a
 * request, with tm_server_result's status = tm_fib_insert's. */
int tm_server_insert(tm_server * srv, const uint32_t * words, uint32_t n, const uint32_t * key_offsets,
					 const uint32_t * keys, const uint32_t * filter_ids, uint32_t maxp, uint64_t * id);
/* Block until request id completed (timeout_ms < 0: no limit); returns its
 * status, or TM_ERR_BUSY on timeout. */
int tm_server_wait(tm_server * srv, uint64_t id, double timeout_ms);
/* Ids of requests completed since the last call, oldest first. */
int tm_server_collect(tm_server * srv, uint64_t * ids, uint32_t capacity, uint32_t * n);
/* Results of a completed request: offsets (n_queries + 1) and values as in
 * tm_read_results; any output may be NULL. */
int tm_server_result(tm_server * srv, uint64_t id, tm_request_info * info, uint64_t * offsets, uint32_t * values,
					 uint64_t capacity);
/* Forget a completed request. */
int tm_server_release(tm_server * srv, uint64_t id);
int tm_server_stats_get(tm_server * srv, tm_server_stats * st);

/* ------------------------------------------------------------------------ */
/* Host: canonical filter table from (filter, keys) entries, as
 * tagmatch::add + consolidate build it (reference tagmatch.cc:22-77,
 * :98-165): identical filters merged with the union of their keys, filters
 * in the reference's lexicographic order (bitvector.hh:389-397), keys
 * ascending, entries without keys dropped.  Outputs are malloc'd; release
 * them with tm_free(). */
int tm_canonicalize(const uint32_t * words, uint64_t n, const uint64_t * key_off, const uint32_t * keys,
					uint32_t ** out_words, uint64_t * out_n, uint64_t ** out_key_off,
					uint32_t ** out_keys);
void tm_free(void * p);

/* ------------------------------------------------------------------------ */
/* Seeded synthetic workload generator (host, multi-threaded).  Not part of
 * the reference interface; it produces the inputs of the tests and bench. */
#define TM_GEN_MAX_TAGS 64
#define TM_GEN_MAX_KEYS 16

typedef struct tm_workload_params {
	uint64_t seed;
	uint32_t n_tags;         /* vocabulary size                              */
	double   zipf_s;         /* Zipf exponent of tag popularity              */
	uint32_t bits_per_tag;   /* Bloom hash functions per tag (k)             */
	uint32_t ftag_min, ftag_max;  /* tags per filter                         */
	uint32_t key_min, key_max;    /* keys per filter                         */
	uint32_t trees, ifaces;       /* key = (tree << 16) | iface              */
	uint32_t qtag_min, qtag_max;  /* tags per query                          */
	double   superset_frac;       /* share of queries built over a filter    */
} tm_workload_params;

typedef struct tm_gen tm_gen;

void tm_workload_default_params(tm_workload_params * p);
tm_gen * tm_gen_create(const tm_workload_params * p);
void tm_gen_destroy(tm_gen * g);
int tm_gen_tag_bits(const tm_gen * g, uint32_t tag, uint8_t * bits);
/* Filters first .. first+n-1: words n x 6, key_counts n, keys n x key_max. */
int tm_gen_filters(const tm_gen * g, uint64_t first, uint64_t n, uint32_t * words,
				   uint32_t * key_counts, uint32_t * keys, int threads);
/* Queries first .. first+n-1 over a table of n_filters generated filters. */
int tm_gen_queries(const tm_gen * g, uint64_t n_filters, uint64_t first, uint64_t n,
				   uint32_t * words, uint8_t * is_superset, int threads);

/* ------------------------------------------------------------------------ */
/* Reference file formats, so that datasets written by the reference's tools
 * can be read into this engine and written back (host only, no device).
 *   TM_IO_BINARY      FIB entries: filter (24 bytes, the six words little-
 *                     endian), u32 key count, u32 keys (reference fib.cc:14-35,
 *                     bitvector.hh:506-540); queries: the 24-byte filter
 *                     (query.cc:11-19).
 *   TM_IO_ASCII       FIB entries "+ <192 bits, bit 0 first> key ...", queries
 *                     "! <192 bits>" (fib.cc:38-60, query.cc:21-40,
 *                     bitvector.hh:543-580); other lines are skipped.
 *   TM_IO_BINARY_PARTITIONED  FIB entries preceded by a u32 partition id
 *                     (fib.hh:41-56, partition_fib_entry).
 *   TM_IO_ASCII_PARTITIONED   FIB entries "f <partition> <192 bits> key ..."
 *                     (fib.cc:100-131, partition_fib_entry::write_ascii).
 * Reading is two calls: scan for the counts, then read into buffers of that
 * size (key_offsets holds n_filters + 1 entries).  Replaces the reference's
 * fib_entry / basic_query read_binary, write_binary, read_ascii and
 * write_ascii and the loops of its workload tools around them. */
#define TM_IO_BINARY              0u
#define TM_IO_ASCII               1u
#define TM_IO_BINARY_PARTITIONED  2u
#define TM_IO_ASCII_PARTITIONED   3u

int tm_io_scan_fib(const char * path, uint32_t format, uint64_t * n_filters, uint64_t * n_keys);
int tm_io_read_fib(const char * path, uint32_t format, uint32_t * words, uint32_t * key_offsets,
				   uint32_t * keys, uint32_t * partitions, uint64_t n_filters, uint64_t n_keys);
/* partitions: n ids for TM_IO_BINARY_PARTITIONED (NULL: 0), else ignored. */
int tm_io_write_fib(const char * path, uint32_t format, const uint32_t * words, uint64_t n,
					const uint32_t * key_offsets, const uint32_t * keys, const uint32_t * partitions);
int tm_io_scan_queries(const char * path, uint32_t format, uint64_t * n);
int tm_io_read_queries(const char * path, uint32_t format, uint32_t * words, uint64_t n);
int tm_io_write_queries(const char * path, uint32_t format, const uint32_t * words, uint64_t n);

/* Partition prefixes (reference partition_prefix, fib.hh:58-81): the first
 * `length` bits of `words` are the prefix of partition `partition`, which
 * holds `size` filters.  TM_IO_BINARY: the 24-byte filter, u8 length, u32
 * partition, u32 size (33 bytes); TM_IO_ASCII: "p <partition> <the first
 * length bits> <size>" (fib.cc:72-98), whose bit string's length is read
 * back as the prefix length. */
typedef struct tm_prefix {
	uint32_t words[6];
	uint32_t length;         /* 0..255, as the reference's uint8_t */
	uint32_t partition;
	uint32_t size;
} tm_prefix;
int tm_io_scan_prefixes(const char * path, uint32_t format, uint64_t * n);
int tm_io_read_prefixes(const char * path, uint32_t format, tm_prefix * prefixes, uint64_t n);
int tm_io_write_prefixes(const char * path, uint32_t format, const tm_prefix * prefixes, uint64_t n);

/* Workload tools: the reference's convert_workload, sample_workload and
 * split_workload programs (convert_workload.cc:12-59, sample_workload.cc:24-88,
 * split_workload.cc:19-119) as library calls, on the same files.  Record kinds: */
#define TM_IO_KIND_FIB      0u   /* FIB entries (any TM_IO_* FIB format)     */
#define TM_IO_KIND_QUERIES  1u   /* queries (TM_IO_BINARY or TM_IO_ASCII)    */
#define TM_IO_KIND_PREFIXES 2u   /* partition prefixes (BINARY or ASCII)     */

/* Rewrite every record of `in` (in_format) to `out` (out_format); *n =
 * records written.  FIB entries keep their partition ids across the
 * partitioned formats (0 when the input has none).  Unlike the reference
 * tool, ASCII lines of another record type are skipped rather than
 * repeating the previous record. */
int tm_io_convert(const char * in, uint32_t in_format, const char * out, uint32_t out_format, uint32_t kind,
				  uint64_t * n);

/* Keep k records of `in`, in input order and in the same format, by the
 * reference tool's selection sampling over n_total records: while fewer
 * are wanted than remain, a record is kept with probability (records still
 * wanted) / (records remaining), else it is kept outright.  As in the
 * reference (sample_workload.cc:24-56), a record kept outright is not
 * counted as seen, so n_total = k does not keep the first k records and a
 * choice can fall past the first n_total.  Choices follow the reference
 * tool's random stream for the same seed (std::default_random_engine, i.e.
 * minstd_rand0, with libstdc++'s uniform_int_distribution), so its output
 * files are reproduced byte for byte.  kind: FIB or QUERIES.  *n_out =
 * records written (k, unless the input runs out first). */
int tm_io_sample(const char * in, uint32_t format, const char * out, uint32_t kind, uint64_t n_total, uint64_t k,
				 uint32_t seed, uint64_t * n_out);

/* Cut a FIB file into prefix partitions as the reference's split_workload
 * does: entries kept in input order when it is strictly ascending, else
 * sorted in descending filter order; then each partition starts at the next
 * entry and takes the following entries while the longest common prefix
 * stays longer than that entry's min_k-th set bit, at most max_size entries
 * (split_workload.cc:37-119).  format TM_IO_BINARY or TM_IO_ASCII for the
 * input and both outputs: prefixes (NULL: not written) and partitioned FIB
 * entries (NULL: not written; TM_IO_BINARY_PARTITIONED or
 * TM_IO_ASCII_PARTITIONED records).  *n_parts = partitions. */
int tm_io_split(const char * in, uint32_t format, const char * prefixes_out, const char * filters_out,
				uint32_t max_size, uint32_t min_k, uint64_t * n_parts);

#ifdef __cplusplus
}
#endif

#endif /* TAGMATCH_AMD_H_INCLUDED */
