/*
 * oracle.c -- CPU restatement of the reference TagMatch algorithms.
 *
 * TEST INFRASTRUCTURE ONLY: the checker of the HIP engine, loaded by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg; never by the
 * engine.  Each function cites the reference file:line it restates.
 *
 * Parity pins (tests/test_oracle.py):
 *   orc_subset, orc_compare, orc_leftmost_diff, orc_from_string
 *       test_bitvector.cc:92-136 (layout), :485-537 (subset truth table),
 *       :542-565 (order), :569-590 (leftmost_diff); test_io.cc:50-68
 *       (byte encoding); and oracle/_ref (reference bitvector.hh compiled).
 *   orc_match      reference test_matcher compiled into oracle/_ref, on
 *                  generated workloads (tests/golden), and the
 *                  test_tagmatch.cc known answers.
 *   orc_partition  reference partitioning.cc compiled into oracle/_ref
 *                  (same positions, same partitions), test_partitioning.cc
 *                  cases, and the partition invariants.
 *   orc_lookup     brute force over masks (every selected mask is a subset,
 *                  every subset mask is selected).
 *   orc_consolidate  test_tagmatch.cc add_same_filter / add_and_clear
 *                  answers and a sort/unique restatement in numpy.
 */
#include "oracle.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ----------------------------------------------------------- bitvector */

static uint64_t block64(const uint32_t * w, int i) {
	return (uint64_t)w[2 * i] | ((uint64_t)w[2 * i + 1] << 32);
}

/* bitvector.hh:290-296 */
int orc_subset(const uint32_t * f, const uint32_t * q) {
	for (int j = 0; j < 6; ++j)
		if (f[j] & ~q[j]) return 0;
	return 1;
}

/* bitvector.hh:389-397: at the first differing 64-bit block, the lowest
 * differing bit decides; the operand holding it is the larger. */
int orc_compare(const uint32_t * a, const uint32_t * b) {
	for (int i = 0; i < 3; ++i) {
		uint64_t x = block64(a, i), y = block64(b, i);
		if (x != y) {
			int pos = __builtin_ctzll(x ^ y);
			return ((y >> pos) & 1) ? -1 : 1;
		}
	}
	return 0;
}

/* bitvector.hh:412-418 */
int orc_leftmost_diff(const uint32_t * a, const uint32_t * b) {
	for (int i = 0; i < 3; ++i) {
		uint64_t x = block64(a, i), y = block64(b, i);
		if (x != y) return i * 64 + __builtin_ctzll(x ^ y);
	}
	return 192;
}

/* bitvector.hh:122-139 */
void orc_from_string(const char * s, uint32_t * out) {
	memset(out, 0, 24);
	for (int p = 0; s[p] && p < 192; ++p)
		if (s[p] == '1') out[p >> 5] |= 1u << (p & 31);
}

void orc_free(void * p) { free(p); }

/* ----------------------------------------------------------- vectors */

typedef struct {
	uint32_t * v;
	uint64_t n, cap;
} vec32;

static void vpush(vec32 * a, uint32_t x) {
	if (a->n == a->cap) {
		a->cap = a->cap ? a->cap * 2 : 1024;
		a->v = (uint32_t *)realloc(a->v, a->cap * sizeof(uint32_t));
	}
	a->v[a->n++] = x;
}

static int cmp_u32(const void * a, const void * b) {
	uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
	return x < y ? -1 : x > y;
}

/* ----------------------------------------------------------- consolidate */

typedef struct {
	uint32_t w[6];
	uint32_t key;
	uint32_t op;
	uint64_t seq;
} op_rec;

static int cmp_op(const void * a, const void * b) {
	const op_rec * x = (const op_rec *)a, * y = (const op_rec *)b;
	int c = orc_compare(x->w, y->w);
	if (c) return c;
	return x->seq < y->seq ? -1 : x->seq > y->seq;
}

static int contains(const uint32_t * v, uint64_t n, uint32_t k) {
	return bsearch(&k, v, n, sizeof(uint32_t), cmp_u32) != NULL;
}

/* tagmatch.cc:22-77 (keyset_change: new = (old - to_remove) + (to_add -
 * old)), :98-165 (merge, drop empty key sets, map order = lexicographic).
 * Without a previous table old = {} and the result is the set of added keys. */
int orc_consolidate(const uint32_t * ops, uint64_t n_ops, uint32_t ** words, uint64_t * n_filters,
					uint64_t ** key_off, uint32_t ** keys) {
	op_rec * r = (op_rec *)malloc((n_ops ? n_ops : 1) * sizeof(op_rec));
	for (uint64_t i = 0; i < n_ops; ++i) {
		memcpy(r[i].w, ops + i * 8, 24);
		r[i].key = ops[i * 8 + 6];
		r[i].op = ops[i * 8 + 7];
		r[i].seq = i;
	}
	qsort(r, n_ops, sizeof(op_rec), cmp_op);
	vec32 W = {0}, K = {0};
	uint64_t * off = (uint64_t *)malloc((n_ops + 1) * sizeof(uint64_t));
	uint64_t nf = 0;
	off[0] = 0;
	vec32 add = {0}, rem = {0};
	for (uint64_t i = 0; i < n_ops;) {
		uint64_t j = i;
		add.n = rem.n = 0;
		while (j < n_ops && orc_compare(r[i].w, r[j].w) == 0) {
			vpush(r[j].op ? &rem : &add, r[j].key);
			++j;
		}
		qsort(add.v, add.n, sizeof(uint32_t), cmp_u32);
		uint64_t u = 0;
		for (uint64_t k = 0; k < add.n; ++k)
			if (k == 0 || add.v[k] != add.v[k - 1]) add.v[u++] = add.v[k];
		/* no previous table: old = {}, removals have nothing to remove */
		if (u) {
			for (int x = 0; x < 6; ++x) vpush(&W, r[i].w[x]);
			for (uint64_t k = 0; k < u; ++k) vpush(&K, add.v[k]);
			off[nf + 1] = K.n;
			++nf;
		}
		i = j;
	}
	(void)contains;
	free(add.v);
	free(rem.v);
	free(r);
	*words = W.v ? W.v : (uint32_t *)malloc(4);
	*n_filters = nf;
	*key_off = off;
	*keys = K.v ? K.v : (uint32_t *)malloc(4);
	return 0;
}

/* ----------------------------------------------------------- brute force */

typedef struct {
	const uint32_t * words;
	uint64_t n;
	const uint64_t * key_off;
	const uint32_t * keys;
	const uint32_t * queries;
	uint64_t qb, qe;
	int mode;
	uint64_t * counts;  /* per query */
	vec32 out;
	uint64_t hits;
	/* schedule form (CPU-B): only the filters of the partitions the lookup
	 * selects; NULL for brute force over the whole table */
	const orc_lookup * lk;
	const uint32_t * part_off;
} match_job;

/* test_matcher.cc:74-81: every filter, in table order, that is a subset of
 * the query contributes its keys; tagmatch_query.hh:105-109 sorts and
 * removes duplicates for match_unique. */
static void collect(const match_job * j, const uint32_t * qw, uint64_t f0, uint64_t f1, vec32 * tmp) {
	for (uint64_t f = f0; f < f1; ++f) {
		if (!orc_subset(j->words + f * 6, qw)) continue;
		if (j->mode == 2) {
			vpush(tmp, (uint32_t)f);
		} else {
			for (uint64_t k = j->key_off[f]; k < j->key_off[f + 1]; ++k) vpush(tmp, j->keys[k]);
		}
	}
}

static void * match_worker(void * arg) {
	match_job * j = (match_job *)arg;
	vec32 tmp = {0};
	uint32_t * sel = j->lk ? (uint32_t *)malloc(((uint64_t)orc_lookup_size(j->lk) + 1) * sizeof(uint32_t)) : NULL;
	for (uint64_t q = j->qb; q < j->qe; ++q) {
		const uint32_t * qw = j->queries + q * 6;
		tmp.n = 0;
		if (j->lk) {
			const uint32_t ns = orc_lookup_query(j->lk, qw, sel);
			for (uint32_t i = 0; i < ns; ++i) collect(j, qw, j->part_off[sel[i]], j->part_off[sel[i] + 1], &tmp);
		} else {
			collect(j, qw, 0, j->n, &tmp);
		}
		if (j->mode != 2) qsort(tmp.v, tmp.n, sizeof(uint32_t), cmp_u32);
		uint64_t m = tmp.n;
		if (j->mode == 0) {
			m = 0;
			for (uint64_t k = 0; k < tmp.n; ++k)
				if (k == 0 || tmp.v[k] != tmp.v[k - 1]) tmp.v[m++] = tmp.v[k];
		}
		j->counts[q] = m;
		for (uint64_t k = 0; k < m; ++k) vpush(&j->out, tmp.v[k]);
	}
	free(tmp.v);
	free(sel);
	return NULL;
}

static int clamp_threads(int t, uint64_t work) {
	if (t <= 0) t = 1;
	if (t > 64) t = 64;
	if ((uint64_t)t > work) t = work ? (int)work : 1;
	return t;
}

static int run_match(const uint32_t * words, uint64_t n, const uint64_t * key_off, const uint32_t * keys,
					 const uint32_t * queries, uint64_t nq, int mode, int threads, uint64_t * off, uint32_t ** vals,
					 const orc_lookup * lk, const uint32_t * part_off);

int orc_match(const uint32_t * words, uint64_t n, const uint64_t * key_off, const uint32_t * keys,
			  const uint32_t * queries, uint64_t nq, int mode, int threads, uint64_t * off,
			  uint32_t ** vals) {
	return run_match(words, n, key_off, keys, queries, nq, mode, threads, off, vals, NULL, NULL);
}

/* The reference schedule on the CPU with the reference's results: each
 * query goes to the partitions whose mask it contains (front_end.cc:465-477)
 * and is brute-forced there (the CPU back end, back_end.cc:490-516); its
 * keys are sorted and deduplicated as match_unique does
 * (tagmatch_query.hh:105-109).  words / key_off / keys are in partitioned
 * order; HITS ids are positions in that order. */
int orc_schedule_match(const uint32_t * words, const uint64_t * key_off, const uint32_t * keys,
					   const uint32_t * part_off, const uint32_t * masks, uint32_t np, const uint32_t * queries,
					   uint64_t nq, int mode, int threads, uint64_t * off, uint32_t ** vals) {
	orc_lookup * lk = orc_lookup_build(masks, np);
	int rc = run_match(words, part_off[np], key_off, keys, queries, nq, mode, threads, off, vals, lk, part_off);
	orc_lookup_free(lk);
	return rc;
}

const char * orc_compiler(void) { return __VERSION__; }

static int run_match(const uint32_t * words, uint64_t n, const uint64_t * key_off, const uint32_t * keys,
					 const uint32_t * queries, uint64_t nq, int mode, int threads, uint64_t * off, uint32_t ** vals,
					 const orc_lookup * lk, const uint32_t * part_off) {
	if (mode < 0 || mode > 2) return -1;
	int T = clamp_threads(threads, nq);
	match_job * jobs = (match_job *)calloc(T, sizeof(match_job));
	pthread_t * th = (pthread_t *)calloc(T, sizeof(pthread_t));
	uint64_t * counts = (uint64_t *)calloc(nq ? nq : 1, sizeof(uint64_t));
	uint64_t chunk = (nq + T - 1) / T;
	for (int t = 0; t < T; ++t) {
		match_job * j = &jobs[t];
		j->words = words; j->n = n; j->key_off = key_off; j->keys = keys; j->queries = queries;
		j->qb = (uint64_t)t * chunk;
		j->qe = j->qb + chunk < nq ? j->qb + chunk : nq;
		if (j->qb > nq) j->qb = nq;
		j->mode = mode;
		j->counts = counts;
		j->lk = lk;
		j->part_off = part_off;
		pthread_create(&th[t], NULL, match_worker, j);
	}
	for (int t = 0; t < T; ++t) pthread_join(th[t], NULL);
	off[0] = 0;
	for (uint64_t q = 0; q < nq; ++q) off[q + 1] = off[q] + counts[q];
	uint32_t * out = (uint32_t *)malloc((off[nq] ? off[nq] : 1) * sizeof(uint32_t));
	for (int t = 0; t < T; ++t) {
		if (jobs[t].qb < jobs[t].qe) memcpy(out + off[jobs[t].qb], jobs[t].out.v, jobs[t].out.n * sizeof(uint32_t));
		free(jobs[t].out.v);
	}
	*vals = out;
	free(counts);
	free(jobs);
	free(th);
	return 0;
}

/* ----------------------------------------------------------- partitioning */

typedef struct {
	uint32_t begin, end;
	uint32_t mask[6];  /* pivot bits on the 1-side path */
	uint32_t used[6];  /* all pivot bits used */
} cand_t;

typedef struct {
	const uint32_t * words;
	uint32_t * fib;     /* permutation being partitioned (partitioning.cc:58 fib_t) */
	uint32_t * tmp;
	uint32_t maxp;
	cand_t * out;       /* finished partitions */
	uint32_t n_out, cap_out;
	cand_t * stack;     /* pending candidates */
	uint32_t n_stack, cap_stack;
	int error;
} part_state;

static int bit_of(const uint32_t * w, unsigned b) { return (w[b >> 5] >> (b & 31)) & 1; }
static void set_bit(uint32_t * w, unsigned b) { w[b >> 5] |= 1u << (b & 31); }
static int is_empty(const uint32_t * w) { return !(w[0] | w[1] | w[2] | w[3] | w[4] | w[5]); }

/* partition_candidate::compute_frequencies, partitioning.cc:92-113 */
static void frequencies(const part_state * S, const cand_t * c, uint32_t * freq) {
	memset(freq, 0, 192 * sizeof(uint32_t));
	for (uint32_t i = c->begin; i < c->end; ++i) {
		const uint32_t * w = S->words + (uint64_t)S->fib[i] * 6;
		for (int j = 0; j < 6; ++j) {
			uint32_t x = w[j];
			while (x) {
				freq[j * 32 + __builtin_ctz(x)]++;
				x &= x - 1;
			}
		}
	}
}

/* partition_candidate::split_zero, partitioning.cc:133-149: stable
 * partition (zeros first); c becomes the 1-side, the 0-side is returned. */
static cand_t split_zero(part_state * S, cand_t * c, unsigned pivot) {
	uint32_t m = c->begin, k = 0;
	for (uint32_t i = c->begin; i < c->end; ++i)
		if (!bit_of(S->words + (uint64_t)S->fib[i] * 6, pivot)) S->fib[m++] = S->fib[i];
		else S->tmp[k++] = S->fib[i];
	memcpy(S->fib + m, S->tmp, k * sizeof(uint32_t));
	cand_t p0 = *c;
	p0.end = m;
	c->begin = m;
	return p0;
}

static void emit(part_state * S, const cand_t * c) {
	if (S->n_out == S->cap_out) {
		S->cap_out = S->cap_out ? S->cap_out * 2 : 64;
		S->out = (cand_t *)realloc(S->out, S->cap_out * sizeof(cand_t));
	}
	S->out[S->n_out++] = *c;
}

static void push(part_state * S, const cand_t * c) {
	if (S->n_stack == S->cap_stack) {
		S->cap_stack = S->cap_stack ? S->cap_stack * 2 : 64;
		S->stack = (cand_t *)realloc(S->stack, S->cap_stack * sizeof(cand_t));
	}
	S->stack[S->n_stack++] = *c;
}

/* nonzero_mask_partitioning, partitioning.cc:230-250 */
static void nonzero_mask(part_state * S, cand_t p) {
	uint32_t freq[192];
	for (;;) {
		frequencies(S, &p, freq);
		uint32_t size = p.end - p.begin;
		uint32_t max_freq = freq[0];
		unsigned pivot = 0;
		for (unsigned b = 1; b < 192; ++b) {
			if (freq[b] > max_freq) {
				pivot = b;
				max_freq = freq[b];
			}
			if (freq[b] == size) set_bit(p.mask, b);
		}
		if (max_freq == size) {
			emit(S, &p);
			return;
		}
		if (max_freq == 0) {  /* all-zero filters: the reference loops forever */
			S->error = -1;
			emit(S, &p);
			return;
		}
		cand_t p0 = split_zero(S, &p, pivot);
		set_bit(p.mask, pivot);
		emit(S, &p);
		p = p0;
	}
}

static uint32_t udist(uint32_t a, uint32_t b) { return a > b ? a - b : b - a; }

/* balancing_pivot, partitioning.cc:262-278 */
static unsigned balancing_pivot(const cand_t * p, const uint32_t * freq) {
	unsigned pivot;
	uint32_t half = (p->end - p->begin) / 2;
	for (pivot = 0; pivot < 192; ++pivot)
		if (!bit_of(p->used, pivot)) break;
	if (pivot >= 192) return 191;
	uint32_t min_dist = udist(half, freq[pivot]);
	for (unsigned b = pivot + 1; b < 192; ++b) {
		uint32_t d = udist(half, freq[b]);
		if (d < min_dist) {
			pivot = b;
			if (d == 0) break;
			min_dist = d;
		}
	}
	return pivot;
}

/* do_balanced_partitioning, partitioning.cc:280-340, with an explicit stack
 * in place of the work queue and thread pool. */
static void balanced(part_state * S) {
	uint32_t freq[192];
	while (S->n_stack) {
		cand_t p = S->stack[--S->n_stack];
		for (;;) {
			frequencies(S, &p, freq);
			unsigned pivot = balancing_pivot(&p, freq);
			cand_t p0 = split_zero(S, &p, pivot);
			set_bit(p.used, pivot);
			memcpy(p0.used, p.used, sizeof(p.used));
			memcpy(p0.mask, p.mask, sizeof(p.mask));
			set_bit(p.mask, pivot);
			uint32_t s0 = p0.end - p0.begin, s1 = p.end - p.begin;
			if (s0 > S->maxp && s1 > S->maxp) {
				push(S, &p0);
			} else if (s1 > S->maxp) {
				if (is_empty(p0.mask)) {
					push(S, &p);
					nonzero_mask(S, p0);
					break;
				}
				emit(S, &p0);
			} else if (s0 > S->maxp) {
				emit(S, &p);
				p = p0;
			} else {
				if (is_empty(p0.mask)) nonzero_mask(S, p0);
				else emit(S, &p0);
				emit(S, &p);
				break;
			}
		}
	}
}

static int cmp_cand(const void * a, const void * b) {
	const cand_t * x = (const cand_t *)a, * y = (const cand_t *)b;
	if (x->begin != y->begin) return x->begin < y->begin ? -1 : 1;
	return x->end < y->end ? -1 : x->end > y->end;
}

/* balanced_partitioning, partitioning.cc:361-411; partitions are reported
 * in position order, masks = AND of members (all-ones if empty). */
int orc_partition(const uint32_t * words, uint32_t n, uint32_t maxp, uint32_t * perm,
				  uint32_t * part_off, uint32_t * masks, uint32_t max_parts, uint32_t * n_parts) {
	part_state S;
	memset(&S, 0, sizeof(S));
	S.words = words;
	S.maxp = maxp;
	S.fib = perm;
	S.tmp = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
	for (uint32_t i = 0; i < n; ++i) perm[i] = i;
	*n_parts = 0;
	part_off[0] = 0;
	if (n == 0) {
		free(S.tmp);
		return 0;
	}
	cand_t root;
	memset(&root, 0, sizeof(root));
	root.begin = 0;
	root.end = n;
	if (n > maxp) {
		push(&S, &root);
		balanced(&S);
	} else {
		nonzero_mask(&S, root);
	}
	qsort(S.out, S.n_out, sizeof(cand_t), cmp_cand);
	int rc = S.error;
	if (S.n_out > max_parts) rc = -2;
	else {
		for (uint32_t p = 0; p < S.n_out; ++p) {
			part_off[p] = S.out[p].begin;
			uint32_t m[6] = {~0u, ~0u, ~0u, ~0u, ~0u, ~0u};
			for (uint32_t i = S.out[p].begin; i < S.out[p].end; ++i)
				for (int j = 0; j < 6; ++j) m[j] &= words[(uint64_t)perm[i] * 6 + j];
			if (masks) memcpy(masks + (uint64_t)p * 6, m, 24);
		}
		part_off[S.n_out] = n;
	}
	*n_parts = S.n_out;
	free(S.out);
	free(S.stack);
	free(S.tmp);
	return rc;
}

/* ----------------------------------------------------------- lookup */

struct orc_lookup {
	uint32_t np;
	uint32_t * masks;
	uint32_t start[194];   /* bucket b: entries [start[b], start[b+1]); 192 = empty masks */
	uint32_t * entries;
};

/* front_end::add_partition / consolidate, front_end.cc:421-461 */
orc_lookup * orc_lookup_build(const uint32_t * masks, uint32_t np) {
	orc_lookup * lk = (orc_lookup *)calloc(1, sizeof(orc_lookup));
	lk->np = np;
	lk->masks = (uint32_t *)malloc((np ? np : 1) * 24);
	memcpy(lk->masks, masks, (size_t)np * 24);
	uint32_t * bucket = (uint32_t *)malloc((np ? np : 1) * sizeof(uint32_t));
	uint32_t size[193];
	memset(size, 0, sizeof(size));
	for (uint32_t p = 0; p < np; ++p) {
		const uint32_t * m = masks + (uint64_t)p * 6;
		if (is_empty(m)) {  /* engine extension: empty mask, always selected */
			bucket[p] = 192;
		} else {
			unsigned min_i = 192;
			for (unsigned b = 0; b < 192; ++b)
				if (bit_of(m, b) && (min_i == 192 || size[b] < size[min_i])) min_i = b;
			bucket[p] = min_i;
		}
		size[bucket[p]]++;
	}
	lk->start[0] = 0;
	for (int b = 0; b < 193; ++b) lk->start[b + 1] = lk->start[b] + size[b];
	lk->entries = (uint32_t *)malloc((np ? np : 1) * sizeof(uint32_t));
	uint32_t fill[193];
	memcpy(fill, lk->start, sizeof(fill));
	for (uint32_t p = 0; p < np; ++p) lk->entries[fill[bucket[p]]++] = p;
	free(bucket);
	return lk;
}

/* do_match, front_end.cc:465-477 */
uint32_t orc_lookup_query(const orc_lookup * lk, const uint32_t * q, uint32_t * sel) {
	uint32_t n = 0;
	for (unsigned b = 0; b < 193; ++b) {
		if (b < 192 && !bit_of(q, b)) continue;
		for (uint32_t e = lk->start[b]; e < lk->start[b + 1]; ++e) {
			uint32_t p = lk->entries[e];
			if (orc_subset(lk->masks + (uint64_t)p * 6, q)) sel[n++] = p;
		}
	}
	qsort(sel, n, sizeof(uint32_t), cmp_u32);
	return n;
}

uint32_t orc_lookup_size(const orc_lookup * lk) { return lk->np; }

void orc_lookup_free(orc_lookup * lk) {
	if (!lk) return;
	free(lk->masks);
	free(lk->entries);
	free(lk);
}

/* ----------------------------------------------------------- CPU baselines */

typedef struct {
	const uint32_t * words;
	uint64_t n;
	const uint32_t * part_off;
	const orc_lookup * lk;
	const uint32_t * queries;
	uint64_t qb, qe;
	uint64_t hits;
} count_job;

static void * brute_worker(void * arg) {
	count_job * j = (count_job *)arg;
	uint64_t h = 0;
	for (uint64_t q = j->qb; q < j->qe; ++q) {
		const uint32_t * qw = j->queries + q * 6;
		for (uint64_t f = 0; f < j->n; ++f) h += orc_subset(j->words + f * 6, qw);
	}
	j->hits = h;
	return NULL;
}

/* lookup then per-partition brute force: front_end.cc:465-477 and the
 * CPU back end of back_end.cc:490-516 */
static void * sched_worker(void * arg) {
	count_job * j = (count_job *)arg;
	uint32_t * sel = (uint32_t *)malloc((j->lk->np ? j->lk->np : 1) * sizeof(uint32_t));
	uint64_t h = 0;
	for (uint64_t q = j->qb; q < j->qe; ++q) {
		const uint32_t * qw = j->queries + q * 6;
		uint32_t ns = orc_lookup_query(j->lk, qw, sel);
		for (uint32_t s = 0; s < ns; ++s)
			for (uint32_t f = j->part_off[sel[s]]; f < j->part_off[sel[s] + 1]; ++f)
				h += orc_subset(j->words + (uint64_t)f * 6, qw);
	}
	free(sel);
	j->hits = h;
	return NULL;
}

static uint64_t run_count(count_job * proto, uint64_t nq, int threads, void * (*fn)(void *)) {
	int T = clamp_threads(threads, nq);
	count_job * jobs = (count_job *)calloc(T, sizeof(count_job));
	pthread_t * th = (pthread_t *)calloc(T, sizeof(pthread_t));
	uint64_t chunk = (nq + T - 1) / T;
	for (int t = 0; t < T; ++t) {
		jobs[t] = *proto;
		jobs[t].qb = (uint64_t)t * chunk < nq ? (uint64_t)t * chunk : nq;
		jobs[t].qe = jobs[t].qb + chunk < nq ? jobs[t].qb + chunk : nq;
		pthread_create(&th[t], NULL, fn, &jobs[t]);
	}
	uint64_t h = 0;
	for (int t = 0; t < T; ++t) {
		pthread_join(th[t], NULL);
		h += jobs[t].hits;
	}
	free(jobs);
	free(th);
	return h;
}

uint64_t orc_bruteforce_count(const uint32_t * words, uint64_t n, const uint32_t * queries, uint64_t nq,
							  int threads) {
	count_job p;
	memset(&p, 0, sizeof(p));
	p.words = words;
	p.n = n;
	p.queries = queries;
	return run_count(&p, nq, threads, brute_worker);
}

uint64_t orc_schedule_count(const uint32_t * words, const uint32_t * part_off, const uint32_t * masks,
							uint32_t np, const uint32_t * queries, uint64_t nq, int threads) {
	orc_lookup * lk = orc_lookup_build(masks, np);
	count_job p;
	memset(&p, 0, sizeof(p));
	p.words = words;
	p.part_off = part_off;
	p.lk = lk;
	p.queries = queries;
	uint64_t h = run_count(&p, nq, threads, sched_worker);
	orc_lookup_free(lk);
	return h;
}
