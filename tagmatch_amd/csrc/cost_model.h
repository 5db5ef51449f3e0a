// Cost model internals shared with the match scheduler.  Internal to
// libtagmatch_amd.so.
//
// A work unit is a (partition, query batch) pair: the plain schedule runs one
// per launch, and the persistent schedule visits one per (query tile,
// partition) with the tile's queries that select the partition as the batch.
// Both are described by the same TM_UNIT_FEATURES features, so a model
// trained on plain-schedule logs can price persistent tiles.
#pragma once

#include "fib.h"

namespace tmx {

// Feature row of one unit (all scaled to roughly [0, 1]):
//   0  log2(1 + partition filters) / 20
//   1  popcount(partition mask) / 64
//   2  log2(1 + B) / 12, B = queries in the batch (0: empty unit)
//   3  mean query popcount / 192
//   4  standard deviation of query popcount / 192
//   5  min query popcount / 192
//   6  max query popcount / 192
//   7..10  fraction of queries with popcount in [0,80) [80,110) [110,140) [140,192]
//   11 mean member-filter popcount / 64
//   12 popcount(OR of the members) / 192
//   13 mean popcount(query AND OR-of-members) / 192
//   14 max popcount(query AND OR-of-members) / 192
//   15 log2(1 + B * partition filters) / 32
// Targets: log2(1 + x) / 16 of hits, key instances of the hit filters, and
// ticks = summed workgroup busy time in 10 ns units.
constexpr int N_FEAT = TM_UNIT_FEATURES;
constexpr int N_TARGET = 3;

// Plain units: units[u] = (partition, batch index, queries in the batch,
// first list entry); the batch's query ids are list[first .. first + queries),
// at most 256 of them.
void launch_unit_features_plain(const dev_fib_view & f, const uint32_t * q, const uint32_t * list,
								const uint4 * units, uint32_t n_units, float * out, hipStream_t s);

// Tile units: row t * P + p for query tile t (queries [t*nqt, (t+1)*nqt) of
// nq) and partition p; the batch is the tile's queries that select p.
void launch_unit_features_tiles(const dev_fib_view & f, const uint32_t * q, uint32_t nq, uint32_t nqt,
								uint32_t n_tiles, float * out, hipStream_t s);

// Batched inference of the fused 16-64-64-3 network on device rows; y gets
// n x 3 floats.  Returns TM_ERR_INVALID for another shape.
bool mlp_fused_shape(const tm_mlp * m);
int mlp_infer(const tm_mlp * m, const float * x, uint32_t n, float * y, hipStream_t s);

// Per tile: sums over partitions of the decoded predicted ticks (cost),
// hits and key instances (empty units count zero).
void launch_tile_costs(const float * x, const float * y, uint32_t n_tiles, uint32_t P, float * cost, float * hits,
					   float * keys, hipStream_t s);

} // namespace tmx
