// Device-resident FIB (filters, block masks, keys, partitions) and the
// device view passed to kernels.  Internal to libtagmatch_amd.so.
#pragma once

#include "hip_util.h"

#include <atomic>
#include <mutex>
#include <vector>

struct tm_fib;

namespace tmx {

// Layout in HBM (all arrays 256-byte aligned by hipMalloc):
//
//   blk_words [nb][6][64] u32   filter word j of lane l of block b at
//                               b*384 + j*64 + l: a wave reads each word
//                               plane with one coalesced 256-byte access.
//   blk_info  [nb][8]     u32   AND of the block's valid filters (6 words),
//                               valid lane count, partition id.  Read with
//                               scalar loads (uniform per wave).
//   slot_keys [nb][64]    u64   up to 4 dense key ids (16 bits each, 0xFFFF
//                               = none) inline; if field 0 is 0xFFFF, bits
//                               16..47 index an overflow list in `ovf`
//                               (count, ids...).
//   slot_fid  [nb][64]    u32   filter id reported in TM_MODE_HITS.
//   key_dict  [K]         u32   sorted distinct keys; dense id = index.
//   part_info [P][8]      u32   partition mask (6 words), first block,
//                               block count.
//   part_stat [P][8]      u32   OR of the member filters (6 words), member
//                               count, total popcount of the members (cost
//                               model features only).
//   sup_info  [nb/8][8]   u32   AND of the masks of blocks 8g..8g+7 (6
//                               words; empty blocks count as all ones): a
//                               query that does not contain it is a
//                               candidate for none of those blocks.
struct dev_fib_view {
	const uint32_t * __restrict__ blk_words;
	const uint32_t * __restrict__ blk_info;
	const uint64_t * __restrict__ slot_keys;
	const uint32_t * __restrict__ slot_fid;
	const uint32_t * __restrict__ ovf;
	uint32_t ovf_n;                    // valid words of ovf when the view was made
	const uint32_t * __restrict__ key_dict;
	const uint32_t * __restrict__ part_info;
	const uint32_t * __restrict__ part_stat;
	const uint32_t * __restrict__ sup_info;
	uint32_t n_parts;
	uint32_t n_blocks;
	uint32_t n_keys;   // K, distinct keys
	uint32_t kw;       // 32-bit words of a per-query key bitmap: ceil(K/32)
};

constexpr uint32_t BLOCK_WORDS = 6 * 64;
constexpr uint32_t SUP_BLOCKS = 8;   // blocks per group of sup_info
constexpr uint32_t KEY_NONE = 0xFFFFu;
constexpr uint32_t MAX_RICH = 1024;
constexpr uint32_t RICH_MIN_KEYS = 32;

// A live insert prepared on the host (partition choice, lane records, host
// copies updated), committed to the device later by the thread that makes
// submissions.  Commits follow prepares in order.
struct live_plan;
int live_prepare(tm_fib * fib, const uint32_t * words, uint32_t n, const uint32_t * key_offsets,
				 const uint32_t * keys, const uint32_t * filter_ids, live_plan ** out);
int live_commit(tm_fib * fib, live_plan * plan);   // frees the plan
void live_discard(live_plan * plan);

} // namespace tmx

struct tm_fib {
	uint32_t * blk_words = nullptr;
	uint32_t * blk_info = nullptr;
	uint64_t * slot_keys = nullptr;
	uint32_t * slot_fid = nullptr;
	uint32_t * ovf = nullptr;
	uint32_t * key_dict = nullptr;
	uint32_t * part_info = nullptr;
	uint32_t * part_stat = nullptr;
	uint32_t * sup_info = nullptr;
	uint64_t sup_groups = 0;             // groups sup_info holds
	uint32_t n_parts = 0, n_blocks = 0, n_keys = 0;
	uint64_t n_filters = 0, n_key_instances = 0, device_bytes = 0, max_part = 0;
	// host copies used by the schedulers and by tm_fib_insert
	std::vector<uint32_t> h_part_info;   // P x 8 (as part_info)
	std::vector<uint32_t> h_part_size;   // P, filters per partition
	std::vector<uint32_t> h_part_stat;   // P x 8 (as part_stat)
	std::vector<uint32_t> h_part_cap;    // P, blocks reserved (used + spare)
	std::vector<uint32_t> h_ovf;         // overflow key lists
	std::vector<uint32_t> h_dict;        // sorted distinct keys
	std::vector<int> h_order;            // bit significance order of the members
	size_t ovf_cap = 0;                  // device words of ovf
	uint32_t reserve_pct = 13;           // spare blocks per partition, percent (+1)
	uint64_t version = 0;                // bumped when P, K or device arrays change
	// live inserts: each lane's insert epoch; a submission sees the lanes
	// whose epoch is at most the FIB's epoch when it was submitted
	uint32_t * slot_epoch = nullptr;     // [nb][64], allocated by the first live insert
	uint32_t epoch = 0;                  // inserts so far
	hipStream_t ins_stream = nullptr;    // live inserts run here, beside the rings' streams
	void * d_recs = nullptr;             // insert records staging
	size_t recs_cap = 0;
	std::atomic<int> in_flight{0};       // ring submissions not yet waited for
	// live inserts prepared on the host ahead of their device commit
	// (tmx::live_prepare / live_commit): host copies change at prepare,
	// under host_mu; the device and `epoch` at commit, in prepare order
	std::mutex host_mu;
	// rich filters: those with the most keys (at least RICH_MIN_KEYS, at most
	// MAX_RICH, from the upload; dropped when an insert grows the dictionary),
	// each with its key set as a dense bitmap.  A query that contains one
	// has all of its keys, whatever else it matches.
	uint32_t * d_rich_words = nullptr;   // n_rich x 8 words (6 + padding)
	uint32_t * d_rich_bm = nullptr;      // n_rich x kw words
	uint32_t n_rich = 0;
	uint32_t n_rich_full = 0;            // the first n_rich_full hold every key
	uint32_t prep_epoch = 0;             // epoch of the last prepared insert
	uint32_t ovf_valid = 1;              // words of ovf written on the device

	tmx::dev_fib_view view() const {
		tmx::dev_fib_view v;
		v.blk_words = blk_words;
		v.blk_info = blk_info;
		v.slot_keys = slot_keys;
		v.slot_fid = slot_fid;
		v.ovf = ovf;
		v.ovf_n = ovf_valid;
		v.key_dict = key_dict;
		v.part_info = part_info;
		v.part_stat = part_stat;
		v.sup_info = sup_info;
		v.n_parts = n_parts;
		v.n_blocks = n_blocks;
		v.n_keys = n_keys;
		v.kw = (n_keys + 31) / 32;
		return v;
	}
	void release();
};
