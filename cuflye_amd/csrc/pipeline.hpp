// pipeline.hpp - launch interface of the index, join, match-sort and group
// kernels (pipeline.hip) and of the device scan.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "chain.hpp"

namespace fh
{

// 2-bit packed sequences: sequence i starts at word wordOff[i] and has len[i]
// bases (16 per word, MSB first, two zero words after each sequence).
struct PackedSeqs
{
	const uint32_t* words;
	const int64_t* wordOff;
	const int32_t* len;
	int32_t n;
};

// Open-addressed k-mer table.  Slot h is one 16-byte pair: the canonical
// k-mer at slots[2h] and its value at slots[2h + 1] (the count while
// counting; then (start << 24 | count), kRepMark or 0), so a probe that hits
// reads one cache line, not one line per array.
struct HashTable
{
	uint64_t* slots;
	uint32_t* cursor;  // fill cursors
	uint64_t mask;     // slots - 1
	__host__ __device__ uint64_t& key(uint64_t h) const { return slots[2 * h]; }
	__host__ __device__ uint64_t& val(uint64_t h) const { return slots[2 * h + 1]; }
};

// ---- scans ----------------------------------------------------------------
// Exclusive scan of n int32 values into int64; out has n + 1 entries (the last
// is the total).  `tmp` needs scanTempWords(n) int64 words.
size_t scanTempWords(int64_t n);
hipError_t exclusiveScan(const int32_t* in, int64_t* out, int64_t n, int64_t* tmp, hipStream_t s);

// ---- index build ------------------------------------------------------------
hipError_t launchCount(const PackedSeqs& reads, const int64_t* winOff, int64_t nWin, int k,
					   HashTable t, hipStream_t s);
// hist[0..hmax) counts, plus acc[0] = sum(count) and acc[1] = distinct for
// counts >= hmax, acc[2] = total distinct.
hipError_t launchHistogram(HashTable t, uint64_t* hist, int hmax, uint64_t* acc, hipStream_t s);
// sizes[h] = count for indexed slots, 0 otherwise; marks repetitive slots.
hipError_t launchClassify(HashTable t, int solidMin, int repMax, int32_t* sizes, uint64_t* counters,
						  hipStream_t s);
hipError_t launchAssign(HashTable t, const int32_t* sizes, const int64_t* starts, hipStream_t s);
// Inserts n given keys with their slot values (an index loaded from the
// caller); *err is set when a key is inserted twice.
hipError_t launchLoadKeys(HashTable t, const uint64_t* keys, const uint64_t* vals, int64_t n, uint32_t* err,
						  hipStream_t s);
// Bucket fill of every window, or of the windows with mask[w] set (null:
// all).
hipError_t launchFill(const PackedSeqs& reads, const int64_t* winOff, int64_t nWin, int k, HashTable t,
					  const uint8_t* mask, uint64_t* ordKeys, hipStream_t s);
// mask[window] = 1 at every minimizer position (window of w k-mers; mask
// zeroed by the caller).
hipError_t launchMinimizerMask(const PackedSeqs& reads, const int64_t* winOff, int64_t nWin, int k, int w,
							   uint8_t* mask, hipStream_t s);
// sizes[h] += 1 per masked window of an indexed k-mer (sizes zeroed by the
// caller).
hipError_t launchMaskedSizes(const PackedSeqs& reads, const int64_t* winOff, int64_t nWin, int k, HashTable t,
							 const uint8_t* mask, int32_t* sizes, hipStream_t s);
// Puts every bucket in (read, pos) order and writes its entries: buckets of at
// most 32 entries by one thread each, larger ones (listed in bigList, at most
// bigCap; *bigCount zeroed by the caller) by one 1,024-thread block each.
hipError_t launchBucketFinish(HashTable t, uint64_t* ordKeys, uint64_t* entries, const int32_t* readLen,
							  int k, uint32_t* bigList, uint64_t* bigCount, uint64_t bigCap, hipStream_t s);

// ---- join ---------------------------------------------------------------------
// One query record of a join batch.
struct JoinQuery
{
	int32_t seq;        // index into the packed sequences it is read from
	int32_t strand;     // 1: forward, 0: reverse complement
	uint32_t iid;       // internal id used by the self-hit test
	int32_t selfMode;   // 0: never, 1: resident read, every position indexed (O(1) test),
						// 2: scan bucket, 3: resident read, binary search (minimizer index)
	int32_t len;
	int32_t reserved;
};

struct JoinArgs
{
	PackedSeqs seqs;           // where query sequences are read from
	const JoinQuery* queries;
	int32_t nQueries;
	int k;
	const int64_t* winOff;     // per query window offsets (nQueries + 1)
	int64_t nWin;
	HashTable table;
	const uint64_t* entries;   // (internal id << 32) | position
	const int32_t* targetLen;  // by read number - 1
	int32_t* winCount;         // per window: matches
	int32_t* winFilt;          // per window: 1 if repetitive
	uint64_t* winVal;          // bucket value (start << 24 | count) | flip << 63 (0 when none)
	const int64_t* matchOff;   // exclusive scan of winCount (nWin + 1)
	const int64_t* filtOff;    // exclusive scan of winFilt (nWin + 1)
	uint64_t* mkey;            // (extIid << 32) | curPos, emission order
	int32_t* mext;             // extPos, emission order
	uint32_t* mval;            // emission index within the query (the stable sort's payload;
							   // null: the fill skips it)
	const int64_t* qMatchOff;  // per query match offsets (nQueries + 1)
	int32_t* filtered;         // filtered positions
};

// 2-bit packs n ASCII sequences (bases[baseOff[i] - baseOff[0] ..], lens[i]
// bases) into words at wordOff (16 bases per word, two spare words per
// sequence); *err gets bit 0 for a byte other than A/C/G/T.
hipError_t launchPackBases(const uint8_t* bases, const int64_t* baseOff, const int64_t* wordOff,
						   const int32_t* lens, int32_t n, int64_t nWords, uint32_t* words, uint32_t* err,
						   hipStream_t s);
hipError_t launchJoinCount(const JoinArgs& a, hipStream_t s);
hipError_t launchJoinFill(const JoinArgs& a, hipStream_t s);

// Candidate rows for the ABI: emission order (sorted == null) or through the
// sort permutation.
hipError_t launchWriteCandidates(const JoinArgs& a, const int64_t* qMatchOff, const int64_t* queryIds,
								 const uint32_t* perm, flyehip_candidate_t* out, int64_t nMatches,
								 hipStream_t s);

// ---- grouping after the exact match sort (msort.hip) --------------------------
// After the sort: hit arrays, group-start flags.
hipError_t launchGatherHits(const int64_t* qMatchOff, int32_t nQueries, const uint64_t* key, const uint32_t* val,
							const int32_t* mext, int32_t* hcur, int32_t* hext, int32_t* groupFlag,
							int64_t nMatches, hipStream_t s);
// hext at the sorted positions flagged in dupFlag from the permutation val.
hipError_t launchFixTiedHits(const int64_t* qMatchOff, int32_t nQueries, const int32_t* dupFlag, const uint32_t* val,
							 const int32_t* mext, int32_t* hext, int64_t nMatches, hipStream_t s);
hipError_t launchBuildGroups(const int64_t* qMatchOff, int32_t nQueries, const uint64_t* key,
							 const int32_t* groupFlag, const int64_t* groupOff, const int32_t* targetLen,
							 GroupDesc* groups, int64_t nMatches, hipStream_t s);

hipError_t launchGroupSizes(GroupDesc* groups, int64_t nGroups, int64_t nMatches, hipStream_t s);

// Per-query divStats windows (patch 0023:245-269) over the written rows.
hipError_t launchDivStats(const flyehip_overlap_t* rows, const int64_t* qRowOff, const int32_t* curLen,
						  int32_t nQueries, int32_t* divCount, float* divOut, const int64_t* divOff,
						  hipStream_t s);

} // namespace fh
