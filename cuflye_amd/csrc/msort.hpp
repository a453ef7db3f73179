// msort.hpp - launch interface of the exact per-query match sort (msort.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fh
{

// The chain's hits in S order, written by the counting sort when hcur is set:
// curPos, extPos (from mext, emission order) and the group-start flags.
struct MsortHits
{
	const int32_t* mext = nullptr;
	int32_t* hcur = nullptr;
	int32_t* hext = nullptr;
	int32_t* groupFlag = nullptr;
};

struct MsortArgs
{
	const int64_t* qOff;   // per query match offsets (nQueries + 1)
	int32_t nQueries;
	const uint64_t* akey;  // emission-order keys: (internal extId << 32) | curPos
	const uint32_t* aval;  // emission index within the query (written by the join)
	uint64_t* skey;        // stable sorted keys
	uint32_t* sval;        // stable sorted permutation (emission index within the query)
	uint32_t* rkey;        // emission-order rank keys (filled here)
	int32_t* dupFlag;      // per sorted position: a neighbouring position holds the same key
	const int32_t* order;  // counting sort: queries by descending match count (null: in order)
	MsortHits hits;        // counting sort (and then the rank keys) only
	uint32_t* ival = nullptr;  // set: the rank keys also write the emission indices
							   // (identity, in emission order) of the queries they cover
};

hipError_t msortTempBytes(int64_t nMatches, int32_t nQueries, int idBits, size_t* bytes);
// Stable order by extId (emission order is already ascending in curPos):
// a per-query counting sort in LDS when the nBins target ids fit it
// (msortCounting; it also writes a.hits when set), else a segmented radix
// sort on the extId bits.
bool msortCounting(int nBins);
hipError_t launchMatchSortStable(const MsortArgs& a, int64_t nMatches, int idBits, int nBins, void* temp,
								 size_t tempBytes, hipStream_t s);
// Tied-position flags of the stable order: a neighbouring position holds the
// same key.
hipError_t launchTieFlags(const MsortArgs& a, int64_t nMatches, hipStream_t s);
// Rank keys (emission order) of the queries that hold a flagged position;
// dup is the exclusive prefix of the flags (nMatches + 1 words).  Keys from
// a.hits (curPos and group starts) when a.hits.hcur is set, else from skey.
hipError_t launchRankKeys(const MsortArgs& a, const int64_t* dup, int64_t nMatches, hipStream_t s);

} // namespace fh
