// msort.hpp - launch interface of the exact per-query match sort (msort.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fh
{

struct MsortArgs
{
	const int64_t* qOff;   // per query match offsets (nQueries + 1)
	int32_t nQueries;
	const int32_t* order;  // replay order (may be null)
	uint64_t* akey;        // emission-order keys; replayed in place for queries with ties
	uint32_t* aval;        // emission index within the query (filled here)
	uint64_t* skey;        // result keys: (internal extId << 32) | curPos
	uint32_t* sval;        // result permutation: emission index within the query
	int32_t* posL;         // partition scratch, one int per match
	int32_t* posR;
	uint8_t* tiedByIdx;    // one byte per match
	uint8_t* fromA;        // one byte per match, zeroed by the caller
	int32_t* qTied;        // per query, zeroed by the caller
	int32_t* queue;        // zeroed by the caller
	int32_t* tieQueries;   // zeroed by the caller
};

hipError_t msortTempBytes(int64_t nMatches, int32_t nQueries, int endBit, size_t* bytes);
hipError_t launchMatchSortExact(const MsortArgs& a, int64_t nMatches, int endBit, void* temp, size_t tempBytes,
								int replayBlocks, hipStream_t s);

} // namespace fh
