// chain.hpp - launch interface of the group-chaining kernels (chain.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/flyehip.h"
#include "segsort.hpp"

namespace fh
{

// One query record of a batch.
struct QueryDesc
{
	int64_t curId;       // signed id
	int64_t filtStart;   // into the batch's filtered-position array
	int32_t curLen;
	int32_t filtCount;
};

// One (query, target) group: hits [hitStart, hitStart + n) of the batch's
// sorted hit arrays, all with the same target strand.
struct GroupDesc
{
	int64_t hitStart;
	int64_t extId;       // signed
	int32_t n;
	int32_t query;       // index into the batch's QueryDesc array
	int32_t extLen;
	int32_t inPlace;     // set by the prefilter: sorted by extPos, but the extPos already
						 // ascend strictly, so the hits are read in place (no sort, no gather)
};

// A chain proposal (an OverlapRange before and after the filters).
struct Proposal
{
	int32_t curBegin;
	int32_t curEnd;
	int32_t extBegin;
	int32_t extEnd;
	int32_t score;
	int32_t chainLength;
	float div;
};

struct ChainArgs
{
	flyehip_overlap_params_t params;
	GroupDesc* groups;         // the prefilter sets inPlace
	const QueryDesc* queries;
	const int32_t* order;      // processing order: size-descending
	int32_t nGroups;
	const int32_t* passOrder;  // groups that pass the prefilters, in processing order
	int32_t nPass;
	int32_t logN;
	const int32_t* cur;        // hits in (extId, curPos) order
	const int32_t* ext;
	const int32_t* filtered;   // per-query ascending filtered positions
	// per-hit scratch (indexed like the hits)
	uint32_t* sortKey;         // 32-bit sort keys (extPos, INT_MAX - score)
	uint32_t* sortVal;         // index within the group, sorted in place with the keys (score order:
							   // bit 31 set on rows with a backtrack pointer)
	int32_t* posL;
	int32_t* posR;
	int32_t* hcur;             // group-order copies (after the extPos re-sort)
	int32_t* hext;
	int32_t* score;
	int32_t* bt;
	Proposal* props;           // per hit (unused with only_max_ext)
	Proposal* outProps;        // per hit; per group with only_max_ext
	// per group
	int32_t* groupPass;
	int32_t* groupCount;
	int32_t* queue;            // work-queue counter of the kernel being launched
	// glibc logf agreement table (sorted by argument bits)
	const uint32_t* logX;
	const uint32_t* logY;
	unsigned long long* dpCells;  // adds the (row, predecessor) pairs the DP evaluates (may be null)
	int32_t passKnown;            // groupPass already holds the prefilters (chainRelevanceKernel)
};

// Walk kernel variants by group size (powers of two, so each is a contiguous
// part of the size-descending processing order): backtrack pointers of up to
// kWalkSmall / kWalkMid / kWalkBig / kWalkHuge hits are staged in LDS as
// 16-bit row indices (4 / 16 / 32 / 64 KiB: the waves per CU fall as the
// stage grows, so each class gets the smallest that holds it); larger groups
// walk them in place.
constexpr int kWalkSmall = 2048;
constexpr int kWalkMid = 8192;
constexpr int kWalkBig = 16384;
constexpr int kWalkHuge = 32768;

// Flye's divergence logf (glibc-exact) for an array of arguments.
hipError_t launchLogf(const float* x, int64_t n, float* out, const uint32_t* tabX, const uint32_t* tabY, int tabN,
					  hipStream_t stream);
hipError_t launchGroupOrder(const GroupDesc* groups, int32_t nGroups, int32_t* bins, int32_t* order,
							hipStream_t stream);
// Before the match sort's tie replay, on the stable order: every group's
// prefilters into groupPass, and groupRel[g] = 1 when the order of the
// group's tied (equal curPos) hits can reach a result (a passing group in
// curPos order, or one sorted by extPos with two hits sharing an extPos);
// tieFlag (cleared by the caller) gets 1 at the tied hits of those groups.
// Uses posL / posR as scratch.
hipError_t launchChainRelevance(const ChainArgs& a, int32_t* groupRel, int32_t* tieFlag, int blocks,
								hipStream_t stream);
// prefilters (or groupPass with passKnown); pushes the extPos sorts to sb
hipError_t launchChainPrefilter(const ChainArgs& a, const SegSortBufs& sb, int blocks, hipStream_t stream);
// passOrder = order filtered by groupPass (flags, scan by the caller, scatter);
// nBig[0] counts passing groups of at least kWalkSmall hits, nBig[1] their hits
// in total (both zeroed by the caller)
hipError_t launchPassFlags(const ChainArgs& a, int32_t* flag, hipStream_t stream);
// tiers (may be null): passing groups of at least kWalkBig hits (tiers[0])
// and of at least kWalkMid hits (tiers[1])
hipError_t launchPassScatter(const ChainArgs& a, const int32_t* flag, const int64_t* pos, int32_t* passOrder,
							 int64_t* nBig, hipStream_t stream, int64_t* tiers = nullptr);
// Groups holding an extPos-sort range whose exact heap sort is still running
// (verify entries with d == 1) get deferred[g] = 1 (deferred zeroed by the
// caller); the split flags then divide the passing groups of the processing
// order into those whose DP can run meanwhile and the deferred ones.
hipError_t launchDeferGroups(const SegRange* verify, int64_t nVerify, const ChainArgs& a, int32_t* deferred,
							 hipStream_t stream);
hipError_t launchSplitPass(const ChainArgs& a, const int32_t* deferred, int32_t* flagReady, int32_t* flagDefer,
						   hipStream_t stream);
// group-order hit copies of passOrder[0, nPass): chunk counts per entry
// (chunks), then, after the caller's exclusive scan of them (chunkOff, nPass
// + 1 words), the copies
hipError_t launchGatherChunks(const ChainArgs& a, int32_t* chunks, hipStream_t stream);
hipError_t launchChainGather(const ChainArgs& a, const int64_t* chunkOff, int blocks, hipStream_t stream);
hipError_t launchChainDp(const ChainArgs& a, int lanes, int blocks, hipStream_t stream);
// writes the score-order keys and values of the groups passOrder[0, nPass)
// into sortKey / sortVal and pushes their ranges to sb
hipError_t launchChainScorePush(const ChainArgs& a, const SegSortBufs& sb, hipStream_t stream);
// backtrack .. selection for passOrder[first, last)
// tier: 0 below kWalkSmall hits, 1 below kWalkMid, 2 below kWalkBig, 3 above
hipError_t launchChainWalk(const ChainArgs& a, int32_t first, int32_t last, int tier, int blocks,
						   hipStream_t stream);
// Backtrack .. selection of passOrder[0, nPass) without the score sort,
// for groups whose walk does not depend on the order of equal scores;
// claimFlag[g] = 1 marks the others (no rows written), which take the score
// sort and launchChainWalk.  Uses posL as scratch.
hipError_t launchChainClaim(const ChainArgs& a, int32_t* claimFlag, int blocks, hipStream_t stream);
hipError_t launchWriteRows(const ChainArgs& a, const int64_t* rowStart, flyehip_overlap_t* rows,
						   hipStream_t stream);

} // namespace fh
