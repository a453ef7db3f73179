// readalign.hpp - launch interface of the read-to-graph chaining kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/flyehip.h"

namespace fh
{

struct ReadAlignArgs
{
	const flyehip_edge_alignment_t* aln;  // all reads' alignments
	const int64_t* alnOff;                // per read (nReads + 1)
	int32_t nReads;
	int32_t longEdge;
	int32_t bigAln;
	flyehip_chain_params_t p;
	// per-alignment scratch (indexed like aln)
	int32_t* kept;        // kept alignment indices (within the read) in sorted order
	uint64_t* key;
	uint32_t* val;
	int32_t* chainScore;  // chain created by kept alignment i
	int32_t* chainPrev;   // chain it extended (-1: none)
	int32_t* chainLen;
	int32_t* active;
	int32_t* frozen;
	int32_t* accepted;
	// per read outputs
	int32_t* keptCount;
	int32_t* acceptedCount;
	int32_t* segCount;
};

hipError_t launchReadAlignChain(const ReadAlignArgs& a, hipStream_t s);
hipError_t launchReadAlignWrite(const ReadAlignArgs& a, const int64_t* segOff, flyehip_chain_segment_t* out,
								hipStream_t s);
// out[keptOff[r] + i] = kept index i of read r (keptOff: exclusive scan of keptCount).
hipError_t launchReadAlignKept(const ReadAlignArgs& a, const int64_t* keptOff, int32_t* out, hipStream_t s);

} // namespace fh
