// segsort.hpp - exact std::sort of many arrays at once (segsort.hip).
//
// Every array of a batch (a query's k-mer matches, a target group's hits) is
// sorted exactly as libstdc++'s std::sort orders it (exsort.hpp).  Introsort
// partitions disjoint ranges independently, so the ranges of all arrays are
// processed level by level instead of array by array:
//   * a "big" range (more than kSegTile elements) is partitioned by many
//     waves at once: its elements are cut into chunks of kSegChunk, the
//     ge / le counts of every chunk are scanned over the whole level, and the
//     swap partners come from the scanned ranks (exsort.hpp's rank form of the
//     Hoare partition), so one level costs a few bandwidth-bound passes over
//     the data of all arrays, not a chain of dependent steps per array;
//   * a range of at most kSegTile elements ("leaf") is finished by one wave in
//     LDS with its remaining depth budget (waveTileSort);
//   * a big range whose depth budget is spent would be heap-sorted by
//     std::sort.  Heapsort is a chain of dependent sift steps, but its result
//     differs from any other sort only in the order of equal keys; so the
//     range is saved ("shadow" copy) and sorted on in free mode (partitions
//     without a depth budget, children tagged with the range); afterwards its
//     sorted keys are checked for an equal neighbour pair, and only a range
//     that has one is heap-sorted from the shadow copy (one lane, in LDS);
//   * inside a leaf, a sub-range of more than kSegLaneMax elements whose
//     budget is spent is handled the same way by the leaf's wave (shadow
//     copy, verify entry, free-mode sort) instead of being heap-sorted by one
//     lane, so the leaves of nearly sorted inputs (where std::sort exhausts its
//     budget everywhere) stay parallel.
// Optionally a range is dropped when the sorted order already known for it is
// final: with `dup` set (exclusive prefix counts over the stable sorted order
// of the positions whose key is shared with a neighbour, "tied" positions), a
// range [f, l) without a tied position holds keys that occur nowhere else in
// the array, so its content and order are those of the stable order.  (An
// equal pair inside the range is not the right test: a tie group split by a
// partition leaves single members in two ranges.)
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "exsort.hpp"
#include "exsort_wave.hpp"
#include "ktimer.hpp"

namespace fh
{

#ifndef FH_SEG_TILE
#define FH_SEG_TILE 1024
#endif
constexpr int kSegTile = FH_SEG_TILE;  // leaf size, elements (LDS tile of one wave)
constexpr int kSegLaneMax = 48;        // leaf sub-ranges finished by one lane
constexpr int kSegRows = 8;           // 64-element rows per chunk
constexpr int kSegChunk = 64 * kSegRows;

constexpr int kSegFreeDepth = 64;      // depth budget of a free-mode range


// Counter words (int64) shared by the init kernels and the rounds.
enum SegCounter
{
	kSegNBig = 0,     // ranges in the list being filled
	kSegBigChunks,    // their chunk total
	kSegBigElems,     // their element total
	kSegNLeaf,
	kSegLeafElems,
	kSegNHeap,        // free-mode ranges out of budget
	kSegHeapElems,
	kSegTieArrays,    // arrays with at least one tied position (pruned sorts)
	kSegError,        // set when a partition produced an impossible cut or swap
	kSegNVerify,      // exact-mode ranges whose depth budget ran out
	kSegExactElems,   // elements of the verify entries that needed the exact heap sort
	kSegHeapCls,      // those entries per heap-sort class: 3 LDS classes, then global memory
	kSegCounters = kSegHeapCls + 4
};

struct SegSortBufs
{
	uint32_t* key;          // sorted in place (big ranges)
	uint32_t* val;
	uint32_t* outVal;       // leaves and heap ranges write their final values here
	uint32_t relevantBit;   // 0, or a value bit marking the elements whose mutual order
							// matters; a free-mode result is then accepted unless a run of
							// equal keys holds two marked elements (outVal == val only)
	const int64_t* dup;     // tied-position prefix counts, n + 1 words (null: no pruning;
							// pruning needs a separate outVal holding the stable order)
	int32_t* posL;          // one int per element
	int32_t* posR;
	SegRange* big[2];       // ping-pong lists of big ranges
	SegRange* leaf;
	SegRange* heap;         // free-mode ranges whose (free) budget ran out
	SegRange* verify;       // exact-mode ranges whose budget ran out
	uint32_t* shadowKey;    // their content when the budget ran out (one per element)
	uint32_t* shadowVal;
	int64_t* counters;      // kSegCounters words
	int64_t* chunkOff;      // per big range of a round: chunk offsets (bigCap + 1)
	int32_t* chunkRange;    // per chunk (chunkCap)
	uint64_t* chunkCnt;     // per chunk: (ge << 32) | le  (chunkCap + 1)
	uint64_t* chunkPre;     // exclusive scan of chunkCnt (max(chunkCap, bigCap) + 1)
	int32_t* rangeM;        // swaps per range (bigCap)
	int32_t* rangeLm;       // first ge element that is not swapped (bigCap)
	void* scanTemp;
	size_t scanTempBytes;
	int64_t bigCap;
	int64_t leafCap;
	int64_t heapCap;
	int64_t verifyCap;      // verify entries: big ranges and leaf sub-ranges, all disjoint
	int64_t chunkCap;
	int64_t* hostCounters;  // pinned, kSegCounters words
	hipStream_t side;       // default-priority side stream of the deferred exact heap sorts
	hipEvent_t forkEv;      // (timing disabled) side stream forked
	hipEvent_t doneEv;      // (timing disabled) deferred exact heap sorts done
	KTimer* timer;          // kernel-family intervals (may be null)
};

struct SegSortCaps
{
	int64_t bigCap;
	int64_t leafCap;
	int64_t heapCap;
	int64_t verifyCap;
	int64_t chunkCap;
	size_t scanTempBytes;
};

// Capacities for arrays totalling n elements in nArrays arrays.
hipError_t segSortCaps(int64_t n, int64_t nArrays, SegSortCaps* caps);

__device__ inline unsigned long long segAppend(int64_t* counter, unsigned long long m)
{
	const int leader = __ffsll(static_cast<long long>(m)) - 1;
	unsigned long long base = 0;
	if (laneId() == leader)
		base = atomicAdd(reinterpret_cast<unsigned long long*>(counter), static_cast<unsigned long long>(__popcll(m)));
	base = __shfl(base, leader, FH_WAVE);
	return base + __popcll(m & laneMaskLt());
}

// Appends a range [f, l) with depth d and tag to the lists: leaf, big
// (bigList), verify + big (an exact-mode range out of budget), or heap (a
// free-mode range out of budget).  Every lane of the wave calls it; lanes with
// push == false add nothing.  Ranges pruned by `dup` (exact mode) are dropped;
// a one-element range is final in place and only copies its value when the
// output array is separate.  One atomic per list and wave.
__device__ inline void segPushWave(const SegSortBufs& b, SegRange* bigList, bool push, int32_t f, int32_t l, int32_t d,
								   int32_t tag)
{
	const int32_t n = l - f;
	if (push && n == 1 && b.outVal != b.val && (tag != 0 || !b.dup || b.dup[l] - b.dup[f] > 0))
		b.outVal[f] = b.val[f];
	if (n < 2) push = false;
	if (push && tag == 0 && b.dup && b.dup[l] - b.dup[f] == 0) push = false;
	const bool toLeaf = push && n <= kSegTile;
	const bool toVerify = push && n > kSegTile && d == 0 && tag == 0;
	const bool toHeap = push && n > kSegTile && d == 0 && tag != 0;
	unsigned long long m = __ballot(toLeaf);
	if (m)
	{
		const unsigned long long i = segAppend(&b.counters[kSegNLeaf], m);
		if (toLeaf) b.leaf[i] = SegRange{f, l, d, tag};
		const long long e = waveSum64(toLeaf ? n : 0);
		if (laneId() == __ffsll(static_cast<long long>(m)) - 1)
			atomicAdd(reinterpret_cast<unsigned long long*>(&b.counters[kSegLeafElems]), static_cast<unsigned long long>(e));
	}
	m = __ballot(toHeap);
	if (m)
	{
		const unsigned long long i = segAppend(&b.counters[kSegNHeap], m);
		if (toHeap) b.heap[i] = SegRange{f, l, 0, tag};
		const long long e = waveSum64(toHeap ? n : 0);
		if (laneId() == __ffsll(static_cast<long long>(m)) - 1)
			atomicAdd(reinterpret_cast<unsigned long long*>(&b.counters[kSegHeapElems]), static_cast<unsigned long long>(e));
	}
	m = __ballot(toVerify);
	if (m)
	{
		const unsigned long long i = segAppend(&b.counters[kSegNVerify], m);
		if (toVerify)
		{
			b.verify[i] = SegRange{f, l, 0, 0};
			d = kSegFreeDepth;
			tag = static_cast<int32_t>(i) + 1;
		}
	}
	const bool toBig = push && n > kSegTile && d > 0;
	m = __ballot(toBig);
	if (m)
	{
		const int64_t chunks = toBig ? (static_cast<int64_t>(n) - 1 + kSegChunk - 1) / kSegChunk : 0;
		const int64_t chunkSum = waveSum64(chunks);
		const int64_t elemSum = waveSum64(toBig ? n : 0);
		const unsigned long long i = segAppend(&b.counters[kSegNBig], m);
		if (laneId() == __ffsll(static_cast<long long>(m)) - 1)
		{
			atomicAdd(reinterpret_cast<unsigned long long*>(&b.counters[kSegBigChunks]),
					  static_cast<unsigned long long>(chunkSum));
			atomicAdd(reinterpret_cast<unsigned long long*>(&b.counters[kSegBigElems]),
					  static_cast<unsigned long long>(elemSum));
		}
		if (toBig) bigList[i] = SegRange{f, l, d, tag};
	}
}

// Zeroes the counters (before the init kernels push the first ranges).
hipError_t segSortReset(const SegSortBufs& b, hipStream_t s);

// Runs the rounds and the leaf / heap passes over what the init kernels
// pushed.  Synchronises the stream once per round (range counts).  With
// exactPending set, the exact heap sorts of the verify entries that need one
// (d == 1 after the verification on s) are left running on the side stream:
// *exactPending tells whether any were launched, and b.doneEv completes with
// them; the caller makes its stream wait on that event before it reads those
// ranges of outVal (or the side stream's timing intervals).
hipError_t segSortRun(SegSortBufs& b, hipStream_t s, int64_t* tieArrays, bool* exactPending = nullptr);

// Pushes every array [off[i], off[i+1]) (i < nArrays), pruned by b.dup.
hipError_t segPushOffsets(const SegSortBufs& b, const int64_t* off, int32_t nArrays, hipStream_t s);

} // namespace fh
