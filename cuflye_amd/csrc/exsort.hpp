// exsort.hpp - exact emulation of libstdc++'s std::sort (introsort) ordering.
//
// Flye sorts with std::sort at every order-sensitive point of the overlap
// path (k-mer matches by (extId, curPos), a target group by extPos, DP rows by
// score, proposals by score, read-to-graph alignments by curBegin, chains by
// score).  std::sort is not stable, and on equal keys its output order is
// whatever introsort's partition swaps leave behind; the chaining DP, the
// backtrack and onlyMaxExt all read that order.  Bit-exact results therefore
// need the same permutation, not just a sorted array.
//
// The steps reproduced are those of libstdc++'s introsort as shipped since
// GCC 4.x (bits/stl_algo.h, unchanged through GCC 13): threshold 16; pivot by
// median of (first+1, mid, last-1) moved to *first; unguarded Hoare partition
// of [first+1, last) around *first; recursion on the right part, loop on the
// left; depth limit 2*floor(log2 n), after which the range is heap-sorted
// (make_heap + sort_heap); a final insertion sort over the whole range.  Each
// step is rewritten here from that published description; the tests compare
// every variant against the real std::sort of the host libstdc++.
//
// Two consequences used by the GPU code:
//  * The final insertion sort is stable and every leaf range left by the
//    introsort loop is bounded by its pivots, so the final pass equals a
//    stable sort of each leaf (equivalently of the whole range).
//  * One unguarded partition is a pure function of two flag sequences:
//    ge[i] = !(a[i] < p) and le[i] = !(p < a[i]) over [first+1, last).  The
//    k-th ge position from the left (L[k]) swaps with the k-th le position from
//    the right (R[k]) while L[k] < R[k]; with m such swaps the cut is L[0]
//    when m == 0, else min(L[m], R[m-1]) (L[m] = +inf when it does not exist).
//    A ge element at i with left rank kl swaps iff kl < (number of le elements
//    right of i); an le element at j with right rank kr swaps iff
//    kr < (number of ge elements left of j).  Both need only prefix counts,
//    so a partition runs as scans over all lanes.
//
// Keys are unsigned integers (uint64 or uint32: every sort of the path maps
// its comparator onto an ascending unsigned key); values are uint32 payloads
// moved with their keys.

#pragma once

#include <stdint.h>

#if defined(__HIPCC__)
#define EXS_HD __host__ __device__ inline
#else
#define EXS_HD inline
#endif

namespace exsort
{

constexpr int kThreshold = 16;

EXS_HD int floorLog2(uint32_t n)
{
	int r = 0;
	while (n > 1)
	{
		n >>= 1;
		++r;
	}
	return r;
}

EXS_HD int depthLimit(uint32_t n) { return 2 * floorLog2(n); }

// Array view over separate key/value arrays (global or LDS memory).
template <class K = uint64_t>
struct Span
{
	K* k;
	uint32_t* v;
	EXS_HD void swap(int a, int b) const
	{
		K tk = k[a];
		k[a] = k[b];
		k[b] = tk;
		uint32_t tv = v[a];
		v[a] = v[b];
		v[b] = tv;
	}
};

// move_median_to_first(result, a, b, c)
template <class K>
EXS_HD void medianToFirst(const Span<K>& s, int result, int a, int b, int c)
{
	const K ka = s.k[a], kb = s.k[b], kc = s.k[c];
	if (ka < kb)
	{
		if (kb < kc) s.swap(result, b);
		else if (ka < kc) s.swap(result, c);
		else s.swap(result, a);
	}
	else if (ka < kc) s.swap(result, a);
	else if (kb < kc) s.swap(result, c);
	else s.swap(result, b);
}

// Hoare partition of [first, last) around the value at `pivot` (unguarded).
template <class K>
EXS_HD int unguardedPartition(const Span<K>& s, int first, int last, int pivot)
{
	const K p = s.k[pivot];
	for (;;)
	{
		while (s.k[first] < p) ++first;
		--last;
		while (p < s.k[last]) --last;
		if (!(first < last)) return first;
		s.swap(first, last);
		++first;
	}
}

template <class K>
EXS_HD int partitionPivot(const Span<K>& s, int first, int last)
{
	int mid = first + (last - first) / 2;
	medianToFirst(s, first, first + 1, mid, last - 1);
	return unguardedPartition(s, first + 1, last, first);
}

// __push_heap: sift `value` up from holeIndex toward topIndex.
template <class K>
EXS_HD void pushHeap(const Span<K>& s, int base, int hole, int top, K vk, uint32_t vv)
{
	int parent = (hole - 1) / 2;
	while (hole > top && s.k[base + parent] < vk)
	{
		s.k[base + hole] = s.k[base + parent];
		s.v[base + hole] = s.v[base + parent];
		hole = parent;
		parent = (hole - 1) / 2;
	}
	s.k[base + hole] = vk;
	s.v[base + hole] = vv;
}

// __adjust_heap over heap [base, base + len).
template <class K>
EXS_HD void adjustHeap(const Span<K>& s, int base, int hole, int len, K vk, uint32_t vv)
{
	const int top = hole;
	int child = hole;
	while (child < (len - 1) / 2)
	{
		child = 2 * (child + 1);
		if (s.k[base + child] < s.k[base + child - 1]) --child;
		s.k[base + hole] = s.k[base + child];
		s.v[base + hole] = s.v[base + child];
		hole = child;
	}
	if ((len & 1) == 0 && child == (len - 2) / 2)
	{
		child = 2 * (child + 1);
		s.k[base + hole] = s.k[base + child - 1];
		s.v[base + hole] = s.v[base + child - 1];
		hole = child - 1;
	}
	pushHeap(s, base, hole, top, vk, vv);
}

// partial_sort(first, last, last): make_heap then sort_heap.
template <class K>
EXS_HD void heapSort(const Span<K>& s, int first, int last)
{
	const int len = last - first;
	if (len < 2) return;
	for (int parent = (len - 2) / 2;; --parent)
	{
		K vk = s.k[first + parent];
		uint32_t vv = s.v[first + parent];
		adjustHeap(s, first, parent, len, vk, vv);
		if (parent == 0) break;
	}
	for (int end = last; end - first > 1;)
	{
		--end;
		K vk = s.k[end];
		uint32_t vv = s.v[end];
		s.k[end] = s.k[first];
		s.v[end] = s.v[first];
		adjustHeap(s, first, 0, end - first, vk, vv);
	}
}

// The same heap sort on packed elements (32-bit key << 32 | 32-bit value),
// compared on the key half only: one word per element move, which halves the
// memory round trips of the dependent sift chain.
EXS_HD bool packedLess(uint64_t a, uint64_t b) { return (a >> 32) < (b >> 32); }

EXS_HD void pushHeapPacked(uint64_t* a, int hole, int top, uint64_t v)
{
	int parent = (hole - 1) / 2;
	while (hole > top && packedLess(a[parent], v))
	{
		a[hole] = a[parent];
		hole = parent;
		parent = (hole - 1) / 2;
	}
	a[hole] = v;
}

EXS_HD void adjustHeapPacked(uint64_t* a, int hole, int len, uint64_t v)
{
	const int top = hole;
	int child = hole;
	while (child < (len - 1) / 2)
	{
		child = 2 * (child + 1);
		const uint64_t r = a[child], l = a[child - 1];
		if (packedLess(r, l))
		{
			--child;
			a[hole] = l;
		}
		else
		{
			a[hole] = r;
		}
		hole = child;
	}
	if ((len & 1) == 0 && child == (len - 2) / 2)
	{
		child = 2 * (child + 1);
		a[hole] = a[child - 1];
		hole = child - 1;
	}
	pushHeapPacked(a, hole, top, v);
}

// partial_sort(first, first + n, first + n) on packed elements.
EXS_HD void heapSortPacked(uint64_t* a, int n)
{
	if (n < 2) return;
	for (int parent = (n - 2) / 2;; --parent)
	{
		adjustHeapPacked(a, parent, n, a[parent]);
		if (parent == 0) break;
	}
	for (int end = n - 1; end > 0; --end)
	{
		const uint64_t v = a[end];
		a[end] = a[0];
		adjustHeapPacked(a, 0, end, v);
	}
}

// ---- pipelined heap sort (exactly partial_sort(first, last, last)) -------------
//
// libstdc++'s __adjust_heap (hole down the larger-child path to a leaf, then
// __push_heap of the value back up) leaves the same array as a top-down sift:
// the larger child (the right one unless right < left) moves up while
// !(child < value); the value settles where that fails.  The path depends only
// on the children, and both forms stop at the deepest path node whose child
// is below the value.
//
// make_heap sifts parents from the last one up to the root; parents of one
// heap level have disjoint subtrees, so a level can be sifted in any order.
// sort_heap pops in sequence, but pop t only reads a level after pop t-1 has
// written it for the last time: started two steps apart, pops advance one
// level per step in lock step, each on its own level.  A pop may start only
// when the previous one has taken two steps, and when no running pop's hole
// is an ancestor of the position the new pop reads its value from and writes
// the popped root to (the last heap position).  The steps run in pairs and a
// pop may start only in the first step of a pair (so two steps apart at
// least; a pop that cannot start waits for the next pair, which the checks
// above make harmless).  heapSortPipelinedModel runs exactly this schedule on
// the host; the GPU kernel runs the same steps with one lane per running pop.

EXS_HD bool isAncestorOrSelf(int h, int x)
{
	// 1-based heap numbering: the ancestor of X k levels up is X >> k
	unsigned H = static_cast<unsigned>(h) + 1u, X = static_cast<unsigned>(x) + 1u;
	if (H > X) return false;
	int dh = 0, dx = 0;
	for (unsigned t = H; t > 1; t >>= 1) ++dh;
	for (unsigned t = X; t > 1; t >>= 1) ++dx;
	return (X >> (dx - dh)) == H;
}

// One top-down sift step at hole h in a heap of m packed elements holding v;
// returns false (and stores v) when the value has settled.
EXS_HD bool siftStepPacked(uint64_t* a, int& h, int m, uint64_t v)
{
	const int c1 = 2 * h + 1;
	if (c1 >= m)
	{
		a[h] = v;
		return false;
	}
	int child = c1;
	uint64_t cv = a[c1];
	if (c1 + 1 < m)
	{
		const uint64_t c2v = a[c1 + 1];
		if (!packedLess(c2v, cv))
		{
			child = c1 + 1;
			cv = c2v;
		}
	}
	if (packedLess(cv, v))
	{
		a[h] = v;
		return false;
	}
	a[h] = cv;
	h = child;
	return true;
}

EXS_HD void siftPacked(uint64_t* a, int h, int m, uint64_t v)
{
	while (siftStepPacked(a, h, m, v))
	{
	}
}

// make_heap, level by level from the deepest parents up.
EXS_HD void makeHeapByLevelsPacked(uint64_t* a, int n)
{
	if (n < 2) return;
	const int lastParent = (n - 2) / 2;
	int top = 0;
	for (unsigned t = static_cast<unsigned>(lastParent) + 1u; t > 1; t >>= 1) ++top;
	for (int lv = top; lv >= 0; --lv)
	{
		const int lo = (1 << lv) - 1;
		const int hi = (lastParent < (2 << lv) - 2) ? lastParent : (2 << lv) - 2;
		for (int p = hi; p >= lo; --p) siftPacked(a, p, n, a[p]);
	}
}

// Host model of the GPU schedule (kSlots pop slots, pops two steps apart).
template <int kSlots = 16>
inline void heapSortPipelinedModel(uint64_t* a, int n)
{
	if (n < 2) return;
	makeHeapByLevelsPacked(a, n);
	bool act[kSlots] = {};
	int hole[kSlots] = {}, size[kSlots] = {};
	uint64_t val[kSlots] = {};
	int L = n, pops = 0;
	for (int cyc = 0;; ++cyc)
	{
		if (L > 1 && (cyc & 1) == 0)
		{
			const int slot = pops % kSlots;
			bool ok = !act[slot];
			for (int j = 0; j < kSlots; ++j)
				if (act[j] && isAncestorOrSelf(hole[j], L - 1)) ok = false;
			if (ok)
			{
				val[slot] = a[L - 1];
				a[L - 1] = a[0];
				hole[slot] = 0;
				size[slot] = L - 1;
				act[slot] = true;
				--L;
				++pops;
			}
		}
		bool any = false;
		for (int j = 0; j < kSlots; ++j) any |= act[j];
		if (!any && L <= 1 && (cyc & 1) == 0) break;
		for (int j = 0; j < kSlots; ++j)
			if (act[j]) act[j] = siftStepPacked(a, hole[j], size[j], val[j]);
	}
}

// Stable insertion sort of [first, last): the final pass of std::sort.
template <class K>
EXS_HD void insertionSort(const Span<K>& s, int first, int last)
{
	for (int i = first + 1; i < last; ++i)
	{
		K vk = s.k[i];
		uint32_t vv = s.v[i];
		int j = i;
		while (j > first && vk < s.k[j - 1])
		{
			s.k[j] = s.k[j - 1];
			s.v[j] = s.v[j - 1];
			--j;
		}
		s.k[j] = vk;
		s.v[j] = vv;
	}
}

// The introsort loop on [first, last) with a given remaining depth budget,
// without the final insertion pass.  libstdc++ recurses on the right part and
// loops on the left; both parts get the decremented budget and are disjoint,
// so the order they are processed in does not change the result.  Here the
// smaller part is processed first and the larger one waits on an explicit
// stack, which bounds the stack by log2(n) + 1 entries.
template <int kStack = 40, class K = uint64_t>
EXS_HD void introLoop(const Span<K>& s, int first, int last, int depth)
{
	int stF[kStack];
	int stL[kStack];
	int stD[kStack];
	int top = 0;
	stF[top] = first;
	stL[top] = last;
	stD[top] = depth;
	++top;
	while (top > 0)
	{
		--top;
		int f = stF[top];
		int l = stL[top];
		int d = stD[top];
		while (l - f > kThreshold)
		{
			if (d == 0)
			{
				heapSort(s, f, l);
				break;
			}
			--d;
			const int cut = partitionPivot(s, f, l);
			if (cut - f < l - cut)
			{
				stF[top] = cut;
				stL[top] = l;
				stD[top] = d;
				++top;
				l = cut;
			}
			else
			{
				stF[top] = f;
				stL[top] = cut;
				stD[top] = d;
				++top;
				f = cut;
			}
		}
	}
}

// Whole std::sort on [first, last): introsort loop + final insertion sort.
template <class K>
EXS_HD void stdSort(const Span<K>& s, int first, int last)
{
	if (last - first < 2) return;
	introLoop(s, first, last, depthLimit(static_cast<uint32_t>(last - first)));
	insertionSort(s, first, last);
}

// The same partition computed from rank arithmetic instead of the scanning
// loop; the GPU runs this formulation with scans over many lanes.  Host
// reference implementation (sequential), used by the tests to check the
// formulation against unguardedPartition.
template <class K>
inline int partitionByRanks(const Span<K>& s, int first, int last, int* posL, int* posR)
{
	// caller has already moved the median to `first`
	const K p = s.k[first];
	const int lo = first + 1;
	int leTotal = 0;
	for (int i = lo; i < last; ++i) leTotal += !(p < s.k[i]);
	int geLeft = 0, leLeft = 0, m = 0, lm = -1;
	for (int i = lo; i < last; ++i)
	{
		const bool ge = !(s.k[i] < p);
		const bool le = !(p < s.k[i]);
		const int leRightOfI = leTotal - leLeft - (le ? 1 : 0);
		if (ge)
		{
			if (geLeft < leRightOfI)
			{
				posL[geLeft] = i;
				++m;
			}
			else if (geLeft == m && lm < 0)
			{
				lm = i;
			}
		}
		if (le)
		{
			const int kr = leRightOfI;
			if (kr < geLeft) posR[kr] = i;
		}
		geLeft += ge;
		leLeft += le;
	}
	int cut;
	if (m == 0)
	{
		cut = lm;
	}
	else
	{
		cut = posR[m - 1];
		if (lm >= 0 && lm < cut) cut = lm;
	}
	for (int t = 0; t < m; ++t) s.swap(posL[t], posR[t]);
	return cut;
}

} // namespace exsort
