// exsort_dev.hpp - block-cooperative exact std::sort emulation on the GPU.
//
// One thread block sorts one array.  Ranges larger than `laneMax` are split by
// cooperative partitions (the rank formulation of exsort.hpp: two counts per
// thread, one block scan, a scatter of swap partners, parallel swaps).
// Ranges of at most `laneMax` elements, and ranges whose depth budget is
// exhausted, are finished serially by one lane each: the introsort loop with
// the remaining budget, then the stable insertion pass.  Because every range
// handed to a lane is bounded by pivots of earlier partitions, the per-range
// insertion passes together equal std::sort's final insertion sort.
#pragma once

#include "dev_common.hpp"
#include "exsort.hpp"

namespace fh
{

struct Seg
{
	int f;
	int l;
	int d;
};

// LDS state a block needs for one cooperative sort.
template <int NT>
struct SortShared
{
	int scan[2 * (NT / FH_WAVE) + 2];
	int lm;
	int m;
	int stackTop;
	int smallCount;
	int overflow;
	Seg stack[80];
};

// Cooperative unguarded partition of s[f, l) (l - f > 16).  Returns the cut.
template <int NT>
__device__ int coopPartition(const exsort::Span& s, int f, int l, int* posL, int* posR,
							 SortShared<NT>& sh)
{
	const int t = static_cast<int>(threadIdx.x);
	if (t == 0)
	{
		exsort::medianToFirst(s, f, f + 1, f + (l - f) / 2, l - 1);
		sh.lm = 0x7fffffff;
	}
	__syncthreads();
	const uint64_t p = s.k[f];
	const int lo = f + 1;
	const int len = l - lo;
	const int chunk = (len + NT - 1) / NT;
	const int b = lo + t * chunk;
	const int e = min(l, b + chunk);
	int ge = 0, le = 0;
	for (int i = b; i < e; ++i)
	{
		const uint64_t v = s.k[i];
		ge += !(v < p);
		le += !(p < v);
	}
	int geBase, leBase, GE, LE;
	blockScan2<NT>(ge, le, geBase, leBase, GE, LE, sh.scan);
	int swaps = 0;
	int geLeft = geBase, leLeft = leBase;
	for (int i = b; i < e; ++i)
	{
		const uint64_t v = s.k[i];
		const bool isGe = !(v < p);
		const bool isLe = !(p < v);
		const int leRight = LE - leLeft - (isLe ? 1 : 0);
		if (isGe)
		{
			if (geLeft < leRight)
			{
				posL[geLeft] = i;
				++swaps;
			}
		}
		if (isLe && leRight < geLeft) posR[leRight] = i;
		geLeft += isGe;
		leLeft += isLe;
	}
	const int m = blockSum<NT>(swaps, sh.scan);
	// L[m]: the ge element whose left rank is m (the first one that does not swap)
	geLeft = geBase;
	for (int i = b; i < e; ++i)
	{
		const uint64_t v = s.k[i];
		if (!(v < p))
		{
			if (geLeft == m) sh.lm = i;
			++geLeft;
		}
	}
	__syncthreads();
	for (int q = t; q < m; q += NT) s.swap(posL[q], posR[q]);
	__syncthreads();
	int cut = sh.lm;
	if (m > 0)
	{
		int r = posR[m - 1];
		if (r < cut) cut = r;
	}
	__syncthreads();
	return cut;
}

// Exact std::sort of s[0, n) by one block of NT threads.  posL/posR need n
// ints each (same memory space as s is fine).  `small` is a list for up to
// `smallCap` ranges handed to lanes.
template <int NT>
__device__ void blockStdSort(const exsort::Span& s, int n, int* posL, int* posR, Seg* small,
							 int smallCap, int laneMax, SortShared<NT>& sh)
{
	if (n < 2) return;
	const int t = static_cast<int>(threadIdx.x);
	if (t == 0)
	{
		sh.stackTop = 0;
		sh.smallCount = 0;
		sh.overflow = 0;
		sh.stack[sh.stackTop++] = Seg{0, n, exsort::depthLimit(static_cast<uint32_t>(n))};
	}
	__syncthreads();
	for (;;)
	{
		const int top = sh.stackTop;
		if (top == 0) break;
		Seg seg = sh.stack[top - 1];
		__syncthreads();
		if (t == 0) sh.stackTop = top - 1;
		const int size = seg.l - seg.f;
		if (size <= laneMax || seg.d == 0)
		{
			if (t == 0)
			{
				if (sh.smallCount < smallCap) small[sh.smallCount++] = seg;
				else sh.overflow = 1;
			}
			__syncthreads();
			continue;
		}
		const int cut = coopPartition<NT>(s, seg.f, seg.l, posL, posR, sh);
		if (t == 0)
		{
			// both halves get the decremented budget (std::__introsort_loop)
			sh.stack[sh.stackTop++] = Seg{seg.f, cut, seg.d - 1};
			sh.stack[sh.stackTop++] = Seg{cut, seg.l, seg.d - 1};
		}
		__syncthreads();
	}
	const int cnt = sh.smallCount;
	for (int q = t; q < cnt; q += NT)
	{
		const Seg g = small[q];
		exsort::introLoop(s, g.f, g.l, g.d);
		exsort::insertionSort(s, g.f, g.l);
	}
	__syncthreads();
}

} // namespace fh
