// exsort_wave.hpp - one wave64 sorts one array exactly like libstdc++'s
// std::sort (see exsort.hpp for the algorithm and its rank formulation).
//
//  * ranges larger than the LDS tile are partitioned in place in global
//    memory by the whole wave: rows of 64 consecutive elements, ranks from
//    ballot + popcount, swap partners scattered to posL/posR, swaps in
//    parallel;
//  * a range of at most kTile elements is copied into the wave's LDS tile,
//    partitioned there the same way down to kLaneMax elements, and the
//    remaining ranges are finished one per lane (introsort loop with the
//    remaining depth budget + insertion pass), 64 ranges at a time;
//  * the heapsort fallback (depth budget exhausted on a large range) runs on
//    lane 0 in global memory; it needs adversarial inputs and is rare.
//
// No block barriers: every wave is independent, so one kernel can give each
// wave a different array (a query's k-mer matches, a target group).  Memory
// written by one lane and read by another is ordered with a workgroup-scope
// fence (all lanes of a wave share the CU's L1) plus a wave barrier.
#pragma once

#include "dev_common.hpp"
#include "exsort.hpp"

namespace fh
{

struct WSeg
{
	int f;
	int l;
	int d;
};

// LDS of one wave's tile sort.  Swap positions inside the tile are 16-bit
// (kTile <= 65536): the smaller footprint raises the waves per CU.
template <class K, int kTile>
struct WaveSortLds
{
	static_assert(kTile <= 65536, "tile positions are 16-bit");
	K k[kTile];
	uint32_t v[kTile];
	uint16_t posL[kTile / 2 + 1];
	uint16_t posR[kTile / 2 + 1];
	WSeg stack[40];
	WSeg lanes[FH_WAVE];
};

// waveStdSort also keeps the stack of ranges larger than the tile.
template <class K, int kTile>
struct WaveStdSortLds : WaveSortLds<K, kTile>
{
	WSeg gstack[40];
};

__device__ inline void waveSync()
{
	__builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
	__builtin_amdgcn_wave_barrier();
}

__device__ inline unsigned long long laneMaskLt()
{
	const int l = laneId();
	return l == 0 ? 0ULL : (~0ULL >> (64 - l));
}

// exsort::introLoop for one lane's range of a tile: at most kLaneMax
// elements, or no depth budget left (straight to the heap sort).  Same
// partitions on the same ranges with the same budgets; the pending larger
// parts are kept in four packed registers instead of a dynamically indexed
// array, which the compiler places in scratch memory.  Continuing with the
// smaller part bounds the pending parts: a part pushed while partitioning a
// range of S elements leaves a current range of at most S / 2, and only
// ranges above exsort::kThreshold (16) are partitioned, so from kLaneMax <= 63
// elements at most two parts are pending.  A pending part packs its tile
// position (16 bits), its length (6 bits: a part is pushed only while the
// budget lasts, when the range holds at most kLaneMax elements) and its
// budget (10 bits).
template <class K, int kLaneMax>
__device__ inline void laneIntroLoop(const exsort::Span<K>& s, int first, int last, int depth)
{
	static_assert(kLaneMax < 4 * exsort::kThreshold, "two pending parts at most");
	static_assert(kLaneMax < 64, "part lengths are 6-bit");
	auto pack = [](int f, int l, int d) -> uint32_t {
		return static_cast<uint32_t>(f) | (static_cast<uint32_t>(l - f) << 16) | (static_cast<uint32_t>(d) << 22);
	};
	uint32_t st0 = 0, st1 = 0, st2 = 0, st3 = 0;
	int top = 0;
	int f = first, l = last, d = depth;
	for (;;)
	{
		while (l - f > exsort::kThreshold)
		{
			if (d == 0)
			{
				exsort::heapSort(s, f, l);
				break;
			}
			--d;
			const int cut = exsort::partitionPivot(s, f, l);
			st3 = st2;
			st2 = st1;
			st1 = st0;
			++top;
			if (cut - f < l - cut)
			{
				st0 = pack(cut, l, d);
				l = cut;
			}
			else
			{
				st0 = pack(f, cut, d);
				f = cut;
			}
		}
		if (top == 0) break;
		f = static_cast<int>(st0 & 0xffffu);
		l = f + static_cast<int>((st0 >> 16) & 0x3fu);
		d = static_cast<int>(st0 >> 22);
		st0 = st1;
		st1 = st2;
		st2 = st3;
		--top;
	}
}

// Partition s[f, l) around the median moved to f; all 64 lanes call it.
// posL / posR need (l - f) / 2 + 1 entries.  freeSeed 0: std::sort's median
// of (f + 1, mid, l - 1); otherwise ("free mode", where any correct sort will
// do) the median of three distinct samples at a pseudo-random offset a third
// of the range apart, which does not degenerate on the inputs that exhausted
// std::sort's depth budget.
template <class K, class P>
__device__ inline int wavePartition(const exsort::Span<K>& s, int f, int l, P* posL, P* posR,
									uint32_t freeSeed = 0)
{
	const int lane = laneId();
	if (lane == 0)
	{
		if (freeSeed == 0)
		{
			exsort::medianToFirst(s, f, f + 1, f + (l - f) / 2, l - 1);
		}
		else
		{
			const uint32_t span = static_cast<uint32_t>(l - f - 1);
			uint64_t h = (static_cast<uint64_t>(freeSeed) << 32) ^ (static_cast<uint64_t>(f) << 16) ^ static_cast<uint32_t>(l);
			h = h * 6364136223846793005ULL + 1442695040888963407ULL;
			const uint32_t o = static_cast<uint32_t>((h >> 33) % span);
			exsort::medianToFirst(s, f, f + 1 + static_cast<int>(o), f + 1 + static_cast<int>((o + span / 3) % span),
								  f + 1 + static_cast<int>((o + 2 * (span / 3)) % span));
		}
	}
	waveSync();
	const K p = s.k[f];
	const int lo = f + 1;
	int LE = 0;
	for (int base = lo; base < l; base += FH_WAVE)
	{
		const int i = base + lane;
		const bool le = i < l && !(p < s.k[i]);
		LE += __popcll(__ballot(le));
	}
	const unsigned long long lt = laneMaskLt();
	int geLeft = 0, leLeft = 0, m = 0;
	int lm = 0x7fffffff;
	for (int base = lo; base < l; base += FH_WAVE)
	{
		const int i = base + lane;
		bool ge = false, le = false;
		if (i < l)
		{
			const K v = s.k[i];
			ge = !(v < p);
			le = !(p < v);
		}
		const unsigned long long bg = __ballot(ge);
		const unsigned long long bl = __ballot(le);
		const int myGe = geLeft + __popcll(bg & lt);
		const int myLe = leLeft + __popcll(bl & lt);
		const int leRight = LE - myLe - (le ? 1 : 0);
		const bool swapGe = ge && myGe < leRight;
		if (swapGe) posL[myGe] = static_cast<P>(i);
		if (le && leRight < myGe) posR[leRight] = static_cast<P>(i);
		const unsigned long long bns = __ballot(ge && !swapGe);
		if (lm == 0x7fffffff && bns) lm = base + __ffsll(static_cast<long long>(bns)) - 1;
		m += __popcll(__ballot(swapGe));
		geLeft += __popcll(bg);
		leLeft += __popcll(bl);
	}
	waveSync();
	for (int q = lane; q < m; q += FH_WAVE) s.swap(posL[q], posR[q]);
	waveSync();
	int cut = lm;
	if (m > 0)
	{
		const int r = static_cast<int>(posR[m - 1]);
		if (r < cut) cut = r;
	}
	waveSync();
	return cut;
}

// Where a tile sort lists the sub-ranges std::sort would heap-sort (depth
// budget spent, more than kLaneMax elements) instead of heap-sorting them one
// lane each: their content goes to the shadow arrays, a verify entry (absolute
// positions) is appended, and the wave sorts them on in free mode.  A later
// pass heap-sorts, from the shadow copy, only the entries whose free result
// holds a tie that matters (segsort.hpp).
// A range of a segmented sort (segsort.hpp).
struct SegRange
{
	int32_t f;    // absolute positions in the key / value arrays
	int32_t l;
	int32_t d;    // remaining depth budget (verify list: 1 = needs the exact heap sort)
	int32_t tag;  // 0: exact mode; t > 0: free mode under verify entry t - 1
};

struct WaveVerify
{
	uint32_t* shadowKey;
	uint32_t* shadowVal;
	SegRange* list;
	unsigned long long* count;
	unsigned long long* error;  // bit 8 set if the list is full (cannot happen when sized as documented)
	int64_t cap;
};

constexpr int kWaveFree = 1 << 30;  // WSeg.d flag: free mode
constexpr int kWaveFreeDepth = 64;   // depth budget of a free-mode range

// Sorts the range g[f, l) (l - f <= kTile) with remaining depth budget d,
// through the wave's LDS tile; ends with the stable insertion pass.  The
// sorted range goes back to g, or, with outV set, its values go to outV[f, l)
// and its keys back to g only when keysBack is set (or when a sub-range was
// listed for verification, which reads the sorted keys).
template <class K, int kTile, int kLaneMax>
__device__ void waveTileSort(const exsort::Span<K>& g, int f, int l, int d, WaveSortLds<K, kTile>& t,
							 uint32_t* outV = nullptr, bool keysBack = false, const WaveVerify* vfy = nullptr)
{
	const int lane = laneId();
	const int n = l - f;
	for (int i = lane; i < n; i += FH_WAVE)
	{
		t.k[i] = g.k[f + i];
		t.v[i] = g.v[f + i];
	}
	waveSync();
	const exsort::Span<K> s{t.k, t.v};
	int top = 0;
	int nl = 0;
	if (lane == 0) t.stack[0] = WSeg{0, n, d};
	top = 1;
	waveSync();
	auto flush = [&]() {
		waveSync();
		if (lane < nl)
		{
			const WSeg w = t.lanes[lane];
			// a lane range has at most kLaneMax elements or no depth budget
			// left (straight to heapsort)
			laneIntroLoop<K, kLaneMax>(s, w.f, w.l, w.d & (kWaveFree - 1));
			exsort::insertionSort(s, w.f, w.l);
		}
		nl = 0;
		waveSync();
	};
	bool listed = false;
	while (top > 0)
	{
		--top;
		const WSeg w = t.stack[top];
		waveSync();
		const bool freeW = (w.d & kWaveFree) != 0;
		const int dw = w.d & (kWaveFree - 1);
		if (vfy && !freeW && dw == 0 && w.l - w.f > kLaneMax)
		{
			for (int i = w.f + lane; i < w.l; i += FH_WAVE)
			{
				vfy->shadowKey[f + i] = static_cast<uint32_t>(t.k[i]);
				vfy->shadowVal[f + i] = t.v[i];
			}
			if (lane == 0)
			{
				const unsigned long long vi = atomicAdd(vfy->count, 1ULL);
				if (static_cast<int64_t>(vi) < vfy->cap)
					vfy->list[vi] = SegRange{f + w.f, f + w.l, 0, 0};
				else
					atomicOr(vfy->error, 256ULL);
				t.stack[top] = WSeg{w.f, w.l, kWaveFreeDepth | kWaveFree};
			}
			top += 1;
			listed = true;
			waveSync();
			continue;
		}
		if (w.l - w.f <= kLaneMax || dw == 0)
		{
			if (lane == 0) t.lanes[nl] = w;
			++nl;
			if (nl == FH_WAVE) flush();
			continue;
		}
		const int cut = wavePartition(s, w.f, w.l, t.posL, t.posR,
									  freeW ? static_cast<uint32_t>(f + w.f) * 2654435761u + 1u : 0u);
		if (lane == 0)
		{
			// larger part below, smaller on top: the stack stays within log2(n) + 1
			// (fields chosen one by one: a choice between two local structs
			// compiles to scratch stores and reloads)
			const int dd = (dw - 1) | (freeW ? kWaveFree : 0);
			const bool leftSmaller = cut - w.f < w.l - cut;
			t.stack[top] = WSeg{leftSmaller ? cut : w.f, leftSmaller ? w.l : cut, dd};
			t.stack[top + 1] = WSeg{leftSmaller ? w.f : cut, leftSmaller ? cut : w.l, dd};
		}
		top += 2;
		waveSync();
	}
	if (nl) flush();
	if (outV)
	{
		const bool kb = keysBack || listed;
		for (int i = lane; i < n; i += FH_WAVE)
		{
			outV[f + i] = t.v[i];
			if (kb) g.k[f + i] = t.k[i];
		}
	}
	else
	{
		for (int i = lane; i < n; i += FH_WAVE)
		{
			g.k[f + i] = t.k[i];
			g.v[f + i] = t.v[i];
		}
	}
	waveSync();
}

// Exact std::sort of g[0, n) by one wave.  posL/posR: global scratch of
// n / 2 + 1 ints each (only touched when n > kTile).
template <class K, int kTile, int kLaneMax>
__device__ void waveStdSort(const exsort::Span<K>& g, int n, int* posL, int* posR, WaveStdSortLds<K, kTile>& t)
{
	if (n < 2) return;
	if (n <= kTile)
	{
		waveTileSort<K, kTile, kLaneMax>(g, 0, n, exsort::depthLimit(static_cast<uint32_t>(n)), t);
		return;
	}
	// global ranges: stack in LDS (a register array indexed at run time would
	// live in scratch memory)
	int top = 0;
	if (laneId() == 0) t.gstack[0] = WSeg{0, n, exsort::depthLimit(static_cast<uint32_t>(n))};
	top = 1;
	waveSync();
	while (top > 0)
	{
		--top;
		const WSeg w = t.gstack[top];
		waveSync();
		if (w.l - w.f <= kTile)
		{
			waveTileSort<K, kTile, kLaneMax>(g, w.f, w.l, w.d, t);
			continue;
		}
		if (w.d == 0)
		{
			if (laneId() == 0) exsort::heapSort(g, w.f, w.l);
			waveSync();
			continue;
		}
		const int cut = wavePartition(g, w.f, w.l, posL, posR);
		if (laneId() == 0)
		{
			const bool leftSmaller = cut - w.f < w.l - cut;
			t.gstack[top] = WSeg{leftSmaller ? cut : w.f, leftSmaller ? w.l : cut, w.d - 1};
			t.gstack[top + 1] = WSeg{leftSmaller ? w.f : cut, leftSmaller ? cut : w.l, w.d - 1};
		}
		top += 2;
		waveSync();
	}
}

} // namespace fh
