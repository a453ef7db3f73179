// segsort.hip - level-synchronous exact introsort over many arrays (see
// segsort.hpp for the scheme and exsort.hpp for the rank form of libstdc++'s
// unguarded partition that the round kernels evaluate).
//
// One round over the current list of big ranges:
//   median  one thread per range: median of (f+1, mid, l-1) moved to f,
//           chunk count of [f+1, l);
//   scan    chunk offsets per range;
//   map     chunk -> range;
//   count   one wave per chunk: ge / le counts against the range's pivot;
//   scan    the packed counts over all chunks of the round;
//   rank    one wave per chunk: left ranks of ge elements and right ranks of
//           le elements; elements that swap scatter their positions to
//           posL[f + rank] / posR[f + rank]; per range the swap count and the
//           first ge element that stays;
//   swap    one lane per swap pair;
//   split   one thread per range: the cut, both children pushed to the next
//           list, the leaf list or the heap list.
#include <hipcub/hipcub.hpp>

#include <cstdio>
#include <cstdlib>

#include "segsort.hpp"

namespace fh
{
namespace
{

constexpr int kWavesPerBlock = 4;

inline unsigned blocksFor(int64_t n, int per) { return static_cast<unsigned>((n + per - 1) / per); }

__device__ inline int waveIndex() { return blockIdx.x * kWavesPerBlock + threadIdx.x / FH_WAVE; }

__global__ void segMedianKernel(SegSortBufs b, const SegRange* rs, int32_t nR)
{
	const int r = blockIdx.x * blockDim.x + threadIdx.x;
	if (r > nR) return;
	if (r == nR)
	{
		b.chunkOff[r] = 0;  // scan input tail
		return;
	}
	const SegRange w = rs[r];
	const exsort::Span<uint32_t> s{b.key, b.val};
	if (w.tag == 0)
	{
		exsort::medianToFirst(s, w.f, w.f + 1, w.f + (w.l - w.f) / 2, w.l - 1);
	}
	else
	{
		// free mode (any correct sort will do): median of three samples at a
		// pseudo-random offset and a third of the range apart, which does not
		// degenerate on the inputs that exhausted the exact budget.  The three
		// positions are distinct, so one element >= and one <= the pivot stay
		// in [f + 1, l) as the partition needs.
		const uint32_t span = static_cast<uint32_t>(w.l - w.f - 1);
		uint64_t h = (static_cast<uint64_t>(w.f) << 32) ^ static_cast<uint32_t>(w.l) ^ (static_cast<uint64_t>(w.d) << 56);
		h = h * 6364136223846793005ULL + 1442695040888963407ULL;
		const uint32_t o = static_cast<uint32_t>((h >> 33) % span);
		const int32_t x0 = w.f + 1 + static_cast<int32_t>(o);
		const int32_t x1 = w.f + 1 + static_cast<int32_t>((o + span / 3) % span);
		const int32_t x2 = w.f + 1 + static_cast<int32_t>((o + 2 * (span / 3)) % span);
		exsort::medianToFirst(s, w.f, x0, x1, x2);
	}
	b.chunkOff[r] = (static_cast<int64_t>(w.l - w.f) - 1 + kSegChunk - 1) / kSegChunk;
	b.rangeM[r] = 0;
	b.rangeLm[r] = 0x7fffffff;
}

__global__ void segMapKernel(SegSortBufs b, int32_t nR)
{
	const int r = waveIndex();
	if (r >= nR) return;
	const int64_t c0 = b.chunkOff[r], c1 = b.chunkOff[r + 1];
	for (int64_t c = c0 + laneId(); c < c1; c += FH_WAVE) b.chunkRange[c] = r;
}

struct ChunkView
{
	int32_t r;
	SegRange w;
	int32_t start;
	int32_t end;
	int64_t j;
};

__device__ inline ChunkView chunkView(const SegSortBufs& b, const SegRange* rs, int64_t c)
{
	ChunkView v;
	v.r = b.chunkRange[c];
	v.w = rs[v.r];
	v.j = c - static_cast<int64_t>(b.chunkOff[v.r]);
	v.start = v.w.f + 1 + static_cast<int32_t>(v.j) * kSegChunk;
	v.end = min(v.w.l, v.start + kSegChunk);
	return v;
}

__global__ void __launch_bounds__(FH_WAVE * kWavesPerBlock) segCountKernel(SegSortBufs b, const SegRange* rs,
																			int64_t nChunks)
{
	const int64_t c = waveIndex();
	if (c > nChunks) return;
	if (c == nChunks)
	{
		if (laneId() == 0) b.chunkCnt[c] = 0;
		return;
	}
	const ChunkView v = chunkView(b, rs, c);
	const uint32_t p = b.key[v.w.f];
	uint32_t x[kSegRows];
#pragma unroll
	for (int u = 0; u < kSegRows; ++u)
	{
		const int i = v.start + u * FH_WAVE + laneId();
		x[u] = i < v.end ? b.key[i] : 0u;
	}
	int ge = 0, le = 0;
#pragma unroll
	for (int u = 0; u < kSegRows; ++u)
	{
		const int i = v.start + u * FH_WAVE + laneId();
		if (i < v.end)
		{
			ge += !(x[u] < p);
			le += !(p < x[u]);
		}
	}
	ge = waveSum(ge);
	le = waveSum(le);
	if (laneId() == 0) b.chunkCnt[c] = (static_cast<uint64_t>(ge) << 32) | static_cast<uint32_t>(le);
}

__global__ void __launch_bounds__(FH_WAVE * kWavesPerBlock) segRankKernel(SegSortBufs b, const SegRange* rs,
																		   int64_t nChunks)
{
	const int64_t c = waveIndex();
	if (c >= nChunks) return;
	const ChunkView v = chunkView(b, rs, c);
	const uint32_t p = b.key[v.w.f];
	const int64_t c0 = b.chunkOff[v.r], c1 = b.chunkOff[v.r + 1];
	const uint64_t base = b.chunkPre[c0];
	const uint64_t pre = b.chunkPre[c] - base;
	const uint64_t tot = b.chunkPre[c1] - base;
	int32_t geLeft = static_cast<int32_t>(pre >> 32);
	int32_t leLeft = static_cast<int32_t>(pre & 0xffffffffULL);
	const int32_t LE = static_cast<int32_t>(tot & 0xffffffffULL);
	uint32_t x[kSegRows];
#pragma unroll
	for (int u = 0; u < kSegRows; ++u)
	{
		const int i = v.start + u * FH_WAVE + laneId();
		x[u] = i < v.end ? b.key[i] : 0u;
	}
	const unsigned long long lt = laneMaskLt();
	int swaps = 0;
	int firstStay = 0x7fffffff;
	int32_t* posL = b.posL + v.w.f;
	int32_t* posR = b.posR + v.w.f;
#pragma unroll
	for (int u = 0; u < kSegRows; ++u)
	{
		const int i = v.start + u * FH_WAVE + laneId();
		const bool in = i < v.end;
		const bool ge = in && !(x[u] < p);
		const bool le = in && !(p < x[u]);
		const unsigned long long bg = __ballot(ge);
		const unsigned long long bl = __ballot(le);
		const int32_t myGe = geLeft + __popcll(bg & lt);
		const int32_t myLe = leLeft + __popcll(bl & lt);
		const int32_t leRight = LE - myLe - (le ? 1 : 0);
		const bool swapGe = ge && myGe < leRight;
		if (swapGe) posL[myGe] = i;
		if (le && leRight < myGe) posR[leRight] = i;
		const unsigned long long stay = __ballot(ge && !swapGe);
		if (firstStay == 0x7fffffff && stay) firstStay = v.start + u * FH_WAVE + __ffsll(static_cast<long long>(stay)) - 1;
		swaps += __popcll(__ballot(swapGe));
		geLeft += __popcll(bg);
		leLeft += __popcll(bl);
	}
	if (laneId() == 0)
	{
		if (swaps) atomicAdd(&b.rangeM[v.r], swaps);
		if (firstStay != 0x7fffffff) atomicMin(&b.rangeLm[v.r], firstStay);
	}
}

__global__ void __launch_bounds__(FH_WAVE * kWavesPerBlock) segSwapKernel(SegSortBufs b, const SegRange* rs,
																		   int64_t nChunks)
{
	const int64_t c = waveIndex();
	if (c >= nChunks) return;
	const ChunkView v = chunkView(b, rs, c);
	const int32_t m = b.rangeM[v.r];
	const int32_t q0 = static_cast<int32_t>(v.j) * kSegChunk;
	if (q0 >= m) return;
	const int32_t* posL = b.posL + v.w.f;
	const int32_t* posR = b.posR + v.w.f;
#pragma unroll
	for (int u = 0; u < kSegRows; ++u)
	{
		const int32_t q = q0 + u * FH_WAVE + laneId();
		if (q < m)
		{
			const int32_t x = posL[q], y = posR[q];
			if (x <= v.w.f || x >= v.w.l || y <= v.w.f || y >= v.w.l || x >= y)
			{
				atomicOr(reinterpret_cast<unsigned long long*>(&b.counters[kSegError]), 1ULL);
				continue;
			}
			const uint32_t kx = b.key[x], ky = b.key[y];
			const uint32_t vx = b.val[x], vy = b.val[y];
			b.key[x] = ky;
			b.key[y] = kx;
			b.val[x] = vy;
			b.val[y] = vx;
		}
	}
}

__global__ void segSplitKernel(SegSortBufs b, const SegRange* rs, int32_t nR, SegRange* next)
{
	const int r = blockIdx.x * blockDim.x + threadIdx.x;
	bool act = r < nR;
	SegRange w{0, 0, 0, 0};
	int32_t cut = 0;
	if (act)
	{
		w = rs[r];
		const int32_t m = b.rangeM[r];
		cut = b.rangeLm[r];
		if (m > 0) cut = min(cut, b.posR[w.f + m - 1]);
		if (cut <= w.f || cut >= w.l)
		{
			// cannot happen for a consistent partition; keeps a broken one in bounds
			atomicOr(reinterpret_cast<unsigned long long*>(&b.counters[kSegError]), 2ULL);
			act = false;
		}
	}
	segPushWave(b, next, act, w.f, cut, w.d - 1, w.tag);
	segPushWave(b, next, act, cut, w.l, w.d - 1, w.tag);
}

__global__ void __launch_bounds__(FH_WAVE) segLeafKernel(SegSortBufs b, int64_t nLeaf)
{
	__shared__ WaveSortLds<uint32_t, kSegTile> lds;
	const int64_t i = blockIdx.x;
	if (i >= nLeaf) return;
	const SegRange w = b.leaf[i];
	// exact-mode sub-ranges out of budget: shadow copy + verify entry + free
	// mode, as for big ranges (verify entries need no shadow kernel)
	const WaveVerify vfy{b.shadowKey, b.shadowVal, b.verify, reinterpret_cast<unsigned long long*>(&b.counters[kSegNVerify]),
						 reinterpret_cast<unsigned long long*>(&b.counters[kSegError]), b.verifyCap};
	// free mode (a leaf of a range listed for verification): free pivots, no
	// nested verify entries, and the sorted keys stay in place for the
	// verification of the enclosing entry
	const bool freeLeaf = w.tag != 0;
	waveTileSort<uint32_t, kSegTile, kSegLaneMax>(exsort::Span<uint32_t>{b.key, b.val}, w.f, w.l,
												  w.d | (freeLeaf ? kWaveFree : 0), lds, b.outVal, freeLeaf, &vfy);
}

// exsort::siftStepPacked without branches: both children are loaded
// together from adjacent words (the first index clamped to the range; hx has
// one spare word after n; a clamped or spare load is not used), keys compared
// on the words' bits from sh up.  Same result: the right child unless it is
// below the left one, moved up unless it is below the value.
template <typename W>
__device__ inline bool heapStep(W* hx, int& h, int m, int n, W v, int sh)
{
	const int c1 = 2 * h + 1;
	const int cc = min(c1, n - 1);
	const W x1 = hx[cc], x2 = hx[cc + 1];
	const bool right = c1 + 1 < m && !((x2 >> sh) < (x1 >> sh));
	const W cv = right ? x2 : x1;
	const bool up = c1 < m && !((cv >> sh) < (v >> sh));
	hx[h] = up ? cv : v;
	h = up ? (right ? c1 + 1 : c1) : h;
	return up;
}

__device__ inline bool heapAncestorOrSelf(int h, int x)
{
	const unsigned H = static_cast<unsigned>(h) + 1u, X = static_cast<unsigned>(x) + 1u;
	if (H > X) return false;
	return (X >> (__clz(H) - __clz(X))) == H;
}

// Heap sort, exactly libstdc++'s make_heap + sort_heap, of the n words hx
// (one wave; keys in the bits from sh up, compared alone); the schedule of
// exsort::heapSortPipelinedModel: make_heap one heap level at a time with a
// lane per parent, then sort_heap with up to kSlots pops in flight, one lane
// each, started two steps apart and advancing one level per step.
template <typename W>
__device__ void waveHeapSort(W* hx, int n, int sh)
{
	constexpr int kSlots = 16;
	const int lane = laneId();
	if (n < 2) return;
	// make_heap
	const int lastParent = (n - 2) / 2;
	const int top = 31 - __clz(static_cast<unsigned>(lastParent) + 1u);
	for (int lv = top; lv >= 0; --lv)
	{
		const int lo = (1 << lv) - 1;
		const int hi = min(lastParent, (2 << lv) - 2);
		for (int p = hi - lane; p >= lo; p -= FH_WAVE)
		{
			int hole = p;
			const W pv = hx[p];
			while (heapStep(hx, hole, n, n, pv, sh))
			{
			}
		}
		waveSync();
	}
	// sort_heap
	bool act = false;
	int h = 0, m = 0;
	W v = 0;
	int L = n, pops = 0;
	// LDS operations of one wave complete in program order, so a cycle's
	// writes reach the next cycle's reads without waiting; the wave barrier
	// only keeps the compiler from moving them across cycles.  Pops start at
	// most every other cycle, so cycles run in pairs and only the first of a
	// pair may start one: the start checks (slot free, no sift above the last
	// element) make the result independent of when a pop starts, and the
	// second cycle is a plain step.
	for (;;)
	{
		unsigned long long running = __ballot(act);
		if (L > 1)
		{
			const int slot = pops % kSlots;
			const bool blocked = __ballot(act && heapAncestorOrSelf(h, L - 1)) != 0;
			if (((running >> slot) & 1ULL) == 0 && !blocked)
			{
				if (lane == slot)
				{
					v = hx[L - 1];
					hx[L - 1] = hx[0];
					h = 0;
					m = L - 1;
					act = true;
				}
				running |= 1ULL << slot;
				--L;
				++pops;
			}
		}
		if (running == 0 && L <= 1) break;
		if (act) act = heapStep(hx, h, m, n, v, sh);
		__builtin_amdgcn_wave_barrier();
		if (act) act = heapStep(hx, h, m, n, v, sh);
		__builtin_amdgcn_wave_barrier();
	}
	waveSync();
}

// Where a range's heap sort runs.  The LDS kernel stages a range as 32-bit
// words, (key - min key) << sh | position (sh the bits of n - 1), when the
// key span fits the bits above sh, else as 64-bit (key << 32 | value) words,
// in the first of three LDS sizes that holds it (16, 64 and 128 KiB, so
// small ranges are not held to a few waves per CU by the largest, and the
// 64 KiB ones fit beside two lanes' DPs); anything larger sorts in global
// memory, one thread per range (class kHeapClasses).  kHeapNoBig: the device
// grants no 128 KiB, so that class falls to global memory.
constexpr int kHeapClasses = 3;
__host__ __device__ constexpr int heapLdsBytes(int cls) { return cls == 0 ? 16 * 1024 : (cls == 1 ? 64 * 1024 : 128 * 1024); }
constexpr int kHeapNoBig = 1;

__device__ inline int heapShift(int n) { return n <= 2 ? 1 : 32 - __clz(static_cast<unsigned>(n - 1)); }
__device__ inline bool heapFits32(uint32_t span, int sh) { return (static_cast<uint64_t>(span) >> (32 - sh)) == 0; }
__device__ inline int heapClass(int n, bool fits32, int flags)
{
	const int top = (flags & kHeapNoBig) ? kHeapClasses - 1 : kHeapClasses;
	for (int c = 0; c < top; ++c)
	{
		// one spare word after the range (heapStep's paired load)
		if ((fits32 && (n + 1) * 4 <= heapLdsBytes(c)) || (n + 1) * 8 <= heapLdsBytes(c)) return c;
	}
	return kHeapClasses;
}

// Heap sorts of the ranges of a list of class cls (an LDS class; only
// entries with d == 1 when onlyFlagged), one wave per range (grid: one block
// per entry, dynamic LDS heapLdsBytes(cls)).  Values go to dstV, keys to dstK
// when it is set.
__global__ void __launch_bounds__(FH_WAVE) segHeapLdsKernel(const SegRange* list, int64_t nList, int onlyFlagged,
															const uint32_t* srcK, const uint32_t* srcV, uint32_t* dstK,
															uint32_t* dstV, int cls, int flags)
{
	extern __shared__ uint64_t hx[];
	const int lane = laneId();
	const int64_t i = blockIdx.x;
	if (i >= nList) return;
	const SegRange w = list[i];
	const int n = w.l - w.f;
	if (n < 1 || (onlyFlagged && w.d != 1)) return;
	const uint32_t* k = srcK + w.f;
	uint32_t mn = 0xffffffffu, mx = 0;
	for (int t = lane; t < n; t += FH_WAVE)
	{
		mn = min(mn, k[t]);
		mx = max(mx, k[t]);
	}
	for (int o = 32; o > 0; o >>= 1)
	{
		mn = min(mn, static_cast<uint32_t>(__shfl_xor(static_cast<int>(mn), o, FH_WAVE)));
		mx = max(mx, static_cast<uint32_t>(__shfl_xor(static_cast<int>(mx), o, FH_WAVE)));
	}
	const int sh = heapShift(n);
	const bool fits = heapFits32(mx - mn, sh);
	if (heapClass(n, fits, flags) != cls) return;
	if (fits && (n + 1) * 4 <= heapLdsBytes(cls))
	{
		uint32_t* hw = reinterpret_cast<uint32_t*>(hx);
		for (int t = lane; t < n; t += FH_WAVE) hw[t] = ((k[t] - mn) << sh) | static_cast<uint32_t>(t);
		waveSync();
		waveHeapSort(hw, n, sh);
		// the values by position, through LDS: every source value is read
		// before any is written (dstV may be srcV)
		const uint32_t mask = (1u << sh) - 1u;
		for (int t = lane; t < n; t += FH_WAVE)
		{
			const uint32_t x = hw[t];
			if (dstK) dstK[w.f + t] = (x >> sh) + mn;
			hw[t] = srcV[w.f + static_cast<int>(x & mask)];
		}
		waveSync();
		for (int t = lane; t < n; t += FH_WAVE) dstV[w.f + t] = hw[t];
	}
	else
	{
		for (int t = lane; t < n; t += FH_WAVE) hx[t] = (static_cast<uint64_t>(k[t]) << 32) | srcV[w.f + t];
		waveSync();
		waveHeapSort(hx, n, 32);
		for (int t = lane; t < n; t += FH_WAVE)
		{
			const uint64_t x = hx[t];
			dstV[w.f + t] = static_cast<uint32_t>(x);
			if (dstK) dstK[w.f + t] = static_cast<uint32_t>(x >> 32);
		}
	}
}

// The ranges beyond the LDS classes in global memory, in place on (k, v),
// one thread per range; the values are then copied to outV when it is a
// separate array.
__global__ void segHeapGlobalKernel(const SegRange* list, int64_t nList, int onlyFlagged, uint32_t* k, uint32_t* v,
									uint32_t* outV, int flags)
{
	const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
	if (i >= nList) return;
	const SegRange w = list[i];
	const int n = w.l - w.f;
	if (n < 1 || (onlyFlagged && w.d != 1)) return;
	// 64-bit words of the largest LDS class hold it whatever its keys
	if ((n + 1) * 8 <= heapLdsBytes(kHeapClasses - ((flags & kHeapNoBig) ? 2 : 1))) return;
	uint32_t mn = 0xffffffffu, mx = 0;
	for (int32_t x = w.f; x < w.l; ++x)
	{
		mn = min(mn, k[x]);
		mx = max(mx, k[x]);
	}
	if (heapClass(n, heapFits32(mx - mn, heapShift(n)), flags) != kHeapClasses) return;
	exsort::heapSort(exsort::Span<uint32_t>{k, v}, w.f, w.l);
	if (outV != v)
		for (int32_t x = w.f; x < w.l; ++x) outV[x] = v[x];
}

// Shadow copies of the verify entries [v0, v1): their content at the moment
// the depth budget ran out (one wave per entry).
__global__ void segShadowKernel(SegSortBufs b, int64_t v0, int64_t v1)
{
	const int64_t i = v0 + waveIndex();
	if (i >= v1) return;
	const SegRange w = b.verify[i];
	for (int32_t x = w.f + laneId(); x < w.l; x += FH_WAVE)
	{
		b.shadowKey[x] = b.key[x];
		b.shadowVal[x] = b.val[x];
	}
}

// A verify entry whose free-mode result has two equal neighbouring keys needs
// the exact heap sort (d = 1); with all keys distinct every sort agrees.  With
// relevantBit set, only a run of equal keys holding two marked values counts
// (the order of the other elements of a run is immaterial to the caller).
__global__ void segVerifyKernel(SegSortBufs b, int64_t nVerify, int heapFlags)
{
	const int64_t i = waveIndex();
	if (i >= nVerify) return;
	const SegRange w = b.verify[i];
	bool need = false;
	for (int32_t x = w.f + laneId(); x < w.l; x += FH_WAVE)
	{
		// x starts a run of at least two equal keys
		if ((x == w.f || b.key[x - 1] != b.key[x]) && x + 1 < w.l && b.key[x + 1] == b.key[x])
		{
			if (!b.relevantBit)
			{
				need = true;
			}
			else
			{
				int marked = 0;
				for (int32_t y = x; y < w.l && b.key[y] == b.key[x] && marked < 2; ++y)
					marked += (b.outVal[y] & b.relevantBit) ? 1 : 0;
				need |= marked >= 2;
			}
		}
	}
	const bool any = __ballot(need) != 0;
	if (laneId() == 0)
	{
		b.verify[i].d = any ? 1 : 0;
		if (any)
		{
			atomicAdd(reinterpret_cast<unsigned long long*>(&b.counters[kSegExactElems]),
					  static_cast<unsigned long long>(w.l - w.f));
			// the heap sort's class, as segHeapLdsKernel / segHeapGlobalKernel
			// decide it from the shadow copy (the same keys, here sorted: the
			// span is last minus first)
			const int n = w.l - w.f;
			const int cls = heapClass(n, heapFits32(b.key[w.l - 1] - b.key[w.f], heapShift(n)), heapFlags);
			atomicAdd(reinterpret_cast<unsigned long long*>(&b.counters[kSegHeapCls + cls]), 1ULL);
		}
	}
}

__global__ void segPushOffsetsKernel(SegSortBufs b, const int64_t* off, int32_t nArrays)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	const bool act = i < nArrays;
	int32_t f = 0, l = 0;
	if (act)
	{
		f = static_cast<int32_t>(off[i]);
		l = static_cast<int32_t>(off[i + 1]);
	}
	const bool tied = act && l - f >= 2 && (!b.dup || b.dup[l] - b.dup[f] > 0);
	const int t = __popcll(__ballot(tied));
	if (laneId() == 0 && t) atomicAdd(reinterpret_cast<unsigned long long*>(&b.counters[kSegTieArrays]),
									  static_cast<unsigned long long>(t));
	segPushWave(b, b.big[0], act, f, l, act ? exsort::depthLimit(static_cast<uint32_t>(l - f)) : 0, 0);
}

hipError_t download(const SegSortBufs& b, hipStream_t s)
{
	hipError_t e = hipMemcpyAsync(b.hostCounters, b.counters, kSegCounters * sizeof(int64_t), hipMemcpyDeviceToHost, s);
	if (e != hipSuccess) return e;
	return hipStreamSynchronize(s);
}

// Heap sorts of a list (heapClass), one launch per class with a block per
// entry, largest class first (it holds the longest chains), one after the
// other; the ranges beyond them sort in place on gk / gv in global memory.  With join they run on the calling stream; otherwise on the
// side stream, forked from it, where doneEv is recorded after the last of
// them.  (One side stream per lane: with the lanes' main streams that is as
// many streams as the runtime's default hardware queues, so no stream shares
// a queue, whose in-order packets would make one lane's heap sorts wait for
// another's.)
int heapFlags()
{
	static const bool bigLds =
		hipFuncSetAttribute(reinterpret_cast<const void*>(segHeapLdsKernel), hipFuncAttributeMaxDynamicSharedMemorySize,
							heapLdsBytes(kHeapClasses - 1)) == hipSuccess;
	return bigLds ? 0 : kHeapNoBig;
}

// clsCount (may be null: launch every class): ranges per class (3 LDS
// classes, then global memory); a class without ranges is not launched, so
// its blocks do not wait for LDS the DP holds.
hipError_t heapTiers(const SegSortBufs& b, const SegRange* list, int64_t n, int onlyFlagged, const uint32_t* srcK,
					 const uint32_t* srcV, uint32_t* dstK, uint32_t* dstV, uint32_t* gk, uint32_t* gv, hipStream_t s,
					 int* firstSlot, bool join = true, const int64_t* clsCount = nullptr)
{
	*firstSlot = -1;
	const int flags = heapFlags();
	const bool bigLds = flags == 0;
	if (n <= 0) return hipSuccess;
	if (n > 0x7fffffffLL) return hipErrorInvalidValue;
	hipError_t e = hipSuccess;
	hipStream_t hs = s;
	if (!join)
	{
		if ((e = hipEventRecord(b.forkEv, s)) != hipSuccess) return e;
		if ((e = hipStreamWaitEvent(b.side, b.forkEv, 0)) != hipSuccess) return e;
		hs = b.side;
	}
	if (!clsCount || clsCount[kHeapClasses] > 0)
	{
		hipLaunchKernelGGL(segHeapGlobalKernel, dim3(blocksFor(n, 64)), dim3(64), 0, hs, list, n, onlyFlagged, gk, gv,
						   dstV, flags);
		if ((e = hipGetLastError()) != hipSuccess) return e;
	}
	for (int cls = kHeapClasses - (bigLds ? 1 : 2); cls >= 0; --cls)
	{
		if (clsCount && clsCount[cls] == 0) continue;
		const int slot = b.timer ? b.timer->begin(FLYEHIP_K_HEAP, hs) : -1;
		if (*firstSlot < 0) *firstSlot = slot;
		hipLaunchKernelGGL(segHeapLdsKernel, dim3(static_cast<unsigned>(n)), dim3(FH_WAVE),
						   static_cast<size_t>(heapLdsBytes(cls)), hs, list, n, onlyFlagged, srcK, srcV, dstK, dstV,
						   cls, flags);
		if (b.timer) b.timer->end(slot, hs);
		if ((e = hipGetLastError()) != hipSuccess) return e;
	}
	return join ? hipSuccess : hipEventRecord(b.doneEv, hs);
}

// Leaf bytes: key and value read, value written (12 per element).  The leaves
// append verify entries with their shadow copies already taken: *nVerified
// moves past them (the round loop copies only the entries its splits add).
hipError_t runLeaves(SegSortBufs& b, hipStream_t s, int64_t* nVerified)
{
	const int64_t nLeaf = b.hostCounters[kSegNLeaf];
	if (nLeaf > 0)
	{
		const int t = b.timer ? b.timer->begin(FLYEHIP_K_LEAF, s, 12.0 * static_cast<double>(b.hostCounters[kSegLeafElems])) : -1;
		hipLaunchKernelGGL(segLeafKernel, dim3(static_cast<unsigned>(nLeaf)), dim3(FH_WAVE), 0, s, b, nLeaf);
		if (b.timer) b.timer->end(t, s);
		hipError_t e = hipGetLastError();
		if (e != hipSuccess) return e;
	}
	hipError_t e = hipMemsetAsync(b.counters + kSegNLeaf, 0, 2 * sizeof(int64_t), s);
	if (e != hipSuccess) return e;
	if (nLeaf > 0)
	{
		if ((e = download(b, s)) != hipSuccess) return e;
		*nVerified = b.hostCounters[kSegNVerify];
	}
	b.hostCounters[kSegNLeaf] = 0;
	b.hostCounters[kSegLeafElems] = 0;
	return hipSuccess;
}

} // namespace

hipError_t segSortCaps(int64_t n, int64_t nArrays, SegSortCaps* caps)
{
	caps->bigCap = n / (kSegTile + 1) + 2;
	caps->leafCap = nArrays + 2 * caps->bigCap + 64;
	caps->heapCap = caps->bigCap;
	// verify entries are disjoint ranges of more than kSegTile (rounds) or
	// kSegLaneMax (leaves) elements
	caps->verifyCap = n / (kSegLaneMax + 1) + caps->bigCap + 64;
	caps->chunkCap = n / kSegChunk + caps->bigCap + 2;
	size_t a = 0, c = 0;
	hipError_t e = hipcub::DeviceScan::ExclusiveSum(nullptr, a, static_cast<const int64_t*>(nullptr),
													static_cast<int64_t*>(nullptr), static_cast<int>(caps->bigCap + 1));
	if (e != hipSuccess) return e;
	e = hipcub::DeviceScan::ExclusiveSum(nullptr, c, static_cast<const uint64_t*>(nullptr),
										 static_cast<uint64_t*>(nullptr), static_cast<int>(caps->chunkCap + 1));
	caps->scanTempBytes = a > c ? a : c;
	return e;
}

hipError_t segSortReset(const SegSortBufs& b, hipStream_t s)
{
	return hipMemsetAsync(b.counters, 0, kSegCounters * sizeof(int64_t), s);
}

hipError_t segPushOffsets(const SegSortBufs& b, const int64_t* off, int32_t nArrays, hipStream_t s)
{
	if (nArrays <= 0) return hipSuccess;
	hipLaunchKernelGGL(segPushOffsetsKernel, dim3(blocksFor(nArrays, 256)), dim3(256), 0, s, b, off, nArrays);
	return hipGetLastError();
}

hipError_t segSortRun(SegSortBufs& b, hipStream_t s, int64_t* tieArrays, bool* exactPending)
{
	if (exactPending) *exactPending = false;
	if (b.dup && b.outVal == b.val) return hipErrorInvalidValue;
	hipError_t e = download(b, s);
	if (e != hipSuccess) return e;
	if (tieArrays) *tieArrays = b.hostCounters[kSegTieArrays];
	int cur = 0;
	int rounds = 0;
	int64_t nVerified = 0;
	const int tpb = FH_WAVE * kWavesPerBlock;
	while (b.hostCounters[kSegNBig] > 0)
	{
		++rounds;
		if (std::getenv("FLYEHIP_SEGSORT_TRACE"))
			std::fprintf(stderr, "[segsort] round %d: big=%lld elems=%lld chunks=%lld leaves=%lld heap=%lld\n", rounds,
						 static_cast<long long>(b.hostCounters[kSegNBig]),
						 static_cast<long long>(b.hostCounters[kSegBigElems]),
						 static_cast<long long>(b.hostCounters[kSegBigChunks]),
						 static_cast<long long>(b.hostCounters[kSegNLeaf]),
						 static_cast<long long>(b.hostCounters[kSegNHeap]));
		const int64_t nR = b.hostCounters[kSegNBig];
		const int64_t nChunks = b.hostCounters[kSegBigChunks];
		if (nR > b.bigCap || nChunks > b.chunkCap) return hipErrorInvalidValue;
		if (b.hostCounters[kSegNLeaf] + 2 * nR > b.leafCap)
		{
			e = runLeaves(b, s, &nVerified);
			if (e != hipSuccess) return e;
		}
		const SegRange* rs = b.big[cur];
		const int32_t r32 = static_cast<int32_t>(nR);
		// round bytes per element of the level's big ranges: keys read by the
		// count pass (4) and by the rank pass (4), swapped pairs read and
		// written (at most half the range: 8); the median, map and split
		// passes and the chunk scans move per-range and per-chunk words only
		const double bigElems = static_cast<double>(b.hostCounters[kSegBigElems]);
		KTimer* const tm = b.timer;
		int tround = tm ? tm->begin(FLYEHIP_K_ROUND_OTHER, s) : -1;
		hipLaunchKernelGGL(segMedianKernel, dim3(blocksFor(nR + 1, 256)), dim3(256), 0, s, b, rs, r32);
		size_t tb = b.scanTempBytes;
		// chunk counts were written to chunkOff; scan them into chunkPre, then
		// copy back so chunkOff holds offsets (chunkPre is reused below)
		e = hipcub::DeviceScan::ExclusiveSum(b.scanTemp, tb, b.chunkOff, reinterpret_cast<int64_t*>(b.chunkPre),
											 static_cast<int>(nR + 1), s);
		if (e != hipSuccess) return e;
		e = hipMemcpyAsync(b.chunkOff, b.chunkPre, static_cast<size_t>(nR + 1) * sizeof(int64_t),
						   hipMemcpyDeviceToDevice, s);
		if (e != hipSuccess) return e;
		hipLaunchKernelGGL(segMapKernel, dim3(blocksFor(nR, kWavesPerBlock)), dim3(tpb), 0, s, b, r32);
		if (tm) tm->end(tround, s);
		tround = tm ? tm->begin(FLYEHIP_K_ROUND_COUNT, s, 4.0 * bigElems) : -1;
		hipLaunchKernelGGL(segCountKernel, dim3(blocksFor(nChunks + 1, kWavesPerBlock)), dim3(tpb), 0, s, b, rs,
						   nChunks);
		if (tm) tm->end(tround, s);
		tround = tm ? tm->begin(FLYEHIP_K_ROUND_OTHER, s) : -1;
		tb = b.scanTempBytes;
		e = hipcub::DeviceScan::ExclusiveSum(b.scanTemp, tb, b.chunkCnt, b.chunkPre, static_cast<int>(nChunks + 1), s);
		if (e != hipSuccess) return e;
		if (tm) tm->end(tround, s);
		tround = tm ? tm->begin(FLYEHIP_K_ROUND_RANK, s, 4.0 * bigElems) : -1;
		hipLaunchKernelGGL(segRankKernel, dim3(blocksFor(nChunks, kWavesPerBlock)), dim3(tpb), 0, s, b, rs, nChunks);
		if (tm) tm->end(tround, s);
		tround = tm ? tm->begin(FLYEHIP_K_ROUND_SWAP, s, 8.0 * bigElems) : -1;
		hipLaunchKernelGGL(segSwapKernel, dim3(blocksFor(nChunks, kWavesPerBlock)), dim3(tpb), 0, s, b, rs, nChunks);
		if (tm) tm->end(tround, s);
		tround = tm ? tm->begin(FLYEHIP_K_ROUND_OTHER, s) : -1;
		e = hipMemsetAsync(b.counters + kSegNBig, 0, 3 * sizeof(int64_t), s);
		if (e != hipSuccess) return e;
		hipLaunchKernelGGL(segSplitKernel, dim3(blocksFor(nR, 256)), dim3(256), 0, s, b, rs, r32, b.big[cur ^ 1]);
		if (tm) tm->end(tround, s);
		e = hipGetLastError();
		if (e != hipSuccess) return e;
		e = download(b, s);
		if (e != hipSuccess) return e;
		const int64_t nv = b.hostCounters[kSegNVerify];
		if (nv > nVerified)
		{
			hipLaunchKernelGGL(segShadowKernel, dim3(blocksFor(nv - nVerified, kWavesPerBlock)), dim3(tpb), 0, s, b,
							   nVerified, nv);
			nVerified = nv;
		}
		cur ^= 1;
	}
	if (b.hostCounters[kSegError]) return hipErrorIllegalState;
	e = runLeaves(b, s, &nVerified);
	if (e != hipSuccess) return e;
	const int64_t nHeap = b.hostCounters[kSegNHeap];
	if (std::getenv("FLYEHIP_SEGSORT_TRACE"))
		std::fprintf(stderr, "[segsort] rounds=%d verify=%lld free-heap=%lld free-heap-elems=%lld\n", rounds,
					 static_cast<long long>(b.hostCounters[kSegNVerify]), static_cast<long long>(nHeap),
					 static_cast<long long>(b.hostCounters[kSegHeapElems]));
	// heap bytes: key and value read, value written (12 per element)
	if (nHeap > 0)
	{
		// free-mode ranges out of budget: any sort will do; keys stay in place
		int slot = -1;
		e = heapTiers(b, b.heap, nHeap, 0, b.key, b.val, b.key, b.outVal, b.key, b.val, s, &slot);
		if (e != hipSuccess) return e;
		if (b.timer) b.timer->addBytes(slot, 12.0 * static_cast<double>(b.hostCounters[kSegHeapElems]));
	}
	const int64_t nVerify = b.hostCounters[kSegNVerify];
	if (nVerify > 0)
	{
		const int tpb = FH_WAVE * kWavesPerBlock;
		hipLaunchKernelGGL(segVerifyKernel, dim3(blocksFor(nVerify, kWavesPerBlock)), dim3(tpb), 0, s, b, nVerify,
						   heapFlags());
		// the exact element count and the ranges per heap class are known
		// once the verification has run
		if ((e = download(b, s)) != hipSuccess) return e;
		int slot = -1;
		e = heapTiers(b, b.verify, nVerify, 1, b.shadowKey, b.shadowVal, nullptr, b.outVal, b.shadowKey, b.shadowVal, s,
					  &slot, exactPending == nullptr, &b.hostCounters[kSegHeapCls]);
		if (e != hipSuccess) return e;
		if (exactPending) *exactPending = true;
		if (b.timer) b.timer->addBytes(slot, 12.0 * static_cast<double>(b.hostCounters[kSegExactElems]));
		if (std::getenv("FLYEHIP_SEGSORT_TRACE"))
		{
			SegRange* hv = new SegRange[static_cast<size_t>(nVerify)];
			(void)hipMemcpyAsync(hv, b.verify, static_cast<size_t>(nVerify) * sizeof(SegRange), hipMemcpyDeviceToHost, s);
			(void)hipStreamSynchronize(s);
			long long exact = 0, elems = 0, largest = 0;
			for (int64_t i = 0; i < nVerify; ++i)
				if (hv[i].d == 1)
				{
					++exact;
					elems += hv[i].l - hv[i].f;
					largest = std::max<long long>(largest, hv[i].l - hv[i].f);
				}
			std::fprintf(stderr, "[segsort] exact heap sorts=%lld elements=%lld largest=%lld\n", exact, elems, largest);
			delete[] hv;
		}
	}
	return hipSuccess;
}

} // namespace fh
