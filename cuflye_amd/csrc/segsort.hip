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
	exsort::medianToFirst(s, w.f, w.f + 1, w.f + (w.l - w.f) / 2, w.l - 1);
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
	segPushWave(b, next, act, w.f, cut, w.d - 1);
	segPushWave(b, next, act, cut, w.l, w.d - 1);
}

__global__ void __launch_bounds__(FH_WAVE) segLeafKernel(SegSortBufs b, int64_t nLeaf)
{
	__shared__ WaveSortLds<uint32_t, kSegTile> lds;
	const int64_t i = blockIdx.x;
	if (i >= nLeaf) return;
	const SegRange w = b.leaf[i];
	waveTileSort<uint32_t, kSegTile, 48>(exsort::Span<uint32_t>{b.key, b.val}, w.f, w.l, w.d, lds, b.outVal);
}

__global__ void segHeapKernel(SegSortBufs b, int64_t nHeap)
{
	const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
	if (i >= nHeap) return;
	const SegRange w = b.heap[i];
	exsort::heapSort(exsort::Span<uint32_t>{b.key, b.val}, w.f, w.l);
	if (b.outVal != b.val)
		for (int32_t x = w.f; x < w.l; ++x) b.outVal[x] = b.val[x];
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
	segPushWave(b, b.big[0], act, f, l, act ? exsort::depthLimit(static_cast<uint32_t>(l - f)) : 0);
}

hipError_t download(const SegSortBufs& b, hipStream_t s)
{
	hipError_t e = hipMemcpyAsync(b.hostCounters, b.counters, kSegCounters * sizeof(int64_t), hipMemcpyDeviceToHost, s);
	if (e != hipSuccess) return e;
	return hipStreamSynchronize(s);
}

hipError_t runLeaves(SegSortBufs& b, hipStream_t s)
{
	const int64_t nLeaf = b.hostCounters[kSegNLeaf];
	if (nLeaf > 0)
	{
		hipLaunchKernelGGL(segLeafKernel, dim3(static_cast<unsigned>(nLeaf)), dim3(FH_WAVE), 0, s, b, nLeaf);
		hipError_t e = hipGetLastError();
		if (e != hipSuccess) return e;
	}
	b.hostCounters[kSegNLeaf] = 0;
	return hipMemsetAsync(b.counters + kSegNLeaf, 0, sizeof(int64_t), s);
}

} // namespace

hipError_t segSortCaps(int64_t n, int64_t nArrays, SegSortCaps* caps)
{
	caps->bigCap = n / (kSegTile + 1) + 2;
	caps->leafCap = nArrays + 2 * caps->bigCap + 64;
	caps->heapCap = caps->bigCap;
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

hipError_t segSortRun(SegSortBufs& b, hipStream_t s, int64_t* tieArrays)
{
	hipError_t e = download(b, s);
	if (e != hipSuccess) return e;
	if (tieArrays) *tieArrays = b.hostCounters[kSegTieArrays];
	int cur = 0;
	const int tpb = FH_WAVE * kWavesPerBlock;
	while (b.hostCounters[kSegNBig] > 0)
	{
		const int64_t nR = b.hostCounters[kSegNBig];
		const int64_t nChunks = b.hostCounters[kSegBigChunks];
		if (nR > b.bigCap || nChunks > b.chunkCap) return hipErrorInvalidValue;
		if (b.hostCounters[kSegNLeaf] + 2 * nR > b.leafCap)
		{
			e = runLeaves(b, s);
			if (e != hipSuccess) return e;
		}
		const SegRange* rs = b.big[cur];
		const int32_t r32 = static_cast<int32_t>(nR);
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
		hipLaunchKernelGGL(segCountKernel, dim3(blocksFor(nChunks + 1, kWavesPerBlock)), dim3(tpb), 0, s, b, rs,
						   nChunks);
		tb = b.scanTempBytes;
		e = hipcub::DeviceScan::ExclusiveSum(b.scanTemp, tb, b.chunkCnt, b.chunkPre, static_cast<int>(nChunks + 1), s);
		if (e != hipSuccess) return e;
		hipLaunchKernelGGL(segRankKernel, dim3(blocksFor(nChunks, kWavesPerBlock)), dim3(tpb), 0, s, b, rs, nChunks);
		hipLaunchKernelGGL(segSwapKernel, dim3(blocksFor(nChunks, kWavesPerBlock)), dim3(tpb), 0, s, b, rs, nChunks);
		e = hipMemsetAsync(b.counters + kSegNBig, 0, 2 * sizeof(int64_t), s);
		if (e != hipSuccess) return e;
		hipLaunchKernelGGL(segSplitKernel, dim3(blocksFor(nR, 256)), dim3(256), 0, s, b, rs, r32, b.big[cur ^ 1]);
		e = hipGetLastError();
		if (e != hipSuccess) return e;
		e = download(b, s);
		if (e != hipSuccess) return e;
		cur ^= 1;
	}
	if (b.hostCounters[kSegError]) return hipErrorIllegalState;
	e = runLeaves(b, s);
	if (e != hipSuccess) return e;
	const int64_t nHeap = b.hostCounters[kSegNHeap];
	if (nHeap > 0)
	{
		hipLaunchKernelGGL(segHeapKernel, dim3(blocksFor(nHeap, 64)), dim3(64), 0, s, b, nHeap);
		e = hipGetLastError();
		if (e != hipSuccess) return e;
	}
	return hipSuccess;
}

} // namespace fh
