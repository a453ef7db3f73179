// msort.hip - the per-query std::sort of k-mer matches by (internal extId,
// curPos) (patches/flye/2.9.6/0001-cuflye-candidate-dump.patch:88-91), exact
// in libstdc++'s tie order, for a whole batch of queries.
//
// Only elements with equal keys can end up in an order other than the sorted
// one, and equal (extId, curPos) keys are rare (a k-mer occurring twice in one
// target read).  So:
//   1. a stable segmented radix sort (one segment per query) gives the sorted
//      order S, which is std::sort's result wherever keys are distinct;
//   2. keys that occur more than once in a query are flagged ("tied");
//   3. for queries with tied keys one wave replays introsort on the emission
//      order, but only inside ranges that still hold a tied element: a range
//      without tied elements holds exactly the ranks it spans, all distinct,
//      so its final content is S on that range and it is not followed
//      further.  Ranges that reach the lane or tile size are finished exactly
//      (exsort_wave.hpp) and marked;
//   4. marked positions take the replayed array, all others take S.
// The replay touches ~n * (log2(tied elements) + 2) elements instead of
// n * log2(n), and queries without ties skip it.
#include <hipcub/hipcub.hpp>

#include "exsort_wave.hpp"
#include "msort.hpp"

namespace fh
{
namespace
{

constexpr int kTile = 1024;
constexpr int kLaneMax = 48;

__device__ inline int findSeg(const int64_t* off, int n, int64_t w)
{
	int lo = 0, hi = n - 1;
	while (lo < hi)
	{
		const int mid = (lo + hi + 1) >> 1;
		if (off[mid] <= w) lo = mid;
		else hi = mid - 1;
	}
	return lo;
}

__global__ void iotaKernel(const int64_t* qOff, int32_t nq, uint32_t* v, int64_t n)
{
	const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
	if (i >= n) return;
	const int q = findSeg(qOff, nq, i);
	v[i] = static_cast<uint32_t>(i - qOff[q]);
}

// Tie flags of sorted positions, scattered to emission indices, and per-query
// counts of tied elements.
__global__ void tieFlagKernel(const int64_t* qOff, int32_t nq, const uint64_t* skey, const uint32_t* sval,
							  uint8_t* tiedByIdx, int32_t* qTied, int64_t n)
{
	const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
	if (i >= n) return;
	const int q = findSeg(qOff, nq, i);
	const int64_t b = qOff[q], e = qOff[q + 1];
	const uint64_t k = skey[i];
	const bool tied = (i > b && skey[i - 1] == k) || (i + 1 < e && skey[i + 1] == k);
	tiedByIdx[b + sval[i]] = tied ? 1 : 0;
	if (tied) atomicAdd(&qTied[q], 1);
}

struct PSeg
{
	int f;
	int l;
	int d;
	int tied;
};

__device__ inline void markRange(uint8_t* fromA, int f, int l)
{
	for (int i = f + laneId(); i < l; i += FH_WAVE) fromA[i] = 1;
}

// Tied elements in a[f, l) (positions hold emission indices in v).
__device__ inline int countTied(const uint32_t* v, const uint8_t* tiedByIdx, int f, int l)
{
	int c = 0;
	for (int i = f + laneId(); i < l; i += FH_WAVE) c += tiedByIdx[v[i]];
	return waveSum(c);
}

__global__ void __launch_bounds__(FH_WAVE) replayKernel(MsortArgs a)
{
	__shared__ WaveSortLds<uint64_t, kTile> lds;
	const int lane = laneId();
	for (;;)
	{
		int qi = 0;
		if (lane == 0) qi = atomicAdd(a.queue, 1);
		qi = __shfl(qi, 0, FH_WAVE);
		if (qi >= a.nQueries) break;
		const int q = a.order ? a.order[qi] : qi;
		const int tiedTotal = a.qTied[q];
		if (tiedTotal == 0) continue;
		const int64_t b = a.qOff[q];
		const int n = static_cast<int>(a.qOff[q + 1] - b);
		uint64_t* k = a.akey + b;
		uint32_t* v = a.aval + b;
		int* posL = a.posL + b;
		int* posR = a.posR + b;
		uint8_t* fromA = a.fromA + b;
		const uint8_t* tb = a.tiedByIdx + b;
		const exsort::Span<uint64_t> g{k, v};
		PSeg st[40];
		int top = 0;
		st[top++] = PSeg{0, n, exsort::depthLimit(static_cast<uint32_t>(n)), tiedTotal};
		while (top > 0)
		{
			const PSeg w = st[--top];
			if (w.tied == 0) continue;  // holds its ranks, all distinct: S is final here
			const int size = w.l - w.f;
			if (size <= kTile || w.d == 0)
			{
				if (size <= kTile) waveTileSort<uint64_t, kTile, kLaneMax>(g, w.f, w.l, w.d, lds);
				else
				{
					if (lane == 0)
					{
						exsort::heapSort(g, w.f, w.l);
					}
					waveSync();
				}
				markRange(fromA, w.f, w.l);
				waveSync();
				continue;
			}
			const int cut = wavePartition(g, w.f, w.l, posL, posR);
			const int tl = countTied(v, tb, w.f, cut);
			const PSeg lft{w.f, cut, w.d - 1, tl}, rgt{cut, w.l, w.d - 1, w.tied - tl};
			const bool leftSmaller = cut - w.f < w.l - cut;
			st[top++] = leftSmaller ? rgt : lft;
			st[top++] = leftSmaller ? lft : rgt;
		}
	}
}

__global__ void composeKernel(const uint64_t* akey, const uint32_t* aval, const uint8_t* fromA, uint64_t* skey,
							  uint32_t* sval, int64_t n)
{
	const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
	if (i >= n || !fromA[i]) return;
	skey[i] = akey[i];
	sval[i] = aval[i];
}

__global__ void countTieQueriesKernel(const int32_t* qTied, int32_t nq, int32_t* out)
{
	const int q = blockIdx.x * blockDim.x + threadIdx.x;
	if (q < nq && qTied[q] > 0) atomicAdd(out, 1);
}

inline dim3 gridFor(int64_t n, int tpb) { return dim3(static_cast<unsigned>((n + tpb - 1) / tpb)); }

} // namespace

hipError_t msortTempBytes(int64_t nMatches, int32_t nQueries, int endBit, size_t* bytes)
{
	*bytes = 0;
	return hipcub::DeviceSegmentedRadixSort::SortPairs(nullptr, *bytes, static_cast<const uint64_t*>(nullptr),
													   static_cast<uint64_t*>(nullptr),
													   static_cast<const uint32_t*>(nullptr),
													   static_cast<uint32_t*>(nullptr), static_cast<int>(nMatches),
													   nQueries, static_cast<const int64_t*>(nullptr),
													   static_cast<const int64_t*>(nullptr), 0, endBit);
}

hipError_t launchMatchSortExact(const MsortArgs& a, int64_t nMatches, int endBit, void* temp, size_t tempBytes,
								int replayBlocks, hipStream_t s)
{
	if (nMatches == 0 || a.nQueries == 0) return hipSuccess;
	const int tpb = 256;
	hipLaunchKernelGGL(iotaKernel, gridFor(nMatches, tpb), dim3(tpb), 0, s, a.qOff, a.nQueries, a.aval, nMatches);
	hipError_t e = hipcub::DeviceSegmentedRadixSort::SortPairs(temp, tempBytes, a.akey, a.skey, a.aval, a.sval,
															   static_cast<int>(nMatches), a.nQueries, a.qOff,
															   a.qOff + 1, 0, endBit, s);
	if (e != hipSuccess) return e;
	hipLaunchKernelGGL(tieFlagKernel, gridFor(nMatches, tpb), dim3(tpb), 0, s, a.qOff, a.nQueries, a.skey, a.sval,
					   a.tiedByIdx, a.qTied, nMatches);
	hipLaunchKernelGGL(replayKernel, dim3(static_cast<unsigned>(replayBlocks)), dim3(FH_WAVE), 0, s, a);
	hipLaunchKernelGGL(composeKernel, gridFor(nMatches, tpb), dim3(tpb), 0, s, a.akey, a.aval, a.fromA, a.skey,
					   a.sval, nMatches);
	hipLaunchKernelGGL(countTieQueriesKernel, gridFor(a.nQueries, tpb), dim3(tpb), 0, s, a.qTied, a.nQueries,
					   a.tieQueries);
	return hipGetLastError();
}

} // namespace fh
