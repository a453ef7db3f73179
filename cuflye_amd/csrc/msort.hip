// msort.hip - the per-query std::sort of k-mer matches by (internal extId,
// curPos) (patches/flye/2.9.6/0001-cuflye-candidate-dump.patch:88-91), exact
// in libstdc++'s tie order, for a whole batch of queries.
//
// Only elements with equal keys can end up in an order other than the sorted
// one, and equal (extId, curPos) keys are rare (a k-mer occurring twice in one
// target read).  So:
//   1. a stable sort of each query's matches by extId gives the sorted order
//      S, which is std::sort's result wherever keys are distinct.  The matches
//      of a query are emitted window by window, so curPos never decreases in
//      emission order and sorting by extId alone is enough.  When the target
//      ids fit an LDS histogram, one block per query counts them and scatters
//      every match to its place in emission order (matchCountSortKernel);
//      otherwise a segmented radix sort over the extId bits does it;
//   2. every match gets a 32-bit rank key (position in S of the first equal
//      key), which orders exactly like the 64-bit key, and the positions of S
//      whose key also sits at a neighbouring position are flagged "tied";
//   3. the segmented introsort (segsort.hip) replays std::sort on the rank
//      keys in emission order, dropping every range that holds no tied
//      position (its keys are unique in the query, so its result is S), and
//      writes the replayed permutation into S.
#include <hipcub/hipcub.hpp>

#include "dev_common.hpp"
#include "exsort_wave.hpp"
#include "msort.hpp"

namespace fh
{
namespace
{

// Per-position kernels over a batch's sorted matches: a block takes
// kPosPerThread x 256 consecutive positions, 256 at a time, so the query of
// a position is searched once per wave and then stepped forward.
constexpr int kPosPerThread = 8;

// Tied-position flags of the stable order: a neighbouring position holds the
// same key.
__global__ void tieFlagsKernel(const int64_t* qOff, int32_t nq, const uint64_t* skey, int32_t* dupFlag, int64_t n)
{
	const int64_t first = static_cast<int64_t>(blockIdx.x) * blockDim.x * kPosPerThread + threadIdx.x;
	if (first - static_cast<int64_t>(laneId()) >= n) return;
	int q = findSegWave(qOff, nq, first < n ? first : n - 1);
	for (int j = 0; j < kPosPerThread; ++j)
	{
		const int64_t i = first + static_cast<int64_t>(j) * blockDim.x;
		if (i >= n) break;
		q = stepSeg(qOff, nq, q, i);
		const int64_t b = qOff[q], e = qOff[q + 1];
		const uint64_t k = skey[i];
		const bool tied = (i > b && skey[i - 1] == k) || (i + 1 < e && skey[i + 1] == k);
		dupFlag[i] = tied ? 1 : 0;
	}
}

// Rank keys (first sorted position of the key in its query) in emission
// order, for the queries that hold a flagged position (dup: exclusive prefix
// of the flags); the replay reads no other query's.  With hcur set, equal
// keys are equal curPos inside a group (no group start between them) and
// skey is not read.  With ival set, the emission indices of those queries
// (the replay's values) are written too.
__global__ void rankKeysKernel(const int64_t* qOff, int32_t nq, const uint64_t* skey, const uint32_t* sval,
							   const int64_t* dup, uint32_t* rkey, int64_t n, const int32_t* hcur,
							   const int32_t* groupFlag, uint32_t* ival)
{
	// rank keys are scattered inside the query's range: one XCD per range
	const int64_t first = static_cast<int64_t>(xcdBlockIndex()) * blockDim.x * kPosPerThread + threadIdx.x;
	if (first - static_cast<int64_t>(laneId()) >= n) return;
	int q = findSegWave(qOff, nq, first < n ? first : n - 1);
	for (int j = 0; j < kPosPerThread; ++j)
	{
		const int64_t i = first + static_cast<int64_t>(j) * blockDim.x;
		if (i >= n) break;
		q = stepSeg(qOff, nq, q, i);
		const int64_t b = qOff[q], e = qOff[q + 1];
		if (dup[e] == dup[b]) continue;
		int64_t f = i;
		if (hcur)
		{
			const int32_t c = hcur[i];
			while (f > b && !groupFlag[f] && hcur[f - 1] == c) --f;
		}
		else
		{
			const uint64_t k = skey[i];
			while (f > b && skey[f - 1] == k) --f;
		}
		rkey[b + sval[i]] = static_cast<uint32_t>(f - b);
		// the emission indices are the identity: written in place, coalesced
		if (ival) ival[i] = static_cast<uint32_t>(i - b);
	}
}

inline dim3 gridForPos(int64_t n, int tpb)
{
	const int64_t per = static_cast<int64_t>(tpb) * kPosPerThread;
	return dim3(static_cast<unsigned>((n + per - 1) / per));
}

// Stable counting sort of one query's matches by internal extId (the high 32
// bits of the key), one 1,024-thread block per query:
//   1. a histogram of the extIds in LDS (LDS atomics);
//   2. its exclusive scan: the first slot of every extId in S;
//   3. the matches in emission order, 6,144 at a time (every wave 384
//      consecutive ones as 6 rows of 64; more rows spill registers): every
//      wave finds, for each lane,
//      the lanes of its row with the same extId (one ballot per id bit), so a
//      lane knows its rank among them and each group has a leader; the
//      waves' leaders claim their slots from the histogram in emission order
//      (one wave at a time, a block barrier between turns), and every match
//      is written to its slot.
// Equal extIds keep their emission order, which is the stable order.  The
// join writes every match's emission index, which is its position in the
// query's emission order, so the values are the positions themselves.  With
// hcur set the kernel writes the chain's hits in S order instead of the
// keys: curPos, the match's extPos (mext in emission order) and the
// group-start flag (the first slot of an extId; the first claim of a bin sets
// bit 31 of its histogram word and passes it on in bit 31 of the leader's
// base); skey is then written at the group starts only.
constexpr int kMsortThreads = 1024;
constexpr int kMsortWaves = kMsortThreads / FH_WAVE;
constexpr int kMsortRows = 6;

// Block barrier that orders LDS accesses only: global loads issued before it
// stay in flight.
__device__ inline void ldsBarrier()
{
	__builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
	__builtin_amdgcn_s_barrier();
	__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Lanes of the wave whose bin equals this lane's (0 for bin < 0, a lane
// without an element): one ballot per bit of the bin.
__device__ inline unsigned long long sameBinMask(int bin, int bits)
{
	const bool act = bin >= 0;
	unsigned long long m = __ballot(act);
	for (int bit = 0; bit < bits; ++bit)
	{
		const bool on = act && ((bin >> bit) & 1);
		const unsigned long long bb = __ballot(on);
		m &= on ? bb : ~bb;
	}
	return act ? m : 0ULL;
}
// LDS: histogram words + one base word per (row, lane) + wave totals
constexpr int kMsortFixedWords = kMsortRows * kMsortThreads + kMsortWaves;

__global__ void __launch_bounds__(kMsortThreads) matchCountSortKernel(const int64_t* qOff, const int32_t* order,
																	  const uint64_t* akey, uint64_t* skey,
																	  uint32_t* sval, int nBins, int bits,
																	  MsortHits hits)
{
	constexpr uint32_t kClaimed = 0x80000000u;
	extern __shared__ uint32_t lds[];
	uint32_t* hist = lds;
	uint32_t* lead = lds + nBins;                       // kMsortRows * kMsortThreads
	uint32_t* waveTot = lead + kMsortRows * kMsortThreads;  // kMsortWaves
	const int t = static_cast<int>(threadIdx.x);
	const int wv = t / FH_WAVE;
	const int lane = laneId();
	// the largest queries first: a block runs a whole query, so the longest
	// one bounds the kernel unless it starts early
	const int q = order ? order[blockIdx.x] : static_cast<int>(blockIdx.x);
	const int64_t b0 = qOff[q];
	const int64_t n = qOff[q + 1] - b0;
	if (n == 0) return;
	for (int i = t; i < nBins; i += kMsortThreads) hist[i] = 0;
	__syncthreads();
	// (per-run aggregation with the bin ballots measured slower than the LDS
	// atomics' serialisation on equal bins)
	for (int64_t i = t; i < n; i += kMsortThreads) atomicAdd(&hist[static_cast<uint32_t>(akey[b0 + i] >> 32)], 1u);
	__syncthreads();
	const unsigned long long lt = laneMaskLt();
	// exclusive scan of the histogram: a contiguous piece per thread
	{
		const int per = (nBins + kMsortThreads - 1) / kMsortThreads;
		const int lo = min(nBins, t * per), hi = min(nBins, lo + per);
		uint32_t sum = 0;
		for (int i = lo; i < hi; ++i) sum += hist[i];
		const int incl = waveInclusiveSum(static_cast<int>(sum));
		if (lane == FH_WAVE - 1) waveTot[wv] = static_cast<uint32_t>(incl);
		__syncthreads();
		uint32_t run = static_cast<uint32_t>(incl) - sum;
		for (int w = 0; w < wv; ++w) run += waveTot[w];
		for (int i = lo; i < hi; ++i)
		{
			const uint32_t c = hist[i];
			hist[i] = run;
			run += c;
		}
		__syncthreads();
	}
	// tile t0 is processed while the next one loads: the barriers of the
	// claiming turns order LDS only, so the next tile's global loads stay in
	// flight across them
	const bool withHits = hits.hcur != nullptr;
	uint64_t key[kMsortRows];
	int32_t ext[kMsortRows];
	int bin[kMsortRows];
	auto load = [&](int64_t base, uint64_t* k, int32_t* e, int* bn) {
#pragma unroll
		for (int r = 0; r < kMsortRows; ++r)
		{
			const int64_t i = base + (wv * kMsortRows + r) * FH_WAVE + lane;
			const bool act = i < n;
			k[r] = act ? akey[b0 + i] : 0;
			e[r] = act && withHits ? hits.mext[b0 + i] : 0;
			bn[r] = act ? static_cast<int>(k[r] >> 32) : -1;
		}
	};
	load(0, key, ext, bin);
	for (int64_t t0 = 0; t0 < n; t0 += kMsortRows * kMsortThreads)
	{
		unsigned long long same[kMsortRows];
#pragma unroll
		for (int r = 0; r < kMsortRows; ++r) same[r] = sameBinMask(bin[r], bits);
		uint64_t nkey[kMsortRows];
		int32_t next[kMsortRows];
		int nbin[kMsortRows];
		load(t0 + kMsortRows * kMsortThreads, nkey, next, nbin);
		// leaders claim slots in emission order: wave w holds the tile's
		// elements [w * 384, (w + 1) * 384) as 6 rows of 64, so the waves take
		// turns in order and each claims its rows in order (LDS operations of
		// one wave complete in order)
		for (int w = 0; w < kMsortWaves; ++w)
		{
			if (wv == w)
			{
#pragma unroll
				for (int r = 0; r < kMsortRows; ++r)
				{
					if (bin[r] >= 0 && (same[r] & lt) == 0)
					{
						const uint32_t h = hist[bin[r]];
						const uint32_t base = h & ~kClaimed;
						hist[bin[r]] = (base + static_cast<uint32_t>(__popcll(same[r]))) | kClaimed;
						// bit 31: this run opens the bin (its first slot starts a group)
						lead[r * kMsortThreads + t] = base | (h & kClaimed ? 0u : kClaimed);
					}
				}
			}
			ldsBarrier();
		}
#pragma unroll
		for (int r = 0; r < kMsortRows; ++r)
		{
			if (bin[r] < 0) continue;
			const int ld = __ffsll(static_cast<long long>(same[r])) - 1;
			const uint32_t lw = lead[r * kMsortThreads + wv * FH_WAVE + ld];
			const int rank = __popcll(same[r] & lt);
			const int64_t d = b0 + (lw & ~kClaimed) + rank;
			sval[d] = static_cast<uint32_t>(t0 + (wv * kMsortRows + r) * FH_WAVE + lane);
			if (withHits)
			{
				// the hits carry the keys: skey only at the group starts
				const bool start = (lw & kClaimed) && rank == 0;
				hits.hcur[d] = static_cast<int32_t>(static_cast<uint32_t>(key[r]));
				hits.hext[d] = ext[r];
				hits.groupFlag[d] = start ? 1 : 0;
				if (start) skey[d] = key[r];
			}
			else
			{
				skey[d] = key[r];
			}
		}
		ldsBarrier();  // lead[] is rewritten by the next tile
#pragma unroll
		for (int r = 0; r < kMsortRows; ++r)
		{
			key[r] = nkey[r];
			ext[r] = next[r];
			bin[r] = nbin[r];
		}
	}
}

int msortMaxBins()
{
	static int maxBins = -1;
	if (maxBins < 0)
	{
		const int maxBytes = 160 * 1024;
		maxBins = hipFuncSetAttribute(reinterpret_cast<const void*>(matchCountSortKernel),
									  hipFuncAttributeMaxDynamicSharedMemorySize, maxBytes) == hipSuccess
					  ? maxBytes / 4 - kMsortFixedWords
					  : 65536 / 4 - kMsortFixedWords;
	}
	return maxBins;
}

} // namespace

hipError_t msortTempBytes(int64_t nMatches, int32_t nQueries, int idBits, size_t* bytes)
{
	*bytes = 0;
	return hipcub::DeviceSegmentedRadixSort::SortPairs(nullptr, *bytes, static_cast<const uint64_t*>(nullptr),
													   static_cast<uint64_t*>(nullptr),
													   static_cast<const uint32_t*>(nullptr),
													   static_cast<uint32_t*>(nullptr), static_cast<int>(nMatches),
													   nQueries, static_cast<const int64_t*>(nullptr),
													   static_cast<const int64_t*>(nullptr), 32, 32 + idBits);
}

bool msortCounting(int nBins) { return nBins <= msortMaxBins(); }

hipError_t launchMatchSortStable(const MsortArgs& a, int64_t nMatches, int idBits, int nBins, void* temp,
								 size_t tempBytes, hipStream_t s)
{
	if (nMatches == 0 || a.nQueries == 0) return hipSuccess;
	if (msortCounting(nBins))
	{
		const size_t ldsBytes = static_cast<size_t>(nBins + kMsortFixedWords) * 4;
		hipLaunchKernelGGL(matchCountSortKernel, dim3(static_cast<unsigned>(a.nQueries)), dim3(kMsortThreads), ldsBytes, s,
						   a.qOff, a.order, a.akey, a.skey, a.sval, nBins, idBits, a.hits);
		return hipGetLastError();
	}
	return hipcub::DeviceSegmentedRadixSort::SortPairs(temp, tempBytes, a.akey, a.skey, a.aval, a.sval,
													   static_cast<int>(nMatches), a.nQueries, a.qOff, a.qOff + 1, 32,
													   32 + idBits, s);
}

hipError_t launchTieFlags(const MsortArgs& a, int64_t nMatches, hipStream_t s)
{
	if (nMatches == 0 || a.nQueries == 0) return hipSuccess;
	const int tpb = 256;
	hipLaunchKernelGGL(tieFlagsKernel, gridForPos(nMatches, tpb), dim3(tpb), 0, s, a.qOff, a.nQueries, a.skey,
					   a.dupFlag, nMatches);
	return hipGetLastError();
}

hipError_t launchRankKeys(const MsortArgs& a, const int64_t* dup, int64_t nMatches, hipStream_t s)
{
	if (nMatches == 0 || a.nQueries == 0) return hipSuccess;
	const int tpb = 256;
	hipLaunchKernelGGL(rankKeysKernel, gridForPos(nMatches, tpb), dim3(tpb), 0, s, a.qOff, a.nQueries, a.skey, a.sval,
					   dup, a.rkey, nMatches, a.hits.hcur, a.hits.groupFlag, a.ival);
	return hipGetLastError();
}

} // namespace fh
