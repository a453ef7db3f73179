// msort.hip - the per-query std::sort of k-mer matches by (internal extId,
// curPos) (patches/flye/2.9.6/0001-cuflye-candidate-dump.patch:88-91), exact
// in libstdc++'s tie order, for a whole batch of queries.
//
// Only elements with equal keys can end up in an order other than the sorted
// one, and equal (extId, curPos) keys are rare (a k-mer occurring twice in one
// target read).  So:
//   1. a stable segmented radix sort (one segment per query) gives the sorted
//      order S, which is std::sort's result wherever keys are distinct.  The
//      matches of a query are emitted window by window, so curPos never
//      decreases in emission order and sorting the extId bits alone is
//      enough;
//   2. every match gets a 32-bit rank key (position in S of the first equal
//      key), which orders exactly like the 64-bit key, and the positions of S
//      whose key also sits at a neighbouring position are flagged "tied";
//   3. the segmented introsort (segsort.hip) replays std::sort on the rank
//      keys in emission order, dropping every range that holds no tied
//      position (its keys are unique in the query, so its result is S), and
//      writes the replayed permutation into S.
#include <hipcub/hipcub.hpp>

#include "dev_common.hpp"
#include "msort.hpp"

namespace fh
{
namespace
{

__global__ void rankKeysKernel(const int64_t* qOff, int32_t nq, const uint64_t* skey, const uint32_t* sval,
							   uint32_t* rkey, int32_t* dupFlag, int64_t n)
{
	// rank keys are scattered inside the query's range: one XCD per range
	const int64_t i = static_cast<int64_t>(xcdBlockIndex()) * blockDim.x + threadIdx.x;
	if (i >= n) return;
	const int q = findSegWave(qOff, nq, i);
	const int64_t b = qOff[q], e = qOff[q + 1];
	const uint64_t k = skey[i];
	int64_t first = i;
	while (first > b && skey[first - 1] == k) --first;
	rkey[b + sval[i]] = static_cast<uint32_t>(first - b);
	dupFlag[i] = (first < i || (i + 1 < e && skey[i + 1] == k)) ? 1 : 0;
}

inline dim3 gridFor(int64_t n, int tpb) { return dim3(static_cast<unsigned>((n + tpb - 1) / tpb)); }

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

hipError_t launchMatchSortStable(const MsortArgs& a, int64_t nMatches, int idBits, void* temp, size_t tempBytes,
								 hipStream_t s)
{
	if (nMatches == 0 || a.nQueries == 0) return hipSuccess;
	const int tpb = 256;
	hipError_t e = hipcub::DeviceSegmentedRadixSort::SortPairs(temp, tempBytes, a.akey, a.skey, a.aval, a.sval,
															   static_cast<int>(nMatches), a.nQueries, a.qOff,
															   a.qOff + 1, 32, 32 + idBits, s);
	(void)tpb;
	return e;
}

hipError_t launchRankKeys(const MsortArgs& a, int64_t nMatches, hipStream_t s)
{
	if (nMatches == 0 || a.nQueries == 0) return hipSuccess;
	const int tpb = 256;
	hipLaunchKernelGGL(rankKeysKernel, gridFor(nMatches, tpb), dim3(tpb), 0, s, a.qOff, a.nQueries, a.skey, a.sval,
					   a.rkey, a.dupFlag, nMatches);
	return hipGetLastError();
}

} // namespace fh
