// chain.hip - per-(query, target) group chaining on gfx950: prefilters, the
// extPos re-sort, the bounded-lookback chaining DP, the score-order sort, the
// backtrack, overlapTest, divergence, primary selection and the divergence
// filter of OverlapDetector::getSeqOverlaps (restated in
// tools/replay_overlap_chains.py:228-380 and
// tools/replay_read_to_graph_source_pack.py:352-491).
//
// Three persistent kernels pull groups from device work queues over a
// size-descending order (big groups first, so the heavy tail starts early):
//   prep    one wave per group: prefilters (wave reductions), the exact
//           std::sort by extPos when extLen > curLen, group-order hit copies;
//   dp      G lanes per group (G in {1, 4, 8, 16, 32, 64}, the lookback tile
//           width the autotuner picks): row i scores G predecessors per step,
//           a prefix max over the tile reproduces the serial strictly-greater
//           running maximum, a ballot finds the first predecessor where the
//           serial loop breaks (an improvement with jumpDiv == 0 && dc < k, or
//           the sorted-axis delta passing maxJump); recent scores stay in an
//           LDS ring per group;
//   claim   one wave per group: the backtrack walks of Flye's score order
//           computed without sorting (subtree maxima of the backtrack forest,
//           below), overlapTest, divergence, selection; a group where equal
//           scores could change the walks is left to
//   finish  one wave per group: exact std::sort of the rows by score, the
//           backtrack (lane 0, O(n)), overlapTest, divergence, selection.
#include "chain.hpp"
#include "exsort_wave.hpp"
#include "segsort.hpp"

namespace fh
{

namespace
{

constexpr int kLaneMax = 48;
constexpr uint32_t kWalkRelevant = 0x80000000u;  // score-order value bit: row has a backtrack pointer

// std::sort(indices, score[a] > score[b]) as an ascending 32-bit key (scores
// are non-negative).
__device__ inline uint32_t scoreOrderKey(int32_t score) { return static_cast<uint32_t>(0x7fffffff - score); }
// LDS ring entries per wave (two rings), split over groups: 4 KiB of LDS per
// wave keeps the DP's occupancy at its VGPR limit (a 16 KiB ring held it to
// about 2.5 waves per SIMD, and the lookback scan is latency-bound)
constexpr int kRingTotal = 512;

__device__ inline float flyeLogf(float x, const uint32_t* tabX, const uint32_t* tabY, int tabN)
{
	const uint32_t u = __float_as_uint(x);
	int lo = 0, hi = tabN;
	while (lo < hi)
	{
		int mid = (lo + hi) >> 1;
		if (tabX[mid] < u) lo = mid + 1;
		else hi = mid;
	}
	if (lo < tabN && tabX[lo] == u) return __uint_as_float(tabY[lo]);
	return static_cast<float>(log(static_cast<double>(x)));
}

__device__ inline int32_t iabs(int32_t v) { return v < 0 ? -v : v; }

__device__ bool overlapTestDev(const Proposal& o, int64_t curId, int64_t extId, int32_t curLen,
							   int32_t extLen, const flyehip_overlap_params_t& p)
{
	const int32_t cr = o.curEnd - o.curBegin;
	const int32_t er = o.extEnd - o.extBegin;
	if (cr < p.minimum_overlap || er < p.minimum_overlap) return false;
	const int32_t lengthDiff = iabs(cr - er);
	// lengthDiff > 0.5 * min(cr, er) and intersect > curRange / 2, on integers
	if (2LL * lengthDiff > static_cast<long long>(min(cr, er))) return false;
	if (curId == extId)
	{
		const int32_t inter = min(o.curEnd, o.extEnd) - max(o.curBegin, o.extBegin);
		if (2LL * inter > static_cast<long long>(cr)) return false;
	}
	if (curId == -extId)
	{
		const int32_t inter = min(o.curEnd, extLen - o.extBegin) - max(o.curBegin, extLen - o.extEnd);
		if (2LL * inter > static_cast<long long>(cr)) return false;
	}
	if (!p.force_local && p.check_overhang)
	{
		const int32_t oh = max(min(o.curBegin, o.extBegin), min(curLen - o.curEnd, extLen - o.extEnd));
		if (oh > p.maximum_overhang) return false;
	}
	return true;
}

__device__ inline int lowerBound(const int32_t* a, int n, int32_t v)
{
	int lo = 0, hi = n;
	while (lo < hi)
	{
		int mid = (lo + hi) >> 1;
		if (a[mid] < v) lo = mid + 1;
		else hi = mid;
	}
	return lo;
}
__device__ inline int upperBound(const int32_t* a, int n, int32_t v)
{
	int lo = 0, hi = n;
	while (lo < hi)
	{
		int mid = (lo + hi) >> 1;
		if (a[mid] <= v) lo = mid + 1;
		else hi = mid;
	}
	return lo;
}

// A group's hits in processing order: the extPos-sorted copy, or the
// (extId, curPos) order itself for groups that keep it.
__device__ inline const int32_t* groupHits(const int32_t* sorted, const int32_t* inOrder, int64_t hb, bool extSorted)
{
	return (extSorted ? sorted : inOrder) + hb;
}

// Whether a group's hits are read from the extPos-sorted copy: sorted by
// extPos, unless their extPos already ascend (inPlace, set by the prefilter).
__device__ inline bool gatheredHits(const GroupDesc& gd, const QueryDesc& qd)
{
	return gd.extLen > qd.curLen && !gd.inPlace;
}

__device__ inline int pullGroup(int32_t* queue)
{
	int g = 0;
	if (laneId() == 0) g = atomicAdd(queue, 1);
	return __shfl(g, 0, FH_WAVE);
}

// ---- prep ------------------------------------------------------------------

// Prefilters (replay_overlap_chains.py:245-270).  Each wave takes 64
// consecutive groups of the processing order: a group whose size is below the
// unique-curPos bound (unique curPos <= n) fails without reading its hits; the
// others are examined by the whole wave one after another (with passKnown,
// groupPass already holds the result and only passing groups are visited).
// Passing groups with extLen > curLen get their extPos sort keys and are
// pushed to the segmented sort.
__global__ void __launch_bounds__(FH_WAVE) chainPrefilterKernel(ChainArgs a, SegSortBufs sb)
{
	const int lane = laneId();
	const flyehip_overlap_params_t& p = a.params;
	const float minUniq = p.min_kmer_survival_rate * static_cast<float>(p.minimum_overlap);
	for (;;)
	{
		int base = 0;
		if (lane == 0) base = atomicAdd(a.queue, FH_WAVE);
		base = __shfl(base, 0, FH_WAVE);
		if (base >= a.nGroups) break;
		const int qi = base + lane;
		int myG = -1;
		bool cand = false;
		if (qi < a.nGroups)
		{
			myG = a.order[qi];
			// with passKnown only the passing groups have work left here
			if (a.passKnown)
			{
				cand = a.groupPass[myG] != 0;
			}
			else
			{
				const int n = a.groups[myG].n;
				cand = n >= 2 && !(static_cast<float>(n) < minUniq);
			}
			if (!cand)
			{
				a.groupPass[myG] = 0;
				a.groupCount[myG] = 0;
			}
		}
		unsigned long long todo = __ballot(cand);
		// the extPos sorts of the 64 groups, pushed together after them (one
		// atomic per list and wave, not per group): lane l holds group l's
		bool doPush = false;
		int32_t pushF = 0, pushL = 0;
		while (todo)
		{
			const int l = __ffsll(static_cast<long long>(todo)) - 1;
			todo &= todo - 1;
			const int g = __shfl(myG, l, FH_WAVE);
			const GroupDesc gd = a.groups[g];
			const QueryDesc qd = a.queries[gd.query];
			const int n = gd.n;
			const int64_t hb = gd.hitStart;
			const int32_t* cur0 = a.cur + hb;
			const int32_t* ext0 = a.ext + hb;
			bool pass;
			if (a.passKnown)
			{
				pass = a.groupPass[g] != 0;  // chainRelevanceKernel
			}
			else
			{
				// unique curPos counted from prevPos = 0 on the (extId, curPos) order
				int uniq = 0;
				int mnE = 0x7fffffff, mxE = -0x7fffffff - 1;
#pragma unroll 4
				for (int i = lane; i < n; i += FH_WAVE)
				{
					const int32_t c = cur0[i];
					const int32_t prev = i == 0 ? 0 : cur0[i - 1];
					uniq += (c != prev);
					const int32_t e = ext0[i];
					mnE = min(mnE, e);
					mxE = max(mxE, e);
				}
				uniq = waveSum(uniq);
				mnE = waveMin(mnE);
				mxE = waveMax(mxE);
				const int32_t mnC = cur0[0];
				const int32_t mxC = cur0[n - 1];
				pass = !(static_cast<float>(uniq) < minUniq);
				if (pass && (mxC - mnC < p.minimum_overlap || mxE - mnE < p.minimum_overlap)) pass = false;
				if (pass && p.check_overhang && !p.force_local)
				{
					if (min(mnC, mnE) > p.maximum_overhang) pass = false;
					if (min(qd.curLen - mxC, gd.extLen - mxE) > p.maximum_overhang) pass = false;
				}
			}
			const bool extSorted = pass && gd.extLen > qd.curLen;
			// extPos already strictly ascending: distinct keys in order, which
			// every sort leaves as they are, so the group needs no sort
			bool ascending = true;
			if (extSorted)
			{
				bool down = false;
#pragma unroll 4
				for (int i = lane + 1; i < n; i += FH_WAVE) down |= !(ext0[i - 1] < ext0[i]);
				ascending = __ballot(down) == 0;
			}
			if (extSorted && !ascending)
			{
				// sort keys and positions only for the groups the sort takes
				uint32_t* sk = a.sortKey + hb;
				uint32_t* sv = a.sortVal + hb;
#pragma unroll 4
				for (int i = lane; i < n; i += FH_WAVE)
				{
					sk[i] = static_cast<uint32_t>(ext0[i]);
					sv[i] = static_cast<uint32_t>(i);
				}
			}
			if (lane == 0)
			{
				a.groupPass[g] = pass ? 1 : 0;
				a.groupCount[g] = 0;
				a.groups[g].inPlace = extSorted && ascending ? 1 : 0;
			}
			if (lane == l)
			{
				doPush = extSorted && !ascending;
				pushF = static_cast<int32_t>(hb);
				pushL = static_cast<int32_t>(hb + n);
			}
		}
		segPushWave(sb, sb.big[0], doPush, pushF, pushL, doPush ? exsort::depthLimit(static_cast<uint32_t>(pushL - pushF)) : 0,
					0);
	}
}

// Before the match sort's tie replay, on the stable (extId, curPos) order:
// the prefilters of every group (they do not depend on the order of tied
// hits) into groupPass, and whether the order of the group's tied hits (equal
// curPos) can reach a result into groupRel: in a passing group chained in
// curPos order it can; in a passing group sorted by extPos only if two of its
// hits share an extPos (otherwise the extPos order is the same whatever order
// the sort starts from).  That test inserts extPos + 1 into a hash table of 2n
// slots in the group's posL / posR scratch (workgroup-scope atomics, the
// group's one wave being the workgroup).  In the groups it marks, the tied
// hits get tieFlag 1 (a group is one extId, so equal keys are equal curPos
// inside it; the caller clears tieFlag first).  Groups are taken 64 at a time
// in index order.
__device__ inline int32_t* hashSlot(int32_t* lo, int32_t* hi, int n, uint32_t s)
{
	return static_cast<int>(s) < n ? lo + s : hi + (static_cast<int>(s) - n);
}

__global__ void __launch_bounds__(FH_WAVE) chainRelevanceKernel(ChainArgs a, int32_t* groupRel, int32_t* tieFlag)
{
	const int lane = laneId();
	const flyehip_overlap_params_t& p = a.params;
	const float minUniq = p.min_kmer_survival_rate * static_cast<float>(p.minimum_overlap);
	for (;;)
	{
		int base = 0;
		if (lane == 0) base = atomicAdd(a.queue, FH_WAVE);
		base = __shfl(base, 0, FH_WAVE);
		if (base >= a.nGroups) break;
		const int myG = base + lane;
		bool cand = false;
		if (myG < a.nGroups)
		{
			const int n = a.groups[myG].n;
			cand = n >= 2 && !(static_cast<float>(n) < minUniq);
			if (!cand)
			{
				a.groupPass[myG] = 0;
				groupRel[myG] = 0;
			}
		}
		unsigned long long todo = __ballot(cand);
		while (todo)
		{
			const int l = __ffsll(static_cast<long long>(todo)) - 1;
			todo &= todo - 1;
			const int g = base + l;
			const GroupDesc gd = a.groups[g];
			const QueryDesc qd = a.queries[gd.query];
			const int n = gd.n;
			const int64_t hb = gd.hitStart;
			const int32_t* cur0 = a.cur + hb;
			const int32_t* ext0 = a.ext + hb;
			// curPos ascends in the group, so its span test needs no scan; a
			// group that fails it neither passes nor has ties that matter
			if (cur0[n - 1] - cur0[0] < p.minimum_overlap)
			{
				if (lane == 0)
				{
					a.groupPass[g] = 0;
					groupRel[g] = 0;
				}
				continue;
			}
			int uniq = 0;
			bool tie = false;
			int mnE = 0x7fffffff, mxE = -0x7fffffff - 1;
#pragma unroll 4
			for (int i = lane; i < n; i += FH_WAVE)
			{
				const int32_t c = cur0[i];
				const int32_t prev = i == 0 ? 0 : cur0[i - 1];
				uniq += (c != prev);
				tie |= i > 0 && c == prev;
				const int32_t e = ext0[i];
				mnE = min(mnE, e);
				mxE = max(mxE, e);
			}
			const bool anyTie = __ballot(tie) != 0;
			uniq = waveSum(uniq);
			mnE = waveMin(mnE);
			mxE = waveMax(mxE);
			const int32_t mnC = cur0[0];
			const int32_t mxC = cur0[n - 1];
			bool pass = !(static_cast<float>(uniq) < minUniq);
			if (pass && (mxC - mnC < p.minimum_overlap || mxE - mnE < p.minimum_overlap)) pass = false;
			if (pass && p.check_overhang && !p.force_local)
			{
				if (min(mnC, mnE) > p.maximum_overhang) pass = false;
				if (min(qd.curLen - mxC, gd.extLen - mxE) > p.maximum_overhang) pass = false;
			}
			bool rel = pass && anyTie;
			if (rel && gd.extLen > qd.curLen)
			{
				// two hits with the same extPos?
				int32_t* lo = a.posL + hb;
				int32_t* hi = a.posR + hb;
				const uint32_t m = 2u * static_cast<uint32_t>(n);
				for (int i = lane; i < n; i += FH_WAVE)
				{
					lo[i] = 0;
					hi[i] = 0;
				}
				__builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
				__builtin_amdgcn_wave_barrier();
				bool dup = false;
				for (int i = lane; i < n; i += FH_WAVE)
				{
					const int32_t key = ext0[i] + 1;
					uint32_t sl = static_cast<uint32_t>((static_cast<uint64_t>(static_cast<uint32_t>(key) * 2654435761u) * m) >> 32);
					for (;;)
					{
						int32_t expected = 0;
						const bool put = __hip_atomic_compare_exchange_strong(
							hashSlot(lo, hi, n, sl), &expected, key, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
							__HIP_MEMORY_SCOPE_WORKGROUP);
						if (put) break;
						if (expected == key)
						{
							dup = true;
							break;
						}
						sl = sl + 1 == m ? 0 : sl + 1;
					}
				}
				rel = __ballot(dup) != 0;
			}
			if (rel)
			{
				for (int i = lane; i < n; i += FH_WAVE)
				{
					const int32_t c = cur0[i];
					if ((i > 0 && cur0[i - 1] == c) || (i + 1 < n && cur0[i + 1] == c)) tieFlag[hb + i] = 1;
				}
			}
			if (lane == 0)
			{
				a.groupPass[g] = pass ? 1 : 0;
				groupRel[g] = rel ? 1 : 0;
			}
		}
	}
}

// Passing groups in processing order (stable compaction), how many of them
// hold at least kWalkSmall hits (a prefix: the order is size-descending by
// floor(log2 n)), and their hits in total.
__global__ void passFlagKernel(const int32_t* order, const int32_t* groupPass, int32_t nGroups, int32_t* flag)
{
	const int qi = blockIdx.x * blockDim.x + threadIdx.x;
	if (qi < nGroups) flag[qi] = groupPass[order[qi]];
}

__global__ void passScatterKernel(const int32_t* order, const int32_t* flag, const int64_t* pos, int32_t nGroups,
								  const GroupDesc* groups, int32_t bigMin, int32_t* passOrder, int64_t* nBig,
								  int64_t* tiers)
{
	const int qi = blockIdx.x * blockDim.x + threadIdx.x;
	bool big = false, mid = false, large = false;
	long long hits = 0;
	if (qi < nGroups && flag[qi])
	{
		const int g = order[qi];
		passOrder[pos[qi]] = g;
		hits = groups[g].n;
		big = hits >= bigMin;
		mid = hits >= kWalkMid;
		large = hits >= kWalkBig;
	}
	const int c = __popcll(__ballot(big));
	const int cm = __popcll(__ballot(mid));
	const int cl = __popcll(__ballot(large));
	hits = waveSum64(hits);
	if (laneId() == 0)
	{
		if (c) atomicAdd(reinterpret_cast<unsigned long long*>(nBig), static_cast<unsigned long long>(c));
		if (hits) atomicAdd(reinterpret_cast<unsigned long long*>(nBig + 1), static_cast<unsigned long long>(hits));
		if (tiers && cl) atomicAdd(reinterpret_cast<unsigned long long*>(tiers), static_cast<unsigned long long>(cl));
		if (tiers && cm) atomicAdd(reinterpret_cast<unsigned long long*>(tiers + 1), static_cast<unsigned long long>(cm));
	}
}

// A verify range lies inside one group's hits; groups are stored in hit
// order, so the group is the last one starting at or before the range.
__global__ void deferGroupsKernel(const SegRange* verify, int64_t nVerify, const GroupDesc* groups, int32_t nGroups,
								  int32_t* deferred)
{
	const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
	if (i >= nVerify) return;
	const SegRange w = verify[i];
	if (w.d != 1) return;
	int lo = 0, hi = nGroups - 1;
	while (lo < hi)
	{
		const int mid = lo + (hi - lo + 1) / 2;
		if (groups[mid].hitStart <= w.f) lo = mid;
		else hi = mid - 1;
	}
	deferred[lo] = 1;
}

__global__ void splitPassKernel(const int32_t* order, const int32_t* groupPass, const int32_t* deferred,
								int32_t nGroups, int32_t* flagReady, int32_t* flagDefer)
{
	const int qi = blockIdx.x * blockDim.x + threadIdx.x;
	if (qi >= nGroups) return;
	const int g = order[qi];
	const bool pass = groupPass[g] != 0, later = deferred[g] != 0;
	flagReady[qi] = pass && !later;
	flagDefer[qi] = pass && later;
}

// Group-order hit copies for the DP: hits in (extId, curPos) order, or in the
// extPos order of the sort when extLen > curLen.  The copies are cut into
// chunks of kGatherChunk hits, one wave per chunk (a group's copy is a chain
// of dependent loads per lane, so one wave per large group would leave it
// alone for milliseconds): gatherChunksKernel counts every passing group's
// chunks, the caller scans the counts, and chainGatherKernel finds each
// chunk's group by binary search over the scan.
constexpr int kGatherChunk = 4096;

__global__ void gatherChunksKernel(ChainArgs a, int32_t* chunks)
{
	const int qi = blockIdx.x * blockDim.x + threadIdx.x;
	if (qi >= a.nPass) return;
	const GroupDesc gd = a.groups[a.passOrder[qi]];
	// groups in curPos order, and extPos-sorted groups already in extPos
	// order, are read in place (groupHits)
	chunks[qi] = gatheredHits(gd, a.queries[gd.query]) ? (gd.n + kGatherChunk - 1) / kGatherChunk : 0;
}

__global__ void __launch_bounds__(FH_WAVE) chainGatherKernel(ChainArgs a, const int64_t* chunkOff)
{
	const int lane = laneId();
	const int64_t total = chunkOff[a.nPass];
	for (int64_t c = blockIdx.x; c < total; c += gridDim.x)
	{
		const int e = findSeg(chunkOff, a.nPass, c);
		const GroupDesc gd = a.groups[a.passOrder[e]];
		const int i0 = static_cast<int>(c - chunkOff[e]) * kGatherChunk;
		const int i1 = min(gd.n, i0 + kGatherChunk);
		const int64_t hb = gd.hitStart;
		const int32_t* cur0 = a.cur + hb;
		const int32_t* ext0 = a.ext + hb;
		int32_t* hc = a.hcur + hb;
		int32_t* he = a.hext + hb;
		const uint32_t* sv = a.sortVal + hb;
#pragma unroll 4
		for (int i = i0 + lane; i < i1; i += FH_WAVE)
		{
			const uint32_t src = sv[i];
			hc[i] = cur0[src];
			he[i] = ext0[src];
		}
	}
}

// ---- dp ----------------------------------------------------------------------

// ---- subgroup steps of the DP ---------------------------------------------
// The lookback tile of G lanes runs one prefix maximum and one maximum per
// step, on DPP moves: LDS permutes (__shfl) would put a round trip each on the
// tile's dependent chain.  For G <= 16 a subgroup lies inside one 16-lane DPP
// row; 32- and 64-lane subgroups add the row broadcasts (row_bcast:15 and
// row_bcast:31, as in a wave64 scan), wave_shr:1 and a readlane of the
// subgroup's last lane.

constexpr int kNegInf = -0x7fffffff - 1;

// max over lanes [row start, l] after the four row shifts, then across rows
// (wide subgroups only)
template <int G>
__device__ inline int widePrefixMax(int v)
{
	static_assert(G == 32 || G == 64, "wide subgroups");
	v = max(v, __builtin_amdgcn_update_dpp(kNegInf, v, 0x111, 0xF, 0xF, false));
	v = max(v, __builtin_amdgcn_update_dpp(kNegInf, v, 0x112, 0xF, 0xF, false));
	v = max(v, __builtin_amdgcn_update_dpp(kNegInf, v, 0x114, 0xF, 0xF, false));
	v = max(v, __builtin_amdgcn_update_dpp(kNegInf, v, 0x118, 0xF, 0xF, false));
	// rows 1 and 3 take the last lane of rows 0 and 2
	v = max(v, __builtin_amdgcn_update_dpp(kNegInf, v, 0x142, 0xA, 0xF, false));
	// rows 2 and 3 take lane 31
	if constexpr (G == 64) v = max(v, __builtin_amdgcn_update_dpp(kNegInf, v, 0x143, 0xC, 0xF, false));
	return v;
}

// Lane l of a row reads lane l - O of the same row (lanes with l < O keep
// their own value; callers ignore the lanes whose source lies in another
// subgroup).
template <int O>
__device__ inline int rowShr(int v)
{
	static_assert(O >= 1 && O <= 15, "row shift");
	return __builtin_amdgcn_update_dpp(v, v, 0x110 + O, 0xF, 0xF, false);
}

// max(v of lanes 0..sl of the subgroup)
template <int G>
__device__ inline int subPrefixMax(int v, int sl)
{
	if constexpr (G <= 16)
	{
		if constexpr (G > 1)
		{
			const int t = rowShr<1>(v);
			if (sl >= 1) v = max(v, t);
		}
		if constexpr (G > 2)
		{
			const int t = rowShr<2>(v);
			if (sl >= 2) v = max(v, t);
		}
		if constexpr (G > 4)
		{
			const int t = rowShr<4>(v);
			if (sl >= 4) v = max(v, t);
		}
		if constexpr (G > 8)
		{
			const int t = rowShr<8>(v);
			if (sl >= 8) v = max(v, t);
		}
	}
	else
	{
		v = widePrefixMax<G>(v);
	}
	return v;
}

// v of lane sl - 1 of the subgroup (undefined for sl == 0)
template <int G>
__device__ inline int subPrev(int v)
{
	if constexpr (G <= 16) return rowShr<1>(v);
	else return __builtin_amdgcn_update_dpp(v, v, 0x138, 0xF, 0xF, false);  // wave_shr:1
}

// max of v over the subgroup, in every lane
template <int G>
__device__ inline int subMax(int v)
{
	if constexpr (G <= 16)
	{
		// quad_perm [1,0,3,2] and [2,3,0,1], then the two quads of an 8-lane
		// half row (row_half_mirror), then the two halves of the row (row_mirror)
		if constexpr (G > 1) v = max(v, __builtin_amdgcn_update_dpp(v, v, 0xB1, 0xF, 0xF, false));
		if constexpr (G > 2) v = max(v, __builtin_amdgcn_update_dpp(v, v, 0x4E, 0xF, 0xF, false));
		if constexpr (G > 4) v = max(v, __builtin_amdgcn_update_dpp(v, v, 0x141, 0xF, 0xF, false));
		if constexpr (G > 8) v = max(v, __builtin_amdgcn_update_dpp(v, v, 0x140, 0xF, 0xF, false));
	}
	else
	{
		const int pre = widePrefixMax<G>(v);
		if constexpr (G == 64)
		{
			v = __builtin_amdgcn_readlane(pre, 63);
		}
		else
		{
			const int lo = __builtin_amdgcn_readlane(pre, 31), hi = __builtin_amdgcn_readlane(pre, 63);
			v = laneId() < 32 ? lo : hi;
		}
	}
	return v;
}

template <int G>
__global__ void __launch_bounds__(FH_WAVE) chainDpKernel(ChainArgs a)
{
	constexpr int kSubs = FH_WAVE / G;
	constexpr int kRing = kRingTotal / kSubs;
	static_assert(kRing >= G, "rows not yet stored (the last G) are read from the ring");
	__shared__ int ring[kSubs][kRing];   // score[j] of the last kRing rows
	__shared__ int ringM[kSubs][kRing];  // max(score[0..j]) of the same rows
	const int lane = laneId();
	const int sub = lane / G;
	const int sl = lane % G;
	const flyehip_overlap_params_t& p = a.params;
	const int32_t k = p.kmer_size;
	const int32_t maxJump = p.maximum_jump;
	const unsigned long long subMask = (G == 64 ? ~0ULL : ((1ULL << G) - 1)) << (sub * G);
	int* rg = ring[sub];
	int* rm = ringM[sub];
	// A predecessor j contributes score[j] + min(dc, de, k) - gapCost, which is
	// at most score[j] + k when the gap penalties are non-negative.  Once
	// max(score[0..j]) + k cannot beat the running maximum, no deeper
	// predecessor can improve it, so the serial scan's result is final and
	// the remaining lookback is skipped (it could neither change the maximum
	// nor trigger the improvement break).
	const bool bounded = p.large_gap_penalty >= 0.0f && p.small_gap_penalty >= 0.0f;

	// per-subgroup state (uniform inside a subgroup)
	int g = -1;
	int n = 0;
	int i = 0;
	bool extSorted = false;
	const int32_t* hc = nullptr;
	const int32_t* he = nullptr;
	int32_t* sc = nullptr;
	int32_t* bt = nullptr;
	int32_t* pmG = nullptr;
	int pm = 0;
	// rows are stored G at a time: lane (r mod G) of the subgroup holds row r's
	// score, backtrack and prefix maximum until its G-row block is complete,
	// then the block is written with one coalesced store per array (single
	// 4-byte stores per row leave L2 lines partly dirty and multiply the
	// written bytes); lookbacks within the LDS ring never read them back
	int hs = 0, hbt = -1, hpm = 0;
	// positions of the previous G rows in registers (lane sl holds row
	// i - 1 - sl: the first lookback tile of a row reads no memory), and the
	// current row's, loaded one row ahead
	int32_t winC = 0, winE = 0, curN = 0, extN = 0;
	// the same rows' scores and prefix maxima (what the LDS rings hold for
	// them): the first tile reads its scores and the bounded-break test its
	// prefix maximum without an LDS round trip
	int32_t winS = 0, winP = 0;
	bool done = false;
	int iter = 0;
	uint32_t cells = 0;  // (row, predecessor) pairs this lane evaluated
	for (;;)
	{
		// subgroups without a group pull the next passing one
		bool need = !done && (g < 0 || i >= n);
		if (g >= 0 && i >= n) g = -1;
		while (need)
		{
			// one atomic per wave for all the subgroups that need a group
			// (lanes of one atomic instruction to one address are not
			// combined: each would be an operation of its own)
			const unsigned long long leaders = __ballot(sl == 0);
			const int first = __ffsll(static_cast<long long>(leaders)) - 1;
			int base = 0;
			if (laneId() == first) base = atomicAdd(a.queue, __popcll(leaders));
			base = __shfl(base, first, FH_WAVE);
			int qi = base + __popcll(leaders & laneMaskLt());
			qi = __shfl(qi, sub * G, FH_WAVE);
			if (qi >= a.nPass)
			{
				done = true;
				break;
			}
			const int gg = a.passOrder[qi];
			const GroupDesc gd = a.groups[gg];
			g = gg;
			n = gd.n;
			i = 1;
			extSorted = gatheredHits(gd, a.queries[gd.query]);
			hc = groupHits(a.hcur, a.cur, gd.hitStart, extSorted);
			he = groupHits(a.hext, a.ext, gd.hitStart, extSorted);
			sc = a.score + gd.hitStart;
			bt = a.bt + gd.hitStart;
			pmG = a.posR + gd.hitStart;  // free during the DP: holds max(score[0..j])
			pm = 0;
			winC = hc[0];  // row 0, for lane 0 of row 1's first tile
			winE = he[0];
			winS = 0;  // row 0's ring entries
			winP = 0;
			curN = n > 1 ? hc[1] : 0;
			extN = n > 1 ? he[1] : 0;
			if (sl == 0)
			{
				rg[0] = 0;
				rm[0] = 0;
				// row 0 is stored now: with one lane per group (G == 1)
				// lane 0 takes row 1 before any block flush, so no flush
				// stores it (wider tiles store it again, unchanged)
				sc[0] = 0;
				bt[0] = -1;
				pmG[0] = 0;
			}
			hs = 0;  // row 0, held by lane 0
			hbt = -1;
			hpm = 0;
			need = false;
		}
		if (__ballot(!done) == 0) break;
		if (!done)
		{
			int carry = 0;
			int carryId = 0;
			for (int tileBase = i - 1; tileBase >= 0; tileBase -= G)
			{
				const int j = tileBase - sl;
				const bool active = j >= 0;
				cells += active ? 1u : 0u;
				int cand = -0x7fffffff - 1;
				bool special = false;
				bool stop = false;
				if (active)
				{
					const bool firstTile = tileBase == i - 1;
					int32_t cj, ej;
					if (firstTile)
					{
						cj = winC;
						ej = winE;
					}
					else
					{
						cj = hc[j];
						ej = he[j];
					}
					const int32_t dc = curN - cj;
					const int32_t de = extN - ej;
					const int32_t jd = iabs(dc - de);
					if (0 < dc && dc < maxJump && 0 < de && de < maxJump && jd <= p.max_jump_gap)
					{
						const int32_t ms = min(min(dc, de), k);
						const float pen = jd > p.gap_jump_threshold ? p.large_gap_penalty : p.small_gap_penalty;
						const int32_t gc = static_cast<int32_t>(pen * static_cast<float>(jd));
						int32_t sj;
						if (firstTile)
						{
							sj = winS;
						}
						else if (i - j < kRing)
						{
							sj = rg[j & (kRing - 1)];
						}
						else
						{
							// far predecessor, from the stored rows; the empty asm on the
							// loaded value keeps this load apart (merged with the ring
							// read it would become a flat load on every step)
							sj = sc[j];
							asm volatile("" : "+v"(sj));
						}
						cand = sj + ms - gc;
						special = (jd == 0 && dc < k);
					}
					stop = extSorted ? (de > maxJump) : (dc > maxJump);
				}
				// running maximum before this lane, in scan order
				const int incl = subPrefixMax<G>(cand, sl);
				int before = subPrev<G>(incl);
				if (sl == 0) before = -0x7fffffff - 1;
				before = max(before, carry);
				const bool improved = active && cand > before;
				const bool brk = active && ((improved && special) || stop);
				const unsigned long long bm = (__ballot(brk) & subMask) >> (sub * G);
				const int lastLane = bm ? (__ffsll(static_cast<long long>(bm)) - 1) : (G - 1);
				const int masked = (active && sl <= lastLane) ? cand : (-0x7fffffff - 1);
				const int mx = subMax<G>(masked);
				if (mx > carry)
				{
					const unsigned long long hit = (__ballot(masked == mx) & subMask) >> (sub * G);
					const int l0 = __ffsll(static_cast<long long>(hit)) - 1;
					carry = mx;
					carryId = tileBase - l0;
				}
				if (bm) break;
				const int jn = tileBase - G;
				if (bounded && jn >= 0 && tileBase == i - 1)
				{
					// after the first tile: lane G - 1's window holds the prefix
					// maximum of row jn + 1, which is at least that of row jn, so
					// the test is the same bound, never looser
					const unsigned long long bb = __ballot(sl == G - 1 && winP + k <= carry) & subMask;
					if (bb) break;
				}
				else if (bounded && jn >= 0)
				{
					int mj;
					if (i - jn < kRing)
					{
						mj = rm[jn & (kRing - 1)];
					}
					else
					{
						mj = pmG[jn];
						asm volatile("" : "+v"(mj));
					}
					if (mj + k <= carry) break;
				}
			}
			const int32_t si = max(carry, k);
			pm = max(pm, si);
			if (sl == (i & (G - 1)))
			{
				hs = si;
				hbt = carry > k ? carryId : -1;
				hpm = pm;
			}
			if (sl == 0)
			{
				rg[i & (kRing - 1)] = si;
				rm[i & (kRing - 1)] = pm;
			}
			if ((i & (G - 1)) == G - 1 || i + 1 >= n)
			{
				const int r = (i & ~(G - 1)) + sl;
				if (r <= i)
				{
					sc[r] = hs;
					bt[r] = hbt;
					pmG[r] = hpm;
				}
			}
			const int32_t nc = i + 1 < n ? hc[i + 1] : 0;
			const int32_t ne = i + 1 < n ? he[i + 1] : 0;
			winC = subPrev<G>(winC);
			winE = subPrev<G>(winE);
			winS = subPrev<G>(winS);
			winP = subPrev<G>(winP);
			if (sl == 0)
			{
				winC = curN;
				winE = extN;
				winS = si;
				winP = pm;
			}
			curN = nc;
			extN = ne;
			++i;
		}
		// The ring write of this row is read by the next row's lanes: LDS
		// operations of one wave execute in order, so only the compiler must
		// not move the next row's loads above it (wave barrier).  Stores to
		// sc[] reach far lookbacks (more than kRing rows back) through the
		// periodic fence, which waits for them.
		++iter;
		if ((iter & 31) == 0) __threadfence_block();
		__builtin_amdgcn_wave_barrier();
	}
	if (a.dpCells)
	{
		const long long total = waveSum64(static_cast<long long>(cells));
		if (lane == 0) atomicAdd(a.dpCells, static_cast<unsigned long long>(total));
	}
}

// ---- finish ----------------------------------------------------------------------


// The score sort's input for the groups of passOrder[0, nPass) (one wave per
// group): std::sort(indices, score desc) as ascending keys, the row index as
// value with the relevance bit on rows that have a backtrack pointer (the
// extPos sort's keys and permutation are dead once the hits are gathered).
__global__ void __launch_bounds__(FH_WAVE) chainScoreKeysKernel(ChainArgs a)
{
	const int qi = blockIdx.x;
	if (qi >= a.nPass) return;
	const GroupDesc gd = a.groups[a.passOrder[qi]];
	const int64_t hb = gd.hitStart;
	const int32_t* sc = a.score + hb;
	const int32_t* bt = a.bt + hb;
	uint32_t* sk = a.sortKey + hb;
	uint32_t* sv = a.sortVal + hb;
	for (int i = laneId(); i < gd.n; i += FH_WAVE)
	{
		sk[i] = scoreOrderKey(sc[i]);
		sv[i] = static_cast<uint32_t>(i) | (bt[i] != -1 ? kWalkRelevant : 0u);
	}
}

// Pushes every passing group's range to the score sort (one lane per group).
__global__ void chainScorePushKernel(ChainArgs a, SegSortBufs sb)
{
	const int qi = blockIdx.x * blockDim.x + threadIdx.x;
	const bool push = qi < a.nPass;
	int32_t f = 0, n = 0;
	if (push)
	{
		const GroupDesc gd = a.groups[a.passOrder[qi]];
		f = static_cast<int32_t>(gd.hitStart);
		n = gd.n;
	}
	segPushWave(sb, sb.big[0], push, f, f + n, exsort::depthLimit(static_cast<uint32_t>(n)), 0);
}

constexpr int kPropTile = 256;  // proposals sorted in LDS up to this count

// Without onlyMaxExt: std::sort(extOverlaps, score desc) of the np proposals
// props[0, np) (listed in walk order), the containment filter and the
// divergence filter; the primaries go to outp.  sk / sv / posL / posR are the
// group's per-hit scratch.  Returns the row count in every lane.
__device__ int selectPrimaries(const Proposal* props, int np, uint32_t* sk, uint32_t* sv, int32_t* posL,
							   int32_t* posR, Proposal* outp, const flyehip_overlap_params_t& p,
							   WaveStdSortLds<uint32_t, kPropTile>& lds, int* red)
{
	const int lane = laneId();
	for (int i = lane; i < np; i += FH_WAVE)
	{
		sk[i] = static_cast<uint32_t>(0x7fffffff - props[i].score);
		sv[i] = static_cast<uint32_t>(i);
	}
	waveSync();
	waveStdSort<uint32_t, kPropTile, kLaneMax>(exsort::Span<uint32_t>{sk, sv}, np, posL, posR, lds);
	if (lane == 0)
	{
		int32_t* prim = posL;  // free again after the sort
		int nprim = 0;
		for (int q = 0; q < np; ++q)
		{
			const Proposal& o = props[sv[q]];
			bool contained = false;
			for (int r = 0; r < nprim; ++r)
			{
				const Proposal& pr = props[prim[r]];
				if (pr.curBegin <= o.curBegin && o.curEnd <= pr.curEnd && pr.extBegin <= o.extBegin &&
					o.extEnd <= pr.extEnd && pr.score > o.score)
				{
					contained = true;
					break;
				}
			}
			if (!contained) prim[nprim++] = static_cast<int32_t>(sv[q]);
		}
		int w = 0;
		for (int r = 0; r < nprim; ++r)
			if (props[prim[r]].div < p.max_divergence) outp[w++] = props[prim[r]];
		red[0] = w;
	}
	waveSync();
	return red[0];
}

// The proposal of a walk that starts at row `start`, reaches row f and
// passes len rows: positions, overlapTest, the filtered-position count and
// the divergence.  Returns whether the proposal survives.
__device__ inline bool makeProposal(Proposal& o, int start, int f, int len, const int32_t* hc, const int32_t* he,
									const int32_t* sc, const QueryDesc& qd, const GroupDesc& gd,
									const int32_t* filt, int filtN, const ChainArgs& a)
{
	const flyehip_overlap_params_t& p = a.params;
	const int32_t k = p.kmer_size;
	o.curBegin = hc[f];
	o.curEnd = hc[start] + k - 1;
	o.extBegin = he[f];
	o.extEnd = he[start] + k - 1;
	o.score = sc[start] - sc[f] + k - 1;
	o.chainLength = len;
	o.div = 0.0f;
	if (!overlapTestDev(o, qd.curId, gd.extId, qd.curLen, gd.extLen, p)) return false;
	const int fc = upperBound(filt, filtN, o.curEnd) - lowerBound(filt, filtN, o.curBegin);
	const float normLen = static_cast<float>(max(o.curEnd - o.curBegin, o.extEnd - o.extBegin) - fc);
	if (!(normLen > 0.0f)) return false;
	float rate = static_cast<float>(len) * p.sample_rate / normLen;
	rate = fminf(rate, 1.0f);
	if (!(rate > 0.0f)) return false;
	const float inv = 1.0f / rate;
	o.div = flyeLogf(inv, a.logX, a.logY, a.logN) / static_cast<float>(k);
	return true;
}

// Lane 0's walks for the live starts m of one round (start rows in
// recStart[lane]): each walk follows the backtrack pointers from its start,
// clearing them, and ends on the first row without one (a chain start, or a
// row an earlier walk passed); (start, first row, length) are recorded in
// walk order.  One loop per pointer storage (16-bit LDS stage, or the rows
// in place), so each step is one load and one store.
template <class T>
__device__ inline int walkStarts(T* bt, unsigned long long m, int* recStart, int* recFirst, int* recLen)
{
	int nrec = 0;
	while (m)
	{
		const int l = __ffsll(static_cast<long long>(m)) - 1;
		m &= m - 1;
		const int start = recStart[l];
		if (bt[start] == -1) continue;
		int f = 0, len = 0, pos = start;
		while (pos != -1)
		{
			f = pos;
			++len;
			const int nxt = bt[pos];
			bt[pos] = -1;
			pos = nxt;
		}
		recFirst[nrec] = f;
		recLen[nrec] = len;
		recStart[nrec] = start;  // slots below l are no longer read
		++nrec;
	}
	return nrec;
}

// Backtrack, overlapTest, divergence and selection of one group per wave,
// groups passOrder[first, last).  The rows are visited in score order 64 at
// a time: the lanes test which starts still have a backtrack pointer, lane 0
// walks those in order through the LDS copy of the pointers (a walk clears
// the pointers it passes, so lane 0 re-checks each start) and records
// (start, first, length); then every lane evaluates one recorded walk
// (positions, overlapTest, the filtered-position count, logf) and the
// surviving proposals are appended in walk order.  kOnlyMax (the onlyMaxExt
// parameter) keeps the best proposal only; the variant without it also holds
// the proposal sort, whose LDS and registers would otherwise cut the
// occupancy of the onlyMaxExt walk.
template <int kBt, bool kOnlyMax>
__global__ void __launch_bounds__(FH_WAVE) chainWalkKernel(ChainArgs a, int32_t first, int32_t last)
{
	// backtrack pointers staged as 16-bit row indices (rows of a staged group
	// are below kBt <= 32768; -1 stays -1)
	static_assert(kBt <= 32768, "staged rows must fit int16_t");
	__shared__ int16_t btS[kBt];
	__shared__ int recStart[FH_WAVE];
	__shared__ int recFirst[FH_WAVE];
	__shared__ int recLen[FH_WAVE];
	__shared__ int red[2];
	const int lane = laneId();
	const flyehip_overlap_params_t& p = a.params;
	for (;;)
	{
		const int qi = first + pullGroup(a.queue);
		if (qi >= last) break;
		const int g = a.passOrder[qi];
		const GroupDesc gd = a.groups[g];
		const QueryDesc qd = a.queries[gd.query];
		// wave-uniform (scalar) size: the staging choice below is a scalar
		// branch, not an exec mask around every step of lane 0's walk
		const int n = __builtin_amdgcn_readfirstlane(gd.n);
		const int64_t hb = gd.hitStart;
		const bool extSorted = gatheredHits(gd, qd);
		const int32_t* hc = groupHits(a.hcur, a.cur, hb, extSorted);
		const int32_t* he = groupHits(a.hext, a.ext, hb, extSorted);
		const int32_t* sc = a.score + hb;
		uint32_t* sk = a.sortKey + hb;
		uint32_t* sv = a.sortVal + hb;
		const bool inLds = n <= kBt;
		int32_t* btG = a.bt + hb;
		auto btGet = [&](int r) -> int { return inLds ? static_cast<int>(btS[r]) : btG[r]; };
		if (inLds)
		{
#pragma unroll 4
			for (int i = lane; i < n; i += FH_WAVE) btS[i] = static_cast<int16_t>(btG[i]);
		}
		waveSync();

		// onlyMaxExt keeps one proposal per group (outProps indexed by group;
		// props unused); otherwise every chain of the group may be listed
		Proposal* props = a.props + hb;
		Proposal* outp = a.outProps + (kOnlyMax ? static_cast<int64_t>(g) : hb);
		const int32_t* filt = a.filtered + qd.filtStart;
		const int filtN = qd.filtCount;
		int np = 0;
		int bestScore = 0;
		bool haveBest = false;
		for (int q0 = 0; q0 < n; q0 += FH_WAVE)
		{
			const int q = q0 + lane;
			const int myStart = q < n ? static_cast<int>(sv[q] & ~kWalkRelevant) : 0;
			const bool live = q < n && btGet(myStart) != -1;
			const unsigned long long m = __ballot(live);
			recStart[lane] = myStart;
			waveSync();
			if (lane == 0)
				red[0] = inLds ? walkStarts(btS, m, recStart, recFirst, recLen)
							   : walkStarts(btG, m, recStart, recFirst, recLen);
			waveSync();
			const int nrec = red[0];
			bool ok = false;
			Proposal o;
			if (lane < nrec)
				ok = makeProposal(o, recStart[lane], recFirst[lane], recLen[lane], hc, he, sc, qd, gd, filt, filtN,
								  a);
			const unsigned long long okm = __ballot(ok);
			if constexpr (kOnlyMax)
			{
				// first strict maximum in walk order
				const int mx = waveMax(ok ? o.score : (-0x7fffffff - 1));
				if (okm && (!haveBest || mx > bestScore))
				{
					const int l = __ffsll(static_cast<long long>(__ballot(ok && o.score == mx))) - 1;
					if (lane == l) outp[0] = o;
					bestScore = mx;
					haveBest = true;
				}
			}
			else if (ok)
			{
				props[np + __popcll(okm & laneMaskLt())] = o;
			}
			np += __popcll(okm);
			waveSync();
		}
		int count = 0;
		if constexpr (kOnlyMax)
		{
			waveSync();
			count = (haveBest && outp[0].div < p.max_divergence) ? 1 : 0;
		}
		else
		{
			__shared__ WaveStdSortLds<uint32_t, kPropTile> lds;
			count = selectPrimaries(props, np, sk, sv, a.posL + hb, a.posR + hb, outp, p, lds, &red[1]);
		}
		if (lane == 0) a.groupCount[g] = count;
		waveSync();
	}
}

// ---- claim walk ------------------------------------------------------------
//
// The walk visits the rows with a backtrack pointer ("relevant" rows; every
// other row is a chain root) in std::sort's score order, follows each start's
// pointers and clears them, so the order only matters where two of those rows
// share a score.  Let D(v) be the largest relevant score in v's subtree of the
// backtrack forest (v included).  Walks run towards lower rows, so a row is
// first reached by the highest-priority relevant row of its subtree: when no
// two relevant rows of equal score can compete (below), v is claimed by the
// row whose score is D(v), a start survives exactly when no descendant
// scores at least as high, and its chain is the claimed path up to its
// lowest row t plus t's parent (reached earlier by a higher walk) unless t is
// a root.  Ties can change the result only between "weakly alive" rows (no
// strictly higher descendant): such a row with an equal-scoring descendant,
// or two surviving starts of equal score, mark the group for the exact
// std::sort path.  Nothing here needs the order, so the score sort is skipped.
//
// D is computed one 64-row chunk at a time from the top: the chunk's rows
// take their own score and what children in higher chunks pushed (acc, in
// the group's free per-hit scratch), pointer doubling over the in-chunk
// parents spreads the maxima up in six LDS rounds, and rows whose parent lies
// below the chunk push their D there.  acc is read and written with
// workgroup-scope atomics (the group's one wave is the workgroup), ordered by
// a workgroup fence between chunks.
constexpr int kClaimMaxAlive = FH_WAVE;  // surviving starts a group may hold (one proposal per lane)

__device__ inline uint32_t accLoad(uint32_t* p)
{
	return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ inline void accMax(uint32_t* p, uint32_t v)
{
	(void)__hip_atomic_fetch_max(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// One wave per group of passOrder[0, nPass); groups it cannot decide get
// claimFlag[g] = 1 (and no rows), the others groupCount and their proposals.
template <bool kOnlyMax>
__global__ void __launch_bounds__(FH_WAVE) chainClaimKernel(ChainArgs a, int32_t* claimFlag)
{
	__shared__ uint32_t val[FH_WAVE];   // D of the chunk's rows (doubling)
	__shared__ int jmp[FH_WAVE];        // 2^r-th in-chunk ancestor, -1 once outside
	__shared__ uint32_t cmax[FH_WAVE];  // max D over a row's children
	__shared__ int aIdx[kClaimMaxAlive];
	__shared__ int aScore[kClaimMaxAlive];
	__shared__ int aCnt[kClaimMaxAlive];
	__shared__ int aTop[kClaimMaxAlive];
	__shared__ int red[2];
	const int lane = laneId();
	const flyehip_overlap_params_t& p = a.params;
	for (;;)
	{
		const int qi = pullGroup(a.queue);
		if (qi >= a.nPass) break;
		const int g = a.passOrder[qi];
		const GroupDesc gd = a.groups[g];
		const QueryDesc qd = a.queries[gd.query];
		const int n = __builtin_amdgcn_readfirstlane(gd.n);
		const int64_t hb = gd.hitStart;
		const bool extSorted = gatheredHits(gd, qd);
		const int32_t* hc = groupHits(a.hcur, a.cur, hb, extSorted);
		const int32_t* he = groupHits(a.hext, a.ext, hb, extSorted);
		const int32_t* sc = a.score + hb;
		const int32_t* bt = a.bt + hb;
		uint32_t* acc = reinterpret_cast<uint32_t*>(a.posL + hb);
		// a group of one chunk has no row below it: no accumulators
		const bool oneChunk = n <= FH_WAVE;
		if (!oneChunk)
		{
			for (int i = lane; i < n; i += FH_WAVE) acc[i] = 0u;
			__builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
		}
		int nAlive = 0;
		bool flagged = false;
		for (int t = (n - 1) & ~(FH_WAVE - 1); t >= 0 && !flagged; t -= FH_WAVE)
		{
			const int v = t + lane;
			const bool valid = v < n;
			int b = -1, s = 0;
			uint32_t a0 = 0;
			if (valid)
			{
				b = bt[v];
				s = sc[v];
				if (!oneChunk) a0 = accLoad(acc + v);
			}
			const bool rel = b != -1;
			const int j0 = (rel && b >= t) ? b - t : -1;
			val[lane] = max(rel ? static_cast<uint32_t>(s) : 0u, a0);
			jmp[lane] = j0;
			cmax[lane] = a0;
			waveSync();
			int jj = j0;
			for (int r = 0; r < 6; ++r)
			{
				if (__ballot(jj >= 0) == 0) break;
				const uint32_t mine = val[lane];
				if (jj >= 0) atomicMax(&val[jj], mine);
				const int nj = jj >= 0 ? jmp[jj] : -1;
				__builtin_amdgcn_wave_barrier();
				jmp[lane] = nj;
				jj = nj;
				__builtin_amdgcn_wave_barrier();
			}
			waveSync();
			const uint32_t d = valid ? val[lane] : 0u;
			if (j0 >= 0) atomicMax(&cmax[j0], d);
			if (rel && b < t) accMax(acc + b, d);
			waveSync();
			const uint32_t dex = cmax[lane];
			// surviving starts (registered before the chunk's claims are counted)
			const bool tie = rel && dex == static_cast<uint32_t>(s);
			const bool alive = rel && dex < static_cast<uint32_t>(s);
			const unsigned long long am = __ballot(alive);
			const int na = __popcll(am);
			if (__ballot(tie) != 0 || nAlive + na > kClaimMaxAlive)
			{
				flagged = true;
				break;
			}
			if (alive)
			{
				const int slot = nAlive + __popcll(am & laneMaskLt());
				aIdx[slot] = v;
				aScore[slot] = s;
				aCnt[slot] = 0;
				aTop[slot] = v;
			}
			nAlive += na;
			waveSync();
			// claims: every row with a relevant row in its subtree, grouped by claimer
			const bool claimed = valid && d != 0u;
			unsigned long long m = __ballot(claimed);
			while (m)
			{
				const int l0 = __ffsll(static_cast<long long>(m)) - 1;
				const uint32_t c = __shfl(d, l0, FH_WAVE);
				const unsigned long long mc = __ballot(claimed && d == c);
				const unsigned long long hm = __ballot(lane < nAlive && static_cast<uint32_t>(aScore[lane]) == c);
				if (hm == 0)
				{
					flagged = true;
					break;
				}
				const int slot = __ffsll(static_cast<long long>(hm)) - 1;
				if (lane == 0)
				{
					aCnt[slot] += __popcll(mc);
					aTop[slot] = t + __ffsll(static_cast<long long>(mc)) - 1;
				}
				m &= ~mc;
				waveSync();
			}
			if (!oneChunk) __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
		}
		// two surviving starts of equal score
		if (!flagged && nAlive > 1)
		{
			bool dup = false;
			if (lane < nAlive)
				for (int r = lane + 1; r < nAlive; ++r) dup |= aScore[r] == aScore[lane];
			flagged = __ballot(dup) != 0;
		}
		if (flagged)
		{
			if (lane == 0) claimFlag[g] = 1;
			continue;
		}
		// one proposal per surviving start; the walk order is score-descending
		const int32_t* filt = a.filtered + qd.filtStart;
		const int filtN = qd.filtCount;
		bool ok = false;
		Proposal o;
		int start = 0;
		if (lane < nAlive)
		{
			start = aIdx[lane];
			const int top = aTop[lane];
			const int pt = bt[top];
			const int f = pt != -1 ? pt : top;
			const int len = aCnt[lane] + (pt != -1 ? 1 : 0);
			ok = makeProposal(o, start, f, len, hc, he, sc, qd, gd, filt, filtN, a);
		}
		const int startScore = ok ? sc[start] : -1;
		int count = 0;
		if constexpr (kOnlyMax)
		{
			// the first proposal of maximal score in walk order: the highest
			// start score among the maximal ones
			const unsigned long long key =
				ok ? ((static_cast<unsigned long long>(static_cast<uint32_t>(o.score) ^ 0x80000000u) << 32) |
					  static_cast<uint32_t>(startScore))
				   : 0ull;
			unsigned long long best = key;
			for (int off = 32; off > 0; off >>= 1)
			{
				const unsigned long long other = __shfl_xor(best, off, FH_WAVE);
				best = other > best ? other : best;
			}
			const unsigned long long bm = __ballot(ok && key == best);
			if (bm)
			{
				const int l = __ffsll(static_cast<long long>(bm)) - 1;
				if (lane == l)
				{
					a.outProps[g] = o;
					red[0] = o.div < p.max_divergence ? 1 : 0;
				}
				waveSync();
				count = red[0];
			}
		}
		else
		{
			// proposals in walk order (start scores are distinct here)
			aScore[lane] = startScore;
			waveSync();
			int rank = 0;
			if (ok)
				for (int r = 0; r < nAlive; ++r) rank += aScore[r] > startScore ? 1 : 0;
			const int np = __popcll(__ballot(ok));
			Proposal* props = a.props + hb;
			if (ok) props[rank] = o;
			waveSync();
			__shared__ WaveStdSortLds<uint32_t, kPropTile> lds;
			count = selectPrimaries(props, np, a.sortKey + hb, a.sortVal + hb, a.posL + hb, a.posR + hb,
									a.outProps + hb, p, lds, &red[1]);
		}
		if (lane == 0)
		{
			a.groupCount[g] = count;
			claimFlag[g] = 0;
		}
		waveSync();
	}
}

// Writes rows: row offsets come from an exclusive scan of groupCount in group
// order; groups are in (query, internal extId) order so rows land in Flye's
// detectedOverlaps order.
__global__ void writeRowsKernel(ChainArgs a, const int64_t* rowStart, flyehip_overlap_t* rows)
{
	const int g = blockIdx.x * blockDim.x + threadIdx.x;
	if (g >= a.nGroups) return;
	const int c = a.groupCount[g];
	if (c == 0) return;
	const GroupDesc gd = a.groups[g];
	const QueryDesc qd = a.queries[gd.query];
	const Proposal* props = a.outProps + (a.params.only_max_ext ? static_cast<int64_t>(g) : gd.hitStart);
	const int64_t o = rowStart[g];
	for (int r = 0; r < c; ++r)
	{
		flyehip_overlap_t row;
		row.cur_id = qd.curId;
		row.ext_id = gd.extId;
		row.cur_begin = props[r].curBegin;
		row.cur_end = props[r].curEnd;
		row.cur_len = qd.curLen;
		row.ext_begin = props[r].extBegin;
		row.ext_end = props[r].extEnd;
		row.ext_len = gd.extLen;
		row.score = props[r].score;
		row.seq_divergence = props[r].div;
		row.chain_length = props[r].chainLength;
		row.reserved = 0;
		rows[o + r] = row;
	}
}

// Size-descending processing order: groups binned by floor(log2 n) (the order
// inside a bin is irrelevant to the results).  Block-local LDS histograms keep
// the global atomics to one per bin and block.
__global__ void binCountKernel(const GroupDesc* groups, int32_t nGroups, int32_t* binCount)
{
	__shared__ int h[33];
	if (threadIdx.x < 33) h[threadIdx.x] = 0;
	__syncthreads();
	const int g = blockIdx.x * blockDim.x + threadIdx.x;
	if (g < nGroups) atomicAdd(&h[__clz(max(1, groups[g].n))], 1);  // 31 - floor(log2 n): big groups first
	__syncthreads();
	if (threadIdx.x < 33 && h[threadIdx.x]) atomicAdd(&binCount[threadIdx.x], h[threadIdx.x]);
}

__global__ void binScatterKernel(const GroupDesc* groups, int32_t nGroups, int32_t* binCursor, int32_t* order)
{
	__shared__ int h[33];
	__shared__ int base[33];
	if (threadIdx.x < 33) h[threadIdx.x] = 0;
	__syncthreads();
	const int g = blockIdx.x * blockDim.x + threadIdx.x;
	int bin = -1, local = 0;
	if (g < nGroups)
	{
		bin = __clz(max(1, groups[g].n));
		local = atomicAdd(&h[bin], 1);
	}
	__syncthreads();
	if (threadIdx.x < 33 && h[threadIdx.x]) base[threadIdx.x] = atomicAdd(&binCursor[threadIdx.x], h[threadIdx.x]);
	__syncthreads();
	if (bin >= 0) order[base[bin] + local] = g;
}

__global__ void binPrefixKernel(int32_t* binCount, int32_t* binCursor)
{
	if (threadIdx.x != 0) return;
	int s = 0;
	for (int b = 0; b < 33; ++b)
	{
		binCursor[b] = s;
		s += binCount[b];
	}
}

__global__ void logfKernel(const float* x, int64_t n, float* out, const uint32_t* tabX, const uint32_t* tabY,
						   int tabN)
{
	const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
	if (i < n) out[i] = flyeLogf(x[i], tabX, tabY, tabN);
}

} // namespace

hipError_t launchLogf(const float* x, int64_t n, float* out, const uint32_t* tabX, const uint32_t* tabY, int tabN,
					  hipStream_t stream)
{
	if (n == 0) return hipSuccess;
	hipLaunchKernelGGL(logfKernel, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, stream, x, n, out, tabX,
					   tabY, tabN);
	return hipGetLastError();
}

hipError_t launchGroupOrder(const GroupDesc* groups, int32_t nGroups, int32_t* bins, int32_t* order,
							hipStream_t stream)
{
	if (nGroups == 0) return hipSuccess;
	const int tpb = 256;
	const dim3 grid((nGroups + tpb - 1) / tpb);
	hipLaunchKernelGGL(binCountKernel, grid, dim3(tpb), 0, stream, groups, nGroups, bins);
	hipLaunchKernelGGL(binPrefixKernel, dim3(1), dim3(64), 0, stream, bins, bins + 40);
	hipLaunchKernelGGL(binScatterKernel, grid, dim3(tpb), 0, stream, groups, nGroups, bins + 40, order);
	return hipGetLastError();
}

hipError_t launchChainRelevance(const ChainArgs& a, int32_t* groupRel, int32_t* tieFlag, int blocks,
								hipStream_t stream)
{
	if (a.nGroups <= 0) return hipSuccess;
	const unsigned grid = static_cast<unsigned>(min(blocks, (a.nGroups + FH_WAVE - 1) / FH_WAVE));
	hipLaunchKernelGGL(chainRelevanceKernel, dim3(grid), dim3(FH_WAVE), 0, stream, a, groupRel, tieFlag);
	return hipGetLastError();
}

hipError_t launchChainPrefilter(const ChainArgs& a, const SegSortBufs& sb, int blocks, hipStream_t stream)
{
	hipLaunchKernelGGL(chainPrefilterKernel, dim3(blocks), dim3(FH_WAVE), 0, stream, a, sb);
	return hipGetLastError();
}

hipError_t launchPassFlags(const ChainArgs& a, int32_t* flag, hipStream_t stream)
{
	if (a.nGroups == 0) return hipSuccess;
	hipLaunchKernelGGL(passFlagKernel, dim3((a.nGroups + 255) / 256), dim3(256), 0, stream, a.order, a.groupPass,
					   a.nGroups, flag);
	return hipGetLastError();
}

hipError_t launchPassScatter(const ChainArgs& a, const int32_t* flag, const int64_t* pos, int32_t* passOrder,
							 int64_t* nBig, hipStream_t stream, int64_t* tiers)
{
	if (a.nGroups == 0) return hipSuccess;
	hipLaunchKernelGGL(passScatterKernel, dim3((a.nGroups + 255) / 256), dim3(256), 0, stream, a.order, flag, pos,
					   a.nGroups, a.groups, kWalkSmall, passOrder, nBig, tiers);
	return hipGetLastError();
}

hipError_t launchDeferGroups(const SegRange* verify, int64_t nVerify, const ChainArgs& a, int32_t* deferred,
							 hipStream_t stream)
{
	if (nVerify == 0 || a.nGroups == 0) return hipSuccess;
	hipLaunchKernelGGL(deferGroupsKernel, dim3(static_cast<unsigned>((nVerify + 255) / 256)), dim3(256), 0, stream,
					   verify, nVerify, a.groups, a.nGroups, deferred);
	return hipGetLastError();
}

hipError_t launchSplitPass(const ChainArgs& a, const int32_t* deferred, int32_t* flagReady, int32_t* flagDefer,
						   hipStream_t stream)
{
	if (a.nGroups == 0) return hipSuccess;
	hipLaunchKernelGGL(splitPassKernel, dim3((a.nGroups + 255) / 256), dim3(256), 0, stream, a.order, a.groupPass,
					   deferred, a.nGroups, flagReady, flagDefer);
	return hipGetLastError();
}

hipError_t launchGatherChunks(const ChainArgs& a, int32_t* chunks, hipStream_t stream)
{
	if (a.nPass <= 0) return hipSuccess;
	hipLaunchKernelGGL(gatherChunksKernel, dim3((a.nPass + 255) / 256), dim3(256), 0, stream, a, chunks);
	return hipGetLastError();
}

hipError_t launchChainGather(const ChainArgs& a, const int64_t* chunkOff, int blocks, hipStream_t stream)
{
	if (a.nPass <= 0) return hipSuccess;
	hipLaunchKernelGGL(chainGatherKernel, dim3(blocks), dim3(FH_WAVE), 0, stream, a, chunkOff);
	return hipGetLastError();
}

hipError_t launchChainScorePush(const ChainArgs& a, const SegSortBufs& sb, hipStream_t stream)
{
	const int tpb = 256;
	if (a.nPass > 0)
	{
		hipLaunchKernelGGL(chainScoreKeysKernel, dim3(static_cast<unsigned>(a.nPass)), dim3(FH_WAVE), 0, stream, a);
		hipLaunchKernelGGL(chainScorePushKernel, dim3((a.nPass + tpb - 1) / tpb), dim3(tpb), 0, stream, a, sb);
	}
	return hipGetLastError();
}

hipError_t launchChainDp(const ChainArgs& a, int lanes, int blocks, hipStream_t stream)
{
	switch (lanes)
	{
	case 1: hipLaunchKernelGGL(chainDpKernel<1>, dim3(blocks), dim3(FH_WAVE), 0, stream, a); break;
	case 4: hipLaunchKernelGGL(chainDpKernel<4>, dim3(blocks), dim3(FH_WAVE), 0, stream, a); break;
	case 8: hipLaunchKernelGGL(chainDpKernel<8>, dim3(blocks), dim3(FH_WAVE), 0, stream, a); break;
	case 16: hipLaunchKernelGGL(chainDpKernel<16>, dim3(blocks), dim3(FH_WAVE), 0, stream, a); break;
	case 32: hipLaunchKernelGGL(chainDpKernel<32>, dim3(blocks), dim3(FH_WAVE), 0, stream, a); break;
	case 64: hipLaunchKernelGGL(chainDpKernel<64>, dim3(blocks), dim3(FH_WAVE), 0, stream, a); break;
	default: return hipErrorInvalidValue;
	}
	return hipGetLastError();
}

template <bool kOnlyMax>
void launchWalkTier(const ChainArgs& a, int32_t first, int32_t last, int tier, unsigned grid, hipStream_t stream)
{
	switch (tier)
	{
	case 3:
		hipLaunchKernelGGL((chainWalkKernel<kWalkHuge, kOnlyMax>), dim3(grid), dim3(FH_WAVE), 0, stream, a, first,
						   last);
		break;
	case 2:
		hipLaunchKernelGGL((chainWalkKernel<kWalkBig, kOnlyMax>), dim3(grid), dim3(FH_WAVE), 0, stream, a, first,
						   last);
		break;
	case 1:
		hipLaunchKernelGGL((chainWalkKernel<kWalkMid, kOnlyMax>), dim3(grid), dim3(FH_WAVE), 0, stream, a, first,
						   last);
		break;
	default:
		hipLaunchKernelGGL((chainWalkKernel<kWalkSmall, kOnlyMax>), dim3(grid), dim3(FH_WAVE), 0, stream, a, first,
						   last);
	}
}

hipError_t launchChainWalk(const ChainArgs& a, int32_t first, int32_t last, int tier, int blocks,
						   hipStream_t stream)
{
	if (last <= first) return hipSuccess;
	// no more waves than groups in the class (the large classes hold few)
	const unsigned grid = static_cast<unsigned>(min(blocks, last - first));
	if (a.params.only_max_ext) launchWalkTier<true>(a, first, last, tier, grid, stream);
	else launchWalkTier<false>(a, first, last, tier, grid, stream);
	return hipGetLastError();
}

hipError_t launchChainClaim(const ChainArgs& a, int32_t* claimFlag, int blocks, hipStream_t stream)
{
	if (a.nPass <= 0) return hipSuccess;
	const unsigned grid = static_cast<unsigned>(min(blocks, a.nPass));
	if (a.params.only_max_ext)
		hipLaunchKernelGGL(chainClaimKernel<true>, dim3(grid), dim3(FH_WAVE), 0, stream, a, claimFlag);
	else
		hipLaunchKernelGGL(chainClaimKernel<false>, dim3(grid), dim3(FH_WAVE), 0, stream, a, claimFlag);
	return hipGetLastError();
}

hipError_t launchWriteRows(const ChainArgs& a, const int64_t* rowStart, flyehip_overlap_t* rows,
						   hipStream_t stream)
{
	if (a.nGroups == 0) return hipSuccess;
	const int tpb = 256;
	hipLaunchKernelGGL(writeRowsKernel, dim3((a.nGroups + tpb - 1) / tpb), dim3(tpb), 0, stream, a, rowStart,
					   rows);
	return hipGetLastError();
}

} // namespace fh
