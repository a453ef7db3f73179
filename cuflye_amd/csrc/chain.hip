// chain.hip - per-(query, target) group chaining on gfx950: prefilters, the
// extPos re-sort, the bounded-lookback chaining DP, the score-order sort, the
// backtrack, overlapTest, divergence, primary selection and the divergence
// filter of OverlapDetector::getSeqOverlaps (restated in
// tools/replay_overlap_chains.py:228-380 and
// tools/replay_read_to_graph_source_pack.py:352-491).
//
// Work distribution: a persistent kernel; each 64-lane block pulls the next
// group from a device work queue (an atomic counter over groups sorted by
// decreasing size, so the heavy tail starts first).  Inside a group:
//   * prefilters are wave reductions;
//   * the two sorts are the exact std::sort emulation of exsort_dev.hpp;
//   * DP row i runs its lookback 64 predecessors per step across the lanes:
//     each lane scores one j, a prefix max over lanes reproduces the serial
//     "strictly greater" running maximum, and a ballot finds the first lane
//     where the serial loop would have broken (an improvement with
//     jumpDiv == 0 && dc < k, or the sorted-axis delta passing maxJump);
//     scores of the last kRing rows live in LDS;
//   * the backtrack and filters run on lane 0 (O(n) per group).
#include "chain.hpp"
#include "exsort_dev.hpp"

namespace fh
{

namespace
{

constexpr int kNT = 64;
constexpr int kRing = 2048;
constexpr int kLaneSortMax = 48;

struct ChainShared
{
	int ring[kRing];
	SortShared<kNT> sort;
	int red[8];
	int group;
};

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
	// 0.5 * min(cr, er) compared in double exactly as the int operands allow
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

__global__ void __launch_bounds__(kNT) chainGroupsKernel(ChainArgs a)
{
	__shared__ ChainShared sh;
	const int lane = static_cast<int>(threadIdx.x);
	const flyehip_overlap_params_t& p = a.params;
	const int32_t k = p.kmer_size;
	for (;;)
	{
		if (lane == 0) sh.group = atomicAdd(a.queue, 1);
		__syncthreads();
		const int qi = sh.group;
		__syncthreads();
		if (qi >= a.nGroups) break;
		const int g = a.order ? a.order[qi] : qi;
		const GroupDesc gd = a.groups[g];
		const QueryDesc qd = a.queries[gd.query];
		const int n = gd.n;
		const int64_t base = gd.hitStart;
		const int32_t* cur0 = a.cur + base;
		const int32_t* ext0 = a.ext + base;
		int32_t outCount = 0;

		// ---- prefilters (replay_overlap_chains.py:245-270) ----
		int uniq = 0;
		int mnE = 0x7fffffff, mxE = -0x7fffffff - 1;
		for (int i = lane; i < n; i += kNT)
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
		bool pass = !(static_cast<float>(uniq) < p.min_kmer_survival_rate * static_cast<float>(p.minimum_overlap));
		if (pass && (mxC - mnC < p.minimum_overlap || mxE - mnE < p.minimum_overlap)) pass = false;
		if (pass && p.check_overhang && !p.force_local)
		{
			if (min(mnC, mnE) > p.maximum_overhang) pass = false;
			if (min(qd.curLen - mxC, gd.extLen - mxE) > p.maximum_overhang) pass = false;
		}
		if (!pass || n < 2)
		{
			if (lane == 0) a.groupCount[g] = 0;
			continue;
		}

		uint64_t* sk = a.sortKey + base;
		uint32_t* sv = a.sortVal + base;
		int* posL = a.posL + base;
		int* posR = a.posR + base;
		Seg* small = a.small + base;
		int32_t* hc = a.hcur + base;
		int32_t* he = a.hext + base;
		int32_t* sc = a.score + base;
		int32_t* bt = a.bt + base;
		const exsort::Span span{sk, sv};

		// ---- extPos re-sort when extLen > curLen (std::sort by extPos) ----
		const bool extSorted = gd.extLen > qd.curLen;
		if (extSorted)
		{
			for (int i = lane; i < n; i += kNT)
			{
				sk[i] = static_cast<uint64_t>(static_cast<uint32_t>(ext0[i]));
				sv[i] = static_cast<uint32_t>(i);
			}
			__syncthreads();
			blockStdSort<kNT>(span, n, posL, posR, small, n, kLaneSortMax, sh.sort);
			for (int i = lane; i < n; i += kNT)
			{
				const uint32_t src = sv[i];
				hc[i] = cur0[src];
				he[i] = ext0[src];
			}
		}
		else
		{
			for (int i = lane; i < n; i += kNT)
			{
				hc[i] = cur0[i];
				he[i] = ext0[i];
			}
		}
		__syncthreads();

		// ---- chaining DP (replay_overlap_chains.py:280-313) ----
		if (lane == 0)
		{
			sc[0] = 0;
			bt[0] = -1;
			sh.ring[0] = 0;
		}
		__syncthreads();
		const int32_t maxJump = p.maximum_jump;
		for (int i = 1; i < n; ++i)
		{
			const int32_t curN = hc[i];
			const int32_t extN = he[i];
			int carry = 0;
			int carryId = 0;
			for (int tileBase = i - 1; tileBase >= 0; tileBase -= kNT)
			{
				const int j = tileBase - lane;
				const bool active = j >= 0;
				int cand = -0x7fffffff - 1;
				bool special = false;
				bool stop = false;
				if (active)
				{
					const int32_t dc = curN - hc[j];
					const int32_t de = extN - he[j];
					const int32_t jd = iabs(dc - de);
					if (0 < dc && dc < maxJump && 0 < de && de < maxJump && jd <= p.max_jump_gap)
					{
						const int32_t ms = min(min(dc, de), k);
						const float pen = jd > p.gap_jump_threshold ? p.large_gap_penalty : p.small_gap_penalty;
						const int32_t gc = static_cast<int32_t>(pen * static_cast<float>(jd));
						const int32_t sj = (i - j <= kRing - 1) ? sh.ring[j & (kRing - 1)] : sc[j];
						cand = sj + ms - gc;
						special = (jd == 0 && dc < k);
					}
					stop = extSorted ? (de > maxJump) : (dc > maxJump);
				}
				// running maximum before this lane, in scan order
				const int incl = waveInclusiveMax(cand);
				int before = __shfl_up(incl, 1, FH_WAVE);
				if (lane == 0) before = -0x7fffffff - 1;
				before = max(before, carry);
				const bool improved = active && cand > before;
				const bool brk = active && ((improved && special) || stop);
				const unsigned long long bm = __ballot(brk);
				const int lastLane = bm ? (__ffsll(static_cast<long long>(bm)) - 1) : (kNT - 1);
				const int masked = (active && lane <= lastLane) ? cand : (-0x7fffffff - 1);
				const int mx = waveMax(masked);
				if (mx > carry)
				{
					const unsigned long long hit = __ballot(masked == mx);
					const int l0 = __ffsll(static_cast<long long>(hit)) - 1;
					carry = mx;
					carryId = tileBase - l0;
				}
				if (bm) break;
			}
			const int32_t si = max(carry, k);
			if (lane == 0)
			{
				sc[i] = si;
				bt[i] = carry > k ? carryId : -1;
				sh.ring[i & (kRing - 1)] = si;
				if ((i & 255) == 0) __threadfence_block();
			}
			// one wave: LDS traffic of a wave is processed in order, so a
			// compiler barrier is enough for lane 0's ring write to reach the
			// next row's readers
			__builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
			__builtin_amdgcn_wave_barrier();
		}
		__threadfence_block();
		__syncthreads();

		// ---- score order: std::sort(indices, score[a] > score[b]) ----
		for (int i = lane; i < n; i += kNT)
		{
			sk[i] = static_cast<uint64_t>(static_cast<uint32_t>(0x7fffffff - sc[i]));
			sv[i] = static_cast<uint32_t>(i);
		}
		__syncthreads();
		blockStdSort<kNT>(span, n, posL, posR, small, n, kLaneSortMax, sh.sort);

		// ---- backtrack, overlapTest, divergence (lane 0) ----
		Proposal* props = a.props + base;
		Proposal* outp = a.outProps + base;
		if (lane == 0)
		{
			const int32_t* filt = a.filtered + qd.filtStart;
			const int filtN = qd.filtCount;
			int np = 0;
			int best = -1;
			for (int q = 0; q < n; ++q)
			{
				const int start = static_cast<int>(sv[q]);
				if (bt[start] == -1) continue;
				int first = 0, len = 0, pos = start;
				while (pos != -1)
				{
					first = pos;
					++len;
					const int nxt = bt[pos];
					bt[pos] = -1;
					pos = nxt;
				}
				Proposal o;
				o.curBegin = hc[first];
				o.curEnd = hc[start] + k - 1;
				o.extBegin = he[first];
				o.extEnd = he[start] + k - 1;
				o.score = sc[start] - sc[first] + k - 1;
				o.chainLength = len;
				o.div = 0.0f;
				if (!overlapTestDev(o, qd.curId, gd.extId, qd.curLen, gd.extLen, p)) continue;
				const int fc = upperBound(filt, filtN, o.curEnd) - lowerBound(filt, filtN, o.curBegin);
				const float normLen = static_cast<float>(max(o.curEnd - o.curBegin, o.extEnd - o.extBegin) - fc);
				if (!(normLen > 0.0f)) continue;
				float rate = static_cast<float>(len) * p.sample_rate / normLen;
				rate = fminf(rate, 1.0f);
				if (!(rate > 0.0f)) continue;
				const float inv = 1.0f / rate;
				const float lg = flyeLogf(inv, a.logX, a.logY, a.logN);
				o.div = lg / static_cast<float>(k);
				if (p.only_max_ext)
				{
					if (best < 0 || o.score > outp[0].score)
					{
						outp[0] = o;
						best = 0;
					}
				}
				else
				{
					props[np++] = o;
				}
			}
			if (p.only_max_ext)
			{
				outCount = (best >= 0 && outp[0].div < p.max_divergence) ? 1 : 0;
			}
			sh.red[0] = np;
			sh.red[1] = outCount;
		}
		__syncthreads();
		if (!p.only_max_ext)
		{
			// std::sort(extOverlaps, score desc), containment filter, divergence filter
			const int np = sh.red[0];
			for (int i = lane; i < np; i += kNT)
			{
				sk[i] = static_cast<uint64_t>(static_cast<uint32_t>(0x7fffffff - props[i].score));
				sv[i] = static_cast<uint32_t>(i);
			}
			__syncthreads();
			blockStdSort<kNT>(span, np, posL, posR, small, np, kLaneSortMax, sh.sort);
			if (lane == 0)
			{
				// primary proposal indices, in sorted order
				int32_t* prim = posL;
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
				{
					if (props[prim[r]].div < p.max_divergence) outp[w++] = props[prim[r]];
				}
				sh.red[1] = w;
			}
			__syncthreads();
		}
		if (lane == 0) a.groupCount[g] = sh.red[1];
		__syncthreads();
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
	const Proposal* props = a.outProps + gd.hitStart;
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

} // namespace

hipError_t launchChainGroups(const ChainArgs& a, int blocks, hipStream_t stream)
{
	hipLaunchKernelGGL(chainGroupsKernel, dim3(blocks), dim3(kNT), 0, stream, a);
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
