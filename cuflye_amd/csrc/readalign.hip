// readalign.hip - read-to-graph alignment chaining (subsystem c) on gfx950.
//
// Per read: keep the edge alignments that pass Flye's input filter
// (patch 0039:72-77: extLen < LONG_EDGE || min(curRange, extRange) > BIG_ALN),
// std::sort them by read_begin (patch 0039:249-251), run
// ReadAligner::chainReadAlignments (tools/replay_read_alignment_chains.py:64-139):
// active/frozen chains, strictly-greater best extension over the active list
// in order, the outdated sweep, std::sort of all chains by score and the
// max-separation acceptance.  Both sorts use the exact std::sort emulation, so
// equal scores keep Flye's order.  Reads hold tens of alignments, so one lane
// owns one read; a chain is stored as (score, previous chain, length) keyed by
// the alignment that created it, which is what Python's `indices + [index]`
// copies amount to.
#include "readalign.hpp"
#include "exsort.hpp"

namespace fh
{
namespace
{

__device__ inline int32_t iabs32(int32_t v) { return v < 0 ? -v : v; }

__global__ void readAlignChainKernel(ReadAlignArgs a)
{
	const int r = blockIdx.x * blockDim.x + threadIdx.x;
	if (r >= a.nReads) return;
	const int64_t b = a.alnOff[r];
	const int m0 = static_cast<int>(a.alnOff[r + 1] - b);
	const flyehip_edge_alignment_t* al = a.aln + b;
	int32_t* kept = a.kept + b;
	uint64_t* key = a.key + b;
	uint32_t* val = a.val + b;
	int32_t* cs = a.chainScore + b;
	int32_t* cp = a.chainPrev + b;
	int32_t* cl = a.chainLen + b;
	int32_t* act = a.active + b;
	int32_t* frz = a.frozen + b;
	int32_t* acc = a.accepted + b;

	// filter, then std::sort by read_begin
	int m = 0;
	for (int i = 0; i < m0; ++i)
	{
		const int32_t cr = al[i].read_end - al[i].read_begin;
		const int32_t er = al[i].edge_end - al[i].edge_begin;
		if (al[i].edge_len < a.longEdge || min(cr, er) > a.bigAln)
		{
			key[m] = static_cast<uint64_t>(static_cast<uint32_t>(al[i].read_begin) ^ 0x80000000U);
			val[m] = static_cast<uint32_t>(i);
			++m;
		}
	}
	exsort::stdSort(exsort::Span<uint64_t>{key, val}, 0, m);
	for (int i = 0; i < m; ++i) kept[i] = static_cast<int32_t>(val[i]);
	a.keptCount[r] = m;

	const flyehip_chain_params_t p = a.p;
	int nAct = 0, nFrz = 0;
	for (int i = 0; i < m; ++i)
	{
		const flyehip_edge_alignment_t& cur = al[kept[i]];
		int32_t maxScore = 0;
		int maxChain = -1;
		int numOutdated = 0;
		const bool canExtend = cur.edge_begin < p.maximum_jump;
		const bool canBeExtended = cur.edge_len - cur.edge_end < p.maximum_jump;
		if (canExtend)
		{
			for (int c = 0; c < nAct; ++c)
			{
				const int ch = act[c];
				const flyehip_edge_alignment_t& prev = al[kept[ch]];
				const int32_t readDiff = cur.read_begin - prev.read_end;
				const int32_t gl = cur.edge_begin;
				const int32_t gr = prev.edge_len - prev.edge_end;
				if (prev.edge_right_node == cur.edge_left_node && p.maximum_jump > readDiff &&
					readDiff > -p.max_read_overlap && gl + gr < p.maximum_jump)
				{
					const int32_t jd = iabs32(readDiff - (gl + gr));
					const int32_t gap = jd > 100 ? jd / 50 : 0;
					const int32_t sc = cs[ch] + cur.score - gap;
					if (sc > maxScore)
					{
						maxScore = sc;
						maxChain = ch;
					}
				}
				if (readDiff > p.maximum_jump) ++numOutdated;
			}
		}
		if (maxChain >= 0)
		{
			cs[i] = maxScore;
			cp[i] = maxChain;
			cl[i] = cl[maxChain] + 1;
			act[nAct++] = i;
		}
		else
		{
			cs[i] = cur.score;
			cp[i] = -1;
			cl[i] = 1;
			if (canBeExtended) act[nAct++] = i;
			else frz[nFrz++] = i;
		}
		if (numOutdated > nAct / 2)
		{
			int w = 0;
			for (int c = 0; c < nAct; ++c)
			{
				const int ch = act[c];
				const flyehip_edge_alignment_t& prev = al[kept[ch]];
				if (cur.read_begin - prev.read_end > p.maximum_jump) frz[nFrz++] = ch;
				else act[w++] = ch;
			}
			nAct = w;
		}
	}
	// active.extend(frozen); std::sort by score desc
	const int nAll = nAct + nFrz;
	for (int c = 0; c < nAct; ++c)
	{
		key[c] = static_cast<uint64_t>(static_cast<int64_t>(0x7fffffff) - cs[act[c]]);
		val[c] = static_cast<uint32_t>(act[c]);
	}
	for (int c = 0; c < nFrz; ++c)
	{
		key[nAct + c] = static_cast<uint64_t>(static_cast<int64_t>(0x7fffffff) - cs[frz[c]]);
		val[nAct + c] = static_cast<uint32_t>(frz[c]);
	}
	exsort::stdSort(exsort::Span<uint64_t>{key, val}, 0, nAll);
	// acceptance: first alignment of a chain is found by walking back
	int nAcc = 0;
	int segs = 0;
	for (int q = 0; q < nAll; ++q)
	{
		const int ch = static_cast<int>(val[q]);
		int first = ch;
		while (cp[first] >= 0) first = cp[first];
		const flyehip_edge_alignment_t& fa = al[kept[first]];
		const flyehip_edge_alignment_t& la = al[kept[ch]];
		if (la.read_end - fa.read_begin < p.minimum_overlap) continue;
		bool clash = false;
		for (int t = 0; t < nAcc; ++t)
		{
			const int ex = acc[t];
			int ef = ex;
			while (cp[ef] >= 0) ef = cp[ef];
			const flyehip_edge_alignment_t& efa = al[kept[ef]];
			const flyehip_edge_alignment_t& ela = al[kept[ex]];
			const int32_t rate = min(la.read_end, ela.read_end) - max(fa.read_begin, efa.read_begin);
			if (rate > p.max_separation)
			{
				clash = true;
				break;
			}
		}
		if (!clash)
		{
			acc[nAcc++] = ch;
			segs += cl[ch];
		}
	}
	a.acceptedCount[r] = nAcc;
	a.segCount[r] = segs;
}

__global__ void readAlignWriteKernel(ReadAlignArgs a, const int64_t* segOff, flyehip_chain_segment_t* out)
{
	const int r = blockIdx.x * blockDim.x + threadIdx.x;
	if (r >= a.nReads) return;
	const int64_t b = a.alnOff[r];
	const int32_t* cs = a.chainScore + b;
	const int32_t* cp = a.chainPrev + b;
	const int32_t* cl = a.chainLen + b;
	const int32_t* acc = a.accepted + b;
	int64_t o = segOff[r];
	for (int t = 0; t < a.acceptedCount[r]; ++t)
	{
		const int ch = acc[t];
		const int len = cl[ch];
		int node = ch;
		for (int s = len - 1; s >= 0; --s)
		{
			flyehip_chain_segment_t g;
			g.read_slot = r;
			g.chain_id = t;
			g.segment_id = s;
			g.input_index = node;
			g.chain_score = cs[ch];
			g.reserved = 0;
			out[o + s] = g;
			node = cp[node];
		}
		o += len;
	}
}

// Compacts each read's kept alignment indices (sorted order) to
// out[keptOff[r] ..), keptOff being the exclusive scan of keptCount.
__global__ void readAlignKeptKernel(ReadAlignArgs a, const int64_t* keptOff, int32_t* out)
{
	const int r = blockIdx.x * blockDim.x + threadIdx.x;
	if (r >= a.nReads) return;
	const int32_t* kept = a.kept + a.alnOff[r];
	const int64_t o = keptOff[r];
	for (int i = 0; i < a.keptCount[r]; ++i) out[o + i] = kept[i];
}

} // namespace

hipError_t launchReadAlignKept(const ReadAlignArgs& a, const int64_t* keptOff, int32_t* out, hipStream_t s)
{
	if (a.nReads == 0) return hipSuccess;
	hipLaunchKernelGGL(readAlignKeptKernel, dim3((a.nReads + 63) / 64), dim3(64), 0, s, a, keptOff, out);
	return hipGetLastError();
}

hipError_t launchReadAlignChain(const ReadAlignArgs& a, hipStream_t s)
{
	if (a.nReads == 0) return hipSuccess;
	hipLaunchKernelGGL(readAlignChainKernel, dim3((a.nReads + 63) / 64), dim3(64), 0, s, a);
	return hipGetLastError();
}

hipError_t launchReadAlignWrite(const ReadAlignArgs& a, const int64_t* segOff, flyehip_chain_segment_t* out,
								hipStream_t s)
{
	if (a.nReads == 0) return hipSuccess;
	hipLaunchKernelGGL(readAlignWriteKernel, dim3((a.nReads + 63) / 64), dim3(64), 0, s, a, segOff, out);
	return hipGetLastError();
}

} // namespace fh
