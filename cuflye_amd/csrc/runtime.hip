// runtime.hip - the C ABI of include/flyehip.h: context, device-memory arena
// with a budget, the resident read store, the index build, and the batched
// overlap pipeline (join -> exact match sort -> groups -> chaining -> rows).
//
// Boundary: this replaces Flye's candidate loop behind
// CUFLYE_CANDIDATE_BACKEND (patches/flye/2.9.6/0002-cuflye-candidate-backend-seam.patch
// file lines 97-102, 0005:245-266) and, for whole overlaps, the overlap-level
// hook of patch 0023:450-462.  Errors never cross the ABI as exceptions; every
// entry point returns a status and leaves a message (adapter, device, budget
// and shape text, as docs/abi/cuda-candidate-backend-v0.md asks).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "chain.hpp"
#include "dev_common.hpp"
#include "msort.hpp"
#include "pipeline.hpp"
#include "readalign.hpp"
#include "segsort.hpp"

extern "C" const unsigned char fh_logf_table[];
extern "C" const unsigned char fh_logf_table_end[];

namespace fh
{
namespace
{

struct FhError
{
	int code;
	std::string msg;
};

[[noreturn]] void fail(int code, const char* fmt, ...)
{
	char buf[1024];
	va_list ap;
	va_start(ap, fmt);
	std::vsnprintf(buf, sizeof(buf), fmt, ap);
	va_end(ap);
	throw FhError{code, buf};
}

#define HIPCHECK(expr)                                                                                   \
	do                                                                                                   \
	{                                                                                                    \
		hipError_t e_ = (expr);                                                                          \
		if (e_ != hipSuccess)                                                                            \
			fail(FLYEHIP_ERR_HIP, "HIP error %d (%s: %s) at %s:%d: %s", static_cast<int>(e_),             \
				 hipGetErrorName(e_), hipGetErrorString(e_), __FILE__, __LINE__, #expr);                 \
	} while (0)

} // namespace
} // namespace fh

struct flyehip_ctx;

namespace fh
{
namespace
{

// A grow-only device buffer charged against the context budget.
struct DevBuf
{
	void* p = nullptr;
	size_t bytes = 0;
};

} // namespace
} // namespace fh

struct flyehip_ctx
{
	int device = 0;
	uint64_t budget = 0;
	uint64_t inUse = 0;
	uint64_t peak = 0;
	hipStream_t stream = nullptr;
	hipEvent_t ev[8] = {};
	std::string err;
	std::mutex mu;
	std::vector<fh::DevBuf*> owned;

	// tuning
	int dpLanes = 64;
	int queriesPerBatch = 2048;
	int groupsPerPull = 1;
	int chainBlocks = 0;

	// read store
	int32_t nReads = 0;
	std::vector<int32_t> readLen;
	std::vector<int64_t> readWordOff;
	fh::DevBuf dWords, dWordOff, dReadLen, dReadWinOff;

	// index
	bool haveIndex = false;
	int k = 0;
	flyehip_index_info_t info{};
	fh::DevBuf dKeys, dVals, dCursor, dEntries;
	uint64_t tableMask = 0;

	// logf table
	fh::DevBuf dLogX, dLogY;
	int logN = 0;

	// batch scratch (grow-only)
	fh::DevBuf bQueries, bWinOff, bWinCount, bWinFilt, bWinInfo, bMatchOff, bFiltOff, bScanTmp;
	fh::DevBuf bMKey, bMExt, bVal, bPosL, bPosR, bHCur, bHExt, bGroupFlag, bGroupOff;
	fh::DevBuf bFiltered, bQMatchOff, bQFiltOff, bQDesc, bGroups, bGroupCount, bRowStart, bRows;
	fh::DevBuf bCScore, bCBt, bCProps, bCOut, bCHCur, bCHExt, bCKey, bCVal, bCPosL, bCPosR;
	fh::DevBuf bGroupPass, bOrder, bPassOrder, bBins, bSKey, bAVal, bRKey, bDupFlag, bDup, bMsortTemp;
	fh::DevBuf bSegBig0, bSegBig1, bSegLeaf, bSegHeap, bSegCounters, bSegChunkOff, bSegChunkRange, bSegChunkCnt;
	fh::DevBuf bSegChunkPre, bSegRangeM, bSegRangeLm, bSegTemp, bSegVerify, bSegShadowK, bSegShadowV;
	int64_t* hostSeg = nullptr;  // pinned counter words of the segmented sort
	fh::DevBuf bQueue, bCounters, bQRowCount, bCurLen, bDivOff, bDivCount, bDivOut, bQueryIds, bCand;
	fh::DevBuf bExtWords, bExtWordOff, bExtLen;
	fh::DevBuf bAln, bAlnOff, bRA;

	flyehip_stats_t stats{};
};

namespace fh
{
namespace
{

void release(flyehip_ctx* c, DevBuf& b)
{
	if (b.p)
	{
		hipFree(b.p);
		c->inUse -= b.bytes;
		b.p = nullptr;
		b.bytes = 0;
	}
}

void* ensure(flyehip_ctx* c, DevBuf& b, size_t bytes)
{
	if (bytes == 0) bytes = 16;
	if (b.bytes >= bytes) return b.p;
	release(c, b);
	size_t want = bytes + bytes / 8;
	want = (want + 255) & ~static_cast<size_t>(255);
	if (c->inUse + want > c->budget)
		fail(FLYEHIP_ERR_BUDGET,
			 "device memory budget exceeded: need %zu more bytes, %llu in use, budget %llu (device %d)", want,
			 static_cast<unsigned long long>(c->inUse), static_cast<unsigned long long>(c->budget), c->device);
	HIPCHECK(hipMalloc(&b.p, want));
	b.bytes = want;
	c->inUse += want;
	c->peak = std::max(c->peak, c->inUse);
	return b.p;
}

template <class T>
T* ensureT(flyehip_ctx* c, DevBuf& b, size_t n)
{
	return static_cast<T*>(ensure(c, b, n * sizeof(T)));
}

void upload(flyehip_ctx* c, void* dst, const void* src, size_t bytes)
{
	if (bytes) HIPCHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
}
void download(flyehip_ctx* c, void* dst, const void* src, size_t bytes)
{
	if (bytes) HIPCHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
}
void sync(flyehip_ctx* c) { HIPCHECK(hipStreamSynchronize(c->stream)); }

int baseCode(uint8_t b)
{
	switch (b)
	{
	case 'A': return 0;
	case 'C': return 1;
	case 'G': return 2;
	case 'T': return 3;
	default: return -1;
	}
}

// Packs sequences 16 bases per word (MSB first) with two spare words each.
void packSeqs(const uint8_t* bases, const int64_t* offsets, int32_t n, std::vector<uint32_t>& words,
			  std::vector<int64_t>& wordOff, std::vector<int32_t>& lens)
{
	words.clear();
	wordOff.assign(static_cast<size_t>(n) + 1, 0);
	lens.assign(static_cast<size_t>(n), 0);
	int64_t total = 0;
	for (int32_t i = 0; i < n; ++i)
	{
		const int64_t len = offsets[i + 1] - offsets[i];
		if (len < 0 || len > 0x7fffffff) fail(FLYEHIP_ERR_ARG, "sequence %d has invalid length %lld", i, static_cast<long long>(len));
		wordOff[static_cast<size_t>(i)] = total;
		total += (len + 15) / 16 + 2;
	}
	wordOff[static_cast<size_t>(n)] = total;
	words.assign(static_cast<size_t>(total), 0);
	for (int32_t i = 0; i < n; ++i)
	{
		const int64_t len = offsets[i + 1] - offsets[i];
		lens[static_cast<size_t>(i)] = static_cast<int32_t>(len);
		const uint8_t* s = bases + offsets[i];
		uint32_t* w = words.data() + wordOff[static_cast<size_t>(i)];
		for (int64_t j = 0; j < len; ++j)
		{
			const int b = baseCode(s[j]);
			if (b < 0) fail(FLYEHIP_ERR_ARG, "sequence %d has a non-ACGT byte 0x%02x at %lld", i, s[j], static_cast<long long>(j));
			w[j >> 4] |= static_cast<uint32_t>(b) << (30 - 2 * (j & 15));
		}
	}
}

void checkParams(const flyehip_overlap_params_t* p, int k)
{
	if (!p) fail(FLYEHIP_ERR_ARG, "null overlap parameters");
	if (p->keep_alignment || p->nucl_alignment || p->partition_bad_mappings)
		fail(FLYEHIP_ERR_UNSUPPORTED,
			 "unsupported overlap shape: keep_alignment=%d nucl_alignment=%d partition_bad_mappings=%d need Flye base "
			 "alignment, which this library does not model",
			 p->keep_alignment, p->nucl_alignment, p->partition_bad_mappings);
	if (p->max_overlaps != 0)
		fail(FLYEHIP_ERR_UNSUPPORTED, "unsupported overlap shape: max_overlaps=%d (only 0 = unlimited)", p->max_overlaps);
	if (k > 0 && p->kmer_size != k) fail(FLYEHIP_ERR_ARG, "kmer_size %d differs from the index k %d", p->kmer_size, k);
	if (p->kmer_size < 1 || p->kmer_size > 31) fail(FLYEHIP_ERR_ARG, "kmer_size %d out of range 1..31", p->kmer_size);
}

int autoSolidMin(const std::vector<uint64_t>& hist)
{
	for (size_t c = 2; c + 1 < hist.size(); ++c)
		if (hist[c] <= hist[c + 1]) return static_cast<int>(c);
	return 2;
}

int autoRepMax(const std::vector<uint64_t>& hist, const uint64_t* acc, int solidMin, float rate)
{
	double total = 0.0, distinct = 0.0;
	for (size_t c = static_cast<size_t>(solidMin); c < hist.size(); ++c)
	{
		total += static_cast<double>(c) * static_cast<double>(hist[c]);
		distinct += static_cast<double>(hist[c]);
	}
	total += static_cast<double>(acc[0]);
	distinct += static_cast<double>(acc[1]);
	const double mean = total / (distinct + 1.0);
	const int rep = static_cast<int>(std::floor(static_cast<double>(rate) * mean));
	return std::max(rep, solidMin);
}

void loadLogTable(flyehip_ctx* c)
{
	const size_t bytes = static_cast<size_t>(fh_logf_table_end - fh_logf_table);
	const size_t n = bytes / 8;
	std::vector<uint32_t> xs(n), ys(n);
	const uint32_t* t = reinterpret_cast<const uint32_t*>(fh_logf_table);
	for (size_t i = 0; i < n; ++i)
	{
		xs[i] = t[2 * i];
		ys[i] = t[2 * i + 1];
	}
	if (n == 0) fail(FLYEHIP_ERR_STATE, "logf agreement table is empty (library built without it)");
	upload(c, ensure(c, c->dLogX, n * 4), xs.data(), n * 4);
	upload(c, ensure(c, c->dLogY, n * 4), ys.data(), n * 4);
	c->logN = static_cast<int>(n);
	sync(c);
}

__global__ void gatherAtKernel(const int64_t* src, const int64_t* idx, int64_t* dst, int32_t n)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i <= n) dst[i] = src[idx[i]];
}

__global__ void makeQueryDescKernel(const int64_t* ids, const int32_t* lens, const int64_t* qFiltOff, int32_t n,
									QueryDesc* out)
{
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n) return;
	QueryDesc d;
	d.curId = ids[i];
	d.curLen = lens[i];
	d.filtStart = qFiltOff[i];
	d.filtCount = static_cast<int32_t>(qFiltOff[i + 1] - qFiltOff[i]);
	out[i] = d;
}

__global__ void queryRowCountKernel(const GroupDesc* groups, const int32_t* groupCount, int32_t nGroups, int32_t* qCount)
{
	const int g = blockIdx.x * blockDim.x + threadIdx.x;
	if (g >= nGroups) return;
	const int c = groupCount[g];
	if (c) atomicAdd(&qCount[groups[g].query], c);
}

__global__ void zeroKernel(int32_t* p, int64_t n)
{
	const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
	if (i < n) p[i] = 0;
}

__global__ void fillU64Kernel(uint64_t* p, int64_t n, uint64_t v)
{
	const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
	if (i < n) p[i] = v;
}

dim3 grid(int64_t n, int tpb) { return dim3(static_cast<unsigned>(std::max<int64_t>(1, (n + tpb - 1) / tpb))); }

void zero32(flyehip_ctx* c, int32_t* p, int64_t n)
{
	if (n <= 0) return;
	hipLaunchKernelGGL(zeroKernel, grid(n, 256), dim3(256), 0, c->stream, p, n);
	HIPCHECK(hipGetLastError());
}

void fill64(flyehip_ctx* c, uint64_t* p, int64_t n, uint64_t v)
{
	if (n <= 0) return;
	hipLaunchKernelGGL(fillU64Kernel, grid(n, 256), dim3(256), 0, c->stream, p, n, v);
	HIPCHECK(hipGetLastError());
}

PackedSeqs storeSeqs(flyehip_ctx* c)
{
	PackedSeqs s;
	s.words = static_cast<const uint32_t*>(c->dWords.p);
	s.wordOff = static_cast<const int64_t*>(c->dWordOff.p);
	s.len = static_cast<const int32_t*>(c->dReadLen.p);
	s.n = c->nReads;
	return s;
}

HashTable table(flyehip_ctx* c)
{
	HashTable t;
	t.keys = static_cast<uint64_t*>(c->dKeys.p);
	t.vals = static_cast<uint64_t*>(c->dVals.p);
	t.cursor = static_cast<uint32_t*>(c->dCursor.p);
	t.mask = c->tableMask;
	return t;
}

int chainBlocks(flyehip_ctx* c)
{
	if (c->chainBlocks > 0) return c->chainBlocks;
	hipDeviceProp_t prop;
	HIPCHECK(hipGetDeviceProperties(&prop, c->device));
	c->chainBlocks = prop.multiProcessorCount * 16;
	return c->chainBlocks;
}

struct Timer
{
	std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
	double ms() const
	{
		return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
	}
};

// Result of one batch through join + sort (+ optional chaining).
struct BatchOut
{
	int64_t matches = 0;
	int64_t filtered = 0;
	int64_t groups = 0;
	int64_t rows = 0;
};

// Join for a batch of queries; leaves matches (emission order) in bMKey/bMExt.
BatchOut runJoin(flyehip_ctx* c, const PackedSeqs& seqs, const std::vector<JoinQuery>& qs, JoinArgs& ja)
{
	const int32_t nq = static_cast<int32_t>(qs.size());
	const int k = c->k;
	std::vector<int64_t> winOff(static_cast<size_t>(nq) + 1, 0);
	for (int32_t i = 0; i < nq; ++i) winOff[static_cast<size_t>(i) + 1] = winOff[static_cast<size_t>(i)] + std::max(0, qs[static_cast<size_t>(i)].len - k + 1);
	const int64_t nWin = winOff.back();
	JoinQuery* dQ = ensureT<JoinQuery>(c, c->bQueries, qs.size());
	int64_t* dWinOff = ensureT<int64_t>(c, c->bWinOff, winOff.size());
	upload(c, dQ, qs.data(), qs.size() * sizeof(JoinQuery));
	upload(c, dWinOff, winOff.data(), winOff.size() * sizeof(int64_t));
	ja = JoinArgs{};
	ja.seqs = seqs;
	ja.queries = dQ;
	ja.nQueries = nq;
	ja.k = k;
	ja.winOff = dWinOff;
	ja.nWin = nWin;
	ja.table = table(c);
	ja.entries = static_cast<const uint64_t*>(c->dEntries.p);
	ja.targetLen = static_cast<const int32_t*>(c->dReadLen.p);
	ja.winCount = ensureT<int32_t>(c, c->bWinCount, static_cast<size_t>(nWin) + 1);
	ja.winFilt = ensureT<int32_t>(c, c->bWinFilt, static_cast<size_t>(nWin) + 1);
	ja.winInfo = ensureT<uint32_t>(c, c->bWinInfo, static_cast<size_t>(nWin) + 1);
	int64_t* matchOff = ensureT<int64_t>(c, c->bMatchOff, static_cast<size_t>(nWin) + 1);
	int64_t* filtOff = ensureT<int64_t>(c, c->bFiltOff, static_cast<size_t>(nWin) + 1);
	int64_t* tmp = ensureT<int64_t>(c, c->bScanTmp, scanTempWords(nWin) + 64);
	ja.matchOff = matchOff;
	ja.filtOff = filtOff;
	HIPCHECK(hipEventRecord(c->ev[0], c->stream));
	HIPCHECK(launchJoinCount(ja, c->stream));
	HIPCHECK(hipEventRecord(c->ev[1], c->stream));
	HIPCHECK(exclusiveScan(ja.winCount, matchOff, nWin, tmp, c->stream));
	HIPCHECK(exclusiveScan(ja.winFilt, filtOff, nWin, tmp, c->stream));
	// per-query offsets
	int64_t* qMatchOff = ensureT<int64_t>(c, c->bQMatchOff, static_cast<size_t>(nq) + 1);
	int64_t* qFiltOff = ensureT<int64_t>(c, c->bQFiltOff, static_cast<size_t>(nq) + 1);
	hipLaunchKernelGGL(gatherAtKernel, grid(nq + 1, 256), dim3(256), 0, c->stream, matchOff, dWinOff, qMatchOff, nq);
	hipLaunchKernelGGL(gatherAtKernel, grid(nq + 1, 256), dim3(256), 0, c->stream, filtOff, dWinOff, qFiltOff, nq);
	HIPCHECK(hipGetLastError());
	int64_t tot[2];
	download(c, &tot[0], matchOff + nWin, sizeof(int64_t));
	download(c, &tot[1], filtOff + nWin, sizeof(int64_t));
	sync(c);
	BatchOut bo;
	bo.matches = tot[0];
	bo.filtered = tot[1];
	ja.mkey = ensureT<uint64_t>(c, c->bMKey, static_cast<size_t>(bo.matches) + 1);
	ja.mext = ensureT<int32_t>(c, c->bMExt, static_cast<size_t>(bo.matches) + 1);
	ja.filtered = ensureT<int32_t>(c, c->bFiltered, static_cast<size_t>(bo.filtered) + 1);
	HIPCHECK(hipEventRecord(c->ev[6], c->stream));
	HIPCHECK(launchJoinFill(ja, c->stream));
	HIPCHECK(hipEventRecord(c->ev[7], c->stream));
	HIPCHECK(hipEventSynchronize(c->ev[7]));
	float e1 = 0, e2 = 0;
	HIPCHECK(hipEventElapsedTime(&e1, c->ev[0], c->ev[1]));
	HIPCHECK(hipEventElapsedTime(&e2, c->ev[6], c->ev[7]));
	c->stats.ms_join += e1 + e2;
	c->stats.windows += nWin;
	c->stats.matches += bo.matches;
	return bo;
}

// Exact per-query std::sort of the batch's matches; the sorted keys land in
// bSKey and the permutation (emission index within the query) in bVal.
// Buffers of the segmented exact sort for n elements in nArrays arrays; the
// caller sets key / val / outVal / dup / posL / posR.
SegSortBufs segBufs(flyehip_ctx* c, int64_t n, int64_t nArrays)
{
	SegSortCaps caps{};
	HIPCHECK(segSortCaps(n, nArrays, &caps));
	SegSortBufs b{};
	b.big[0] = ensureT<SegRange>(c, c->bSegBig0, static_cast<size_t>(caps.bigCap));
	b.big[1] = ensureT<SegRange>(c, c->bSegBig1, static_cast<size_t>(caps.bigCap));
	b.leaf = ensureT<SegRange>(c, c->bSegLeaf, static_cast<size_t>(caps.leafCap));
	b.heap = ensureT<SegRange>(c, c->bSegHeap, static_cast<size_t>(caps.heapCap));
	b.verify = ensureT<SegRange>(c, c->bSegVerify, static_cast<size_t>(caps.heapCap));
	b.shadowKey = ensureT<uint32_t>(c, c->bSegShadowK, static_cast<size_t>(n) + 1);
	b.shadowVal = ensureT<uint32_t>(c, c->bSegShadowV, static_cast<size_t>(n) + 1);
	b.counters = ensureT<int64_t>(c, c->bSegCounters, kSegCounters);
	b.chunkOff = ensureT<int64_t>(c, c->bSegChunkOff, static_cast<size_t>(caps.bigCap) + 1);
	b.chunkRange = ensureT<int32_t>(c, c->bSegChunkRange, static_cast<size_t>(caps.chunkCap));
	b.chunkCnt = ensureT<uint64_t>(c, c->bSegChunkCnt, static_cast<size_t>(caps.chunkCap) + 1);
	b.chunkPre = ensureT<uint64_t>(c, c->bSegChunkPre, static_cast<size_t>(std::max(caps.chunkCap, caps.bigCap)) + 1);
	b.rangeM = ensureT<int32_t>(c, c->bSegRangeM, static_cast<size_t>(caps.bigCap));
	b.rangeLm = ensureT<int32_t>(c, c->bSegRangeLm, static_cast<size_t>(caps.bigCap));
	b.scanTemp = ensure(c, c->bSegTemp, caps.scanTempBytes + 256);
	b.scanTempBytes = caps.scanTempBytes;
	b.bigCap = caps.bigCap;
	b.leafCap = caps.leafCap;
	b.heapCap = caps.heapCap;
	b.chunkCap = caps.chunkCap;
	b.hostCounters = c->hostSeg;
	return b;
}

void runSegSort(flyehip_ctx* c, SegSortBufs& b, int64_t* tieArrays)
{
	const hipError_t e = segSortRun(b, c->stream, tieArrays);
	if (e == hipErrorIllegalState) fail(FLYEHIP_ERR_HIP, "segmented sort: inconsistent partition");
	HIPCHECK(e);
}

// Exact std::sort of every query's matches: stable order, then the tie replay.
// Leaves the sorted keys in bSKey and the permutation (emission index within
// the query) in bVal.
void runMatchSort(flyehip_ctx* c, const JoinArgs& ja, int64_t nMatches)
{
	if (nMatches >= 0x7fffffffLL) fail(FLYEHIP_ERR_BUDGET, "batch of %lld matches exceeds 2^31 - 1", static_cast<long long>(nMatches));
	const size_t nm = static_cast<size_t>(nMatches) + 1;
	const int32_t nq = ja.nQueries;
	MsortArgs ms{};
	ms.qOff = static_cast<const int64_t*>(c->bQMatchOff.p);
	ms.nQueries = nq;
	ms.akey = ja.mkey;
	ms.aval = ensureT<uint32_t>(c, c->bAVal, nm);
	ms.skey = ensureT<uint64_t>(c, c->bSKey, nm);
	ms.sval = ensureT<uint32_t>(c, c->bVal, nm);
	ms.rkey = ensureT<uint32_t>(c, c->bRKey, nm);
	ms.dupFlag = ensureT<int32_t>(c, c->bDupFlag, nm);
	int64_t* dup = ensureT<int64_t>(c, c->bDup, nm);
	int64_t* tmp = ensureT<int64_t>(c, c->bScanTmp, scanTempWords(nMatches) + 64);
	int bits = 1;
	while ((int64_t{1} << bits) < 2 * static_cast<int64_t>(std::max(1, c->nReads))) ++bits;
	size_t tempBytes = 0;
	HIPCHECK(msortTempBytes(nMatches, nq, bits, &tempBytes));
	void* temp = ensure(c, c->bMsortTemp, tempBytes + 256);
	SegSortBufs sb = segBufs(c, nMatches, nq);
	sb.key = ms.rkey;
	sb.val = ms.aval;
	sb.outVal = ms.sval;
	sb.dup = dup;
	sb.posL = ensureT<int32_t>(c, c->bPosL, nm);
	sb.posR = ensureT<int32_t>(c, c->bPosR, nm);
	HIPCHECK(hipEventRecord(c->ev[0], c->stream));
	HIPCHECK(launchMatchSortStable(ms, nMatches, bits, temp, tempBytes, c->stream));
	HIPCHECK(exclusiveScan(ms.dupFlag, dup, nMatches, tmp, c->stream));
	HIPCHECK(segSortReset(sb, c->stream));
	HIPCHECK(segPushOffsets(sb, ms.qOff, nq, c->stream));
	int64_t ties = 0;
	runSegSort(c, sb, &ties);
	HIPCHECK(hipEventRecord(c->ev[1], c->stream));
	HIPCHECK(hipEventSynchronize(c->ev[1]));
	float es = 0;
	HIPCHECK(hipEventElapsedTime(&es, c->ev[0], c->ev[1]));
	c->stats.ms_sort += es;
	c->stats.tie_queries += ties;
}

// Order, prefilters + extPos sort, DP, score sort + walks over ca.nGroups
// groups whose hits total nHits; leaves groupCount set.
void runChainKernels(flyehip_ctx* c, ChainArgs& ca, int64_t nHits)
{
	const int64_t ng = ca.nGroups;
	zero32(c, ca.groupCount, ng + 1);
	if (ng == 0) return;
	if (nHits >= 0x7fffffffLL) fail(FLYEHIP_ERR_BUDGET, "batch of %lld hits exceeds 2^31 - 1", static_cast<long long>(nHits));
	int32_t* bins = ensureT<int32_t>(c, c->bBins, 96);
	int32_t* order = ensureT<int32_t>(c, c->bOrder, static_cast<size_t>(ng) + 1);
	int32_t* passOrder = ensureT<int32_t>(c, c->bPassOrder, static_cast<size_t>(ng) + 1);
	int32_t* flag = ensureT<int32_t>(c, c->bGroupFlag, static_cast<size_t>(ng) + 1);
	int64_t* pos = ensureT<int64_t>(c, c->bGroupOff, static_cast<size_t>(ng) + 1);
	int64_t* tmp = ensureT<int64_t>(c, c->bScanTmp, scanTempWords(std::max(ng, nHits)) + 64);
	int32_t* queues = ensureT<int32_t>(c, c->bQueue, 16);
	int64_t* nBig = ensureT<int64_t>(c, c->bCounters, 16);
	zero32(c, bins, 96);
	zero32(c, queues, 16);
	HIPCHECK(hipMemsetAsync(nBig, 0, sizeof(int64_t), c->stream));
	SegSortBufs sb = segBufs(c, nHits, ng);
	sb.key = ca.sortKey;
	sb.val = ca.sortVal;
	sb.outVal = ca.sortVal;
	sb.dup = nullptr;
	sb.posL = ca.posL;
	sb.posR = ca.posR;
	const int blocks = chainBlocks(c);

	HIPCHECK(hipEventRecord(c->ev[2], c->stream));
	HIPCHECK(launchGroupOrder(ca.groups, ca.nGroups, bins, order, c->stream));
	ca.order = order;
	HIPCHECK(segSortReset(sb, c->stream));
	ca.queue = queues + 0;
	HIPCHECK(launchChainPrefilter(ca, sb, blocks, c->stream));
	runSegSort(c, sb, nullptr);
	HIPCHECK(launchPassFlags(ca, flag, c->stream));
	HIPCHECK(exclusiveScan(flag, pos, ng, tmp, c->stream));
	HIPCHECK(launchPassScatter(ca, flag, pos, passOrder, nBig, c->stream));
	int64_t counts[2] = {0, 0};
	download(c, &counts[0], pos + ng, sizeof(int64_t));
	download(c, &counts[1], nBig, sizeof(int64_t));
	sync(c);
	ca.passOrder = passOrder;
	ca.nPass = static_cast<int32_t>(counts[0]);
	const int32_t nBigPass = static_cast<int32_t>(counts[1]);
	ca.queue = queues + 1;
	HIPCHECK(launchChainGather(ca, blocks, c->stream));
	HIPCHECK(hipEventRecord(c->ev[3], c->stream));
	ca.queue = queues + 2;
	HIPCHECK(launchChainDp(ca, c->dpLanes, blocks, c->stream));
	HIPCHECK(hipEventRecord(c->ev[4], c->stream));
	HIPCHECK(segSortReset(sb, c->stream));
	ca.queue = queues + 3;
	HIPCHECK(launchChainScoreKeys(ca, sb, blocks, c->stream));
	runSegSort(c, sb, nullptr);
	ca.queue = queues + 4;
	HIPCHECK(launchChainWalk(ca, 0, nBigPass, true, blocks, c->stream));
	ca.queue = queues + 5;
	HIPCHECK(launchChainWalk(ca, nBigPass, ca.nPass, false, blocks, c->stream));
	HIPCHECK(hipEventRecord(c->ev[5], c->stream));
	HIPCHECK(hipEventSynchronize(c->ev[5]));
	float a = 0, b = 0, d = 0;
	HIPCHECK(hipEventElapsedTime(&a, c->ev[2], c->ev[3]));
	HIPCHECK(hipEventElapsedTime(&b, c->ev[3], c->ev[4]));
	HIPCHECK(hipEventElapsedTime(&d, c->ev[4], c->ev[5]));
	c->stats.ms_group += a;
	c->stats.ms_chain += b;
	c->stats.ms_extract += d;
}

// Groups + chaining for a batch whose sorted matches sit in bMKey/bVal.
// Returns rows on the device in bRows and per-query row counts in bQRowCount.
BatchOut runChain(flyehip_ctx* c, const JoinArgs& ja, int64_t nMatches, const std::vector<int64_t>& qIds,
				  const std::vector<int32_t>& qLens, const flyehip_overlap_params_t& p, double* msGroup, double* msChain)
{
	BatchOut bo;
	const int32_t nq = ja.nQueries;
	Timer tg;
	int32_t* hcur = ensureT<int32_t>(c, c->bHCur, static_cast<size_t>(nMatches) + 1);
	int32_t* hext = ensureT<int32_t>(c, c->bHExt, static_cast<size_t>(nMatches) + 1);
	int32_t* gflag = ensureT<int32_t>(c, c->bGroupFlag, static_cast<size_t>(nMatches) + 1);
	int64_t* goff = ensureT<int64_t>(c, c->bGroupOff, static_cast<size_t>(nMatches) + 1);
	int64_t* tmp = ensureT<int64_t>(c, c->bScanTmp, scanTempWords(nMatches) + 64);
	const int64_t* qMatchOff = static_cast<const int64_t*>(c->bQMatchOff.p);
	const uint64_t* skey = static_cast<const uint64_t*>(c->bSKey.p);
	HIPCHECK(launchGatherHits(qMatchOff, nq, skey, static_cast<const uint32_t*>(c->bVal.p), ja.mext, hcur, hext, gflag,
							  nMatches, c->stream));
	HIPCHECK(exclusiveScan(gflag, goff, nMatches, tmp, c->stream));
	int64_t nGroups = 0;
	download(c, &nGroups, goff + nMatches, sizeof(int64_t));
	sync(c);
	GroupDesc* groups = ensureT<GroupDesc>(c, c->bGroups, static_cast<size_t>(nGroups) + 1);
	HIPCHECK(launchBuildGroups(qMatchOff, nq, skey, gflag, goff, ja.targetLen, groups, nMatches, c->stream));
	HIPCHECK(launchGroupSizes(groups, nGroups, nMatches, c->stream));
	// query descriptors
	int64_t* dIds = ensureT<int64_t>(c, c->bQueryIds, static_cast<size_t>(nq) + 1);
	int32_t* dLens = ensureT<int32_t>(c, c->bCurLen, static_cast<size_t>(nq) + 1);
	upload(c, dIds, qIds.data(), qIds.size() * sizeof(int64_t));
	upload(c, dLens, qLens.data(), qLens.size() * sizeof(int32_t));
	QueryDesc* qd = ensureT<QueryDesc>(c, c->bQDesc, static_cast<size_t>(nq) + 1);
	hipLaunchKernelGGL(makeQueryDescKernel, grid(nq, 256), dim3(256), 0, c->stream, dIds, dLens,
					   static_cast<const int64_t*>(c->bQFiltOff.p), nq, qd);
	HIPCHECK(hipGetLastError());
	sync(c);
	(void)msGroup;
	(void)tg;

	Timer tc;
	ChainArgs ca{};
	ca.params = p;
	ca.groups = groups;
	ca.queries = qd;
	ca.order = nullptr;
	ca.nGroups = static_cast<int32_t>(nGroups);
	ca.logN = c->logN;
	ca.cur = hcur;
	ca.ext = hext;
	ca.filtered = ja.filtered;
	const size_t nm = static_cast<size_t>(nMatches) + 1;
	ca.sortKey = ensureT<uint32_t>(c, c->bCKey, nm);
	ca.sortVal = ensureT<uint32_t>(c, c->bCVal, nm);
	ca.posL = ensureT<int32_t>(c, c->bCPosL, nm);
	ca.posR = ensureT<int32_t>(c, c->bCPosR, nm);
	ca.hcur = ensureT<int32_t>(c, c->bCHCur, nm);
	ca.hext = ensureT<int32_t>(c, c->bCHExt, nm);
	ca.score = ensureT<int32_t>(c, c->bCScore, nm);
	ca.bt = ensureT<int32_t>(c, c->bCBt, nm);
	ca.props = ensureT<Proposal>(c, c->bCProps, nm);
	ca.outProps = ensureT<Proposal>(c, c->bCOut, nm);
	ca.groupCount = ensureT<int32_t>(c, c->bGroupCount, static_cast<size_t>(nGroups) + 1);
	ca.groupPass = ensureT<int32_t>(c, c->bGroupPass, static_cast<size_t>(nGroups) + 1);
	ca.logX = static_cast<const uint32_t*>(c->dLogX.p);
	ca.logY = static_cast<const uint32_t*>(c->dLogY.p);
	runChainKernels(c, ca, nMatches);
	int64_t* rowStart = ensureT<int64_t>(c, c->bRowStart, static_cast<size_t>(nGroups) + 1);
	HIPCHECK(exclusiveScan(ca.groupCount, rowStart, nGroups, tmp, c->stream));
	int64_t nRows = 0;
	download(c, &nRows, rowStart + nGroups, sizeof(int64_t));
	sync(c);
	flyehip_overlap_t* rows = ensureT<flyehip_overlap_t>(c, c->bRows, static_cast<size_t>(nRows) + 1);
	HIPCHECK(launchWriteRows(ca, rowStart, rows, c->stream));
	int32_t* qCount = ensureT<int32_t>(c, c->bQRowCount, static_cast<size_t>(nq) + 1);
	zero32(c, qCount, nq + 1);
	if (nGroups > 0)
	{
		hipLaunchKernelGGL(queryRowCountKernel, grid(nGroups, 256), dim3(256), 0, c->stream, groups, ca.groupCount,
						   static_cast<int32_t>(nGroups), qCount);
		HIPCHECK(hipGetLastError());
	}
	sync(c);
	(void)msChain;
	(void)tc;
	bo.groups = nGroups;
	bo.rows = nRows;
	c->stats.groups += nGroups;
	c->stats.overlaps += nRows;
	return bo;
}

void resetStats(flyehip_ctx* c)
{
	c->stats = flyehip_stats_t{};
}

template <class F>
int guarded(flyehip_ctx* c, F f)
{
	if (!c) return FLYEHIP_ERR_ARG;
	std::lock_guard<std::mutex> lock(c->mu);
	try
	{
		HIPCHECK(hipSetDevice(c->device));
		f();
		c->err.clear();
		return FLYEHIP_OK;
	}
	catch (const FhError& e)
	{
		c->err = e.msg;
		return e.code;
	}
	catch (const std::bad_alloc&)
	{
		c->err = "host allocation failed";
		return FLYEHIP_ERR_BUDGET;
	}
}

// Batched overlap pipeline over JoinQuery lists (resident or external).
void overlapsImpl(flyehip_ctx* c, const PackedSeqs& seqs, const std::vector<JoinQuery>& allQ,
				  const int64_t* queryIds, const flyehip_overlap_params_t* p, flyehip_overlap_t* rows,
				  int64_t rowCap, int64_t* rowOffsets, int64_t* rowsNeeded, float* divStats, int64_t divCap,
				  int64_t* divOffsets, int batchQueries)
{
	checkParams(p, c->k);
	resetStats(c);
	Timer total;
	const int32_t n = static_cast<int32_t>(allQ.size());
	int64_t written = 0;
	int64_t divWritten = 0;
	rowOffsets[0] = 0;
	if (divOffsets) divOffsets[0] = 0;
	bool capacityShort = false;
	std::vector<flyehip_overlap_t> hostRows;
	std::vector<int32_t> hostCount;
	for (int32_t b0 = 0; b0 < n; b0 += batchQueries)
	{
		const int32_t b1 = std::min(n, b0 + batchQueries);
		std::vector<JoinQuery> qs(allQ.begin() + b0, allQ.begin() + b1);
		std::vector<int64_t> ids(queryIds + b0, queryIds + b1);
		std::vector<int32_t> lens;
		for (auto& q : qs) lens.push_back(q.len);
		JoinArgs ja;
		BatchOut jo = runJoin(c, seqs, qs, ja);
		sync(c);

		runMatchSort(c, ja, jo.matches);

		BatchOut co = runChain(c, ja, jo.matches, ids, lens, *p, &c->stats.ms_group, &c->stats.ms_chain);
		Timer te;
		const int32_t nq = b1 - b0;
		hostCount.assign(static_cast<size_t>(nq) + 1, 0);
		download(c, hostCount.data(), c->bQRowCount.p, static_cast<size_t>(nq) * sizeof(int32_t));
		if (written + co.rows <= rowCap)
		{
			download(c, rows + written, c->bRows.p, static_cast<size_t>(co.rows) * sizeof(flyehip_overlap_t));
		}
		else
		{
			capacityShort = true;
		}
		if (divOffsets)
		{
			// divStats windows, computed on the device from the batch rows
			std::vector<int64_t> dOff(static_cast<size_t>(nq) + 1, 0);
			std::vector<int64_t> qRowOff(static_cast<size_t>(nq) + 1, 0);
			for (int32_t i = 0; i < nq; ++i)
			{
				dOff[static_cast<size_t>(i) + 1] = dOff[static_cast<size_t>(i)] + lens[static_cast<size_t>(i)] / 10000 + 1;
			}
			sync(c);
			for (int32_t i = 0; i < nq; ++i) qRowOff[static_cast<size_t>(i) + 1] = qRowOff[static_cast<size_t>(i)] + hostCount[static_cast<size_t>(i)];
			int64_t* dDivOff = ensureT<int64_t>(c, c->bDivOff, dOff.size());
			int64_t* dQRowOff = ensureT<int64_t>(c, c->bRowStart, std::max(qRowOff.size(), static_cast<size_t>(co.groups) + 1));
			int32_t* dDivCount = ensureT<int32_t>(c, c->bDivCount, static_cast<size_t>(nq) + 1);
			float* dDivOut = ensureT<float>(c, c->bDivOut, static_cast<size_t>(dOff.back()) + 1);
			upload(c, dDivOff, dOff.data(), dOff.size() * sizeof(int64_t));
			upload(c, dQRowOff, qRowOff.data(), qRowOff.size() * sizeof(int64_t));
			HIPCHECK(launchDivStats(static_cast<const flyehip_overlap_t*>(c->bRows.p), dQRowOff,
									static_cast<const int32_t*>(c->bCurLen.p), nq, dDivCount, dDivOut, dDivOff, c->stream));
			std::vector<int32_t> dc(static_cast<size_t>(nq));
			std::vector<float> dv(static_cast<size_t>(dOff.back()) + 1);
			download(c, dc.data(), dDivCount, static_cast<size_t>(nq) * sizeof(int32_t));
			download(c, dv.data(), dDivOut, static_cast<size_t>(dOff.back()) * sizeof(float));
			sync(c);
			for (int32_t i = 0; i < nq; ++i)
			{
				for (int32_t t = 0; t < dc[static_cast<size_t>(i)]; ++t)
				{
					if (divStats && divWritten < divCap) divStats[divWritten] = dv[static_cast<size_t>(dOff[static_cast<size_t>(i)] + t)];
					++divWritten;
				}
				divOffsets[b0 + i + 1] = divWritten;
			}
		}
		sync(c);
		for (int32_t i = 0; i < nq; ++i)
		{
			written += hostCount[static_cast<size_t>(i)];
			rowOffsets[b0 + i + 1] = written;
		}
		(void)te;
		c->stats.batches += 1;
	}
	c->stats.queries = n;
	c->stats.ms_total = total.ms();
	c->stats.device_bytes_peak = static_cast<int64_t>(c->peak);
	if (rowsNeeded) *rowsNeeded = written;
	if (capacityShort) fail(FLYEHIP_ERR_CAPACITY, "row capacity %lld too small: %lld rows needed",
							static_cast<long long>(rowCap), static_cast<long long>(written));
}

std::vector<JoinQuery> residentQueries(flyehip_ctx* c, const int64_t* ids, int32_t n)
{
	if (!c->haveIndex) fail(FLYEHIP_ERR_STATE, "no index: call flyehip_store_reads and flyehip_build_index first");
	std::vector<JoinQuery> qs(static_cast<size_t>(n));
	for (int32_t i = 0; i < n; ++i)
	{
		const int64_t id = ids[i];
		const int64_t r = id > 0 ? id : -id;
		if (id == 0 || r > c->nReads) fail(FLYEHIP_ERR_ARG, "query id %lld does not name a resident read (1..%d)", static_cast<long long>(id), c->nReads);
		JoinQuery q;
		q.seq = static_cast<int32_t>(r - 1);
		q.strand = id > 0 ? 1 : 0;
		q.iid = toInternal(id);
		q.selfMode = 1;
		q.len = c->readLen[static_cast<size_t>(r - 1)];
		q.reserved = 0;
		qs[static_cast<size_t>(i)] = q;
	}
	return qs;
}

} // namespace
} // namespace fh

using namespace fh;

extern "C" {

int flyehip_abi_version(void) { return FLYEHIP_ABI_VERSION; }

int flyehip_open(int device, uint64_t memory_budget_bytes, flyehip_ctx** out_ctx)
{
	if (!out_ctx) return FLYEHIP_ERR_ARG;
	*out_ctx = nullptr;
	flyehip_ctx* c = new (std::nothrow) flyehip_ctx();
	if (!c) return FLYEHIP_ERR_BUDGET;
	c->device = device;
	int rc = FLYEHIP_OK;
	try
	{
		int count = 0;
		HIPCHECK(hipGetDeviceCount(&count));
		if (device < 0 || device >= count) fail(FLYEHIP_ERR_ARG, "device %d not present (%d HIP devices)", device, count);
		HIPCHECK(hipSetDevice(device));
		size_t freeB = 0, totalB = 0;
		HIPCHECK(hipMemGetInfo(&freeB, &totalB));
		c->budget = memory_budget_bytes ? memory_budget_bytes : static_cast<uint64_t>(static_cast<double>(freeB) * 0.9);
		HIPCHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
		for (auto& e : c->ev) HIPCHECK(hipEventCreate(&e));
		HIPCHECK(hipHostMalloc(reinterpret_cast<void**>(&c->hostSeg), fh::kSegCounters * sizeof(int64_t)));
		loadLogTable(c);
	}
	catch (const FhError& e)
	{
		rc = e.code;
		// keep the context so the caller can read the message, then fail
		c->err = e.msg;
		*out_ctx = c;
		return rc;
	}
	*out_ctx = c;
	return FLYEHIP_OK;
}

int flyehip_close(flyehip_ctx* c)
{
	if (!c) return FLYEHIP_OK;
	hipSetDevice(c->device);
	if (c->stream) hipStreamSynchronize(c->stream);
	DevBuf* all[] = {&c->dWords, &c->dWordOff, &c->dReadLen, &c->dReadWinOff, &c->dKeys, &c->dVals, &c->dCursor,
					 &c->dEntries, &c->dLogX, &c->dLogY, &c->bQueries, &c->bWinOff, &c->bWinCount, &c->bWinFilt,
					 &c->bWinInfo, &c->bMatchOff, &c->bFiltOff, &c->bScanTmp, &c->bMKey, &c->bMExt, &c->bVal,
					 &c->bPosL, &c->bPosR, &c->bHCur, &c->bHExt, &c->bGroupFlag, &c->bGroupOff,
					 &c->bFiltered, &c->bQMatchOff, &c->bQFiltOff, &c->bQDesc, &c->bGroups, &c->bGroupCount,
					 &c->bRowStart, &c->bRows, &c->bCScore, &c->bCBt, &c->bCProps, &c->bCOut, &c->bCHCur,
					 &c->bCHExt, &c->bCKey, &c->bCVal, &c->bCPosL, &c->bCPosR, &c->bGroupPass, &c->bOrder, &c->bPassOrder, &c->bBins, &c->bSKey, &c->bAVal, &c->bRKey, &c->bDupFlag, &c->bDup, &c->bMsortTemp, &c->bQueue,
					 &c->bSegBig0, &c->bSegBig1, &c->bSegLeaf, &c->bSegHeap, &c->bSegCounters, &c->bSegChunkOff,
					 &c->bSegChunkRange, &c->bSegChunkCnt, &c->bSegChunkPre, &c->bSegRangeM, &c->bSegRangeLm, &c->bSegTemp,
					 &c->bSegVerify, &c->bSegShadowK, &c->bSegShadowV,
					 &c->bCounters, &c->bQRowCount, &c->bCurLen, &c->bDivOff, &c->bDivCount, &c->bDivOut,
					 &c->bQueryIds, &c->bCand, &c->bExtWords, &c->bExtWordOff, &c->bExtLen, &c->bAln, &c->bAlnOff,
					 &c->bRA};
	for (DevBuf* b : all) release(c, *b);
	if (c->hostSeg) hipHostFree(c->hostSeg);
	for (auto& e : c->ev)
		if (e) hipEventDestroy(e);
	if (c->stream) hipStreamDestroy(c->stream);
	delete c;
	return FLYEHIP_OK;
}

const char* flyehip_last_error(const flyehip_ctx* c) { return c ? c->err.c_str() : "null context"; }

int flyehip_get_stats(const flyehip_ctx* c, flyehip_stats_t* out)
{
	if (!c || !out) return FLYEHIP_ERR_ARG;
	*out = c->stats;
	return FLYEHIP_OK;
}

int flyehip_set_tuning(flyehip_ctx* c, int32_t lanes, int32_t qpb, int32_t gpp)
{
	return guarded(c, [&]() {
		if (lanes)
		{
			if (lanes != 1 && lanes != 4 && lanes != 8 && lanes != 16 && lanes != 32 && lanes != 64)
				fail(FLYEHIP_ERR_ARG, "dp_lanes_per_group %d not supported (1, 4, 8, 16, 32, 64)", lanes);
			c->dpLanes = lanes;
		}
		if (qpb)
		{
			if (qpb < 1) fail(FLYEHIP_ERR_ARG, "queries_per_batch must be positive");
			c->queriesPerBatch = qpb;
		}
		if (gpp) c->groupsPerPull = gpp;
	});
}

int flyehip_store_reads(flyehip_ctx* c, const uint8_t* bases, const int64_t* offsets, int32_t n_reads)
{
	return guarded(c, [&]() {
		if (!bases || !offsets || n_reads <= 0) fail(FLYEHIP_ERR_ARG, "empty read store");
		std::vector<uint32_t> words;
		std::vector<int64_t> wordOff;
		std::vector<int32_t> lens;
		packSeqs(bases, offsets, n_reads, words, wordOff, lens);
		c->haveIndex = false;
		c->nReads = n_reads;
		c->readLen = lens;
		c->readWordOff = wordOff;
		upload(c, ensure(c, c->dWords, words.size() * 4), words.data(), words.size() * 4);
		upload(c, ensure(c, c->dWordOff, wordOff.size() * 8), wordOff.data(), wordOff.size() * 8);
		upload(c, ensure(c, c->dReadLen, lens.size() * 4), lens.data(), lens.size() * 4);
		sync(c);
	});
}

int flyehip_build_index(flyehip_ctx* c, const flyehip_index_params_t* ip, flyehip_index_info_t* out)
{
	return guarded(c, [&]() {
		if (!ip) fail(FLYEHIP_ERR_ARG, "null index parameters");
		if (c->nReads <= 0) fail(FLYEHIP_ERR_STATE, "no reads stored");
		const int k = ip->kmer_size;
		if (k < 1 || k > 31) fail(FLYEHIP_ERR_ARG, "kmer_size %d out of range 1..31", k);
		c->k = k;
		c->haveIndex = false;
		std::vector<int64_t> winOff(static_cast<size_t>(c->nReads) + 1, 0);
		for (int32_t i = 0; i < c->nReads; ++i)
			winOff[static_cast<size_t>(i) + 1] = winOff[static_cast<size_t>(i)] + std::max(0, c->readLen[static_cast<size_t>(i)] - k + 1);
		const int64_t nWin = winOff.back();
		int64_t* dWinOff = ensureT<int64_t>(c, c->dReadWinOff, winOff.size());
		upload(c, dWinOff, winOff.data(), winOff.size() * 8);
		uint64_t slots = 1024;
		while (slots < static_cast<uint64_t>(nWin) * 3 / 2 + 1024) slots <<= 1;
		c->tableMask = slots - 1;
		ensure(c, c->dKeys, slots * 8);
		ensure(c, c->dVals, slots * 8);
		ensure(c, c->dCursor, slots * 4);
		fill64(c, static_cast<uint64_t*>(c->dKeys.p), static_cast<int64_t>(slots), kEmptyKey);
		fill64(c, static_cast<uint64_t*>(c->dVals.p), static_cast<int64_t>(slots), 0);
		const PackedSeqs rs = storeSeqs(c);
		HashTable t = table(c);
		HIPCHECK(launchCount(rs, dWinOff, nWin, k, t, c->stream));
		const int hmax = 1 << 16;
		uint64_t* dHist = ensureT<uint64_t>(c, c->bCKey, static_cast<size_t>(hmax) + 8);
		fill64(c, dHist, hmax + 8, 0);
		HIPCHECK(launchHistogram(t, dHist, hmax, dHist + hmax, c->stream));
		std::vector<uint64_t> hist(static_cast<size_t>(hmax) + 8);
		download(c, hist.data(), dHist, hist.size() * 8);
		sync(c);
		uint64_t acc[3] = {hist[static_cast<size_t>(hmax)], hist[static_cast<size_t>(hmax) + 1], hist[static_cast<size_t>(hmax) + 2]};
		hist.resize(static_cast<size_t>(hmax));
		// trim to the last non-zero count so the valley scan matches the oracle's histogram
		size_t last = 0;
		for (size_t i = 0; i < hist.size(); ++i)
			if (hist[i]) last = i;
		hist.resize(last + 2);
		const int solidMin = ip->solid_min > 0 ? ip->solid_min : autoSolidMin(hist);
		const int repMax = ip->rep_max > 0 ? ip->rep_max : autoRepMax(hist, acc, solidMin, ip->repeat_kmer_rate);
		if (repMax >= (1 << kCountBits)) fail(FLYEHIP_ERR_ARG, "rep_max %d exceeds the bucket size limit %d", repMax, (1 << kCountBits) - 1);
		int32_t* sizes = ensureT<int32_t>(c, c->bWinCount, slots + 1);
		uint64_t* counters = ensureT<uint64_t>(c, c->bCounters, 16);
		fill64(c, counters, 16, 0);
		HIPCHECK(launchClassify(t, solidMin, repMax, sizes, counters, c->stream));
		int64_t* starts = ensureT<int64_t>(c, c->bMatchOff, slots + 1);
		int64_t* tmp = ensureT<int64_t>(c, c->bScanTmp, scanTempWords(static_cast<int64_t>(slots)) + 64);
		HIPCHECK(exclusiveScan(sizes, starts, static_cast<int64_t>(slots), tmp, c->stream));
		int64_t nEntries = 0;
		uint64_t cnt[2];
		download(c, &nEntries, starts + slots, 8);
		download(c, cnt, counters, 16);
		sync(c);
		HIPCHECK(launchAssign(t, sizes, starts, c->stream));
		uint64_t* ord = ensureT<uint64_t>(c, c->bMKey, static_cast<size_t>(nEntries) + 1);
		ensure(c, c->dEntries, static_cast<size_t>(nEntries) * 8 + 8);
		HIPCHECK(launchFill(rs, dWinOff, nWin, k, t, ord, c->stream));
		HIPCHECK(launchBucketFinish(t, ord, static_cast<uint64_t*>(c->dEntries.p), static_cast<const int32_t*>(c->dReadLen.p), k, c->stream));
		sync(c);
		c->info = flyehip_index_info_t{};
		c->info.kmer_size = k;
		c->info.solid_min = solidMin;
		c->info.rep_max = repMax;
		c->info.distinct_kmers = static_cast<int64_t>(acc[2]);
		c->info.solid_kmers = static_cast<int64_t>(cnt[0]);
		c->info.repetitive_kmers = static_cast<int64_t>(cnt[1]);
		c->info.indexed_positions = nEntries;
		c->info.hash_slots = static_cast<int64_t>(slots);
		c->info.device_bytes = static_cast<int64_t>(c->dKeys.bytes + c->dVals.bytes + c->dCursor.bytes + c->dEntries.bytes + c->dWords.bytes);
		c->haveIndex = true;
		if (out) *out = c->info;
	});
}

int flyehip_collect_candidates(flyehip_ctx* c, const int64_t* query_ids, int32_t n, flyehip_candidate_t* matches,
							   int64_t match_cap, int64_t* match_offsets, int64_t* matches_needed, int32_t* filtered,
							   int64_t filtered_cap, int64_t* filtered_offsets, int64_t* filtered_needed)
{
	return guarded(c, [&]() {
		std::vector<JoinQuery> qs = residentQueries(c, query_ids, n);
		JoinArgs ja;
		BatchOut bo = runJoin(c, storeSeqs(c), qs, ja);
		std::vector<int64_t> qmo(static_cast<size_t>(n) + 1), qfo(static_cast<size_t>(n) + 1);
		download(c, qmo.data(), c->bQMatchOff.p, qmo.size() * 8);
		download(c, qfo.data(), c->bQFiltOff.p, qfo.size() * 8);
		sync(c);
		if (matches_needed) *matches_needed = bo.matches;
		if (filtered_needed) *filtered_needed = bo.filtered;
		for (int32_t i = 0; i <= n; ++i)
		{
			match_offsets[i] = qmo[static_cast<size_t>(i)];
			filtered_offsets[i] = qfo[static_cast<size_t>(i)];
		}
		if (bo.matches > match_cap || bo.filtered > filtered_cap)
			fail(FLYEHIP_ERR_CAPACITY, "capacity too small: %lld matches / %lld filtered needed",
				 static_cast<long long>(bo.matches), static_cast<long long>(bo.filtered));
		int64_t* dIds = ensureT<int64_t>(c, c->bQueryIds, static_cast<size_t>(n) + 1);
		upload(c, dIds, query_ids, static_cast<size_t>(n) * 8);
		flyehip_candidate_t* dc = ensureT<flyehip_candidate_t>(c, c->bCand, static_cast<size_t>(bo.matches) + 1);
		HIPCHECK(launchWriteCandidates(ja, static_cast<const int64_t*>(c->bQMatchOff.p), dIds, nullptr, dc, bo.matches, c->stream));
		download(c, matches, dc, static_cast<size_t>(bo.matches) * sizeof(flyehip_candidate_t));
		download(c, filtered, ja.filtered, static_cast<size_t>(bo.filtered) * 4);
		sync(c);
	});
}

int flyehip_collect_sorted_candidates(flyehip_ctx* c, const int64_t* query_ids, int32_t n, flyehip_candidate_t* matches,
									  int64_t match_cap, int64_t* match_offsets, int64_t* matches_needed)
{
	return guarded(c, [&]() {
		resetStats(c);
		std::vector<JoinQuery> qs = residentQueries(c, query_ids, n);
		JoinArgs ja;
		BatchOut bo = runJoin(c, storeSeqs(c), qs, ja);
		std::vector<int64_t> qmo(static_cast<size_t>(n) + 1);
		download(c, qmo.data(), c->bQMatchOff.p, qmo.size() * 8);
		sync(c);
		if (matches_needed) *matches_needed = bo.matches;
		for (int32_t i = 0; i <= n; ++i) match_offsets[i] = qmo[static_cast<size_t>(i)];
		if (bo.matches > match_cap)
			fail(FLYEHIP_ERR_CAPACITY, "capacity too small: %lld matches needed", static_cast<long long>(bo.matches));
		runMatchSort(c, ja, bo.matches);
		int64_t* dIds = ensureT<int64_t>(c, c->bQueryIds, static_cast<size_t>(n) + 1);
		upload(c, dIds, query_ids, static_cast<size_t>(n) * 8);
		flyehip_candidate_t* dc = ensureT<flyehip_candidate_t>(c, c->bCand, static_cast<size_t>(bo.matches) + 1);
		JoinArgs sorted = ja;
		sorted.mkey = static_cast<uint64_t*>(c->bSKey.p);
		HIPCHECK(launchWriteCandidates(sorted, static_cast<const int64_t*>(c->bQMatchOff.p), dIds,
									   static_cast<const uint32_t*>(c->bVal.p), dc, bo.matches, c->stream));
		download(c, matches, dc, static_cast<size_t>(bo.matches) * sizeof(flyehip_candidate_t));
		sync(c);
	});
}

int flyehip_overlaps(flyehip_ctx* c, const int64_t* query_ids, int32_t n, const flyehip_overlap_params_t* p,
					 flyehip_overlap_t* rows, int64_t row_cap, int64_t* row_offsets, int64_t* rows_needed,
					 float* div_stats, int64_t div_cap, int64_t* div_offsets)
{
	return guarded(c, [&]() {
		std::vector<JoinQuery> qs = residentQueries(c, query_ids, n);
		overlapsImpl(c, storeSeqs(c), qs, query_ids, p, rows, row_cap, row_offsets, rows_needed, div_stats, div_cap,
					 div_offsets, c->queriesPerBatch);
	});
}

int flyehip_overlaps_per_query(flyehip_ctx* c, const int64_t* query_ids, int32_t n, const flyehip_overlap_params_t* p,
							   flyehip_overlap_t* rows, int64_t row_cap, int64_t* row_offsets, int64_t* rows_needed)
{
	return guarded(c, [&]() {
		std::vector<JoinQuery> qs = residentQueries(c, query_ids, n);
		overlapsImpl(c, storeSeqs(c), qs, query_ids, p, rows, row_cap, row_offsets, rows_needed, nullptr, 0, nullptr, 1);
	});
}

int flyehip_overlaps_external(flyehip_ctx* c, const uint8_t* bases, const int64_t* offsets, const int64_t* query_ids,
							  int32_t n, const flyehip_overlap_params_t* p, flyehip_overlap_t* rows, int64_t row_cap,
							  int64_t* row_offsets, int64_t* rows_needed)
{
	return guarded(c, [&]() {
		if (!c->haveIndex) fail(FLYEHIP_ERR_STATE, "no index: call flyehip_store_reads and flyehip_build_index first");
		std::vector<uint32_t> words;
		std::vector<int64_t> wordOff;
		std::vector<int32_t> lens;
		packSeqs(bases, offsets, n, words, wordOff, lens);
		upload(c, ensure(c, c->bExtWords, words.size() * 4), words.data(), words.size() * 4);
		upload(c, ensure(c, c->bExtWordOff, wordOff.size() * 8), wordOff.data(), wordOff.size() * 8);
		upload(c, ensure(c, c->bExtLen, lens.size() * 4 + 4), lens.data(), lens.size() * 4);
		PackedSeqs s;
		s.words = static_cast<const uint32_t*>(c->bExtWords.p);
		s.wordOff = static_cast<const int64_t*>(c->bExtWordOff.p);
		s.len = static_cast<const int32_t*>(c->bExtLen.p);
		s.n = n;
		std::vector<JoinQuery> qs(static_cast<size_t>(n));
		for (int32_t i = 0; i < n; ++i)
		{
			const int64_t id = query_ids[i];
			if (id == 0) fail(FLYEHIP_ERR_ARG, "query id 0 is not a Flye id");
			const int64_t r = id > 0 ? id : -id;
			JoinQuery q;
			q.seq = i;
			q.strand = 1;
			q.iid = toInternal(id);
			q.selfMode = (r <= c->nReads) ? 2 : 0;
			q.len = lens[static_cast<size_t>(i)];
			q.reserved = 0;
			qs[static_cast<size_t>(i)] = q;
		}
		overlapsImpl(c, s, qs, query_ids, p, rows, row_cap, row_offsets, rows_needed, nullptr, 0, nullptr, c->queriesPerBatch);
	});
}

int flyehip_chain_lists(flyehip_ctx* c, const int32_t* cur_pos, const int32_t* ext_pos, const int64_t* list_offsets,
						int32_t n_lists, const int64_t* cur_ids, const int64_t* ext_ids, const int32_t* cur_lens,
						const int32_t* ext_lens, const flyehip_overlap_params_t* p, flyehip_overlap_t* rows,
						int64_t row_cap, int64_t* row_offsets, int64_t* rows_needed)
{
	return guarded(c, [&]() {
		checkParams(p, 0);
		resetStats(c);
		if (!c->dLogX.p) fail(FLYEHIP_ERR_STATE, "context not initialised");
		Timer total;
		const int64_t nHits = list_offsets[n_lists] - list_offsets[0];
		const int64_t h0 = list_offsets[0];
		// groups: one per non-empty list
		std::vector<GroupDesc> gs;
		std::vector<QueryDesc> qd(static_cast<size_t>(n_lists));
		for (int32_t i = 0; i < n_lists; ++i)
		{
			const int64_t b = list_offsets[i] - h0, e = list_offsets[i + 1] - h0;
			if (e < b) fail(FLYEHIP_ERR_ARG, "list %d has negative length", i);
			qd[static_cast<size_t>(i)] = QueryDesc{cur_ids[i], 0, cur_lens[i], 0};
			if (e == b) continue;
			if (e - b > 0x7fffffff) fail(FLYEHIP_ERR_ARG, "list %d too long", i);
			GroupDesc g{};
			g.hitStart = b;
			g.extId = ext_ids[i];
			g.n = static_cast<int32_t>(e - b);
			g.query = i;
			g.extLen = ext_lens[i];
			gs.push_back(g);
		}
		const size_t nm = static_cast<size_t>(nHits) + 1;
		int32_t* dc = ensureT<int32_t>(c, c->bHCur, nm);
		int32_t* de = ensureT<int32_t>(c, c->bHExt, nm);
		upload(c, dc, cur_pos + h0, static_cast<size_t>(nHits) * 4);
		upload(c, de, ext_pos + h0, static_cast<size_t>(nHits) * 4);
		GroupDesc* dg = ensureT<GroupDesc>(c, c->bGroups, gs.size() + 1);
		upload(c, dg, gs.data(), gs.size() * sizeof(GroupDesc));
		QueryDesc* dq = ensureT<QueryDesc>(c, c->bQDesc, qd.size() + 1);
		upload(c, dq, qd.data(), qd.size() * sizeof(QueryDesc));
		int32_t* dFilt = ensureT<int32_t>(c, c->bFiltered, 4);
		ChainArgs ca{};
		ca.params = *p;
		ca.groups = dg;
		ca.queries = dq;
		ca.order = nullptr;
		ca.nGroups = static_cast<int32_t>(gs.size());
		ca.logN = c->logN;
		ca.cur = dc;
		ca.ext = de;
		ca.filtered = dFilt;
		ca.sortKey = ensureT<uint32_t>(c, c->bCKey, nm);
		ca.sortVal = ensureT<uint32_t>(c, c->bCVal, nm);
		ca.posL = ensureT<int32_t>(c, c->bCPosL, nm);
		ca.posR = ensureT<int32_t>(c, c->bCPosR, nm);
		ca.hcur = ensureT<int32_t>(c, c->bCHCur, nm);
		ca.hext = ensureT<int32_t>(c, c->bCHExt, nm);
		ca.score = ensureT<int32_t>(c, c->bCScore, nm);
		ca.bt = ensureT<int32_t>(c, c->bCBt, nm);
		ca.props = ensureT<Proposal>(c, c->bCProps, nm);
		ca.outProps = ensureT<Proposal>(c, c->bCOut, nm);
		ca.groupCount = ensureT<int32_t>(c, c->bGroupCount, gs.size() + 1);
		ca.groupPass = ensureT<int32_t>(c, c->bGroupPass, gs.size() + 1);
		ca.logX = static_cast<const uint32_t*>(c->dLogX.p);
		ca.logY = static_cast<const uint32_t*>(c->dLogY.p);
		runChainKernels(c, ca, nHits);
		int64_t* rowStart = ensureT<int64_t>(c, c->bRowStart, gs.size() + 1);
		int64_t* tmp = ensureT<int64_t>(c, c->bScanTmp, scanTempWords(static_cast<int64_t>(gs.size())) + 64);
		HIPCHECK(exclusiveScan(ca.groupCount, rowStart, static_cast<int64_t>(gs.size()), tmp, c->stream));
		int64_t nRows = 0;
		download(c, &nRows, rowStart + gs.size(), 8);
		sync(c);
		flyehip_overlap_t* dr = ensureT<flyehip_overlap_t>(c, c->bRows, static_cast<size_t>(nRows) + 1);
		HIPCHECK(launchWriteRows(ca, rowStart, dr, c->stream));
		std::vector<int32_t> gc(gs.size());
		download(c, gc.data(), ca.groupCount, gs.size() * 4);
		if (nRows <= row_cap) download(c, rows, dr, static_cast<size_t>(nRows) * sizeof(flyehip_overlap_t));
		sync(c);
		if (rows_needed) *rows_needed = nRows;
		std::vector<int64_t> per(static_cast<size_t>(n_lists), 0);
		for (size_t g = 0; g < gs.size(); ++g) per[static_cast<size_t>(gs[g].query)] += gc[g];
		row_offsets[0] = 0;
		for (int32_t i = 0; i < n_lists; ++i) row_offsets[i + 1] = row_offsets[i] + per[static_cast<size_t>(i)];
		c->stats.groups = static_cast<int64_t>(gs.size());
		c->stats.matches = nHits;
		c->stats.overlaps = nRows;
		c->stats.ms_total = total.ms();
		c->stats.device_bytes_peak = static_cast<int64_t>(c->peak);
		if (nRows > row_cap) fail(FLYEHIP_ERR_CAPACITY, "row capacity %lld too small: %lld needed", static_cast<long long>(row_cap), static_cast<long long>(nRows));
	});
}

int flyehip_divergence_logf(flyehip_ctx* c, const float* x, int64_t n, float* out)
{
	return guarded(c, [&]() {
		if (n < 0 || (n > 0 && (!x || !out))) fail(FLYEHIP_ERR_ARG, "invalid logf arguments");
		const int64_t chunk = 1 << 26;
		float* dx = ensureT<float>(c, c->bDivOut, static_cast<size_t>(std::min(n, chunk)) + 1);
		float* dy = ensureT<float>(c, c->bCKey, static_cast<size_t>(std::min(n, chunk)) + 1);
		for (int64_t b = 0; b < n; b += chunk)
		{
			const int64_t m = std::min(chunk, n - b);
			upload(c, dx, x + b, static_cast<size_t>(m) * 4);
			HIPCHECK(launchLogf(dx, m, dy, static_cast<const uint32_t*>(c->dLogX.p),
								static_cast<const uint32_t*>(c->dLogY.p), c->logN, c->stream));
			download(c, out + b, dy, static_cast<size_t>(m) * 4);
			sync(c);
		}
	});
}

int flyehip_read_align_chains(flyehip_ctx* c, const flyehip_edge_alignment_t* alignments, const int64_t* aln_offsets,
							  int32_t n_reads, int32_t long_edge, int32_t big_aln, const flyehip_chain_params_t* params,
							  int32_t* sorted_index, int64_t* kept_offsets, flyehip_chain_segment_t* segments,
							  int64_t segment_cap, int64_t* segment_offsets, int64_t* segments_needed)
{
	return guarded(c, [&]() {
		if (!params) fail(FLYEHIP_ERR_ARG, "null chain parameters");
		resetStats(c);
		const int64_t m = aln_offsets[n_reads] - aln_offsets[0];
		const int64_t a0 = aln_offsets[0];
		std::vector<int64_t> off(static_cast<size_t>(n_reads) + 1);
		for (int32_t r = 0; r <= n_reads; ++r) off[static_cast<size_t>(r)] = aln_offsets[r] - a0;
		const size_t nm = static_cast<size_t>(m) + 1;
		flyehip_edge_alignment_t* dA = ensureT<flyehip_edge_alignment_t>(c, c->bAln, nm);
		int64_t* dOff = ensureT<int64_t>(c, c->bAlnOff, off.size());
		upload(c, dA, alignments + a0, static_cast<size_t>(m) * sizeof(flyehip_edge_alignment_t));
		upload(c, dOff, off.data(), off.size() * 8);
		ReadAlignArgs ra{};
		ra.aln = dA;
		ra.alnOff = dOff;
		ra.nReads = n_reads;
		ra.longEdge = long_edge;
		ra.bigAln = big_aln;
		ra.p = *params;
		ra.kept = ensureT<int32_t>(c, c->bCPosL, nm);
		ra.key = ensureT<uint64_t>(c, c->bCKey, nm);
		ra.val = ensureT<uint32_t>(c, c->bCVal, nm);
		ra.chainScore = ensureT<int32_t>(c, c->bCScore, nm);
		ra.chainPrev = ensureT<int32_t>(c, c->bCBt, nm);
		ra.chainLen = ensureT<int32_t>(c, c->bCPosR, nm);
		ra.active = ensureT<int32_t>(c, c->bCHCur, nm);
		ra.frozen = ensureT<int32_t>(c, c->bCHExt, nm);
		ra.accepted = ensureT<int32_t>(c, c->bHCur, nm);
		int32_t* counts = ensureT<int32_t>(c, c->bRA, 3 * static_cast<size_t>(n_reads) + 3);
		ra.keptCount = counts;
		ra.acceptedCount = counts + n_reads + 1;
		ra.segCount = counts + 2 * (n_reads + 1);
		Timer tc;
		HIPCHECK(launchReadAlignChain(ra, c->stream));
		int64_t* segOff = ensureT<int64_t>(c, c->bRowStart, static_cast<size_t>(n_reads) + 1);
		int64_t* tmp = ensureT<int64_t>(c, c->bScanTmp, scanTempWords(n_reads) + 64);
		HIPCHECK(exclusiveScan(ra.segCount, segOff, n_reads, tmp, c->stream));
		std::vector<int64_t> so(static_cast<size_t>(n_reads) + 1);
		std::vector<int32_t> kc(static_cast<size_t>(n_reads));
		download(c, so.data(), segOff, so.size() * 8);
		download(c, kc.data(), ra.keptCount, kc.size() * 4);
		sync(c);
		const int64_t nSeg = so.back();
		if (segments_needed) *segments_needed = nSeg;
		flyehip_chain_segment_t* dS = ensureT<flyehip_chain_segment_t>(c, c->bRows, static_cast<size_t>(nSeg) + 1);
		HIPCHECK(launchReadAlignWrite(ra, segOff, dS, c->stream));
		std::vector<int32_t> keptAll(static_cast<size_t>(m) + 1);
		download(c, keptAll.data(), ra.kept, static_cast<size_t>(m) * 4);
		if (nSeg <= segment_cap) download(c, segments, dS, static_cast<size_t>(nSeg) * sizeof(flyehip_chain_segment_t));
		sync(c);
		c->stats.ms_chain = tc.ms();
		int64_t kw = 0;
		kept_offsets[0] = 0;
		for (int32_t r = 0; r < n_reads; ++r)
		{
			for (int32_t i = 0; i < kc[static_cast<size_t>(r)]; ++i) sorted_index[kw++] = keptAll[static_cast<size_t>(off[static_cast<size_t>(r)] + i)];
			kept_offsets[r + 1] = kw;
		}
		for (int32_t r = 0; r <= n_reads; ++r) segment_offsets[r] = so[static_cast<size_t>(r)];
		c->stats.queries = n_reads;
		c->stats.matches = m;
		c->stats.overlaps = nSeg;
		if (nSeg > segment_cap) fail(FLYEHIP_ERR_CAPACITY, "segment capacity %lld too small: %lld needed", static_cast<long long>(segment_cap), static_cast<long long>(nSeg));
	});
}

} // extern "C"
