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
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <condition_variable>
#include <mutex>
#include <exception>
#include <new>
#include <string>
#include <thread>
#include <deque>
#include <unordered_map>
#include <vector>

#include "chain.hpp"
#include "dev_common.hpp"
#include "msort.hpp"
#include "pipeline.hpp"
#include "readalign.hpp"
#include "ktimer.hpp"
#include "segsort.hpp"
#include "wire.hpp"

extern "C" const unsigned char fh_logf_table[];
extern "C" const unsigned char fh_logf_table_end[];

// ---- no-allocation scratch ------------------------------------------------
// flyehip_read_align_chains_noalloc keeps every intermediate buffer in one
// file-scope device array, so a call allocates nothing.  Limits: 2^20
// alignments (aln_offsets[n_reads]) and 2^17 reads per call; the array
// (about 43 MiB) is held on every device the library runs on until the
// process ends.  All calls on one device share it: calls that overlap on one
// device race.  Every word a call reads is written by the same call first
// (per-read counts, scans, per-alignment chain state), so nothing needs
// clearing between calls.  Host code never passes `scratch` itself to a
// kernel: it uses the device address hipGetSymbolAddress returns for the
// current device (scratchOn).
namespace fh
{
constexpr int64_t kNoAllocMaxAlignments = int64_t{1} << 20;
constexpr int32_t kNoAllocMaxReads = 1 << 17;
constexpr size_t na256(size_t x) { return (x + 255) & ~static_cast<size_t>(255); }
constexpr size_t kNaI32 = na256(static_cast<size_t>(kNoAllocMaxAlignments) * 4);    // one int32 per alignment
constexpr size_t kNaU64 = na256(static_cast<size_t>(kNoAllocMaxAlignments) * 8);    // sort keys
constexpr size_t kNaRead = na256((static_cast<size_t>(kNoAllocMaxReads) + 1) * 4);  // one int32 per read
constexpr size_t kNaScan = na256((static_cast<size_t>(kNoAllocMaxReads) / 4096 + 2 + 64) * 8);
constexpr size_t kNoAllocScratchBytes = kNaU64 + 8 * kNaI32 + 3 * kNaRead + kNaScan;
} // namespace fh

static __device__ __align__(256) unsigned char scratch[fh::kNoAllocScratchBytes];

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

// Keeps device allocation out of the region where several lanes have work in
// flight.  A lane holds the gate (enter .. leave) while it runs a batch; a
// lane whose buffer must grow drains its own streams, gives up its hold and
// waits until no lane holds the gate (every other lane is between batches,
// its streams idle), allocates, and takes its hold back.  Waiting writers keep
// new batches from starting, so a growth waits for at most one batch per
// other lane.
struct AllocGate
{
	std::mutex m;
	std::condition_variable cv;
	int active = 0;   // lanes holding the gate
	int writers = 0;  // lanes waiting to allocate or allocating
	bool writing = false;

	void enter()
	{
		std::unique_lock<std::mutex> l(m);
		cv.wait(l, [&] { return writers == 0; });
		++active;
	}
	void leave()
	{
		std::lock_guard<std::mutex> l(m);
		--active;
		cv.notify_all();
	}
	void growBegin()
	{
		std::unique_lock<std::mutex> l(m);
		--active;
		++writers;
		cv.notify_all();
		cv.wait(l, [&] { return active == 0 && !writing; });
		writing = true;
	}
	void growEnd()
	{
		std::unique_lock<std::mutex> l(m);
		writing = false;
		--writers;
		cv.notify_all();
		cv.wait(l, [&] { return writers == 0; });
		++active;
	}
};

// One batch pipeline: a stream (plus a side stream) and the grow-only batch
// buffers.  Batches on different lanes run concurrently; they share the read
// store, the index and the memory budget of the context.
struct Lane
{
	flyehip_ctx* c = nullptr;
	hipStream_t stream = nullptr;
	hipStream_t side = nullptr;  // side stream (default priority)
	hipEvent_t sideEv[2] = {};   // fork, done (timing disabled)
	hipEvent_t ev[8] = {};
	int64_t* hostSeg = nullptr;  // pinned counter words of the segmented sort
	flyehip_stats_t stats{};     // stage timings and counts of this lane's batches
	KTimer timer;                // kernel-family intervals of the batch being enqueued
	DevBuf bQueries, bWinOff, bWinCount, bWinFilt, bWinInfo, bMatchOff, bFiltOff, bScanTmp;
	DevBuf bMKey, bMExt, bVal, bPosL, bPosR, bHCur, bHExt, bGroupFlag, bGroupOff;
	DevBuf bFiltered, bQMatchOff, bQFiltOff, bQDesc, bGroups, bGroupCount, bRowStart, bRows;
	DevBuf bCScore, bCBt, bCProps, bCOut, bCHCur, bCHExt, bCKey, bCVal, bCPosL, bCPosR;
	DevBuf bGroupDefer, bFlagReady, bFlagDefer, bPosReady, bPosDefer, bReadyOrder, bDeferOrder;
	DevBuf bClaimFlag, bFlagOrder, bGroupRel, bGatherCnt, bGatherOff;
	DevBuf bGroupPass, bOrder, bPassOrder, bBins, bSKey, bAVal, bRKey, bDupFlag, bDup, bMsortTemp;
	DevBuf bSegBig0, bSegBig1, bSegLeaf, bSegHeap, bSegCounters, bSegChunkOff, bSegChunkRange, bSegChunkCnt;
	DevBuf bSegChunkPre, bSegRangeM, bSegRangeLm, bSegTemp, bSegVerify, bSegShadowK, bSegShadowV;
	DevBuf bQueue, bCounters, bQRowCount, bCurLen, bDivOff, bDivCount, bDivOut, bQueryIds, bCand;
	DevBuf bExtWords, bExtWordOff, bExtLen;
	DevBuf bAln, bAlnOff, bRA;
	DevBuf bRawBases, bRawOff, bPackErr;  // ASCII bases being packed (packSeqsDevice)
	DevBuf bGroupStart;                   // group starts of a batch's hits (score keys)
	DevBuf bMinMask;                      // minimizer flags of the index build (one byte per window)
	DevBuf bQOrder;                       // queries by descending match count (stable match sort)
};

// One flyehip_overlaps_submit call: its query records and, as batches finish,
// their rows in query order (a prefix of `done` queries is complete).
struct AsyncJob
{
	std::vector<int64_t> ids;
	flyehip_overlap_params_t p{};
	bool withDiv = false;
	int32_t done = 0;              // queries whose results are in place
	int code = FLYEHIP_OK;         // first error (then every query is done)
	std::string msg;
	std::vector<flyehip_overlap_t> rows;
	std::vector<int64_t> rowOff{0};
	std::vector<float> divs;
	std::vector<int64_t> divOff{0};
};

// Submitted jobs, run in submission order by one worker thread; collectors
// wait on the condition variable for their query's batch.  Lock order: the
// worker takes the context mutex for a chunk and this mutex afterwards,
// never both at once.
struct AsyncQueue
{
	std::mutex m;
	std::condition_variable cv;
	std::deque<std::shared_ptr<AsyncJob>> pending;
	std::unordered_map<uint64_t, std::shared_ptr<AsyncJob>> jobs;
	std::thread worker;
	bool stop = false;
	uint64_t nextTicket = 1;
};

} // namespace
} // namespace fh

struct flyehip_ctx
{
	int device = 0;
	uint64_t budget = 0;
	uint64_t inUse = 0;
	uint64_t peak = 0;
	std::mutex memMu;  // budget accounting (lanes allocate from their own threads)
	fh::AllocGate gate;           // allocation while several lanes run
	int lanesRunning = 0;         // lanes of the current multi-lane call
	std::string err;
	std::mutex mu;
	std::vector<std::unique_ptr<fh::Lane>> lanes;  // lanes[0] serves every single-stream call

	// tuning
	int dpLanes = 8;
	int queriesPerBatch = 2048;
	int batchesInFlight = 2;
	int numCUs = 0;

	// read store
	int32_t nReads = 0;
	std::vector<int32_t> readLen;
	std::vector<int64_t> readWordOff;
	fh::DevBuf dWords, dWordOff, dReadLen, dReadWinOff;

	// index
	bool haveIndex = false;
	bool indexLoaded = false;  // buckets given by the caller (flyehip_load_index), in the caller's order
	int k = 0;
	flyehip_index_info_t info{};
	fh::DevBuf dKeys, dCursor, dEntries;  // dKeys: (key, value) slot pairs
	uint64_t tableMask = 0;

	// logf table
	fh::DevBuf dLogX, dLogY;
	int logN = 0;

	// device address of this device's no-allocation scratch array
	unsigned char* scratch = nullptr;

	// flyehip_overlaps_submit / _collect (started on the first submit)
	std::unique_ptr<fh::AsyncQueue> async;

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
		(void)hipFree(b.p);
		std::lock_guard<std::mutex> lock(c->memMu);
		c->inUse -= b.bytes;
		b.p = nullptr;
		b.bytes = 0;
	}
}

// The lane the calling host thread drives inside a multi-lane call.
thread_local Lane* tlsLane = nullptr;

// Waits for everything a lane has enqueued (its stream and side stream).
void quiesce(Lane* L)
{
	HIPCHECK(hipStreamSynchronize(L->stream));
	HIPCHECK(hipStreamSynchronize(L->side));
}

void* ensure(flyehip_ctx* c, DevBuf& b, size_t bytes)
{
	if (bytes == 0) bytes = 16;
	if (b.bytes >= bytes) return b.p;
	// inside a multi-lane call: allocate only while no other lane has work
	// in flight (AllocGate)
	Lane* const L = tlsLane;
	const bool gated = L != nullptr && c->lanesRunning > 1;
	if (gated)
	{
		quiesce(L);
		c->gate.growBegin();
	}
	struct Regate
	{
		flyehip_ctx* c;
		bool on;
		~Regate()
		{
			if (on) c->gate.growEnd();
		}
	} regate{c, gated};
	release(c, b);
	// headroom for a somewhat larger next batch: 1/8, 1/32 from 1 GiB up
	// (the per-match buffers, where 1/8 of three lanes is tens of GB)
	size_t want = bytes + (bytes >= (size_t{1} << 30) ? bytes / 32 : bytes / 8);
	want = (want + 255) & ~static_cast<size_t>(255);
	{
		// reserve under the lock, allocate outside it
		std::lock_guard<std::mutex> lock(c->memMu);
		if (c->inUse + want > c->budget)
			fail(FLYEHIP_ERR_BUDGET,
				 "device memory budget exceeded: need %zu more bytes, %llu in use, budget %llu (device %d)", want,
				 static_cast<unsigned long long>(c->inUse), static_cast<unsigned long long>(c->budget), c->device);
		c->inUse += want;
		c->peak = std::max(c->peak, c->inUse);
	}
	const hipError_t e = hipMalloc(&b.p, want);
	if (e != hipSuccess)
	{
		b.p = nullptr;
		std::lock_guard<std::mutex> lock(c->memMu);
		c->inUse -= want;
	}
	HIPCHECK(e);
	b.bytes = want;
	return b.p;
}

template <class T>
T* ensureT(flyehip_ctx* c, DevBuf& b, size_t n)
{
	return static_cast<T*>(ensure(c, b, n * sizeof(T)));
}

void upload(Lane* L, void* dst, const void* src, size_t bytes)
{
	if (bytes) HIPCHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, L->stream));
}
void download(Lane* L, void* dst, const void* src, size_t bytes)
{
	if (bytes) HIPCHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, L->stream));
}
void sync(Lane* L) { HIPCHECK(hipStreamSynchronize(L->stream)); }

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
// 2-bit packing of sequences on the device: the host computes the word offsets and lengths
// (one pass over the sequences, not the bases), the ASCII bytes cross PCIe as
// they are and packBasesKernel packs them; only a failed check goes back to
// the host copy to name the first byte other than A/C/G/T.
void packSeqsDevice(flyehip_ctx* c, Lane* L, const uint8_t* bases, const int64_t* offsets, int32_t n,
					DevBuf& dWords, DevBuf& dWordOff, DevBuf& dLens, std::vector<int64_t>& wordOff,
					std::vector<int32_t>& lens)
{
	wordOff.assign(static_cast<size_t>(n) + 1, 0);
	lens.assign(static_cast<size_t>(n), 0);
	int64_t total = 0;
	for (int32_t i = 0; i < n; ++i)
	{
		const int64_t len = offsets[i + 1] - offsets[i];
		if (len < 0 || len > 0x7fffffff) fail(FLYEHIP_ERR_ARG, "sequence %d has invalid length %lld", i, static_cast<long long>(len));
		wordOff[static_cast<size_t>(i)] = total;
		lens[static_cast<size_t>(i)] = static_cast<int32_t>(len);
		total += (len + 15) / 16 + 2;
	}
	wordOff[static_cast<size_t>(n)] = total;
	const int64_t nb = offsets[n] - offsets[0];
	uint8_t* dB = ensureT<uint8_t>(c, L->bRawBases, static_cast<size_t>(nb) + 1);
	int64_t* dBO = ensureT<int64_t>(c, L->bRawOff, static_cast<size_t>(n) + 1);
	uint32_t* dErr = ensureT<uint32_t>(c, L->bPackErr, 1);
	uint32_t* dW = ensureT<uint32_t>(c, dWords, static_cast<size_t>(total));
	int64_t* dWO = ensureT<int64_t>(c, dWordOff, static_cast<size_t>(n) + 1);
	int32_t* dL = ensureT<int32_t>(c, dLens, static_cast<size_t>(n) + 1);
	upload(L, dB, bases + offsets[0], static_cast<size_t>(nb));
	upload(L, dBO, offsets, (static_cast<size_t>(n) + 1) * 8);
	upload(L, dWO, wordOff.data(), wordOff.size() * 8);
	upload(L, dL, lens.data(), lens.size() * 4);
	HIPCHECK(hipMemsetAsync(dErr, 0, 4, L->stream));
	HIPCHECK(launchPackBases(dB, dBO, dWO, dL, n, total, dW, dErr, L->stream));
	uint32_t err = 0;
	download(L, &err, dErr, 4);
	sync(L);
	if (err)
		for (int32_t i = 0; i < n; ++i)
			for (int64_t j = offsets[i]; j < offsets[i + 1]; ++j)
				if (baseCode(bases[j]) < 0)
					fail(FLYEHIP_ERR_ARG, "sequence %d has a non-ACGT byte 0x%02x at %lld", i, bases[j],
						 static_cast<long long>(j - offsets[i]));
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
	Lane* L = c->lanes[0].get();
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
	upload(L, ensure(c, c->dLogX, n * 4), xs.data(), n * 4);
	upload(L, ensure(c, c->dLogY, n * 4), ys.data(), n * 4);
	c->logN = static_cast<int>(n);
	sync(L);
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

// Four words per thread: one 16-byte store where p is 16-byte aligned.
__global__ void zeroKernel(int32_t* p, int64_t n)
{
	const int64_t j = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) * 4;
	if (j >= n) return;
	if (j + 4 <= n && (reinterpret_cast<uintptr_t>(p) & 15) == 0)
	{
		*reinterpret_cast<int4*>(p + j) = make_int4(0, 0, 0, 0);
		return;
	}
	for (int64_t k = j; k < n && k < j + 4; ++k) p[k] = 0;
}

__global__ void fillU64Kernel(uint64_t* p, int64_t n, uint64_t v)
{
	const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
	if (i < n) p[i] = v;
}

// n (key, value) pairs set to (k, v): an empty hash table
__global__ void fillPairsKernel(uint64_t* p, int64_t n, uint64_t k, uint64_t v)
{
	const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
	if (i < n) *reinterpret_cast<ulonglong2*>(p + 2 * i) = make_ulonglong2(k, v);
}

dim3 grid(int64_t n, int tpb) { return dim3(static_cast<unsigned>(std::max<int64_t>(1, (n + tpb - 1) / tpb))); }

void zero32(Lane* L, int32_t* p, int64_t n)
{
	if (n <= 0) return;
	hipLaunchKernelGGL(zeroKernel, grid((n + 3) / 4, 256), dim3(256), 0, L->stream, p, n);
	HIPCHECK(hipGetLastError());
}

void fill64(Lane* L, uint64_t* p, int64_t n, uint64_t v)
{
	if (n <= 0) return;
	hipLaunchKernelGGL(fillU64Kernel, grid(n, 256), dim3(256), 0, L->stream, p, n, v);
	HIPCHECK(hipGetLastError());
}

void emptyTable(Lane* L, uint64_t* slots, int64_t n)
{
	hipLaunchKernelGGL(fillPairsKernel, grid(n, 256), dim3(256), 0, L->stream, slots, n, kEmptyKey, uint64_t{0});
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
	t.slots = static_cast<uint64_t*>(c->dKeys.p);
	t.cursor = static_cast<uint32_t*>(c->dCursor.p);
	t.mask = c->tableMask;
	return t;
}

// Single-wave blocks of the chain's work-queue kernels: 16 per CU (half the
// CU's wave slots at their register use) for one lane; with several lanes
// running, 24 per CU shared among them (at least 4 each), so two lanes' DPs
// leave a quarter of the slots and 64 KiB of LDS per CU to the other kernels
// in flight, above all the deferred heap sorts, whose 64 KiB blocks would
// otherwise wait for the DPs to end.
int chainBlocks(flyehip_ctx* c)
{
	if (c->numCUs == 0)
	{
		hipDeviceProp_t prop;
		HIPCHECK(hipGetDeviceProperties(&prop, c->device));
		c->numCUs = prop.multiProcessorCount;
	}
	const int lanes = c->lanesRunning;
	return c->numCUs * (lanes > 1 ? std::max(4, 24 / lanes) : 16);
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
// withVal: the fill writes every match's emission index (mval); without it
// the caller fills mval where it needs it (the tie replay's queries).
BatchOut runJoin(Lane* L, const PackedSeqs& seqs, const std::vector<JoinQuery>& qs, JoinArgs& ja, bool withVal = true)
{
	flyehip_ctx* c = L->c;
	const int32_t nq = static_cast<int32_t>(qs.size());
	const int k = c->k;
	std::vector<int64_t> winOff(static_cast<size_t>(nq) + 1, 0);
	for (int32_t i = 0; i < nq; ++i) winOff[static_cast<size_t>(i) + 1] = winOff[static_cast<size_t>(i)] + std::max(0, qs[static_cast<size_t>(i)].len - k + 1);
	const int64_t nWin = winOff.back();
	JoinQuery* dQ = ensureT<JoinQuery>(c, L->bQueries, qs.size());
	int64_t* dWinOff = ensureT<int64_t>(c, L->bWinOff, winOff.size());
	upload(L, dQ, qs.data(), qs.size() * sizeof(JoinQuery));
	upload(L, dWinOff, winOff.data(), winOff.size() * sizeof(int64_t));
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
	ja.winCount = ensureT<int32_t>(c, L->bWinCount, static_cast<size_t>(nWin) + 1);
	ja.winFilt = ensureT<int32_t>(c, L->bWinFilt, static_cast<size_t>(nWin) + 1);
	ja.winVal = ensureT<uint64_t>(c, L->bWinInfo, static_cast<size_t>(nWin) + 1);
	int64_t* matchOff = ensureT<int64_t>(c, L->bMatchOff, static_cast<size_t>(nWin) + 1);
	int64_t* filtOff = ensureT<int64_t>(c, L->bFiltOff, static_cast<size_t>(nWin) + 1);
	int64_t* tmp = ensureT<int64_t>(c, L->bScanTmp, scanTempWords(nWin) + 64);
	ja.matchOff = matchOff;
	ja.filtOff = filtOff;
	// join bytes: 2-bit read (L / 4 per query) + one 8-byte slot probe per
	// window; the fill adds 8 bytes per bucket entry read and 16 per match
	// written (u64 key, i32 extPos, u32 emission index)
	double seqBytes = 0.0;
	for (const auto& q : qs) seqBytes += static_cast<double>((q.len + 3) / 4);
	HIPCHECK(hipEventRecord(L->ev[0], L->stream));
	const int tj = L->timer.begin(FLYEHIP_K_JOIN_COUNT, L->stream, seqBytes + 8.0 * static_cast<double>(nWin));
	HIPCHECK(launchJoinCount(ja, L->stream));
	L->timer.end(tj, L->stream);
	HIPCHECK(hipEventRecord(L->ev[1], L->stream));
	HIPCHECK(exclusiveScan(ja.winCount, matchOff, nWin, tmp, L->stream));
	HIPCHECK(exclusiveScan(ja.winFilt, filtOff, nWin, tmp, L->stream));
	// per-query offsets
	int64_t* qMatchOff = ensureT<int64_t>(c, L->bQMatchOff, static_cast<size_t>(nq) + 1);
	int64_t* qFiltOff = ensureT<int64_t>(c, L->bQFiltOff, static_cast<size_t>(nq) + 1);
	hipLaunchKernelGGL(gatherAtKernel, grid(nq + 1, 256), dim3(256), 0, L->stream, matchOff, dWinOff, qMatchOff, nq);
	hipLaunchKernelGGL(gatherAtKernel, grid(nq + 1, 256), dim3(256), 0, L->stream, filtOff, dWinOff, qFiltOff, nq);
	HIPCHECK(hipGetLastError());
	int64_t tot[2];
	download(L, &tot[0], matchOff + nWin, sizeof(int64_t));
	download(L, &tot[1], filtOff + nWin, sizeof(int64_t));
	sync(L);
	BatchOut bo;
	bo.matches = tot[0];
	bo.filtered = tot[1];
	ja.mkey = ensureT<uint64_t>(c, L->bMKey, static_cast<size_t>(bo.matches) + 1);
	ja.mext = ensureT<int32_t>(c, L->bMExt, static_cast<size_t>(bo.matches) + 1);
	ja.mval = ensureT<uint32_t>(c, L->bAVal, static_cast<size_t>(bo.matches) + 1);
	ja.qMatchOff = qMatchOff;
	ja.filtered = ensureT<int32_t>(c, L->bFiltered, static_cast<size_t>(bo.filtered) + 1);
	HIPCHECK(hipEventRecord(L->ev[6], L->stream));
	// per match: index entry read, key + extPos (+ emission index) written
	const int tf = L->timer.begin(FLYEHIP_K_JOIN_FILL, L->stream,
								  seqBytes + 8.0 * static_cast<double>(nWin) +
									  (withVal ? 24.0 : 20.0) * static_cast<double>(bo.matches));
	JoinArgs fa = ja;
	if (!withVal) fa.mval = nullptr;
	HIPCHECK(launchJoinFill(fa, L->stream));
	L->timer.end(tf, L->stream);
	HIPCHECK(hipEventRecord(L->ev[7], L->stream));
	HIPCHECK(hipEventSynchronize(L->ev[7]));
	float e1 = 0, e2 = 0;
	HIPCHECK(hipEventElapsedTime(&e1, L->ev[0], L->ev[1]));
	HIPCHECK(hipEventElapsedTime(&e2, L->ev[6], L->ev[7]));
	L->stats.ms_join += e1 + e2;
	L->stats.windows += nWin;
	L->stats.matches += bo.matches;
	return bo;
}

// Exact per-query std::sort of the batch's matches; the sorted keys land in
// bSKey and the permutation (emission index within the query) in bVal.
// Buffers of the segmented exact sort for n elements in nArrays arrays; the
// caller sets key / val / outVal / dup / posL / posR.
SegSortBufs segBufs(Lane* L, int64_t n, int64_t nArrays)
{
	flyehip_ctx* c = L->c;
	SegSortCaps caps{};
	HIPCHECK(segSortCaps(n, nArrays, &caps));
	SegSortBufs b{};
	b.big[0] = ensureT<SegRange>(c, L->bSegBig0, static_cast<size_t>(caps.bigCap));
	b.big[1] = ensureT<SegRange>(c, L->bSegBig1, static_cast<size_t>(caps.bigCap));
	b.leaf = ensureT<SegRange>(c, L->bSegLeaf, static_cast<size_t>(caps.leafCap));
	b.heap = ensureT<SegRange>(c, L->bSegHeap, static_cast<size_t>(caps.heapCap));
	b.verify = ensureT<SegRange>(c, L->bSegVerify, static_cast<size_t>(caps.verifyCap));
	b.shadowKey = ensureT<uint32_t>(c, L->bSegShadowK, static_cast<size_t>(n) + 1);
	b.shadowVal = ensureT<uint32_t>(c, L->bSegShadowV, static_cast<size_t>(n) + 1);
	b.counters = ensureT<int64_t>(c, L->bSegCounters, kSegCounters);
	b.chunkOff = ensureT<int64_t>(c, L->bSegChunkOff, static_cast<size_t>(caps.bigCap) + 1);
	b.chunkRange = ensureT<int32_t>(c, L->bSegChunkRange, static_cast<size_t>(caps.chunkCap));
	b.chunkCnt = ensureT<uint64_t>(c, L->bSegChunkCnt, static_cast<size_t>(caps.chunkCap) + 1);
	b.chunkPre = ensureT<uint64_t>(c, L->bSegChunkPre, static_cast<size_t>(std::max(caps.chunkCap, caps.bigCap)) + 1);
	b.rangeM = ensureT<int32_t>(c, L->bSegRangeM, static_cast<size_t>(caps.bigCap));
	b.rangeLm = ensureT<int32_t>(c, L->bSegRangeLm, static_cast<size_t>(caps.bigCap));
	b.scanTemp = ensure(c, L->bSegTemp, caps.scanTempBytes + 256);
	b.scanTempBytes = caps.scanTempBytes;
	b.bigCap = caps.bigCap;
	b.leafCap = caps.leafCap;
	b.heapCap = caps.heapCap;
	b.verifyCap = caps.verifyCap;
	b.chunkCap = caps.chunkCap;
	b.hostCounters = L->hostSeg;
	b.side = L->side;
	b.forkEv = L->sideEv[0];
	b.doneEv = L->sideEv[1];
	b.timer = &L->timer;
	return b;
}

void runSegSort(Lane* L, SegSortBufs& b, int64_t* tieArrays, bool* exactPending = nullptr)
{
	const hipError_t e = segSortRun(b, L->stream, tieArrays, exactPending);
	if (e == hipErrorIllegalState) fail(FLYEHIP_ERR_HIP, "segmented sort: inconsistent partition");
	HIPCHECK(e);
}

// The match sort's tie replay after the stable order (launchMatchSortStable):
// rank keys and tied-position flags, then the pruned exact introsort of every
// query holding a flagged position, values into bVal.  Without a relevance
// mask (ms.groupRel null) every tie counts and bVal becomes Flye's order.
// With it only the ties of the groups marked relevant are flagged: ranges
// without a flagged position are dropped, so bVal holds Flye's order at the
// flagged positions (their groups lie wholly in replayed ranges) and at every
// position of a distinct key; positions of other ties may hold any member
// (even one twice, where a tie straddles a dropped and a replayed range), so
// the caller takes those from the stable order (their order cannot reach a
// result).
void runTieReplay(Lane* L, const JoinArgs& ja, int64_t nMatches, MsortArgs& ms, bool flagsReady = false)
{
	flyehip_ctx* c = L->c;
	const size_t nm = static_cast<size_t>(nMatches) + 1;
	const int32_t nq = ja.nQueries;
	int64_t* dup = ensureT<int64_t>(c, L->bDup, nm);
	int64_t* tmp = ensureT<int64_t>(c, L->bScanTmp, scanTempWords(nMatches) + 64);
	SegSortBufs sb = segBufs(L, nMatches, nq);
	sb.key = ms.rkey;
	sb.val = ja.mval;
	sb.outVal = ms.sval;
	sb.dup = dup;
	sb.posL = ensureT<int32_t>(c, L->bPosL, nm);
	sb.posR = ensureT<int32_t>(c, L->bPosR, nm);
	HIPCHECK(hipEventRecord(L->ev[0], L->stream));
	// tie flags (unless the caller wrote them): sorted key read, flag written
	// (12 bytes per match); rank keys of the queries with a flagged position:
	// sorted key and value read, rank key written (16 bytes per match of
	// those queries, not counted)
	if (!flagsReady)
	{
		const int tk = L->timer.begin(FLYEHIP_K_RANK_KEYS, L->stream, 12.0 * static_cast<double>(nMatches));
		HIPCHECK(launchTieFlags(ms, nMatches, L->stream));
		L->timer.end(tk, L->stream);
	}
	HIPCHECK(exclusiveScan(ms.dupFlag, dup, nMatches, tmp, L->stream));
	const int tr = L->timer.begin(FLYEHIP_K_RANK_KEYS, L->stream, 0.0);
	HIPCHECK(launchRankKeys(ms, dup, nMatches, L->stream));
	L->timer.end(tr, L->stream);
	HIPCHECK(segSortReset(sb, L->stream));
	HIPCHECK(segPushOffsets(sb, ms.qOff, nq, L->stream));
	int64_t ties = 0;
	runSegSort(L, sb, &ties);
	HIPCHECK(hipEventRecord(L->ev[1], L->stream));
	HIPCHECK(hipEventSynchronize(L->ev[1]));
	float es = 0;
	HIPCHECK(hipEventElapsedTime(&es, L->ev[0], L->ev[1]));
	L->stats.ms_sort += es;
	L->stats.tie_queries += ties;
}

// The match sort's arguments for the batch's buffers.
MsortArgs matchSortArgs(Lane* L, const JoinArgs& ja, int64_t nMatches)
{
	flyehip_ctx* c = L->c;
	const size_t nm = static_cast<size_t>(nMatches) + 1;
	MsortArgs ms{};
	ms.qOff = static_cast<const int64_t*>(L->bQMatchOff.p);
	ms.nQueries = ja.nQueries;
	ms.akey = ja.mkey;
	ms.aval = ja.mval;
	ms.skey = ensureT<uint64_t>(c, L->bSKey, nm);
	ms.sval = ensureT<uint32_t>(c, L->bVal, nm);
	ms.rkey = ensureT<uint32_t>(c, L->bRKey, nm);
	ms.dupFlag = ensureT<int32_t>(c, L->bDupFlag, nm);
	return ms;
}

// Stable (extId, curPos) order of every query's matches (bSKey, bVal); with
// replay the tie replay follows, so bVal becomes std::sort's permutation
// (emission index within the query).  With withHits the counting sort also
// writes the chain's hits in that order (bHCur, bHExt, bGroupFlag); returns
// whether it did (the radix fallback leaves them to gatherHitsKernel).
bool runMatchSort(Lane* L, const JoinArgs& ja, int64_t nMatches, bool replay = true, bool withHits = false)
{
	flyehip_ctx* c = L->c;
	if (nMatches >= 0x7fffffffLL) fail(FLYEHIP_ERR_BUDGET, "batch of %lld matches exceeds 2^31 - 1", static_cast<long long>(nMatches));
	const int32_t nq = ja.nQueries;
	MsortArgs ms = matchSortArgs(L, ja, nMatches);
	int bits = 1;
	while ((int64_t{1} << bits) < 2 * static_cast<int64_t>(std::max(1, c->nReads))) ++bits;
	const int nBins = 2 * std::max(1, c->nReads);  // internal target ids
	const bool counting = msortCounting(nBins);
	if (counting && nq > 1)
	{
		// queries by descending match count for the counting sort's blocks
		std::vector<int64_t> qo(static_cast<size_t>(nq) + 1);
		download(L, qo.data(), ms.qOff, qo.size() * sizeof(int64_t));
		sync(L);
		std::vector<int32_t> order(static_cast<size_t>(nq));
		for (int32_t i = 0; i < nq; ++i) order[static_cast<size_t>(i)] = i;
		std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) {
			return qo[static_cast<size_t>(x) + 1] - qo[static_cast<size_t>(x)] > qo[static_cast<size_t>(y) + 1] - qo[static_cast<size_t>(y)];
		});
		int32_t* dOrder = ensureT<int32_t>(c, L->bQOrder, static_cast<size_t>(nq));
		upload(L, dOrder, order.data(), order.size() * sizeof(int32_t));
		ms.order = dOrder;
	}
	const bool hits = counting && withHits;
	if (hits)
	{
		const size_t nm = static_cast<size_t>(nMatches) + 1;
		ms.hits.mext = ja.mext;
		ms.hits.hcur = ensureT<int32_t>(c, L->bHCur, nm);
		ms.hits.hext = ensureT<int32_t>(c, L->bHExt, nm);
		ms.hits.groupFlag = ensureT<int32_t>(c, L->bGroupFlag, nm);
	}
	size_t tempBytes = 0;
	if (!counting) HIPCHECK(msortTempBytes(nMatches, nq, bits, &tempBytes));
	void* temp = counting ? nullptr : ensure(c, L->bMsortTemp, tempBytes + 256);
	HIPCHECK(hipEventRecord(L->ev[0], L->stream));
	// stable order: the counting sort reads the keys twice (histogram, then
	// to place them) and writes keys and values once (28 bytes per match);
	// with the hits it reads extPos and writes curPos, extPos, group flag and
	// value instead of key and value (36); the radix fallback makes 2 passes
	// x 24 bytes
	const double perMatch = counting ? (hits ? 36.0 : 28.0) : 48.0;
	const int tr = L->timer.begin(FLYEHIP_K_RADIX, L->stream, perMatch * static_cast<double>(nMatches));
	HIPCHECK(launchMatchSortStable(ms, nMatches, bits, nBins, temp, tempBytes, L->stream));
	L->timer.end(tr, L->stream);
	HIPCHECK(hipEventRecord(L->ev[1], L->stream));
	HIPCHECK(hipEventSynchronize(L->ev[1]));
	float es = 0;
	HIPCHECK(hipEventElapsedTime(&es, L->ev[0], L->ev[1]));
	L->stats.ms_sort += es;
	if (replay) runTieReplay(L, ja, nMatches, ms);
	return hits;
}

// Group-order hit copies of ca.passOrder[0, ca.nPass), one wave per chunk.
void runGather(Lane* L, const ChainArgs& ca, int64_t* tmp, int blocks)
{
	flyehip_ctx* c = L->c;
	const size_t np = static_cast<size_t>(std::max(ca.nPass, 0)) + 1;
	int32_t* chunks = ensureT<int32_t>(c, L->bGatherCnt, np);
	int64_t* chunkOff = ensureT<int64_t>(c, L->bGatherOff, np);
	HIPCHECK(launchGatherChunks(ca, chunks, L->stream));
	HIPCHECK(exclusiveScan(chunks, chunkOff, ca.nPass, tmp, L->stream));
	HIPCHECK(launchChainGather(ca, chunkOff, blocks, L->stream));
}

// Order, prefilters + extPos sort, DP, score sort + walks over ca.nGroups
// groups whose hits total nHits; leaves groupCount set.
void runChainKernels(Lane* L, ChainArgs& ca, int64_t nHits)
{
	flyehip_ctx* c = L->c;
	const int64_t ng = ca.nGroups;
	zero32(L, ca.groupCount, ng + 1);
	if (ng == 0) return;
	if (nHits >= 0x7fffffffLL) fail(FLYEHIP_ERR_BUDGET, "batch of %lld hits exceeds 2^31 - 1", static_cast<long long>(nHits));
	int32_t* bins = ensureT<int32_t>(c, L->bBins, 96);
	int32_t* order = ensureT<int32_t>(c, L->bOrder, static_cast<size_t>(ng) + 1);
	int32_t* passOrder = ensureT<int32_t>(c, L->bPassOrder, static_cast<size_t>(ng) + 1);
	int32_t* flag = ensureT<int32_t>(c, L->bGroupFlag, static_cast<size_t>(ng) + 1);
	int64_t* pos = ensureT<int64_t>(c, L->bGroupOff, static_cast<size_t>(ng) + 1);
	int64_t* tmp = ensureT<int64_t>(c, L->bScanTmp, scanTempWords(std::max(ng, nHits)) + 64);
	int32_t* queues = ensureT<int32_t>(c, L->bQueue, 16);  // work-queue counters 0-10
	int64_t* nBig = ensureT<int64_t>(c, L->bCounters, 16);
	zero32(L, bins, 96);
	zero32(L, queues, 16);
	// nBig[8]: DP cells; nBig[9..12]: groups the claim walk leaves to the ordered walk
	HIPCHECK(hipMemsetAsync(nBig, 0, 13 * sizeof(int64_t), L->stream));
	SegSortBufs sb = segBufs(L, nHits, ng);
	sb.key = ca.sortKey;
	sb.val = ca.sortVal;
	sb.outVal = ca.sortVal;
	sb.dup = nullptr;
	sb.posL = ca.posL;
	sb.posR = ca.posR;
	const int blocks = chainBlocks(c);

	HIPCHECK(hipEventRecord(L->ev[2], L->stream));
	HIPCHECK(launchGroupOrder(ca.groups, ca.nGroups, bins, order, L->stream));
	ca.order = order;
	HIPCHECK(segSortReset(sb, L->stream));
	ca.queue = queues + 0;
	const int tp = L->timer.begin(FLYEHIP_K_PREFILTER, L->stream, 8.0 * static_cast<double>(nHits));
	if (ca.passKnown)
	{
		// the pass is known (chainRelevanceKernel): the passing groups in
		// processing order first, then the prefilter visits only them
		HIPCHECK(launchPassFlags(ca, flag, L->stream));
		HIPCHECK(exclusiveScan(flag, pos, ng, tmp, L->stream));
		HIPCHECK(launchPassScatter(ca, flag, pos, passOrder, nBig, L->stream));
		int64_t np = 0;
		download(L, &np, pos + ng, sizeof(int64_t));
		sync(L);
		ChainArgs pa = ca;
		pa.order = passOrder;
		pa.nGroups = static_cast<int32_t>(np);
		HIPCHECK(launchChainPrefilter(pa, sb, blocks, L->stream));
	}
	else
	{
		HIPCHECK(launchChainPrefilter(ca, sb, blocks, L->stream));
	}
	L->timer.end(tp, L->stream);
	// the extPos sort leaves its exact heap sorts (libstdc++'s heapsort
	// fallback, a dependent chain per range) running on the side stream: the
	// groups that hold such a range are deferred, and the gather + DP of all
	// other passing groups run meanwhile
	bool exactPending = false;
	// should anything below throw, the lane stream still waits for the
	// pending heap sorts before its next batch reuses their buffers
	struct JoinOnExit
	{
		hipStream_t s;
		hipEvent_t e;
		bool* pending;
		~JoinOnExit()
		{
			if (*pending) (void)hipStreamWaitEvent(s, e, 0);
		}
	} joinOnExit{L->stream, L->sideEv[1], &exactPending};
	runSegSort(L, sb, nullptr, &exactPending);
	if (!ca.passKnown)
	{
		HIPCHECK(launchPassFlags(ca, flag, L->stream));
		HIPCHECK(exclusiveScan(flag, pos, ng, tmp, L->stream));
		HIPCHECK(launchPassScatter(ca, flag, pos, passOrder, nBig, L->stream));
	}
	int32_t* readyOrder = passOrder;
	int32_t* deferOrder = nullptr;
	int64_t* posReady = nullptr;
	if (exactPending)
	{
		int32_t* deferred = ensureT<int32_t>(c, L->bGroupDefer, static_cast<size_t>(ng) + 1);
		int32_t* flagReady = ensureT<int32_t>(c, L->bFlagReady, static_cast<size_t>(ng) + 1);
		int32_t* flagDefer = ensureT<int32_t>(c, L->bFlagDefer, static_cast<size_t>(ng) + 1);
		int64_t* posDefer = ensureT<int64_t>(c, L->bPosDefer, static_cast<size_t>(ng) + 1);
		posReady = ensureT<int64_t>(c, L->bPosReady, static_cast<size_t>(ng) + 1);
		readyOrder = ensureT<int32_t>(c, L->bReadyOrder, static_cast<size_t>(ng) + 1);
		deferOrder = ensureT<int32_t>(c, L->bDeferOrder, static_cast<size_t>(ng) + 1);
		zero32(L, deferred, ng);
		HIPCHECK(launchDeferGroups(sb.verify, sb.hostCounters[kSegNVerify], ca, deferred, L->stream));
		HIPCHECK(launchSplitPass(ca, deferred, flagReady, flagDefer, L->stream));
		HIPCHECK(exclusiveScan(flagReady, posReady, ng, tmp, L->stream));
		HIPCHECK(launchPassScatter(ca, flagReady, posReady, readyOrder, nBig + 2, L->stream));
		HIPCHECK(exclusiveScan(flagDefer, posDefer, ng, tmp, L->stream));
		HIPCHECK(launchPassScatter(ca, flagDefer, posDefer, deferOrder, nBig + 4, L->stream));
	}
	int64_t counts[3] = {0, 0, 0};
	int64_t split[5] = {0, 0, 0, 0, 0};  // ready (big groups, hits), deferred (big groups, hits), ready groups
	download(L, &counts[0], pos + ng, sizeof(int64_t));
	download(L, &counts[1], nBig, 2 * sizeof(int64_t));
	if (exactPending)
	{
		download(L, split, nBig + 2, 4 * sizeof(int64_t));
		download(L, &split[4], posReady + ng, sizeof(int64_t));
	}
	sync(L);
	const int32_t nPassAll = static_cast<int32_t>(counts[0]);
	const double passHits = static_cast<double>(counts[2]);
	const int32_t nReady = exactPending ? static_cast<int32_t>(split[4]) : nPassAll;
	const double readyHits = exactPending ? static_cast<double>(split[1]) : passHits;
	const double deferHits = static_cast<double>(split[3]);
	ca.passOrder = readyOrder;
	ca.nPass = nReady;
	ca.dpCells = reinterpret_cast<unsigned long long*>(nBig + 8);
	runGather(L, ca, tmp, blocks);
	HIPCHECK(hipEventRecord(L->ev[3], L->stream));
	ca.queue = queues + 2;
	// DP: hits read (hcur, hext) and score / backtrack written, 16 bytes per
	// hit (plus the prefix maxima it keeps for its far lookbacks)
	const int td = L->timer.begin(FLYEHIP_K_DP, L->stream, 16.0 * readyHits);
	HIPCHECK(launchChainDp(ca, c->dpLanes, blocks, L->stream));
	L->timer.end(td, L->stream);
	// backtrack .. selection without the score sort for every passing group
	// whose walk cannot depend on the order of equal scores; the flagged
	// groups (claimFlag) take the exact score sort and the ordered walk.  The
	// ready groups' claim runs while the deferred groups' heap sorts finish.
	int32_t* claimFlag = ensureT<int32_t>(c, L->bClaimFlag, static_cast<size_t>(ng) + 1);
	ca.queue = queues + 10;
	// claim: score and backtrack read, 8 bytes per passing hit
	const int tcl = L->timer.begin(FLYEHIP_K_CLAIM, L->stream, 8.0 * readyHits);
	HIPCHECK(launchChainClaim(ca, claimFlag, blocks, L->stream));
	L->timer.end(tcl, L->stream);
	if (deferOrder)
	{
		exactPending = false;
		HIPCHECK(hipStreamWaitEvent(L->stream, L->sideEv[1], 0));
		ca.passOrder = deferOrder;
		ca.nPass = nPassAll - nReady;
		runGather(L, ca, tmp, blocks);
		ca.queue = queues + 7;
		// the deferred groups are the few largest (they hold the ranges that
		// exhausted the sort's depth budget): a whole wave per group's
		// lookback finishes them sooner than the batch's tile width (the rows
		// do not depend on it)
		const int td2 = L->timer.begin(FLYEHIP_K_DP, L->stream, 16.0 * deferHits);
		HIPCHECK(launchChainDp(ca, std::max(c->dpLanes, 64), blocks, L->stream));
		L->timer.end(td2, L->stream);
		ca.queue = queues + 12;
		const int tcl2 = L->timer.begin(FLYEHIP_K_CLAIM, L->stream, 8.0 * deferHits);
		HIPCHECK(launchChainClaim(ca, claimFlag, blocks, L->stream));
		L->timer.end(tcl2, L->stream);
		ca.passOrder = passOrder;
		ca.nPass = nPassAll;
	}
	HIPCHECK(hipEventRecord(L->ev[4], L->stream));
	int32_t* flagOrder = ensureT<int32_t>(c, L->bFlagOrder, static_cast<size_t>(ng) + 1);
	{
		ChainArgs fa = ca;
		fa.order = passOrder;
		fa.groupPass = claimFlag;
		fa.nGroups = nPassAll;
		HIPCHECK(launchPassFlags(fa, flag, L->stream));
		HIPCHECK(exclusiveScan(flag, pos, nPassAll, tmp, L->stream));
		HIPCHECK(launchPassScatter(fa, flag, pos, flagOrder, nBig + 9, L->stream, nBig + 11));
	}
	int64_t fc[4] = {0, 0, 0, 0};  // flagged groups of at least kWalkSmall hits, their hits, tiers
	int64_t nFlagged = 0;
	download(L, fc, nBig + 9, 4 * sizeof(int64_t));
	download(L, &nFlagged, pos + nPassAll, sizeof(int64_t));
	sync(L);
	L->stats.exact_walk_groups += nFlagged;
	if (nFlagged > 0)
	{
		ca.passOrder = flagOrder;
		ca.nPass = static_cast<int32_t>(nFlagged);
		HIPCHECK(segSortReset(sb, L->stream));
		sb.relevantBit = 0x80000000u;  // rows with a backtrack pointer (chain.hip)
		ca.queue = queues + 3;
		// score-order keys from the DP's rows, then the ranges
		const int tsk = L->timer.begin(FLYEHIP_K_SCORE_KEYS, L->stream);
		HIPCHECK(launchChainScorePush(ca, sb, L->stream));
		L->timer.end(tsk, L->stream);
		runSegSort(L, sb, nullptr);
		ca.queue = queues + 4;
		// walk: score order and backtrack pointers read, 8 bytes per hit
		const int tw = L->timer.begin(FLYEHIP_K_WALK, L->stream, 8.0 * static_cast<double>(fc[1]));
		// size classes of the processing order (size-descending by power of two)
		const int32_t nLarge = static_cast<int32_t>(fc[2]), nMid = static_cast<int32_t>(fc[3]);
		const int32_t nBigF = static_cast<int32_t>(fc[0]);
		HIPCHECK(launchChainWalk(ca, 0, nLarge, 3, blocks, L->stream));
		ca.queue = queues + 8;
		HIPCHECK(launchChainWalk(ca, nLarge, nMid, 2, blocks, L->stream));
		ca.queue = queues + 9;
		HIPCHECK(launchChainWalk(ca, nMid, nBigF, 1, blocks, L->stream));
		ca.queue = queues + 5;
		HIPCHECK(launchChainWalk(ca, nBigF, ca.nPass, 0, blocks, L->stream));
		L->timer.end(tw, L->stream);
		ca.passOrder = passOrder;
		ca.nPass = nPassAll;
	}
	int64_t dpCells = 0;
	download(L, &dpCells, nBig + 8, sizeof(int64_t));
	HIPCHECK(hipEventRecord(L->ev[5], L->stream));
	HIPCHECK(hipEventSynchronize(L->ev[5]));
	L->stats.dp_cells += dpCells;
	float a = 0, b = 0, d = 0;
	HIPCHECK(hipEventElapsedTime(&a, L->ev[2], L->ev[3]));
	HIPCHECK(hipEventElapsedTime(&b, L->ev[3], L->ev[4]));
	HIPCHECK(hipEventElapsedTime(&d, L->ev[4], L->ev[5]));
	L->stats.ms_group += a;
	L->stats.ms_chain += b;
	L->stats.ms_extract += d;
}

// Groups + chaining for a batch whose sorted matches sit in bMKey/bVal.
// Returns rows on the device in bRows and per-query row counts in bQRowCount.
// With replayTies the matches come in the stable order (runMatchSort without
// replay): after the groups are built, chainRelevanceKernel marks the groups
// whose tied hits' order can reach a result, the tie replay runs for those
// ties alone, and the hits at their positions take the replayed order; every
// other tie keeps the stable order, which no result can tell from Flye's.
BatchOut runChain(Lane* L, const JoinArgs& ja, int64_t nMatches, const std::vector<int64_t>& qIds,
				  const std::vector<int32_t>& qLens, const flyehip_overlap_params_t& p, double* msGroup, double* msChain,
				  bool replayTies = false, bool hitsReady = false)
{
	flyehip_ctx* c = L->c;
	BatchOut bo;
	const int32_t nq = ja.nQueries;
	Timer tg;
	int32_t* hcur = ensureT<int32_t>(c, L->bHCur, static_cast<size_t>(nMatches) + 1);
	int32_t* hext = ensureT<int32_t>(c, L->bHExt, static_cast<size_t>(nMatches) + 1);
	int32_t* gflag = ensureT<int32_t>(c, L->bGroupFlag, static_cast<size_t>(nMatches) + 1);
	int64_t* goff = ensureT<int64_t>(c, L->bGroupOff, static_cast<size_t>(nMatches) + 1);
	int64_t* tmp = ensureT<int64_t>(c, L->bScanTmp, scanTempWords(nMatches) + 64);
	const int64_t* qMatchOff = static_cast<const int64_t*>(L->bQMatchOff.p);
	const uint64_t* skey = static_cast<const uint64_t*>(L->bSKey.p);
	if (!hitsReady)
	{
		// gather: sorted key, permutation and extPos read, hits and group
		// flags written (8 + 4 + 4 + 8 + 4 bytes per match)
		const int tgk = L->timer.begin(FLYEHIP_K_GATHER, L->stream, 28.0 * static_cast<double>(nMatches));
		HIPCHECK(launchGatherHits(qMatchOff, nq, skey, static_cast<const uint32_t*>(L->bVal.p), ja.mext, hcur, hext,
								  gflag, nMatches, L->stream));
		L->timer.end(tgk, L->stream);
	}
	HIPCHECK(exclusiveScan(gflag, goff, nMatches, tmp, L->stream));
	int64_t nGroups = 0;
	download(L, &nGroups, goff + nMatches, sizeof(int64_t));
	sync(L);
	GroupDesc* groups = ensureT<GroupDesc>(c, L->bGroups, static_cast<size_t>(nGroups) + 1);
	HIPCHECK(launchBuildGroups(qMatchOff, nq, skey, gflag, goff, ja.targetLen, groups, nMatches, L->stream));
	HIPCHECK(launchGroupSizes(groups, nGroups, nMatches, L->stream));
	// query descriptors
	int64_t* dIds = ensureT<int64_t>(c, L->bQueryIds, static_cast<size_t>(nq) + 1);
	int32_t* dLens = ensureT<int32_t>(c, L->bCurLen, static_cast<size_t>(nq) + 1);
	upload(L, dIds, qIds.data(), qIds.size() * sizeof(int64_t));
	upload(L, dLens, qLens.data(), qLens.size() * sizeof(int32_t));
	QueryDesc* qd = ensureT<QueryDesc>(c, L->bQDesc, static_cast<size_t>(nq) + 1);
	hipLaunchKernelGGL(makeQueryDescKernel, grid(nq, 256), dim3(256), 0, L->stream, dIds, dLens,
					   static_cast<const int64_t*>(L->bQFiltOff.p), nq, qd);
	HIPCHECK(hipGetLastError());
	if (replayTies)
	{
		const size_t nmr = static_cast<size_t>(nMatches) + 1;
		ChainArgs ra{};
		ra.params = p;
		ra.groups = groups;
		ra.queries = qd;
		ra.nGroups = static_cast<int32_t>(nGroups);
		ra.cur = hcur;
		ra.ext = hext;
		ra.posL = ensureT<int32_t>(c, L->bPosL, nmr);
		ra.posR = ensureT<int32_t>(c, L->bPosR, nmr);
		ra.groupPass = ensureT<int32_t>(c, L->bGroupPass, static_cast<size_t>(nGroups) + 1);
		int32_t* queues = ensureT<int32_t>(c, L->bQueue, 16);
		zero32(L, queues, 16);
		ra.queue = queues + 15;
		int32_t* groupRel = ensureT<int32_t>(c, L->bGroupRel, static_cast<size_t>(nGroups) + 1);
		MsortArgs ms = matchSortArgs(L, ja, nMatches);
		if (hitsReady)
		{
			// the counting sort wrote skey at the group starts only, and the
			// join left the emission indices to the rank keys
			ms.hits.hcur = hcur;
			ms.hits.groupFlag = gflag;
			ms.ival = ja.mval;
		}
		// the tie flags start cleared: the relevance pass sets those of the
		// tied hits of the groups it marks
		zero32(L, ms.dupFlag, nMatches);
		// the prefilters (hits read, 8 bytes each) and the extPos duplicate test
		const int tp = L->timer.begin(FLYEHIP_K_PREFILTER, L->stream, 8.0 * static_cast<double>(nMatches));
		HIPCHECK(launchChainRelevance(ra, groupRel, ms.dupFlag, chainBlocks(c), L->stream));
		L->timer.end(tp, L->stream);
		runTieReplay(L, ja, nMatches, ms, true);
		HIPCHECK(launchFixTiedHits(qMatchOff, nq, ms.dupFlag, ms.sval, ja.mext, hext, nMatches, L->stream));
	}
	sync(L);
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
	// The match sort's per-match buffers are free once the groups are built
	// and hold the chain's: its partition scratch, the rank keys and tie
	// flags (sort keys and values), the tied-position prefix and the
	// emission-order keys (8 bytes a match each: two int32 arrays, the second
	// at a 256-byte boundary).  A lane then holds 24 bytes a match less.
	const size_t half = (nm + 63) & ~static_cast<size_t>(63);
	ca.sortKey = ensureT<uint32_t>(c, L->bRKey, nm);
	ca.sortVal = ensureT<uint32_t>(c, L->bDupFlag, nm);
	ca.posL = ensureT<int32_t>(c, L->bPosL, nm);
	ca.posR = ensureT<int32_t>(c, L->bPosR, nm);
	ca.hcur = ensureT<int32_t>(c, L->bDup, 2 * half);
	ca.hext = ca.hcur + half;
	ca.score = ensureT<int32_t>(c, L->bMKey, 2 * half);
	ca.bt = ca.score + half;
	// with onlyMaxExt a group keeps one proposal: per-group output, no list
	ca.props = ensureT<Proposal>(c, L->bCProps, p.only_max_ext ? 1 : nm);
	ca.outProps = ensureT<Proposal>(c, L->bCOut, p.only_max_ext ? static_cast<size_t>(nGroups) + 1 : nm);
	ca.groupCount = ensureT<int32_t>(c, L->bGroupCount, static_cast<size_t>(nGroups) + 1);
	ca.groupPass = ensureT<int32_t>(c, L->bGroupPass, static_cast<size_t>(nGroups) + 1);
	ca.logX = static_cast<const uint32_t*>(c->dLogX.p);
	ca.logY = static_cast<const uint32_t*>(c->dLogY.p);
	ca.passKnown = replayTies ? 1 : 0;
	runChainKernels(L, ca, nMatches);
	int64_t* rowStart = ensureT<int64_t>(c, L->bRowStart, static_cast<size_t>(nGroups) + 1);
	HIPCHECK(exclusiveScan(ca.groupCount, rowStart, nGroups, tmp, L->stream));
	int64_t nRows = 0;
	download(L, &nRows, rowStart + nGroups, sizeof(int64_t));
	sync(L);
	flyehip_overlap_t* rows = ensureT<flyehip_overlap_t>(c, L->bRows, static_cast<size_t>(nRows) + 1);
	HIPCHECK(launchWriteRows(ca, rowStart, rows, L->stream));
	int32_t* qCount = ensureT<int32_t>(c, L->bQRowCount, static_cast<size_t>(nq) + 1);
	zero32(L, qCount, nq + 1);
	if (nGroups > 0)
	{
		hipLaunchKernelGGL(queryRowCountKernel, grid(nGroups, 256), dim3(256), 0, L->stream, groups, ca.groupCount,
						   static_cast<int32_t>(nGroups), qCount);
		HIPCHECK(hipGetLastError());
	}
	sync(L);
	(void)msChain;
	(void)tc;
	bo.groups = nGroups;
	bo.rows = nRows;
	L->stats.groups += nGroups;
	L->stats.overlaps += nRows;
	return bo;
}

void resetStats(flyehip_ctx* c)
{
	c->stats = flyehip_stats_t{};
	for (auto& l : c->lanes) l->stats = flyehip_stats_t{};
}

// Adds the lanes' stage timings and counts into the context's stats.
void foldStats(flyehip_ctx* c)
{
	for (auto& l : c->lanes)
	{
		l->timer.collect(l->stats);  // every call ends with its streams synchronised
		const flyehip_stats_t& s = l->stats;
		c->stats.ms_join += s.ms_join;
		c->stats.ms_sort += s.ms_sort;
		c->stats.ms_group += s.ms_group;
		c->stats.ms_chain += s.ms_chain;
		c->stats.ms_extract += s.ms_extract;
		c->stats.windows += s.windows;
		c->stats.matches += s.matches;
		c->stats.groups += s.groups;
		c->stats.overlaps += s.overlaps;
		c->stats.batches += s.batches;
		c->stats.tie_queries += s.tie_queries;
		c->stats.exact_walk_groups += s.exact_walk_groups;
		c->stats.dp_cells += s.dp_cells;
		for (int f = 0; f < FLYEHIP_K_COUNT; ++f)
		{
			c->stats.kernel_ms[f] += s.kernel_ms[f];
			c->stats.kernel_launches[f] += s.kernel_launches[f];
			c->stats.kernel_bytes[f] += s.kernel_bytes[f];
		}
		l->stats = flyehip_stats_t{};
	}
	c->stats.device_bytes_peak = static_cast<int64_t>(c->peak);
}

void createLane(flyehip_ctx* c)
{
	auto l = std::make_unique<Lane>();
	l->c = c;
	HIPCHECK(hipStreamCreateWithFlags(&l->stream, hipStreamNonBlocking));
	for (auto& e : l->ev) HIPCHECK(hipEventCreate(&e));
	HIPCHECK(hipStreamCreateWithFlags(&l->side, hipStreamNonBlocking));
	for (auto& e : l->sideEv) HIPCHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
	HIPCHECK(hipHostMalloc(reinterpret_cast<void**>(&l->hostSeg), kSegCounters * sizeof(int64_t)));
	HIPCHECK(l->timer.create());
	c->lanes.push_back(std::move(l));
}

// Releases a lane's per-batch device buffers; the next call sizes them again.
void releaseBatchBuffers(flyehip_ctx* c, Lane& l)
{
	if (l.stream) (void)hipStreamSynchronize(l.stream);
	for (DevBuf* b : {&l.bQueries, &l.bWinOff, &l.bWinCount, &l.bWinFilt, &l.bWinInfo, &l.bMatchOff, &l.bFiltOff,
					  &l.bScanTmp, &l.bMKey, &l.bMExt, &l.bVal, &l.bPosL, &l.bPosR, &l.bHCur, &l.bHExt, &l.bGroupFlag,
					  &l.bGroupOff, &l.bFiltered, &l.bQMatchOff, &l.bQFiltOff, &l.bQDesc, &l.bGroups, &l.bGroupCount,
					  &l.bRowStart, &l.bRows, &l.bCScore, &l.bCBt, &l.bCProps, &l.bCOut, &l.bCHCur, &l.bCHExt, &l.bCKey,
					  &l.bCVal, &l.bCPosL, &l.bCPosR, &l.bGroupPass, &l.bOrder, &l.bPassOrder, &l.bBins, &l.bSKey,
					  &l.bAVal, &l.bRKey, &l.bDupFlag, &l.bDup, &l.bMsortTemp, &l.bSegBig0, &l.bSegBig1, &l.bSegLeaf,
					  &l.bSegHeap, &l.bSegCounters, &l.bSegChunkOff, &l.bSegChunkRange, &l.bSegChunkCnt,
					  &l.bSegChunkPre, &l.bSegRangeM, &l.bSegRangeLm, &l.bSegTemp, &l.bSegVerify, &l.bSegShadowK,
					  &l.bSegShadowV, &l.bQueue, &l.bCounters, &l.bQRowCount, &l.bCurLen, &l.bDivOff, &l.bDivCount,
					  &l.bDivOut, &l.bQueryIds, &l.bCand, &l.bExtWords, &l.bExtWordOff, &l.bExtLen, &l.bAln,
					  &l.bAlnOff, &l.bRA, &l.bGroupDefer, &l.bFlagReady, &l.bFlagDefer, &l.bPosReady, &l.bPosDefer,
					  &l.bReadyOrder, &l.bDeferOrder, &l.bClaimFlag, &l.bFlagOrder, &l.bGroupRel, &l.bGatherCnt, &l.bGatherOff, &l.bRawBases, &l.bRawOff, &l.bPackErr, &l.bGroupStart, &l.bMinMask, &l.bQOrder})
		release(c, *b);
}

void destroyLane(flyehip_ctx* c, Lane& l)
{
	releaseBatchBuffers(c, l);
	if (l.hostSeg) (void)hipHostFree(l.hostSeg);
	l.timer.destroy();
	for (auto& e : l.ev)
		if (e) (void)hipEventDestroy(e);
	for (auto& e : l.sideEv)
		if (e) (void)hipEventDestroy(e);
	if (l.side) (void)hipStreamDestroy(l.side);
	if (l.stream) (void)hipStreamDestroy(l.stream);
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
		foldStats(c);
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

// Host copy of one batch's results, in query order.
struct BatchResult
{
	std::vector<flyehip_overlap_t> rows;
	std::vector<int32_t> count;     // rows per query
	std::vector<int32_t> divCount;  // divStats windows per query
	std::vector<float> divs;
};

// One batch allQ[b0, b1) on a lane: join, match sort (stable order, then the
// tie replay for the ties whose order can reach a result), chaining; rows
// (and divStats windows, computed on the device) copied to the host.
void runBatch(Lane* L, const PackedSeqs& seqs, const std::vector<JoinQuery>& allQ, const int64_t* queryIds,
			  const flyehip_overlap_params_t& p, int32_t b0, int32_t b1, bool withDiv, BatchResult& out)
{
	flyehip_ctx* c = L->c;
	std::vector<JoinQuery> qs(allQ.begin() + b0, allQ.begin() + b1);
	std::vector<int64_t> ids(queryIds + b0, queryIds + b1);
	std::vector<int32_t> lens;
	for (auto& q : qs) lens.push_back(q.len);
	JoinArgs ja;
	// with the counting sort the emission indices are the positions: the
	// fill skips them and the tie replay writes its queries'
	const bool counting = msortCounting(2 * std::max(1, c->nReads));
	BatchOut jo = runJoin(L, seqs, qs, ja, !counting);
	sync(L);
	const bool hitsReady = runMatchSort(L, ja, jo.matches, false, true);
	BatchOut co = runChain(L, ja, jo.matches, ids, lens, p, nullptr, nullptr, true, hitsReady);
	const int32_t nq = b1 - b0;
	out.count.assign(static_cast<size_t>(nq), 0);
	out.rows.resize(static_cast<size_t>(co.rows));
	download(L, out.count.data(), L->bQRowCount.p, static_cast<size_t>(nq) * sizeof(int32_t));
	if (co.rows) download(L, out.rows.data(), L->bRows.p, static_cast<size_t>(co.rows) * sizeof(flyehip_overlap_t));
	sync(L);
	if (withDiv)
	{
		// divStats windows, computed on the device from the batch rows
		std::vector<int64_t> dOff(static_cast<size_t>(nq) + 1, 0);
		std::vector<int64_t> qRowOff(static_cast<size_t>(nq) + 1, 0);
		for (int32_t i = 0; i < nq; ++i)
		{
			dOff[static_cast<size_t>(i) + 1] = dOff[static_cast<size_t>(i)] + lens[static_cast<size_t>(i)] / 10000 + 1;
			qRowOff[static_cast<size_t>(i) + 1] = qRowOff[static_cast<size_t>(i)] + out.count[static_cast<size_t>(i)];
		}
		int64_t* dDivOff = ensureT<int64_t>(c, L->bDivOff, dOff.size());
		int64_t* dQRowOff = ensureT<int64_t>(c, L->bRowStart, std::max(qRowOff.size(), static_cast<size_t>(co.groups) + 1));
		int32_t* dDivCount = ensureT<int32_t>(c, L->bDivCount, static_cast<size_t>(nq) + 1);
		float* dDivOut = ensureT<float>(c, L->bDivOut, static_cast<size_t>(dOff.back()) + 1);
		upload(L, dDivOff, dOff.data(), dOff.size() * sizeof(int64_t));
		upload(L, dQRowOff, qRowOff.data(), qRowOff.size() * sizeof(int64_t));
		HIPCHECK(launchDivStats(static_cast<const flyehip_overlap_t*>(L->bRows.p), dQRowOff,
								static_cast<const int32_t*>(L->bCurLen.p), nq, dDivCount, dDivOut, dDivOff, L->stream));
		out.divCount.assign(static_cast<size_t>(nq), 0);
		std::vector<float> dv(static_cast<size_t>(dOff.back()) + 1);
		download(L, out.divCount.data(), dDivCount, static_cast<size_t>(nq) * sizeof(int32_t));
		download(L, dv.data(), dDivOut, static_cast<size_t>(dOff.back()) * sizeof(float));
		sync(L);
		for (int32_t i = 0; i < nq; ++i)
			for (int32_t t = 0; t < out.divCount[static_cast<size_t>(i)]; ++t)
				out.divs.push_back(dv[static_cast<size_t>(dOff[static_cast<size_t>(i)] + t)]);
	}
	L->timer.collect(L->stats);
	L->stats.batches += 1;
}

// Batched overlap pipeline over JoinQuery lists (resident or external).
// Batches are dealt to up to batchesInFlight lanes, each driven by its own
// host thread on its own stream, so one batch's long tails (the last groups
// of a kernel, a heap-sort fallback) overlap the next batch's work; results
// are assembled in query order afterwards.
std::vector<BatchResult> runBatches(flyehip_ctx* c, const PackedSeqs& seqs, const std::vector<JoinQuery>& allQ,
									const int64_t* queryIds, const flyehip_overlap_params_t* p, bool withDiv,
									int batchQueries, int maxLanes)
{
	const int32_t n = static_cast<int32_t>(allQ.size());
	const int32_t nBatches = (n + batchQueries - 1) / batchQueries;
	const int nl = std::max(1, std::min(std::min(c->batchesInFlight, maxLanes), nBatches));
	while (static_cast<int>(c->lanes.size()) < nl) createLane(c);
	std::vector<BatchResult> res(static_cast<size_t>(nBatches));
	// lane i takes batch i first (the same lane gets the same batch on every
	// call, so its grow-only buffers fit it after the first call), then the
	// rest first-come
	std::atomic<int32_t> next{nl};
	std::vector<std::exception_ptr> errs(static_cast<size_t>(nl));
	c->lanesRunning = nl;
	auto worker = [&](int li) {
		Lane* L = c->lanes[static_cast<size_t>(li)].get();
		tlsLane = L;
		try
		{
			HIPCHECK(hipSetDevice(c->device));
			for (int32_t b = li; b < nBatches; b = next.fetch_add(1))
			{
				const int32_t b0 = b * batchQueries;
				// hold the allocation gate for the batch; leave it with the
				// lane's streams idle
				struct Hold
				{
					flyehip_ctx* c;
					Lane* L;
					Hold(flyehip_ctx* c_, Lane* L_) : c(c_), L(L_) { if (c->lanesRunning > 1) c->gate.enter(); }
					~Hold()
					{
						if (c->lanesRunning > 1)
						{
							(void)hipStreamSynchronize(L->stream);
							(void)hipStreamSynchronize(L->side);
							c->gate.leave();
						}
					}
				} hold(c, L);
				runBatch(L, seqs, allQ, queryIds, *p, b0, std::min(n, b0 + batchQueries), withDiv,
						 res[static_cast<size_t>(b)]);
			}
		}
		catch (...)
		{
			errs[static_cast<size_t>(li)] = std::current_exception();
			next = nBatches;
		}
		tlsLane = nullptr;
	};
	std::vector<std::thread> th;
	for (int li = 1; li < nl; ++li) th.emplace_back(worker, li);
	worker(0);
	for (auto& t : th) t.join();
	c->lanesRunning = 0;
	for (auto& e : errs)
		if (e) std::rethrow_exception(e);
	return res;
}

// Batched overlap pipeline over JoinQuery lists (resident or external) with
// the rows copied into the caller's arrays.
void overlapsImpl(flyehip_ctx* c, const PackedSeqs& seqs, const std::vector<JoinQuery>& allQ,
				  const int64_t* queryIds, const flyehip_overlap_params_t* p, flyehip_overlap_t* rows,
				  int64_t rowCap, int64_t* rowOffsets, int64_t* rowsNeeded, float* divStats, int64_t divCap,
				  int64_t* divOffsets, int batchQueries, int maxLanes)
{
	checkParams(p, c->k);
	resetStats(c);
	Timer total;
	const int32_t n = static_cast<int32_t>(allQ.size());
	const int32_t nBatches = (n + batchQueries - 1) / batchQueries;
	std::vector<BatchResult> res =
		runBatches(c, seqs, allQ, queryIds, p, divOffsets != nullptr, batchQueries, maxLanes);

	int64_t written = 0, divWritten = 0;
	bool capacityShort = false;
	rowOffsets[0] = 0;
	if (divOffsets) divOffsets[0] = 0;
	for (int32_t b = 0; b < nBatches; ++b)
	{
		const BatchResult& r = res[static_cast<size_t>(b)];
		const int32_t b0 = b * batchQueries;
		const int64_t nr = static_cast<int64_t>(r.rows.size());
		if (written + nr <= rowCap)
		{
			if (nr) std::memcpy(rows + written, r.rows.data(), static_cast<size_t>(nr) * sizeof(flyehip_overlap_t));
		}
		else
		{
			capacityShort = true;
		}
		size_t d = 0;
		for (size_t i = 0; i < r.count.size(); ++i)
		{
			written += r.count[i];
			rowOffsets[b0 + static_cast<int64_t>(i) + 1] = written;
			if (divOffsets)
			{
				for (int32_t t = 0; t < r.divCount[i]; ++t, ++d)
				{
					if (divStats && divWritten < divCap) divStats[divWritten] = r.divs[d];
					++divWritten;
				}
				divOffsets[b0 + static_cast<int64_t>(i) + 1] = divWritten;
			}
		}
	}
	foldStats(c);
	c->stats.queries = n;
	c->stats.ms_total = total.ms();
	if (rowsNeeded) *rowsNeeded = written;
	// div_offsets[n] already holds the divStats values needed
	const bool divShort = divStats && divWritten > divCap;
	if (capacityShort || divShort)
		fail(FLYEHIP_ERR_CAPACITY, "capacity too small: %lld rows needed (capacity %lld), %lld divStats values needed "
				 "(capacity %lld)",
			 static_cast<long long>(written), static_cast<long long>(rowCap), static_cast<long long>(divWritten),
			 static_cast<long long>(divCap));
}

// ---- submitted jobs -----------------------------------------------------------

std::vector<JoinQuery> residentQueries(flyehip_ctx* c, const int64_t* ids, int32_t n);

// One chunk of a job under the context mutex: the batched pipeline over the
// chunk's query records.
std::vector<BatchResult> asyncRunChunk(flyehip_ctx* c, const AsyncJob& job, int32_t b0, int32_t b1)
{
	std::lock_guard<std::mutex> lock(c->mu);
	HIPCHECK(hipSetDevice(c->device));
	std::vector<JoinQuery> qs = residentQueries(c, job.ids.data() + b0, b1 - b0);
	checkParams(&job.p, c->k);
	resetStats(c);
	std::vector<BatchResult> res =
		runBatches(c, storeSeqs(c), qs, job.ids.data() + b0, &job.p, job.withDiv, c->queriesPerBatch, 4);
	foldStats(c);
	return res;
}

// Runs submitted jobs in order, a chunk of queries_per_batch x
// batches_in_flight query records at a time, and publishes each chunk's rows.
void asyncWorker(flyehip_ctx* c)
{
	AsyncQueue& q = *c->async;
	for (;;)
	{
		std::shared_ptr<AsyncJob> job;
		{
			std::unique_lock<std::mutex> l(q.m);
			q.cv.wait(l, [&] { return q.stop || !q.pending.empty(); });
			if (q.stop) return;
			job = q.pending.front();
		}
		const int32_t n = static_cast<int32_t>(job->ids.size());
		for (int32_t b0 = 0; b0 < n;)
		{
			int32_t chunk;
			{
				std::lock_guard<std::mutex> l(c->mu);
				chunk = std::max(1, c->queriesPerBatch * std::max(1, c->batchesInFlight));
			}
			const int32_t b1 = std::min(n, b0 + chunk);
			std::vector<BatchResult> res;
			int code = FLYEHIP_OK;
			std::string msg;
			try
			{
				res = asyncRunChunk(c, *job, b0, b1);
			}
			catch (const FhError& e)
			{
				code = e.code;
				msg = e.msg;
			}
			catch (const std::bad_alloc&)
			{
				code = FLYEHIP_ERR_BUDGET;
				msg = "host allocation failed";
			}
			catch (...)
			{
				code = FLYEHIP_ERR_STATE;
				msg = "unexpected exception in the job worker";
			}
			std::lock_guard<std::mutex> l(q.m);
			if (code != FLYEHIP_OK)
			{
				job->code = code;
				job->msg = msg;
				job->done = n;
			}
			else
			{
				for (const BatchResult& r : res)
				{
					job->rows.insert(job->rows.end(), r.rows.begin(), r.rows.end());
					job->divs.insert(job->divs.end(), r.divs.begin(), r.divs.end());
					for (size_t i = 0; i < r.count.size(); ++i)
					{
						job->rowOff.push_back(job->rowOff.back() + r.count[i]);
						job->divOff.push_back(job->divOff.back() + (job->withDiv ? r.divCount[i] : 0));
					}
				}
				job->done = b1;
			}
			q.cv.notify_all();
			if (q.stop) return;
			if (code != FLYEHIP_OK) break;
			b0 = b1;
		}
		std::lock_guard<std::mutex> l(q.m);
		q.pending.pop_front();
		q.cv.notify_all();
	}
}

// Stops the worker; jobs not finished fail, so no collector waits forever.
void asyncStop(flyehip_ctx* c)
{
	if (!c->async) return;
	AsyncQueue& q = *c->async;
	{
		std::lock_guard<std::mutex> l(q.m);
		q.stop = true;
		for (auto& kv : q.jobs)
		{
			AsyncJob& j = *kv.second;
			if (j.done < static_cast<int32_t>(j.ids.size()))
			{
				j.code = FLYEHIP_ERR_STATE;
				j.msg = "context closed before the job finished";
				j.done = static_cast<int32_t>(j.ids.size());
			}
		}
		q.cv.notify_all();
	}
	if (q.worker.joinable()) q.worker.join();
}

void setError(flyehip_ctx* c, const std::string& msg)
{
	std::lock_guard<std::mutex> lock(c->mu);
	c->err = msg;
}

// ---- no-allocation read-to-graph chaining (the scratch array is at the top) ----
constexpr int kMaxDevices = 64;
std::mutex gScratchMu;
unsigned char* gScratch[kMaxDevices] = {};

// Device address of `scratch` on `device` (current on the calling thread),
// looked up once per device.
unsigned char* scratchOn(int device)
{
	if (device < 0 || device >= kMaxDevices) fail(FLYEHIP_ERR_ARG, "device %d beyond %d", device, kMaxDevices);
	std::lock_guard<std::mutex> lock(gScratchMu);
	if (!gScratch[device])
	{
		void* p = nullptr;
		HIPCHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(::scratch)));
		gScratch[device] = static_cast<unsigned char*>(p);
	}
	return gScratch[device];
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
		// a loaded index need not hold the query's own window, nor keep
		// buckets in (read, position) order: scan the bucket for it
		q.selfMode = c->indexLoaded ? 2 : (c->info.minimizer_window > 0 ? 3 : 1);
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
		c->scratch = scratchOn(device);
		createLane(c);
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
	asyncStop(c);
	(void)hipSetDevice(c->device);
	for (auto& l : c->lanes) destroyLane(c, *l);
	for (DevBuf* b : {&c->dWords, &c->dWordOff, &c->dReadLen, &c->dReadWinOff, &c->dKeys, &c->dCursor,
					  &c->dEntries, &c->dLogX, &c->dLogY})
		release(c, *b);
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
		}
		if (gpp)
		{
			if (gpp < 1 || gpp > 4) fail(FLYEHIP_ERR_ARG, "batches_in_flight %d not supported (1 .. 4)", gpp);
		}
		// buffers sized for another batch size or number of batches in flight
		// would hold device memory the new setting may need
		const bool resize = (qpb && qpb != c->queriesPerBatch) || (gpp && gpp != c->batchesInFlight);
		if (qpb) c->queriesPerBatch = qpb;
		if (gpp) c->batchesInFlight = gpp;
		if (resize)
			for (auto& l : c->lanes) releaseBatchBuffers(c, *l);
	});
}

int flyehip_store_reads(flyehip_ctx* c, const uint8_t* bases, const int64_t* offsets, int32_t n_reads)
{
	return guarded(c, [&]() {
		Lane* L = c->lanes[0].get();
		if (!bases || !offsets || n_reads <= 0) fail(FLYEHIP_ERR_ARG, "empty read store");
		std::vector<int64_t> wordOff;
		std::vector<int32_t> lens;
		c->haveIndex = false;
		packSeqsDevice(c, L, bases, offsets, n_reads, c->dWords, c->dWordOff, c->dReadLen, wordOff, lens);
		c->nReads = n_reads;
		c->readLen = lens;
		c->readWordOff = wordOff;
	});
}

namespace
{

// Counts every canonical k-mer of the stored reads, fixes the thresholds and
// fills the buckets: with every position, or (window > 0) with the minimizer
// positions only (minimizerMaskKernel).
void buildIndexImpl(flyehip_ctx* c, const flyehip_index_params_t* ip, int32_t window, flyehip_index_info_t* out)
{
		Lane* L = c->lanes[0].get();
		if (!ip) fail(FLYEHIP_ERR_ARG, "null index parameters");
		if (c->nReads <= 0) fail(FLYEHIP_ERR_STATE, "no reads stored");
		const int k = ip->kmer_size;
		if (k < 1 || k > 31) fail(FLYEHIP_ERR_ARG, "kmer_size %d out of range 1..31", k);
		if (window < 0 || window > 256) fail(FLYEHIP_ERR_ARG, "minimizer window %d out of range 1..256", window);
		c->k = k;
		c->haveIndex = false;
		c->indexLoaded = false;
		std::vector<int64_t> winOff(static_cast<size_t>(c->nReads) + 1, 0);
		for (int32_t i = 0; i < c->nReads; ++i)
			winOff[static_cast<size_t>(i) + 1] = winOff[static_cast<size_t>(i)] + std::max(0, c->readLen[static_cast<size_t>(i)] - k + 1);
		const int64_t nWin = winOff.back();
		int64_t* dWinOff = ensureT<int64_t>(c, c->dReadWinOff, winOff.size());
		upload(L, dWinOff, winOff.data(), winOff.size() * 8);
		uint64_t slots = 1024;
		while (slots < static_cast<uint64_t>(nWin) * 3 / 2 + 1024) slots <<= 1;
		c->tableMask = slots - 1;
		ensure(c, c->dKeys, slots * 16);
		ensure(c, c->dCursor, slots * 4);
		emptyTable(L, static_cast<uint64_t*>(c->dKeys.p), static_cast<int64_t>(slots));
		const PackedSeqs rs = storeSeqs(c);
		HashTable t = table(c);
		HIPCHECK(launchCount(rs, dWinOff, nWin, k, t, L->stream));
		const int hmax = 1 << 16;
		uint64_t* dHist = ensureT<uint64_t>(c, L->bCKey, static_cast<size_t>(hmax) + 8);
		fill64(L, dHist, hmax + 8, 0);
		HIPCHECK(launchHistogram(t, dHist, hmax, dHist + hmax, L->stream));
		std::vector<uint64_t> hist(static_cast<size_t>(hmax) + 8);
		download(L, hist.data(), dHist, hist.size() * 8);
		sync(L);
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
		int32_t* sizes = ensureT<int32_t>(c, L->bWinCount, slots + 1);
		uint64_t* counters = ensureT<uint64_t>(c, L->bCounters, 16);
		fill64(L, counters, 16, 0);
		HIPCHECK(launchClassify(t, solidMin, repMax, sizes, counters, L->stream));
		uint8_t* mask = nullptr;
		if (window > 0)
		{
			// bucket sizes from the minimizer positions alone
			mask = ensureT<uint8_t>(c, L->bMinMask, static_cast<size_t>(nWin) + 4);
			zero32(L, reinterpret_cast<int32_t*>(mask), (nWin + 3) / 4);
			HIPCHECK(launchMinimizerMask(rs, dWinOff, nWin, k, window, mask, L->stream));
			zero32(L, sizes, static_cast<int64_t>(slots));
			HIPCHECK(launchMaskedSizes(rs, dWinOff, nWin, k, t, mask, sizes, L->stream));
		}
		int64_t* starts = ensureT<int64_t>(c, L->bMatchOff, slots + 1);
		int64_t* tmp = ensureT<int64_t>(c, L->bScanTmp, scanTempWords(static_cast<int64_t>(slots)) + 64);
		HIPCHECK(exclusiveScan(sizes, starts, static_cast<int64_t>(slots), tmp, L->stream));
		int64_t nEntries = 0;
		uint64_t cnt[2];
		download(L, &nEntries, starts + slots, 8);
		download(L, cnt, counters, 16);
		sync(L);
		HIPCHECK(launchAssign(t, sizes, starts, L->stream));
		uint64_t* ord = ensureT<uint64_t>(c, L->bMKey, static_cast<size_t>(nEntries) + 1);
		ensure(c, c->dEntries, static_cast<size_t>(nEntries) * 8 + 8);
		HIPCHECK(launchFill(rs, dWinOff, nWin, k, t, mask, ord, L->stream));
		// buckets above 32 entries (at most nEntries / 33 of them) are sorted
		// by a block each; counters[2] counts them (zeroed above)
		const uint64_t bigCap = static_cast<uint64_t>(nEntries) / 33 + 1;
		uint32_t* bigList = ensureT<uint32_t>(c, L->bCVal, static_cast<size_t>(bigCap));
		HIPCHECK(launchBucketFinish(t, ord, static_cast<uint64_t*>(c->dEntries.p),
									static_cast<const int32_t*>(c->dReadLen.p), k, bigList, counters + 2, bigCap,
									L->stream));
		sync(L);
		c->info = flyehip_index_info_t{};
		c->info.kmer_size = k;
		c->info.solid_min = solidMin;
		c->info.rep_max = repMax;
		c->info.distinct_kmers = static_cast<int64_t>(acc[2]);
		c->info.solid_kmers = static_cast<int64_t>(cnt[0]);
		c->info.repetitive_kmers = static_cast<int64_t>(cnt[1]);
		c->info.indexed_positions = nEntries;
		c->info.hash_slots = static_cast<int64_t>(slots);
		c->info.device_bytes = static_cast<int64_t>(c->dKeys.bytes + c->dCursor.bytes + c->dEntries.bytes + c->dWords.bytes);
		c->info.minimizer_window = window;
		c->haveIndex = true;
		if (out) *out = c->info;
}

} // namespace

int flyehip_build_index(flyehip_ctx* c, const flyehip_index_params_t* ip, flyehip_index_info_t* out)
{
	return guarded(c, [&]() { buildIndexImpl(c, ip, 0, out); });
}

int flyehip_build_index_minimizers(flyehip_ctx* c, const flyehip_index_params_t* ip, int32_t window,
								   flyehip_index_info_t* out)
{
	return guarded(c, [&]() {
		if (window < 1) fail(FLYEHIP_ERR_ARG, "minimizer window %d out of range 1..256", window);
		buildIndexImpl(c, ip, window, out);
	});
}

int flyehip_load_index(flyehip_ctx* c, int32_t kmer_size, const uint64_t* kmers, int64_t n_kmers,
					   const int64_t* bucket_offsets, const int64_t* entry_ids, const int32_t* entry_pos,
					   const uint64_t* repetitive, int64_t n_repetitive, flyehip_index_info_t* out)
{
	return guarded(c, [&]() {
		Lane* L = c->lanes[0].get();
		if (c->nReads <= 0) fail(FLYEHIP_ERR_STATE, "no reads stored");
		const int k = kmer_size;
		if (k < 1 || k > 31) fail(FLYEHIP_ERR_ARG, "kmer_size %d out of range 1..31", k);
		if (n_kmers < 0 || n_repetitive < 0 || (n_kmers > 0 && (!kmers || !bucket_offsets)) ||
			(n_repetitive > 0 && !repetitive))
			fail(FLYEHIP_ERR_ARG, "invalid index tables");
		c->haveIndex = false;
		c->indexLoaded = false;
		const int64_t nEntries = n_kmers > 0 ? bucket_offsets[n_kmers] - bucket_offsets[0] : 0;
		if (nEntries > 0 && (!entry_ids || !entry_pos)) fail(FLYEHIP_ERR_ARG, "null bucket entries");
		const uint64_t kmask = (uint64_t{1} << (2 * k)) - 1;
		auto rc = [&](uint64_t x) {
			uint64_t r = 0;
			for (int i = 0; i < k; ++i)
			{
				r = (r << 2) | (~x & 3ULL);
				x >>= 2;
			}
			return r;
		};
		// slot values: (start << 24 | count) for a bucket, kRepMark for a
		// repetitive k-mer; every k-mer must be in standard form
		const int64_t nKeys = n_kmers + n_repetitive;
		std::vector<uint64_t> keys(static_cast<size_t>(nKeys)), vals(static_cast<size_t>(nKeys));
		for (int64_t i = 0; i < n_kmers; ++i)
		{
			const uint64_t x = kmers[i];
			if (x > kmask || rc(x) < x) fail(FLYEHIP_ERR_ARG, "k-mer %lld (%llu) is not a canonical %d-mer",
										   static_cast<long long>(i), static_cast<unsigned long long>(x), k);
			const int64_t b = bucket_offsets[i] - bucket_offsets[0], e = bucket_offsets[i + 1] - bucket_offsets[0];
			if (e <= b || e - b > static_cast<int64_t>(kCountMask))
				fail(FLYEHIP_ERR_ARG, "bucket %lld has %lld entries (1 .. %llu)", static_cast<long long>(i),
					 static_cast<long long>(e - b), static_cast<unsigned long long>(kCountMask));
			keys[static_cast<size_t>(i)] = x;
			vals[static_cast<size_t>(i)] = (static_cast<uint64_t>(b) << kCountBits) | static_cast<uint64_t>(e - b);
		}
		for (int64_t i = 0; i < n_repetitive; ++i)
		{
			const uint64_t x = repetitive[i];
			if (x > kmask || rc(x) < x) fail(FLYEHIP_ERR_ARG, "repetitive k-mer %lld is not a canonical %d-mer",
										   static_cast<long long>(i), k);
			keys[static_cast<size_t>(n_kmers + i)] = x;
			vals[static_cast<size_t>(n_kmers + i)] = kRepMark;
		}
		// entries: (internal id << 32 | position) where the canonical k-mer
		// is spelled on that strand
		std::vector<uint64_t> ent(static_cast<size_t>(nEntries));
		for (int64_t j = 0; j < nEntries; ++j)
		{
			const int64_t id = entry_ids[j];
			const int64_t r = id > 0 ? id : -id;
			if (id == 0 || r > c->nReads) fail(FLYEHIP_ERR_ARG, "entry %lld names read %lld (1..%d stored)",
											   static_cast<long long>(j), static_cast<long long>(id), c->nReads);
			const int32_t p = entry_pos[j];
			if (p < 0 || p > c->readLen[static_cast<size_t>(r - 1)] - k)
				fail(FLYEHIP_ERR_ARG, "entry %lld: position %d outside read %lld (length %d)", static_cast<long long>(j),
					 p, static_cast<long long>(id), c->readLen[static_cast<size_t>(r - 1)]);
			ent[static_cast<size_t>(j)] = (static_cast<uint64_t>(toInternal(id)) << 32) | static_cast<uint32_t>(p);
		}
		uint64_t slots = 1024;
		while (slots < static_cast<uint64_t>(nKeys) * 3 / 2 + 1024) slots <<= 1;
		c->k = k;
		c->tableMask = slots - 1;
		ensure(c, c->dKeys, slots * 16);
		ensure(c, c->dCursor, slots * 4);
		ensure(c, c->dEntries, static_cast<size_t>(nEntries) * 8 + 8);
		emptyTable(L, static_cast<uint64_t*>(c->dKeys.p), static_cast<int64_t>(slots));
		uint64_t* dk = ensureT<uint64_t>(c, L->bCKey, static_cast<size_t>(nKeys) + 1);
		uint64_t* dv = ensureT<uint64_t>(c, L->bSKey, static_cast<size_t>(nKeys) + 1);
		uint32_t* dErr = ensureT<uint32_t>(c, L->bPackErr, 1);
		upload(L, dk, keys.data(), keys.size() * 8);
		upload(L, dv, vals.data(), vals.size() * 8);
		upload(L, c->dEntries.p, ent.data(), ent.size() * 8);
		HIPCHECK(hipMemsetAsync(dErr, 0, 4, L->stream));
		HIPCHECK(launchLoadKeys(table(c), dk, dv, nKeys, dErr, L->stream));
		uint32_t err = 0;
		download(L, &err, dErr, 4);
		sync(L);
		if (err) fail(FLYEHIP_ERR_ARG, "a k-mer appears twice in the index tables");
		c->info = flyehip_index_info_t{};
		c->info.kmer_size = k;
		c->info.distinct_kmers = nKeys;
		c->info.solid_kmers = n_kmers;
		c->info.repetitive_kmers = n_repetitive;
		c->info.indexed_positions = nEntries;
		c->info.hash_slots = static_cast<int64_t>(slots);
		c->info.device_bytes = static_cast<int64_t>(c->dKeys.bytes + c->dCursor.bytes + c->dEntries.bytes + c->dWords.bytes);
		c->haveIndex = true;
		c->indexLoaded = true;
		if (out) *out = c->info;
	});
}

int flyehip_collect_candidates(flyehip_ctx* c, const int64_t* query_ids, int32_t n, flyehip_candidate_t* matches,
							   int64_t match_cap, int64_t* match_offsets, int64_t* matches_needed, int32_t* filtered,
							   int64_t filtered_cap, int64_t* filtered_offsets, int64_t* filtered_needed)
{
	return guarded(c, [&]() {
		Lane* L = c->lanes[0].get();
		std::vector<JoinQuery> qs = residentQueries(c, query_ids, n);
		JoinArgs ja;
		BatchOut bo = runJoin(L, storeSeqs(c), qs, ja);
		std::vector<int64_t> qmo(static_cast<size_t>(n) + 1), qfo(static_cast<size_t>(n) + 1);
		download(L, qmo.data(), L->bQMatchOff.p, qmo.size() * 8);
		download(L, qfo.data(), L->bQFiltOff.p, qfo.size() * 8);
		sync(L);
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
		int64_t* dIds = ensureT<int64_t>(c, L->bQueryIds, static_cast<size_t>(n) + 1);
		upload(L, dIds, query_ids, static_cast<size_t>(n) * 8);
		flyehip_candidate_t* dc = ensureT<flyehip_candidate_t>(c, L->bCand, static_cast<size_t>(bo.matches) + 1);
		HIPCHECK(launchWriteCandidates(ja, static_cast<const int64_t*>(L->bQMatchOff.p), dIds, nullptr, dc, bo.matches, L->stream));
		download(L, matches, dc, static_cast<size_t>(bo.matches) * sizeof(flyehip_candidate_t));
		download(L, filtered, ja.filtered, static_cast<size_t>(bo.filtered) * 4);
		sync(L);
	});
}

int flyehip_candidates_canonical_sha256(flyehip_ctx* c, const int64_t* query_ids, int32_t n, uint8_t* sha256,
										int64_t* records)
{
	return guarded(c, [&]() {
		Lane* L = c->lanes[0].get();
		if (!sha256 || n < 0 || (n > 0 && !query_ids)) fail(FLYEHIP_ERR_ARG, "invalid arguments");
		resetStats(c);
		std::vector<int64_t> ids(query_ids, query_ids + n);
		std::sort(ids.begin(), ids.end());
		(void)residentQueries(c, ids.data(), n);
		CanonSha256 h;
		// query records per join batch: about 30 M matches (1 GB of records) at
		// the 20 kb ONT-like reads' 115 k matches per record
		const int32_t kBatch = 256;
		const int threads = std::max(1, std::min(16, static_cast<int>(std::thread::hardware_concurrency())));
		std::vector<flyehip_candidate_t> host;
		std::vector<int64_t> qmo;
		for (int32_t b0 = 0; b0 < n; b0 += kBatch)
		{
			const int32_t nb = std::min(kBatch, n - b0);
			std::vector<JoinQuery> qs = residentQueries(c, ids.data() + b0, nb);
			JoinArgs ja;
			const BatchOut bo = runJoin(L, storeSeqs(c), qs, ja);
			qmo.resize(static_cast<size_t>(nb) + 1);
			download(L, qmo.data(), L->bQMatchOff.p, qmo.size() * sizeof(int64_t));
			int64_t* dIds = ensureT<int64_t>(c, L->bQueryIds, static_cast<size_t>(nb) + 1);
			upload(L, dIds, ids.data() + b0, static_cast<size_t>(nb) * sizeof(int64_t));
			flyehip_candidate_t* dc = ensureT<flyehip_candidate_t>(c, L->bCand, static_cast<size_t>(bo.matches) + 1);
			HIPCHECK(launchWriteCandidates(ja, static_cast<const int64_t*>(L->bQMatchOff.p), dIds, nullptr, dc,
										   bo.matches, L->stream));
			host.resize(static_cast<size_t>(bo.matches));
			download(L, host.data(), dc, host.size() * sizeof(flyehip_candidate_t));
			sync(L);
			h.addQueries(host.data(), qmo.data(), nb, threads);
		}
		if (!h.finish(sha256)) fail(FLYEHIP_ERR_STATE, "canonical hash: %s", h.error().c_str());
		if (records) *records = h.records();
		c->stats.queries = n;
	});
}

int flyehip_collect_sorted_candidates(flyehip_ctx* c, const int64_t* query_ids, int32_t n, flyehip_candidate_t* matches,
									  int64_t match_cap, int64_t* match_offsets, int64_t* matches_needed)
{
	return guarded(c, [&]() {
		Lane* L = c->lanes[0].get();
		resetStats(c);
		std::vector<JoinQuery> qs = residentQueries(c, query_ids, n);
		JoinArgs ja;
		BatchOut bo = runJoin(L, storeSeqs(c), qs, ja);
		std::vector<int64_t> qmo(static_cast<size_t>(n) + 1);
		download(L, qmo.data(), L->bQMatchOff.p, qmo.size() * 8);
		sync(L);
		if (matches_needed) *matches_needed = bo.matches;
		for (int32_t i = 0; i <= n; ++i) match_offsets[i] = qmo[static_cast<size_t>(i)];
		if (bo.matches > match_cap)
			fail(FLYEHIP_ERR_CAPACITY, "capacity too small: %lld matches needed", static_cast<long long>(bo.matches));
		runMatchSort(L, ja, bo.matches);
		int64_t* dIds = ensureT<int64_t>(c, L->bQueryIds, static_cast<size_t>(n) + 1);
		upload(L, dIds, query_ids, static_cast<size_t>(n) * 8);
		flyehip_candidate_t* dc = ensureT<flyehip_candidate_t>(c, L->bCand, static_cast<size_t>(bo.matches) + 1);
		JoinArgs sorted = ja;
		sorted.mkey = static_cast<uint64_t*>(L->bSKey.p);
		HIPCHECK(launchWriteCandidates(sorted, static_cast<const int64_t*>(L->bQMatchOff.p), dIds,
									   static_cast<const uint32_t*>(L->bVal.p), dc, bo.matches, L->stream));
		download(L, matches, dc, static_cast<size_t>(bo.matches) * sizeof(flyehip_candidate_t));
		sync(L);
	});
}

int flyehip_overlaps(flyehip_ctx* c, const int64_t* query_ids, int32_t n, const flyehip_overlap_params_t* p,
					 flyehip_overlap_t* rows, int64_t row_cap, int64_t* row_offsets, int64_t* rows_needed,
					 float* div_stats, int64_t div_cap, int64_t* div_offsets)
{
	return guarded(c, [&]() {
		std::vector<JoinQuery> qs = residentQueries(c, query_ids, n);
		overlapsImpl(c, storeSeqs(c), qs, query_ids, p, rows, row_cap, row_offsets, rows_needed, div_stats, div_cap,
					 div_offsets, c->queriesPerBatch, 4);
	});
}

int flyehip_overlaps_submit(flyehip_ctx* c, const int64_t* query_ids, int32_t n, const flyehip_overlap_params_t* p,
							int32_t with_div_stats, uint64_t* ticket)
{
	return guarded(c, [&]() {
		if (!ticket) fail(FLYEHIP_ERR_ARG, "null ticket");
		if (n < 0 || (n > 0 && !query_ids)) fail(FLYEHIP_ERR_ARG, "invalid query list");
		checkParams(p, c->k);
		(void)residentQueries(c, query_ids, n);  // the ids name resident reads
		if (!c->async)
		{
			c->async = std::make_unique<AsyncQueue>();
			c->async->worker = std::thread(asyncWorker, c);
		}
		auto job = std::make_shared<AsyncJob>();
		job->ids.assign(query_ids, query_ids + n);
		job->p = *p;
		job->withDiv = with_div_stats != 0;
		AsyncQueue& q = *c->async;
		std::lock_guard<std::mutex> l(q.m);
		const uint64_t t = q.nextTicket++;
		q.jobs.emplace(t, job);
		q.pending.push_back(job);
		q.cv.notify_all();
		*ticket = t;
	});
}

int flyehip_overlaps_collect(flyehip_ctx* c, uint64_t ticket, int32_t query_index, flyehip_overlap_t* rows,
							 int64_t row_cap, int64_t* rows_needed, float* div_stats, int64_t div_cap,
							 int64_t* divs_needed)
{
	if (!c) return FLYEHIP_ERR_ARG;
	if (!c->async)
	{
		setError(c, "no job submitted");
		return FLYEHIP_ERR_STATE;
	}
	AsyncQueue& q = *c->async;
	int code = FLYEHIP_OK;
	std::string msg;
	{
		std::unique_lock<std::mutex> l(q.m);
		auto it = q.jobs.find(ticket);
		if (it == q.jobs.end())
		{
			code = FLYEHIP_ERR_ARG;
			msg = "unknown or released ticket";
		}
		else
		{
			std::shared_ptr<AsyncJob> job = it->second;
			const int32_t n = static_cast<int32_t>(job->ids.size());
			if (query_index < 0 || query_index >= n)
			{
				code = FLYEHIP_ERR_ARG;
				msg = "query index outside the submitted list";
			}
			else
			{
				q.cv.wait(l, [&] { return job->done > query_index; });
				if (static_cast<int32_t>(job->rowOff.size()) <= query_index + 1)
				{
					code = job->code;
					msg = job->msg;
				}
				else
				{
					const size_t i = static_cast<size_t>(query_index);
					const int64_t r0 = job->rowOff[i], nr = job->rowOff[i + 1] - r0;
					const int64_t d0 = job->divOff[i], nd = job->divOff[i + 1] - d0;
					if (rows_needed) *rows_needed = nr;
					if (divs_needed) *divs_needed = nd;
					if (nr > row_cap || (div_stats && nd > div_cap))
					{
						code = FLYEHIP_ERR_CAPACITY;
						msg = "capacity too small for the query's rows or divStats values";
					}
					else
					{
						if (nr) std::memcpy(rows, job->rows.data() + r0, static_cast<size_t>(nr) * sizeof(flyehip_overlap_t));
						if (div_stats && nd) std::memcpy(div_stats, job->divs.data() + d0, static_cast<size_t>(nd) * sizeof(float));
					}
				}
			}
		}
	}
	if (code != FLYEHIP_OK) setError(c, msg);
	return code;
}

int flyehip_overlaps_release(flyehip_ctx* c, uint64_t ticket)
{
	if (!c) return FLYEHIP_ERR_ARG;
	if (!c->async) return FLYEHIP_OK;
	AsyncQueue& q = *c->async;
	std::unique_lock<std::mutex> l(q.m);
	auto it = q.jobs.find(ticket);
	if (it == q.jobs.end()) return FLYEHIP_OK;
	std::shared_ptr<AsyncJob> job = it->second;
	q.cv.wait(l, [&] { return job->done >= static_cast<int32_t>(job->ids.size()); });
	q.jobs.erase(ticket);
	return FLYEHIP_OK;
}

int flyehip_overlaps_per_query(flyehip_ctx* c, const int64_t* query_ids, int32_t n, const flyehip_overlap_params_t* p,
							   flyehip_overlap_t* rows, int64_t row_cap, int64_t* row_offsets, int64_t* rows_needed)
{
	return guarded(c, [&]() {
		std::vector<JoinQuery> qs = residentQueries(c, query_ids, n);
		overlapsImpl(c, storeSeqs(c), qs, query_ids, p, rows, row_cap, row_offsets, rows_needed, nullptr, 0, nullptr, 1, 1);
	});
}

int flyehip_overlaps_external(flyehip_ctx* c, const uint8_t* bases, const int64_t* offsets, const int64_t* query_ids,
							  int32_t n, const flyehip_overlap_params_t* p, flyehip_overlap_t* rows, int64_t row_cap,
							  int64_t* row_offsets, int64_t* rows_needed)
{
	return guarded(c, [&]() {
		Lane* L = c->lanes[0].get();
		if (!c->haveIndex) fail(FLYEHIP_ERR_STATE, "no index: call flyehip_store_reads and flyehip_build_index first");
		std::vector<int64_t> wordOff;
		std::vector<int32_t> lens;
		packSeqsDevice(c, L, bases, offsets, n, L->bExtWords, L->bExtWordOff, L->bExtLen, wordOff, lens);
		PackedSeqs s;
		s.words = static_cast<const uint32_t*>(L->bExtWords.p);
		s.wordOff = static_cast<const int64_t*>(L->bExtWordOff.p);
		s.len = static_cast<const int32_t*>(L->bExtLen.p);
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
		overlapsImpl(c, s, qs, query_ids, p, rows, row_cap, row_offsets, rows_needed, nullptr, 0, nullptr, c->queriesPerBatch, 4);
	});
}

int flyehip_chain_lists(flyehip_ctx* c, const int32_t* cur_pos, const int32_t* ext_pos, const int64_t* list_offsets,
						int32_t n_lists, const int64_t* cur_ids, const int64_t* ext_ids, const int32_t* cur_lens,
						const int32_t* ext_lens, const flyehip_overlap_params_t* p, flyehip_overlap_t* rows,
						int64_t row_cap, int64_t* row_offsets, int64_t* rows_needed)
{
	return guarded(c, [&]() {
		Lane* L = c->lanes[0].get();
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
		int32_t* dc = ensureT<int32_t>(c, L->bHCur, nm);
		int32_t* de = ensureT<int32_t>(c, L->bHExt, nm);
		upload(L, dc, cur_pos + h0, static_cast<size_t>(nHits) * 4);
		upload(L, de, ext_pos + h0, static_cast<size_t>(nHits) * 4);
		GroupDesc* dg = ensureT<GroupDesc>(c, L->bGroups, gs.size() + 1);
		upload(L, dg, gs.data(), gs.size() * sizeof(GroupDesc));
		QueryDesc* dq = ensureT<QueryDesc>(c, L->bQDesc, qd.size() + 1);
		upload(L, dq, qd.data(), qd.size() * sizeof(QueryDesc));
		int32_t* dFilt = ensureT<int32_t>(c, L->bFiltered, 4);
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
		ca.sortKey = ensureT<uint32_t>(c, L->bCKey, nm);
		ca.sortVal = ensureT<uint32_t>(c, L->bCVal, nm);
		ca.posL = ensureT<int32_t>(c, L->bCPosL, nm);
		ca.posR = ensureT<int32_t>(c, L->bCPosR, nm);
		ca.hcur = ensureT<int32_t>(c, L->bCHCur, nm);
		ca.hext = ensureT<int32_t>(c, L->bCHExt, nm);
		ca.score = ensureT<int32_t>(c, L->bCScore, nm);
		ca.bt = ensureT<int32_t>(c, L->bCBt, nm);
		ca.props = ensureT<Proposal>(c, L->bCProps, p->only_max_ext ? 1 : nm);
		ca.outProps = ensureT<Proposal>(c, L->bCOut, p->only_max_ext ? gs.size() + 1 : nm);
		ca.groupCount = ensureT<int32_t>(c, L->bGroupCount, gs.size() + 1);
		ca.groupPass = ensureT<int32_t>(c, L->bGroupPass, gs.size() + 1);
		ca.logX = static_cast<const uint32_t*>(c->dLogX.p);
		ca.logY = static_cast<const uint32_t*>(c->dLogY.p);
		runChainKernels(L, ca, nHits);
		int64_t* rowStart = ensureT<int64_t>(c, L->bRowStart, gs.size() + 1);
		int64_t* tmp = ensureT<int64_t>(c, L->bScanTmp, scanTempWords(static_cast<int64_t>(gs.size())) + 64);
		HIPCHECK(exclusiveScan(ca.groupCount, rowStart, static_cast<int64_t>(gs.size()), tmp, L->stream));
		int64_t nRows = 0;
		download(L, &nRows, rowStart + gs.size(), 8);
		sync(L);
		flyehip_overlap_t* dr = ensureT<flyehip_overlap_t>(c, L->bRows, static_cast<size_t>(nRows) + 1);
		HIPCHECK(launchWriteRows(ca, rowStart, dr, L->stream));
		std::vector<int32_t> gc(gs.size());
		download(L, gc.data(), ca.groupCount, gs.size() * 4);
		if (nRows <= row_cap) download(L, rows, dr, static_cast<size_t>(nRows) * sizeof(flyehip_overlap_t));
		sync(L);
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
		Lane* L = c->lanes[0].get();
		if (n < 0 || (n > 0 && (!x || !out))) fail(FLYEHIP_ERR_ARG, "invalid logf arguments");
		const int64_t chunk = 1 << 26;
		float* dx = ensureT<float>(c, L->bDivOut, static_cast<size_t>(std::min(n, chunk)) + 1);
		float* dy = ensureT<float>(c, L->bCKey, static_cast<size_t>(std::min(n, chunk)) + 1);
		for (int64_t b = 0; b < n; b += chunk)
		{
			const int64_t m = std::min(chunk, n - b);
			upload(L, dx, x + b, static_cast<size_t>(m) * 4);
			HIPCHECK(launchLogf(dx, m, dy, static_cast<const uint32_t*>(c->dLogX.p),
								static_cast<const uint32_t*>(c->dLogY.p), c->logN, L->stream));
			download(L, out + b, dy, static_cast<size_t>(m) * 4);
			sync(L);
		}
	});
}

int flyehip_read_align_chains(flyehip_ctx* c, const flyehip_edge_alignment_t* alignments, const int64_t* aln_offsets,
							  int32_t n_reads, int32_t long_edge, int32_t big_aln, const flyehip_chain_params_t* params,
							  int32_t* sorted_index, int64_t* kept_offsets, flyehip_chain_segment_t* segments,
							  int64_t segment_cap, int64_t* segment_offsets, int64_t* segments_needed)
{
	return guarded(c, [&]() {
		Lane* L = c->lanes[0].get();
		if (!params) fail(FLYEHIP_ERR_ARG, "null chain parameters");
		resetStats(c);
		const int64_t m = aln_offsets[n_reads] - aln_offsets[0];
		const int64_t a0 = aln_offsets[0];
		std::vector<int64_t> off(static_cast<size_t>(n_reads) + 1);
		for (int32_t r = 0; r <= n_reads; ++r) off[static_cast<size_t>(r)] = aln_offsets[r] - a0;
		const size_t nm = static_cast<size_t>(m) + 1;
		flyehip_edge_alignment_t* dA = ensureT<flyehip_edge_alignment_t>(c, L->bAln, nm);
		int64_t* dOff = ensureT<int64_t>(c, L->bAlnOff, off.size());
		upload(L, dA, alignments + a0, static_cast<size_t>(m) * sizeof(flyehip_edge_alignment_t));
		upload(L, dOff, off.data(), off.size() * 8);
		ReadAlignArgs ra{};
		ra.aln = dA;
		ra.alnOff = dOff;
		ra.nReads = n_reads;
		ra.longEdge = long_edge;
		ra.bigAln = big_aln;
		ra.p = *params;
		ra.kept = ensureT<int32_t>(c, L->bCPosL, nm);
		ra.key = ensureT<uint64_t>(c, L->bCKey, nm);
		ra.val = ensureT<uint32_t>(c, L->bCVal, nm);
		ra.chainScore = ensureT<int32_t>(c, L->bCScore, nm);
		ra.chainPrev = ensureT<int32_t>(c, L->bCBt, nm);
		ra.chainLen = ensureT<int32_t>(c, L->bCPosR, nm);
		ra.active = ensureT<int32_t>(c, L->bCHCur, nm);
		ra.frozen = ensureT<int32_t>(c, L->bCHExt, nm);
		ra.accepted = ensureT<int32_t>(c, L->bHCur, nm);
		int32_t* counts = ensureT<int32_t>(c, L->bRA, 3 * static_cast<size_t>(n_reads) + 3);
		ra.keptCount = counts;
		ra.acceptedCount = counts + n_reads + 1;
		ra.segCount = counts + 2 * (n_reads + 1);
		Timer tc;
		HIPCHECK(launchReadAlignChain(ra, L->stream));
		int64_t* segOff = ensureT<int64_t>(c, L->bRowStart, static_cast<size_t>(n_reads) + 1);
		int64_t* tmp = ensureT<int64_t>(c, L->bScanTmp, scanTempWords(n_reads) + 64);
		HIPCHECK(exclusiveScan(ra.segCount, segOff, n_reads, tmp, L->stream));
		std::vector<int64_t> so(static_cast<size_t>(n_reads) + 1);
		std::vector<int32_t> kc(static_cast<size_t>(n_reads));
		download(L, so.data(), segOff, so.size() * 8);
		download(L, kc.data(), ra.keptCount, kc.size() * 4);
		sync(L);
		const int64_t nSeg = so.back();
		if (segments_needed) *segments_needed = nSeg;
		flyehip_chain_segment_t* dS = ensureT<flyehip_chain_segment_t>(c, L->bRows, static_cast<size_t>(nSeg) + 1);
		HIPCHECK(launchReadAlignWrite(ra, segOff, dS, L->stream));
		std::vector<int32_t> keptAll(static_cast<size_t>(m) + 1);
		download(L, keptAll.data(), ra.kept, static_cast<size_t>(m) * 4);
		if (nSeg <= segment_cap) download(L, segments, dS, static_cast<size_t>(nSeg) * sizeof(flyehip_chain_segment_t));
		sync(L);
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

int flyehip_read_align_chains_noalloc(flyehip_ctx* c, const flyehip_edge_alignment_t* d_alignments,
									  const int64_t* d_aln_offsets, int32_t n_reads, int32_t long_edge,
									  int32_t big_aln, const flyehip_chain_params_t* params, int32_t* d_sorted_index,
									  int64_t* d_kept_offsets, flyehip_chain_segment_t* d_segments,
									  int64_t segment_cap, int64_t* d_segment_offsets, int64_t* segments_needed)
{
	return guarded(c, [&]() {
		Lane* L = c->lanes[0].get();
		if (!params) fail(FLYEHIP_ERR_ARG, "null chain parameters");
		if (n_reads < 0 || n_reads > kNoAllocMaxReads)
			fail(FLYEHIP_ERR_ARG, "%d reads exceed the no-allocation limit of %d reads per call", n_reads,
				 kNoAllocMaxReads);
		if (n_reads > 0 && (!d_alignments || !d_aln_offsets || !d_sorted_index || !d_kept_offsets ||
							!d_segment_offsets || (segment_cap > 0 && !d_segments)))
			fail(FLYEHIP_ERR_ARG, "null device buffer");
		if (!c->scratch) fail(FLYEHIP_ERR_STATE, "no-allocation scratch not set up");
		resetStats(c);
		Timer tc;
		int64_t total = 0;
		if (n_reads > 0)
		{
			download(L, &total, d_aln_offsets + n_reads, sizeof(int64_t));
			sync(L);
		}
		if (total < 0 || total > kNoAllocMaxAlignments)
			fail(FLYEHIP_ERR_ARG, "aln_offsets[n_reads] = %lld exceeds the no-allocation limit of %lld alignments",
				 static_cast<long long>(total), static_cast<long long>(kNoAllocMaxAlignments));
		// carve the scratch array (offsets fixed at compile time)
		unsigned char* s = c->scratch;
		auto take = [&s](size_t bytes) {
			unsigned char* p = s;
			s += bytes;
			return p;
		};
		ReadAlignArgs ra{};
		ra.aln = d_alignments;
		ra.alnOff = d_aln_offsets;
		ra.nReads = n_reads;
		ra.longEdge = long_edge;
		ra.bigAln = big_aln;
		ra.p = *params;
		ra.key = reinterpret_cast<uint64_t*>(take(kNaU64));
		ra.val = reinterpret_cast<uint32_t*>(take(kNaI32));
		ra.kept = reinterpret_cast<int32_t*>(take(kNaI32));
		ra.chainScore = reinterpret_cast<int32_t*>(take(kNaI32));
		ra.chainPrev = reinterpret_cast<int32_t*>(take(kNaI32));
		ra.chainLen = reinterpret_cast<int32_t*>(take(kNaI32));
		ra.active = reinterpret_cast<int32_t*>(take(kNaI32));
		ra.frozen = reinterpret_cast<int32_t*>(take(kNaI32));
		ra.accepted = reinterpret_cast<int32_t*>(take(kNaI32));
		ra.keptCount = reinterpret_cast<int32_t*>(take(kNaRead));
		ra.acceptedCount = reinterpret_cast<int32_t*>(take(kNaRead));
		ra.segCount = reinterpret_cast<int32_t*>(take(kNaRead));
		int64_t* tmp = reinterpret_cast<int64_t*>(take(kNaScan));
		if (s != c->scratch + kNoAllocScratchBytes) fail(FLYEHIP_ERR_STATE, "scratch layout mismatch");
		HIPCHECK(launchReadAlignChain(ra, L->stream));
		HIPCHECK(exclusiveScan(ra.keptCount, d_kept_offsets, n_reads, tmp, L->stream));
		HIPCHECK(launchReadAlignKept(ra, d_kept_offsets, d_sorted_index, L->stream));
		HIPCHECK(exclusiveScan(ra.segCount, d_segment_offsets, n_reads, tmp, L->stream));
		int64_t nSeg = 0;
		download(L, &nSeg, d_segment_offsets + n_reads, sizeof(int64_t));
		sync(L);
		if (segments_needed) *segments_needed = nSeg;
		if (nSeg > segment_cap)
			fail(FLYEHIP_ERR_CAPACITY, "segment capacity %lld too small: %lld needed",
				 static_cast<long long>(segment_cap), static_cast<long long>(nSeg));
		HIPCHECK(launchReadAlignWrite(ra, d_segment_offsets, d_segments, L->stream));
		sync(L);
		c->stats.ms_chain = tc.ms();
		c->stats.queries = n_reads;
		c->stats.matches = total;
		c->stats.overlaps = nSeg;
	});
}

} // extern "C"
