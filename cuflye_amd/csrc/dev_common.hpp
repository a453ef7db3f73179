// dev_common.hpp - shared device helpers for the gfx950 kernels: wave64
// scans and reductions, k-mer arithmetic on the 2-bit packed read store, the
// hash used by the open-addressed k-mer table, and the id mapping.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/flyehip.h"

#define FH_WAVE 64

namespace fh
{

// ---- ids (tools/replay_read_to_graph_source_pack.py:85-88) -------------
__host__ __device__ inline uint32_t toInternal(int64_t sid)
{
	return sid > 0 ? static_cast<uint32_t>(2 * (sid - 1)) : static_cast<uint32_t>(2 * (-sid) - 1);
}
__host__ __device__ inline int64_t toSigned(uint32_t iid)
{
	return (iid & 1U) ? -static_cast<int64_t>((iid + 1) / 2) : static_cast<int64_t>(iid / 2 + 1);
}

// ---- wave64 helpers ------------------------------------------------------
__device__ inline int laneId() { return static_cast<int>(__lane_id()); }

// ---- XCD-aware block order ------------------------------------------------
// Workgroups are dealt to the 8 XCDs round robin by block index, and every
// XCD has its own L2.  For a kernel whose accesses are random inside a
// contiguous region of its elements (a query's matches), this index gives the
// blocks one XCD runs one contiguous eighth of the grid, so a region's lines
// are fetched and written back through one L2 instead of up to eight (partial
// lines from several L2s multiply the bytes written).  A bijection on
// [0, gridDim.x).
constexpr uint32_t kXcds = 8;
__device__ inline uint32_t xcdBlockIndex()
{
	const uint32_t n = gridDim.x, b = blockIdx.x;
	const uint32_t q = n / kXcds, r = n % kXcds;
	const uint32_t x = b % kXcds, k = b / kXcds;
	return x * q + (x < r ? x : r) + k;
}

// ---- segment lookup ------------------------------------------------------
// Largest r in [lo, n) with off[r] <= w (off ascending, empty segments
// allowed; off[lo] <= w).
__device__ inline int findSegFrom(const int64_t* off, int lo, int n, int64_t w)
{
	int hi = n - 1;
	while (lo < hi)
	{
		const int mid = (lo + hi + 1) >> 1;
		if (off[mid] <= w) lo = mid;
		else hi = mid - 1;
	}
	return lo;
}
__device__ inline int findSeg(const int64_t* off, int n, int64_t w) { return findSegFrom(off, 0, n, w); }

// findSeg for kernels whose lanes hold ascending elements (one element per
// thread in launch order): the binary search runs once per wave, for the
// first active lane's element, on wave-uniform (scalar) loads, and every lane
// steps forward over the few segment starts between that element and its
// own; a lane still short after four steps (a run of empty segments) finishes
// with a binary search over the rest.
__device__ inline int findSegWave(const int64_t* off, int n, int64_t w)
{
	const uint64_t u = static_cast<uint64_t>(w);
	const uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(u));
	const uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(u >> 32));
	const int64_t w0 = static_cast<int64_t>((static_cast<uint64_t>(hi) << 32) | lo);
	int r = findSeg(off, n, w0);
	for (int s = 0; s < 4 && r + 1 < n && off[r + 1] <= w; ++s) ++r;
	if (r + 1 < n && off[r + 1] <= w) r = findSegFrom(off, r + 1, n, w);
	return r;
}

// Segment of w, given the segment r of a position at or before w: steps
// forward over the segment starts in between (a binary search over the rest
// after four).
__device__ inline int stepSeg(const int64_t* off, int n, int r, int64_t w)
{
	for (int s = 0; s < 4 && r + 1 < n && off[r + 1] <= w; ++s) ++r;
	if (r + 1 < n && off[r + 1] <= w) r = findSegFrom(off, r + 1, n, w);
	return r;
}

__device__ inline int waveSum(int v)
{
	for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, FH_WAVE);
	return v;
}
__device__ inline long long waveSum64(long long v)
{
	for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, FH_WAVE);
	return v;
}
__device__ inline int waveMax(int v)
{
	for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, FH_WAVE));
	return v;
}
__device__ inline int waveMin(int v)
{
	for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, FH_WAVE));
	return v;
}
// inclusive prefix sum over lanes 0..63
__device__ inline int waveInclusiveSum(int v)
{
	const int l = laneId();
	for (int o = 1; o < FH_WAVE; o <<= 1)
	{
		int t = __shfl_up(v, o, FH_WAVE);
		if (l >= o) v += t;
	}
	return v;
}
// inclusive prefix max over lanes 0..63
__device__ inline int waveInclusiveMax(int v)
{
	const int l = laneId();
	for (int o = 1; o < FH_WAVE; o <<= 1)
	{
		int t = __shfl_up(v, o, FH_WAVE);
		if (l >= o) v = max(v, t);
	}
	return v;
}
__device__ inline long long waveInclusiveSum64(long long v)
{
	const int l = laneId();
	for (int o = 1; o < FH_WAVE; o <<= 1)
	{
		long long t = __shfl_up(v, o, FH_WAVE);
		if (l >= o) v += t;
	}
	return v;
}

// Exclusive scan of two counters over a block of NT threads (NT % 64 == 0).
// `sh` must hold 2 * (NT / 64) ints of LDS.  Returns the exclusive prefix of
// (a, b) and the block totals.
template <int NT>
__device__ inline void blockScan2(int a, int b, int& exA, int& exB, int& totA, int& totB, int* sh)
{
	const int w = static_cast<int>(threadIdx.x) / FH_WAVE;
	int ia = waveInclusiveSum(a);
	int ib = waveInclusiveSum(b);
	if (laneId() == FH_WAVE - 1)
	{
		sh[2 * w] = ia;
		sh[2 * w + 1] = ib;
	}
	__syncthreads();
	int baseA = 0, baseB = 0;
	totA = 0;
	totB = 0;
	for (int i = 0; i < NT / FH_WAVE; ++i)
	{
		int sa = sh[2 * i], sb = sh[2 * i + 1];
		if (i < w)
		{
			baseA += sa;
			baseB += sb;
		}
		totA += sa;
		totB += sb;
	}
	exA = baseA + ia - a;
	exB = baseB + ib - b;
	__syncthreads();
}

template <int NT>
__device__ inline int blockSum(int a, int* sh)
{
	const int w = static_cast<int>(threadIdx.x) / FH_WAVE;
	int s = waveSum(a);
	if (laneId() == 0) sh[w] = s;
	__syncthreads();
	int t = 0;
	for (int i = 0; i < NT / FH_WAVE; ++i) t += sh[i];
	__syncthreads();
	return t;
}

// ---- k-mers on the packed store -------------------------------------------
// Read n's forward strand is stored 16 bases per uint32 word, MSB first
// (base 0 in bits 31..30), starting at word wordOff[n]; two spare words
// follow every read so a window never reads past the allocation.
__device__ inline uint64_t kmerAt(const uint32_t* words, int64_t wordOff, int32_t pos, int k)
{
	const uint32_t* w = words + wordOff + (pos >> 4);
	const int off = pos & 15;  // bases into the first word
	uint64_t hi = (static_cast<uint64_t>(w[0]) << 32) | w[1];
	if (off + k <= 32)
	{
		uint64_t v = hi >> (2 * (32 - off - k));
		return k == 32 ? v : (v & ((1ULL << (2 * k)) - 1));
	}
	// k up to 31 with off up to 15: take a third word
	const int need = off + k - 32;  // bases from w[2]
	uint64_t a = (hi << (2 * need)) | (static_cast<uint64_t>(w[2]) >> (2 * (16 - need)));
	return a & ((1ULL << (2 * k)) - 1);
}

__host__ __device__ inline uint64_t rcKmer(uint64_t x, int k)
{
#if defined(__HIP_DEVICE_COMPILE__)
	uint64_t v = ~x;
	uint32_t lo = static_cast<uint32_t>(v), hi = static_cast<uint32_t>(v >> 32);
	uint64_t r = (static_cast<uint64_t>(__builtin_bitreverse32(lo)) << 32) | __builtin_bitreverse32(hi);
	r = ((r >> 1) & 0x5555555555555555ULL) | ((r & 0x5555555555555555ULL) << 1);
	return r >> (64 - 2 * k);
#else
	uint64_t r = 0;
	for (int i = 0; i < k; ++i)
	{
		r = (r << 2) | (~x & 3ULL);
		x >>= 2;
	}
	return r;
#endif
}

// Kmer::standardForm: canonical = min(x, rc); flipped unless x < rc.
__host__ __device__ inline uint64_t canonical(uint64_t x, int k, bool& flip)
{
	uint64_t r = rcKmer(x, k);
	flip = !(x < r);
	return flip ? r : x;
}

// 64-bit mixer (splitmix64 finaliser) for the open-addressed table.
__host__ __device__ inline uint64_t hashKmer(uint64_t x)
{
	x ^= x >> 30;
	x *= 0xbf58476d1ce4e5b9ULL;
	x ^= x >> 27;
	x *= 0x94d049bb133111ebULL;
	x ^= x >> 31;
	return x;
}

// Minimizer order of a canonical k-mer: the same mixer, so the order is a
// fixed pseudo-random permutation of k-mers (ties only between equal k-mers).
__host__ __device__ inline uint64_t minimizerOrder(uint64_t canonicalKmer) { return hashKmer(canonicalKmer); }

constexpr uint64_t kEmptyKey = ~0ULL;
constexpr uint64_t kRepMark = ~0ULL;
constexpr int kCountBits = 24;
constexpr uint64_t kCountMask = (1ULL << kCountBits) - 1;
constexpr uint64_t kWinFlipBit = 1ULL << 63;  // join windows: flip flag beside a bucket value

// Probe an interleaved (key, value) slot array for `key`: the slot, or -1
// when absent; *value receives the slot's value (read with the key, one
// 16-byte load per probe).
__device__ inline int64_t findSlot(const uint64_t* slots, uint64_t mask, uint64_t key, uint64_t* value = nullptr)
{
	uint64_t h = hashKmer(key) & mask;
	for (;;)
	{
		const ulonglong2 kv = *reinterpret_cast<const ulonglong2*>(slots + 2 * h);
		if (kv.x == key)
		{
			if (value) *value = kv.y;
			return static_cast<int64_t>(h);
		}
		if (kv.x == kEmptyKey) return -1;
		h = (h + 1) & mask;
	}
}

} // namespace fh
