// pipeline.hip - gfx950 kernels of the k-mer stage of getSeqOverlaps:
// the VertexIndex build (count in an open-addressed HBM hash table, histogram,
// thresholds applied, bucket fill in (read, position) order), the query k-mer
// join in Flye's emission order (window count, scan, fill), the group split,
// divStats, and the device scan.
//
// Data layout in HBM (all sizes per element):
//   packed reads   2 bits/base, 16 bases per uint32 word, each read word aligned
//   hash table     keys u64 + vals u64 + fill cursor u32 per slot (power-of-two
//                  slots, linear probing, splitmix64 finaliser as the hash)
//   bucket entries u64 (internal id << 32 | position), bucket = contiguous run
//   join windows   count i32 + filtered i32 + slot/flip u32 per query window
//   matches        key u64 (internal extId << 32 | curPos) + extPos i32
#include <cstring>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/iterator/transform_iterator.hpp>

#include "pipeline.hpp"
#include "dev_common.hpp"

namespace fh
{

namespace
{

constexpr int kScanNT = 256;
constexpr int kScanIT = 16;
constexpr int kScanTile = kScanNT * kScanIT;

__device__ inline int64_t blockExclusiveScan64(int64_t v, int64_t& total, int64_t* sh)
{
	const int w = static_cast<int>(threadIdx.x) / FH_WAVE;
	const int nw = static_cast<int>(blockDim.x) / FH_WAVE;
	int64_t incl = waveInclusiveSum64(v);
	if (laneId() == FH_WAVE - 1) sh[w] = incl;
	__syncthreads();
	int64_t base = 0;
	total = 0;
	for (int i = 0; i < nw; ++i)
	{
		if (i < w) base += sh[i];
		total += sh[i];
	}
	__syncthreads();
	return base + incl - v;
}

__global__ void __launch_bounds__(kScanNT) scanReduceKernel(const int32_t* in, int64_t n, int64_t* blockSums)
{
	__shared__ int64_t sh[kScanNT / FH_WAVE];
	const int64_t b0 = static_cast<int64_t>(blockIdx.x) * kScanTile;
	int64_t s = 0;
	for (int i = 0; i < kScanIT; ++i)
	{
		const int64_t idx = b0 + static_cast<int64_t>(i) * kScanNT + threadIdx.x;
		if (idx < n) s += in[idx];
	}
	int64_t total;
	blockExclusiveScan64(s, total, sh);
	if (threadIdx.x == 0) blockSums[blockIdx.x] = total;
}

__global__ void __launch_bounds__(1024) scanBlockSumsKernel(int64_t* sums, int64_t nb)
{
	__shared__ int64_t sh[1024 / FH_WAVE];
	const int64_t per = (nb + 1023) / 1024;
	const int64_t b = static_cast<int64_t>(threadIdx.x) * per;
	const int64_t e = min(nb, b + per);
	int64_t s = 0;
	for (int64_t i = b; i < e; ++i) s += sums[i];
	int64_t total;
	int64_t ex = blockExclusiveScan64(s, total, sh);
	for (int64_t i = b; i < e; ++i)
	{
		const int64_t v = sums[i];
		sums[i] = ex;
		ex += v;
	}
	if (threadIdx.x == 0) sums[nb] = total;
}

// LDS slot of tile position p: one pad word per 16, so a thread reading its
// kScanIT consecutive positions and the block reading a row of kScanNT
// positions both spread over the banks.
__device__ inline int scanPad(int p) { return p + (p >> 4); }

// Tile-wise exclusive scan with the tile's offset from scanBlockSumsKernel.
// Loads and stores are row-major across the block (each wave touches
// consecutive words); the per-thread runs of kScanIT consecutive values the
// scan works on are transposed through LDS.
__global__ void __launch_bounds__(kScanNT) scanDownKernel(const int32_t* in, int64_t* out, int64_t n,
														  const int64_t* blockSums, int64_t nb)
{
	__shared__ int64_t sh[kScanNT / FH_WAVE];
	__shared__ int64_t tile[kScanTile + kScanTile / 16];
	const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanTile;
	const int t = static_cast<int>(threadIdx.x);
	for (int i = 0; i < kScanIT; ++i)
	{
		const int p = i * kScanNT + t;
		const int64_t idx = base + p;
		tile[scanPad(p)] = idx < n ? in[idx] : 0;
	}
	__syncthreads();
	int64_t v[kScanIT];
	int64_t s = 0;
	for (int i = 0; i < kScanIT; ++i)
	{
		v[i] = tile[scanPad(t * kScanIT + i)];
		s += v[i];
	}
	int64_t total;
	int64_t ex = blockExclusiveScan64(s, total, sh) + blockSums[blockIdx.x];
	for (int i = 0; i < kScanIT; ++i)
	{
		tile[scanPad(t * kScanIT + i)] = ex;
		ex += v[i];
	}
	__syncthreads();
	for (int i = 0; i < kScanIT; ++i)
	{
		const int p = i * kScanNT + t;
		const int64_t idx = base + p;
		if (idx < n) out[idx] = tile[scanPad(p)];
	}
	if (blockIdx.x == 0 && threadIdx.x == 0) out[n] = blockSums[nb];
}

__global__ void countKernel(PackedSeqs reads, const int64_t* winOff, int64_t nWin, int k, HashTable t)
{
	const int64_t w = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
	if (w >= nWin) return;
	const int r = findSegWave(winOff, reads.n, w);
	const int32_t pos = static_cast<int32_t>(w - winOff[r]);
	bool flip;
	const uint64_t c = canonical(kmerAt(reads.words, reads.wordOff[r], pos, k), k, flip);
	uint64_t h = hashKmer(c) & t.mask;
	for (;;)
	{
		const uint64_t key = t.key(h);
		if (key == c)
		{
			atomicAdd(reinterpret_cast<unsigned long long*>(&t.val(h)), 1ULL);
			return;
		}
		if (key == kEmptyKey)
		{
			const uint64_t prev = atomicCAS(reinterpret_cast<unsigned long long*>(&t.key(h)),
											static_cast<unsigned long long>(kEmptyKey),
											static_cast<unsigned long long>(c));
			if (prev == kEmptyKey || prev == c)
			{
				atomicAdd(reinterpret_cast<unsigned long long*>(&t.val(h)), 1ULL);
				return;
			}
		}
		h = (h + 1) & t.mask;
	}
}

constexpr int kHistLds = 1024;

__global__ void __launch_bounds__(256) histKernel(HashTable t, uint64_t* hist, int hmax, uint64_t* acc)
{
	__shared__ unsigned int lh[kHistLds];
	__shared__ unsigned int distinct;
	for (int i = threadIdx.x; i < kHistLds; i += blockDim.x) lh[i] = 0;
	if (threadIdx.x == 0) distinct = 0;
	__syncthreads();
	const uint64_t slots = t.mask + 1;
	unsigned int myDistinct = 0;
	for (uint64_t h = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; h < slots;
		 h += static_cast<uint64_t>(gridDim.x) * blockDim.x)
	{
		if (t.key(h) == kEmptyKey) continue;
		++myDistinct;
		const uint64_t c = t.val(h);
		if (c < static_cast<uint64_t>(kHistLds)) atomicAdd(&lh[c], 1U);
		else if (c < static_cast<uint64_t>(hmax)) atomicAdd(reinterpret_cast<unsigned long long*>(&hist[c]), 1ULL);
		else
		{
			atomicAdd(reinterpret_cast<unsigned long long*>(&acc[0]), static_cast<unsigned long long>(c));
			atomicAdd(reinterpret_cast<unsigned long long*>(&acc[1]), 1ULL);
		}
	}
	atomicAdd(&distinct, myDistinct);
	__syncthreads();
	for (int i = threadIdx.x; i < kHistLds && i < hmax; i += blockDim.x)
		if (lh[i]) atomicAdd(reinterpret_cast<unsigned long long*>(&hist[i]), static_cast<unsigned long long>(lh[i]));
	if (threadIdx.x == 0) atomicAdd(reinterpret_cast<unsigned long long*>(&acc[2]), static_cast<unsigned long long>(distinct));
}

__global__ void classifyKernel(HashTable t, int solidMin, int repMax, int32_t* sizes, uint64_t* counters)
{
	const uint64_t h = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
	if (h > t.mask) return;
	int32_t size = 0;
	if (t.key(h) != kEmptyKey)
	{
		const uint64_t c = t.val(h);
		if (c >= static_cast<uint64_t>(solidMin) && c <= static_cast<uint64_t>(repMax))
		{
			size = static_cast<int32_t>(c);
			atomicAdd(reinterpret_cast<unsigned long long*>(&counters[0]), 1ULL);
		}
		else if (c >= static_cast<uint64_t>(solidMin))
		{
			t.val(h) = kRepMark;
			atomicAdd(reinterpret_cast<unsigned long long*>(&counters[1]), 1ULL);
		}
		else
		{
			t.val(h) = 0;
		}
	}
	sizes[h] = size;
	t.cursor[h] = 0;
}

__device__ inline bool indexed(uint64_t v) { return v != 0 && v != kRepMark; }

// A solid, non-repetitive k-mer with no bucket entry (a minimizer index where
// the k-mer is nowhere a minimizer) is not in the index: value 0.
__global__ void assignKernel(HashTable t, const int32_t* sizes, const int64_t* starts)
{
	const uint64_t h = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
	if (h > t.mask) return;
	const int32_t s = sizes[h];
	if (s > 0) t.val(h) = (static_cast<uint64_t>(starts[h]) << kCountBits) | static_cast<uint64_t>(s);
	else if (indexed(t.val(h))) t.val(h) = 0;
}

// Minimizer windows (buildIndexMinimizers' selection, as defined by this
// library): over every run of w consecutive k-mer positions of a read (one
// run of all its positions when it has fewer than w), the position whose
// canonical k-mer has the smallest minimizerOrder, the leftmost on equal
// values, is a minimizer.  One thread per window start; mask[window] = 1 at
// every selected position (equal stores from several threads).
__global__ void minimizerMaskKernel(PackedSeqs reads, const int64_t* winOff, int64_t nWin, int k, int w, uint8_t* mask)
{
	const int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
	if (g >= nWin) return;
	const int r = findSegWave(winOff, reads.n, g);
	const int64_t w0 = winOff[r];
	const int32_t nk = static_cast<int32_t>(winOff[r + 1] - w0);
	const int32_t s = static_cast<int32_t>(g - w0);
	const int32_t last = nk >= w ? nk - w : 0;  // last window start
	if (s > last) return;
	const int32_t e = min(nk, s + w);
	int32_t best = s;
	uint64_t bestOrd = ~0ULL;
	for (int32_t p = s; p < e; ++p)
	{
		bool flip;
		const uint64_t o = minimizerOrder(canonical(kmerAt(reads.words, reads.wordOff[r], p, k), k, flip));
		if (p == s || o < bestOrd)
		{
			bestOrd = o;
			best = p;
		}
	}
	mask[w0 + best] = 1;
}

// Bucket sizes of a minimizer index: selected positions of indexed k-mers.
__global__ void maskedSizesKernel(PackedSeqs reads, const int64_t* winOff, int64_t nWin, int k, HashTable t,
								  const uint8_t* mask, int32_t* sizes)
{
	const int64_t w = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
	if (w >= nWin || !mask[w]) return;
	const int r = findSegWave(winOff, reads.n, w);
	const int32_t pos = static_cast<int32_t>(w - winOff[r]);
	bool flip;
	const uint64_t c = canonical(kmerAt(reads.words, reads.wordOff[r], pos, k), k, flip);
	uint64_t v = 0;
	const int64_t h = findSlot(t.slots, t.mask, c, &v);
	if (h >= 0 && indexed(v)) atomicAdd(&sizes[h], 1);
}

__global__ void loadKeysKernel(HashTable t, const uint64_t* keys, const uint64_t* vals, int64_t n, uint32_t* err)
{
	const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
	if (i >= n) return;
	const uint64_t c = keys[i];
	uint64_t h = hashKmer(c) & t.mask;
	for (;;)
	{
		const uint64_t prev = atomicCAS(reinterpret_cast<unsigned long long*>(&t.key(h)),
										static_cast<unsigned long long>(kEmptyKey), static_cast<unsigned long long>(c));
		if (prev == kEmptyKey)
		{
			t.val(h) = vals[i];
			t.cursor[h] = 0;
			return;
		}
		if (prev == c)
		{
			atomicOr(err, 1u);
			return;
		}
		h = (h + 1) & t.mask;
	}
}

__global__ void fillKernel(PackedSeqs reads, const int64_t* winOff, int64_t nWin, int k, HashTable t,
						   const uint8_t* mask, uint64_t* ordKeys)
{
	const int64_t w = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
	if (w >= nWin || (mask && !mask[w])) return;
	const int r = findSegWave(winOff, reads.n, w);
	const int32_t pos = static_cast<int32_t>(w - winOff[r]);
	bool flip;
	const uint64_t c = canonical(kmerAt(reads.words, reads.wordOff[r], pos, k), k, flip);
	uint64_t v = 0;
	const int64_t h = findSlot(t.slots, t.mask, c, &v);
	if (h < 0) return;
	if (!indexed(v)) return;
	const uint32_t slot = atomicAdd(&t.cursor[h], 1U);
	const uint64_t start = v >> kCountBits;
	ordKeys[start + slot] = (static_cast<uint64_t>(r) << 33) | (static_cast<uint64_t>(pos) << 1) | (flip ? 1ULL : 0ULL);
}

// Bucket entry of one fill key (read << 33 | pos << 1 | flip): the stored
// strand's internal id and the position where it spells the canonical k-mer.
__device__ inline uint64_t bucketEntry(uint64_t x, const int32_t* readLen, int k)
{
	const uint32_t r = static_cast<uint32_t>(x >> 33);
	const int32_t pos = static_cast<int32_t>((x >> 1) & 0xffffffffULL);
	if (x & 1ULL) return (static_cast<uint64_t>(2 * r + 1) << 32) | static_cast<uint32_t>(readLen[r] - pos - k);
	return (static_cast<uint64_t>(2 * r) << 32) | static_cast<uint32_t>(pos);
}

// Buckets of at most kBucketSerial entries are put in (read, pos) order by
// their own thread (insertion sort); larger ones are listed for
// bucketBlockSortKernel.
constexpr int kBucketSerial = 32;

__global__ void bucketFinishKernel(HashTable t, uint64_t* ordKeys, uint64_t* entries, const int32_t* readLen, int k,
								   uint32_t* bigList, unsigned long long* bigCount, uint64_t bigCap)
{
	const uint64_t h = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
	if (h > t.mask) return;
	const uint64_t v = t.val(h);
	if (!indexed(v)) return;
	const uint64_t start = v >> kCountBits;
	const int n = static_cast<int>(v & kCountMask);
	if (n > kBucketSerial)
	{
		const unsigned long long i = atomicAdd(bigCount, 1ULL);
		if (i < bigCap) bigList[i] = static_cast<uint32_t>(h);
		return;
	}
	uint64_t* o = ordKeys + start;
	for (int i = 1; i < n; ++i)
	{
		const uint64_t x = o[i];
		int j = i;
		while (j > 0 && o[j - 1] > x)
		{
			o[j] = o[j - 1];
			--j;
		}
		o[j] = x;
	}
	for (int i = 0; i < n; ++i) entries[start + i] = bucketEntry(o[i], readLen, k);
}

// One block per listed bucket (grid-stride over the list): an ascending
// bitonic network over the bucket in place, the positions past its end taken
// as +infinity (every compare-exchange puts the minimum at the lower index,
// so a pair with its upper index past the end never swaps and is skipped).
// The keys of a bucket are distinct, so the result is the (read, pos) order.
__global__ void __launch_bounds__(1024) bucketBlockSortKernel(HashTable t, uint64_t* ordKeys, uint64_t* entries,
															   const int32_t* readLen, int k, const uint32_t* bigList,
															   const unsigned long long* bigCount)
{
	const uint64_t nb = *bigCount;
	for (uint64_t b = blockIdx.x; b < nb; b += gridDim.x)
	{
		const uint64_t v = t.val(bigList[b]);
		const uint64_t start = v >> kCountBits;
		const int64_t n = static_cast<int64_t>(v & kCountMask);
		uint64_t* o = ordKeys + start;
		int64_t np2 = 1;
		while (np2 < n) np2 <<= 1;
		for (int64_t kk = 2; kk <= np2; kk <<= 1)
		{
			const int64_t half = kk >> 1;
			for (int64_t p = threadIdx.x; p < (np2 >> 1); p += blockDim.x)
			{
				const int64_t i = (p / half) * kk + (p % half);
				const int64_t j = (p / half) * kk + kk - 1 - (p % half);
				if (j < n)
				{
					const uint64_t a = o[i], c = o[j];
					if (c < a)
					{
						o[i] = c;
						o[j] = a;
					}
				}
			}
			__syncthreads();
			for (int64_t jj = half >> 1; jj >= 1; jj >>= 1)
			{
				for (int64_t p = threadIdx.x; p < (np2 >> 1); p += blockDim.x)
				{
					const int64_t i = (p / jj) * (2 * jj) + (p % jj);
					const int64_t j = i + jj;
					if (j < n)
					{
						const uint64_t a = o[i], c = o[j];
						if (c < a)
						{
							o[i] = c;
							o[j] = a;
						}
					}
				}
				__syncthreads();
			}
		}
		for (int64_t i = threadIdx.x; i < n; i += blockDim.x) entries[start + i] = bucketEntry(o[i], readLen, k);
		__syncthreads();
	}
}

// ---- join -----------------------------------------------------------------

struct WindowKmer
{
	int q;
	int32_t pos;
	uint64_t y;     // query k-mer (numRepr)
	uint64_t c;     // canonical
	bool g;         // flipped
	uint64_t xf;    // forward-strand k-mer of the same window
};

__device__ inline WindowKmer windowAt(const JoinArgs& a, int64_t w)
{
	WindowKmer r;
	r.q = findSegWave(a.winOff, a.nQueries, w);
	r.pos = static_cast<int32_t>(w - a.winOff[r.q]);
	const JoinQuery Q = a.queries[r.q];
	const int64_t wo = a.seqs.wordOff[Q.seq];
	if (Q.strand)
	{
		r.xf = kmerAt(a.seqs.words, wo, r.pos, a.k);
		r.y = r.xf;
	}
	else
	{
		r.xf = kmerAt(a.seqs.words, wo, Q.len - r.pos - a.k, a.k);
		r.y = rcKmer(r.xf, a.k);
	}
	r.c = canonical(r.y, a.k, r.g);
	return r;
}

__device__ inline void transformEntry(uint64_t e, bool g, const int32_t* targetLen, int k, uint32_t& id, int32_t& pos)
{
	id = static_cast<uint32_t>(e >> 32);
	pos = static_cast<int32_t>(static_cast<uint32_t>(e));
	if (g)
	{
		id ^= 1U;
		pos = targetLen[id >> 1] - pos - k;
	}
}

// Self entry of a window inside its bucket: -1 when the window cannot hit
// itself.  Resident reads: the bucket is in (read, forward position) order, so
// the entry stored for the query read's own window is found by binary search
// (absent when the index holds minimizer positions only and this one is
// not); other queries scan the bucket.
__device__ inline int selfIndex(const JoinArgs& a, const JoinQuery& Q, int32_t pos, bool g, uint64_t start, int n)
{
	if (Q.selfMode == 0) return -1;
	if (Q.selfMode == 2)
	{
		for (int i = 0; i < n; ++i)
		{
			uint32_t id;
			int32_t p;
			transformEntry(a.entries[start + i], g, a.targetLen, a.k, id, p);
			if (id == Q.iid && p == pos) return i;
		}
		return -1;
	}
	const uint32_t r = static_cast<uint32_t>(Q.seq);
	const int32_t pf = Q.strand ? pos : Q.len - pos - a.k;
	int lo = 0, hi = n;
	while (lo < hi)
	{
		const int mid = (lo + hi) >> 1;
		const uint64_t e = a.entries[start + mid];
		const uint32_t id = static_cast<uint32_t>(e >> 32);
		const uint32_t rd = id >> 1;
		const int32_t ep = static_cast<int32_t>(static_cast<uint32_t>(e));
		const int32_t fp = (id & 1U) ? a.targetLen[rd] - ep - a.k : ep;
		if (rd < r || (rd == r && fp < pf)) lo = mid + 1;
		else hi = mid;
	}
	if (lo >= n) return -1;
	uint32_t id;
	int32_t p;
	transformEntry(a.entries[start + lo], g, a.targetLen, a.k, id, p);
	return (id == Q.iid && p == pos) ? lo : -1;
}

__global__ void joinCountKernel(JoinArgs a)
{
	const int64_t w = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
	if (w >= a.nWin) return;
	const WindowKmer wk = windowAt(a, w);
	int32_t count = 0;
	int32_t filt = 0;
	uint64_t info = 0;
	uint64_t v = 0;
	const int64_t h = findSlot(a.table.slots, a.table.mask, wk.c, &v);
	if (h >= 0)
	{
		if (v == kRepMark) filt = 1;
		else if (v != 0)
		{
			count = static_cast<int32_t>(v & kCountMask);
			const JoinQuery Q = a.queries[wk.q];
			int self = 0;
			if (Q.selfMode == 1)
			{
				// the query read's own stored entry for this window
				const int32_t pf = Q.strand ? wk.pos : Q.len - wk.pos - a.k;
				bool ff;
				canonical(wk.xf, a.k, ff);
				uint64_t e = ff ? ((static_cast<uint64_t>(2 * Q.seq + 1) << 32) | static_cast<uint32_t>(Q.len - pf - a.k))
								: ((static_cast<uint64_t>(2 * Q.seq) << 32) | static_cast<uint32_t>(pf));
				uint32_t id;
				int32_t p;
				transformEntry(e, wk.g, a.targetLen, a.k, id, p);
				self = (id == Q.iid && p == wk.pos) ? 1 : 0;
			}
			else if (Q.selfMode == 2)
			{
				const uint64_t start = v >> kCountBits;
				for (int i = 0; i < count; ++i)
				{
					uint32_t id;
					int32_t p;
					transformEntry(a.entries[start + i], wk.g, a.targetLen, a.k, id, p);
					if (id == Q.iid && p == wk.pos) self = 1;
				}
			}
			else if (Q.selfMode == 3)
			{
				self = selfIndex(a, Q, wk.pos, wk.g, v >> kCountBits, count) >= 0 ? 1 : 0;
			}
			count -= self;
			// the bucket (start << 24 | count) and the window's flip in bit 63
			// (bucket starts stay below 2^39): the fill reads it here instead
			// of probing the table again
			info = v | (wk.g ? kWinFlipBit : 0ULL);
		}
	}
	a.winCount[w] = count;
	a.winFilt[w] = filt;
	a.winVal[w] = info;
}

// One wave fills the matches of 64 consecutive windows together: the raw
// bucket entries of the 64 windows are enumerated across the lanes (entry t
// of the wave goes to lane t % 64), so bucket reads and match writes are
// contiguous runs instead of one run per lane.
__global__ void __launch_bounds__(256) joinFillKernel(JoinArgs a)
{
	const int lane = laneId();
	const int64_t w0 = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) - lane;
	const int64_t w = w0 + lane;
	int raw = 0;
	uint64_t start = 0;
	int self = -1;
	int64_t outBase = 0;
	int64_t qBase = 0;
	uint32_t qiid = 0;
	int32_t pos = 0;
	int gflag = 0;
	if (w < a.nWin)
	{
		const int q = findSegWave(a.winOff, a.nQueries, w);
		pos = static_cast<int32_t>(w - a.winOff[q]);
		if (a.winFilt[w]) a.filtered[a.filtOff[w]] = pos;
		if (a.winCount[w] > 0)
		{
			const uint64_t info = a.winVal[w];
			const bool g = (info & kWinFlipBit) != 0;
			const uint64_t v = info & ~kWinFlipBit;
			start = v >> kCountBits;
			raw = static_cast<int>(v & kCountMask);
			const JoinQuery Q = a.queries[q];
			qiid = Q.iid;
			gflag = g ? 1 : 0;
			self = selfIndex(a, Q, pos, g, start, raw);
			outBase = a.matchOff[w];
			qBase = a.qMatchOff[q];
		}
	}
	const int incl = waveInclusiveSum(raw);
	const int total = __shfl(incl, FH_WAVE - 1, FH_WAVE);
	const int excl = incl - raw;
	for (int t0 = 0; t0 < total; t0 += FH_WAVE)
	{
		const int t = t0 + lane;
		// window of entry t: largest lane l with excl(l) <= t
		int lo = 0;
		for (int step = 32; step > 0; step >>= 1)
		{
			const int cand = lo + step;
			const int e = __shfl(excl, cand < FH_WAVE ? cand : FH_WAVE - 1, FH_WAVE);
			if (cand < FH_WAVE && e <= t) lo = cand;
		}
		const int l = lo;
		const int rawL = __shfl(raw, l, FH_WAVE);
		const int exL = __shfl(excl, l, FH_WAVE);
		const uint64_t stL = __shfl(start, l, FH_WAVE);
		const int selfL = __shfl(self, l, FH_WAVE);
		const int64_t obL = __shfl(outBase, l, FH_WAVE);
		const int64_t qbL = __shfl(qBase, l, FH_WAVE);
		const int gL = __shfl(gflag, l, FH_WAVE);
		const int32_t posL = __shfl(pos, l, FH_WAVE);
		if (t < total)
		{
			const int j = t - exL;
			(void)rawL;
			if (j != selfL)
			{
				uint32_t id;
				int32_t p;
				transformEntry(a.entries[stL + j], gL != 0, a.targetLen, a.k, id, p);
				const int64_t o = obL + j - ((selfL >= 0 && j > selfL) ? 1 : 0);
				a.mkey[o] = (static_cast<uint64_t>(id) << 32) | static_cast<uint32_t>(posL);
				a.mext[o] = p;
				if (a.mval) a.mval[o] = static_cast<uint32_t>(o - qbL);
			}
		}
	}
	(void)qiid;
}

__global__ void writeCandidatesKernel(JoinArgs a, const int64_t* qMatchOff, const int64_t* queryIds,
									  const uint32_t* perm, flyehip_candidate_t* out, int64_t nMatches)
{
	const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
	if (i >= nMatches) return;
	const int q = findSegWave(qMatchOff, a.nQueries, i);
	const int64_t qb = qMatchOff[q];
	const uint64_t key = a.mkey[i];
	const int32_t ext = perm ? a.mext[qb + perm[i]] : a.mext[i];
	const uint32_t eid = static_cast<uint32_t>(key >> 32);
	const int32_t pos = static_cast<int32_t>(static_cast<uint32_t>(key));
	const JoinQuery Q = a.queries[q];
	const int64_t wo = a.seqs.wordOff[Q.seq];
	uint64_t y = Q.strand ? kmerAt(a.seqs.words, wo, pos, a.k) : rcKmer(kmerAt(a.seqs.words, wo, Q.len - pos - a.k, a.k), a.k);
	flyehip_candidate_t c;
	c.query_id = queryIds[q];
	c.target_id = toSigned(eid);
	c.kmer = y;
	c.query_pos = pos;
	c.target_pos = ext;
	out[i] = c;
}

// ---- per-query match sort -------------------------------------------------------

__global__ void gatherHitsKernel(const int64_t* qMatchOff, int32_t nQueries, const uint64_t* key, const uint32_t* val,
								 const int32_t* mext, int32_t* hcur, int32_t* hext, int32_t* groupFlag, int64_t nMatches)
{
	// extPos is gathered inside the query's range: one XCD per range
	const int64_t i = static_cast<int64_t>(xcdBlockIndex()) * blockDim.x + threadIdx.x;
	if (i >= nMatches) return;
	const int q = findSegWave(qMatchOff, nQueries, i);
	const int64_t qb = qMatchOff[q];
	const uint64_t k = key[i];
	hcur[i] = static_cast<int32_t>(static_cast<uint32_t>(k));
	hext[i] = mext[qb + val[i]];
	groupFlag[i] = (i == qb || (key[i - 1] >> 32) != (k >> 32)) ? 1 : 0;
}

// After the tie replay: the hits at flagged (tied, relevant) sorted positions
// take the extPos of the match the replay put there (curPos is the key's, the
// same for every member of a tie).
// The flags of positions [4t, 4t + 4) for thread t: one 16-byte load (flag
// arrays come from the 256-byte-aligned batch buffers), zero past n.
__device__ inline int4 flags4(const int32_t* f, int64_t j, int64_t n)
{
	if (j + 4 <= n) return *reinterpret_cast<const int4*>(f + j);
	int4 r = make_int4(0, 0, 0, 0);
	if (j < n) r.x = f[j];
	if (j + 1 < n) r.y = f[j + 1];
	if (j + 2 < n) r.z = f[j + 2];
	return r;
}

__device__ inline int flagAt(const int4& v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : (k == 2 ? v.z : v.w)); }

// Four positions per thread (the flags are sparse: most threads read their
// 16 bytes and stop).
__global__ void fixTiedHitsKernel(const int64_t* qMatchOff, int32_t nQueries, const int32_t* dupFlag,
								  const uint32_t* val, const int32_t* mext, int32_t* hext, int64_t nMatches)
{
	const int64_t j = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) * 4;
	if (j >= nMatches) return;
	const int4 f = flags4(dupFlag, j, nMatches);
	for (int k = 0; k < 4; ++k)
	{
		if (!flagAt(f, k)) continue;
		const int64_t i = j + k;
		const int q = findSegWave(qMatchOff, nQueries, i);
		hext[i] = mext[qMatchOff[q] + val[i]];
	}
}

// One position per thread: group starts are dense (one in about 40
// positions), and four per thread measured slower (a query search per
// position round instead of one per wave).
__global__ void buildGroupsKernel(const int64_t* qMatchOff, int32_t nQueries, const uint64_t* key, const int32_t* groupFlag,
								  const int64_t* groupOff, const int32_t* targetLen, GroupDesc* groups, int64_t nMatches)
{
	const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
	if (i >= nMatches) return;
	if (!groupFlag[i]) return;
	const int64_t g = groupOff[i];
	const int q = findSegWave(qMatchOff, nQueries, i);
	const uint32_t eid = static_cast<uint32_t>(key[i] >> 32);
	GroupDesc d;
	d.hitStart = i;
	d.extId = toSigned(eid);
	d.n = 0;  // set by groupSizeKernel from the next group's start
	d.query = q;
	d.extLen = targetLen[eid >> 1];
	d.inPlace = 0;
	groups[g] = d;
}

__global__ void groupSizeKernel(GroupDesc* groups, int64_t nGroups, int64_t nMatches)
{
	const int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
	if (g >= nGroups) return;
	const int64_t end = g + 1 < nGroups ? groups[g + 1].hitStart : nMatches;
	groups[g].n = static_cast<int32_t>(end - groups[g].hitStart);
}

__global__ void divStatsKernel(const flyehip_overlap_t* rows, const int64_t* qRowOff, const int32_t* curLen,
							   int32_t nQueries, int32_t* divCount, float* divOut, const int64_t* divOff)
{
	const int q = blockIdx.x * blockDim.x + threadIdx.x;
	if (q >= nQueries) return;
	const int kWnd = 10000;
	const int nw = curLen[q] / kWnd + 1;
	const int64_t rb = qRowOff[q], re = qRowOff[q + 1];
	int cnt = 0;
	for (int w = 0; w < nw; ++w)
	{
		int bestRange = 0;
		float bestDiv = 0.0f;
		for (int64_t r = rb; r < re; ++r)
		{
			const flyehip_overlap_t& o = rows[r];
			if (o.cur_begin < 0) continue;
			if (o.cur_begin / kWnd != w) continue;
			const int range = o.cur_end - o.cur_begin;
			if (range > bestRange)
			{
				bestRange = range;
				bestDiv = o.seq_divergence;
			}
		}
		if (bestRange > 0) divOut[divOff[q] + cnt++] = bestDiv;
	}
	divCount[q] = cnt;
}

// 2-bit packing of ASCII sequences: one thread per output word (16 bases,
// most significant first; a sequence starts on a word boundary and ends with
// two spare zero words).  A byte other than A/C/G/T sets *err.
__global__ void packBasesKernel(const uint8_t* bases, const int64_t* baseOff, const int64_t* wordOff,
								const int32_t* lens, int32_t n, int64_t nWords, uint32_t* words, uint32_t* err)
{
	const int64_t w = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
	if (w >= nWords) return;
	const int q = findSegWave(wordOff, n, w);
	const int64_t j0 = (w - wordOff[q]) * 16;
	const int64_t len = lens[q];
	const uint8_t* src = bases + (baseOff[q] - baseOff[0]);
	uint32_t v = 0;
	bool bad = false;
	for (int t = 0; t < 16; ++t)
	{
		const int64_t j = j0 + t;
		if (j >= len) break;
		const uint8_t b = src[j];
		const int code = b == 'A' ? 0 : b == 'C' ? 1 : b == 'G' ? 2 : b == 'T' ? 3 : -1;
		bad |= code < 0;
		v |= static_cast<uint32_t>(code & 3) << (30 - 2 * t);
	}
	words[w] = v;
	if (bad) atomicOr(err, 1u);
}

inline dim3 gridFor(int64_t n, int tpb) { return dim3(static_cast<unsigned>((n + tpb - 1) / tpb)); }

} // namespace

// rocPRIM's single-pass (decoupled look-back) scan reads the input once and
// writes the output once; the three-kernel scan above reads it twice.
struct WidenI32
{
	__host__ __device__ int64_t operator()(int32_t x) const { return x; }
};

size_t lookbackScanBytes(int64_t n)
{
	size_t bytes = 0;
	if (n <= 0) return 0;
	auto it = rocprim::make_transform_iterator(static_cast<const int32_t*>(nullptr), WidenI32{});
	if (rocprim::exclusive_scan(nullptr, bytes, it, static_cast<int64_t*>(nullptr), int64_t{0},
								static_cast<size_t>(n), rocprim::plus<int64_t>()) != hipSuccess)
		return 0;
	return bytes;
}

__global__ void scanTotalKernel(const int32_t* in, int64_t* out, int64_t n) { out[n] = out[n - 1] + in[n - 1]; }

size_t scanTempWords(int64_t n)
{
	const size_t three = static_cast<size_t>((n + kScanTile - 1) / kScanTile + 1);
	const size_t look = (lookbackScanBytes(n) + 7) / 8 + 1;
	return three > look ? three : look;
}

hipError_t exclusiveScan(const int32_t* in, int64_t* out, int64_t n, int64_t* tmp, hipStream_t s)
{
	const int64_t nb = (n + kScanTile - 1) / kScanTile;
	if (nb == 0)
	{
		return hipMemsetAsync(out, 0, sizeof(int64_t), s);
	}
	size_t bytes = lookbackScanBytes(n);
	if (bytes > 0)
	{
		auto it = rocprim::make_transform_iterator(in, WidenI32{});
		hipError_t e = rocprim::exclusive_scan(static_cast<void*>(tmp), bytes, it, out, int64_t{0},
											   static_cast<size_t>(n), rocprim::plus<int64_t>(), s);
		if (e != hipSuccess) return e;
		hipLaunchKernelGGL(scanTotalKernel, dim3(1), dim3(1), 0, s, in, out, n);
		return hipGetLastError();
	}
	hipLaunchKernelGGL(scanReduceKernel, dim3(static_cast<unsigned>(nb)), dim3(kScanNT), 0, s, in, n, tmp);
	hipLaunchKernelGGL(scanBlockSumsKernel, dim3(1), dim3(1024), 0, s, tmp, nb);
	hipLaunchKernelGGL(scanDownKernel, dim3(static_cast<unsigned>(nb)), dim3(kScanNT), 0, s, in, out, n, tmp, nb);
	return hipGetLastError();
}

hipError_t launchCount(const PackedSeqs& reads, const int64_t* winOff, int64_t nWin, int k, HashTable t, hipStream_t s)
{
	if (nWin == 0) return hipSuccess;
	hipLaunchKernelGGL(countKernel, gridFor(nWin, 256), dim3(256), 0, s, reads, winOff, nWin, k, t);
	return hipGetLastError();
}

hipError_t launchHistogram(HashTable t, uint64_t* hist, int hmax, uint64_t* acc, hipStream_t s)
{
	hipLaunchKernelGGL(histKernel, dim3(2048), dim3(256), 0, s, t, hist, hmax, acc);
	return hipGetLastError();
}

hipError_t launchClassify(HashTable t, int solidMin, int repMax, int32_t* sizes, uint64_t* counters, hipStream_t s)
{
	hipLaunchKernelGGL(classifyKernel, gridFor(static_cast<int64_t>(t.mask + 1), 256), dim3(256), 0, s, t, solidMin,
					   repMax, sizes, counters);
	return hipGetLastError();
}

hipError_t launchAssign(HashTable t, const int32_t* sizes, const int64_t* starts, hipStream_t s)
{
	hipLaunchKernelGGL(assignKernel, gridFor(static_cast<int64_t>(t.mask + 1), 256), dim3(256), 0, s, t, sizes, starts);
	return hipGetLastError();
}

hipError_t launchFill(const PackedSeqs& reads, const int64_t* winOff, int64_t nWin, int k, HashTable t,
					  const uint8_t* mask, uint64_t* ordKeys, hipStream_t s)
{
	if (nWin == 0) return hipSuccess;
	hipLaunchKernelGGL(fillKernel, gridFor(nWin, 256), dim3(256), 0, s, reads, winOff, nWin, k, t, mask, ordKeys);
	return hipGetLastError();
}

hipError_t launchMinimizerMask(const PackedSeqs& reads, const int64_t* winOff, int64_t nWin, int k, int w,
							   uint8_t* mask, hipStream_t s)
{
	if (nWin == 0) return hipSuccess;
	hipLaunchKernelGGL(minimizerMaskKernel, gridFor(nWin, 256), dim3(256), 0, s, reads, winOff, nWin, k, w, mask);
	return hipGetLastError();
}

hipError_t launchMaskedSizes(const PackedSeqs& reads, const int64_t* winOff, int64_t nWin, int k, HashTable t,
							 const uint8_t* mask, int32_t* sizes, hipStream_t s)
{
	if (nWin == 0) return hipSuccess;
	hipLaunchKernelGGL(maskedSizesKernel, gridFor(nWin, 256), dim3(256), 0, s, reads, winOff, nWin, k, t, mask, sizes);
	return hipGetLastError();
}

hipError_t launchLoadKeys(HashTable t, const uint64_t* keys, const uint64_t* vals, int64_t n, uint32_t* err,
						  hipStream_t s)
{
	if (n == 0) return hipSuccess;
	hipLaunchKernelGGL(loadKeysKernel, gridFor(n, 256), dim3(256), 0, s, t, keys, vals, n, err);
	return hipGetLastError();
}

hipError_t launchBucketFinish(HashTable t, uint64_t* ordKeys, uint64_t* entries, const int32_t* readLen, int k,
							  uint32_t* bigList, uint64_t* bigCount, uint64_t bigCap, hipStream_t s)
{
	auto* cnt = reinterpret_cast<unsigned long long*>(bigCount);
	hipLaunchKernelGGL(bucketFinishKernel, gridFor(static_cast<int64_t>(t.mask + 1), 256), dim3(256), 0, s, t, ordKeys,
					   entries, readLen, k, bigList, cnt, bigCap);
	hipError_t e = hipGetLastError();
	if (e != hipSuccess) return e;
	hipLaunchKernelGGL(bucketBlockSortKernel, dim3(512), dim3(1024), 0, s, t, ordKeys, entries, readLen, k,
					   static_cast<const uint32_t*>(bigList), static_cast<const unsigned long long*>(cnt));
	return hipGetLastError();
}

hipError_t launchPackBases(const uint8_t* bases, const int64_t* baseOff, const int64_t* wordOff,
						   const int32_t* lens, int32_t n, int64_t nWords, uint32_t* words, uint32_t* err,
						   hipStream_t s)
{
	if (nWords == 0) return hipSuccess;
	hipLaunchKernelGGL(packBasesKernel, gridFor(nWords, 256), dim3(256), 0, s, bases, baseOff, wordOff, lens, n,
					   nWords, words, err);
	return hipGetLastError();
}

hipError_t launchJoinCount(const JoinArgs& a, hipStream_t s)
{
	if (a.nWin == 0) return hipSuccess;
	hipLaunchKernelGGL(joinCountKernel, gridFor(a.nWin, 256), dim3(256), 0, s, a);
	return hipGetLastError();
}

hipError_t launchJoinFill(const JoinArgs& a, hipStream_t s)
{
	if (a.nWin == 0) return hipSuccess;
	hipLaunchKernelGGL(joinFillKernel, gridFor(a.nWin, 256), dim3(256), 0, s, a);
	return hipGetLastError();
}

hipError_t launchWriteCandidates(const JoinArgs& a, const int64_t* qMatchOff, const int64_t* queryIds,
								 const uint32_t* perm, flyehip_candidate_t* out, int64_t nMatches, hipStream_t s)
{
	if (nMatches == 0) return hipSuccess;
	hipLaunchKernelGGL(writeCandidatesKernel, gridFor(nMatches, 256), dim3(256), 0, s, a, qMatchOff, queryIds, perm,
					   out, nMatches);
	return hipGetLastError();
}

hipError_t launchGatherHits(const int64_t* qMatchOff, int32_t nQueries, const uint64_t* key, const uint32_t* val,
							const int32_t* mext, int32_t* hcur, int32_t* hext, int32_t* groupFlag, int64_t nMatches,
							hipStream_t s)
{
	if (nMatches == 0) return hipSuccess;
	hipLaunchKernelGGL(gatherHitsKernel, gridFor(nMatches, 256), dim3(256), 0, s, qMatchOff, nQueries, key, val, mext,
					   hcur, hext, groupFlag, nMatches);
	return hipGetLastError();
}

hipError_t launchFixTiedHits(const int64_t* qMatchOff, int32_t nQueries, const int32_t* dupFlag, const uint32_t* val,
							 const int32_t* mext, int32_t* hext, int64_t nMatches, hipStream_t s)
{
	if (nMatches == 0) return hipSuccess;
	hipLaunchKernelGGL(fixTiedHitsKernel, gridFor((nMatches + 3) / 4, 256), dim3(256), 0, s, qMatchOff, nQueries,
					   dupFlag, val, mext, hext, nMatches);
	return hipGetLastError();
}

hipError_t launchBuildGroups(const int64_t* qMatchOff, int32_t nQueries, const uint64_t* key, const int32_t* groupFlag,
							 const int64_t* groupOff, const int32_t* targetLen, GroupDesc* groups, int64_t nMatches,
							 hipStream_t s)
{
	if (nMatches == 0) return hipSuccess;
	hipLaunchKernelGGL(buildGroupsKernel, gridFor(nMatches, 256), dim3(256), 0, s, qMatchOff, nQueries, key, groupFlag,
					   groupOff, targetLen, groups, nMatches);
	const int64_t nGroups = 0;  // filled below from the scan total
	(void)nGroups;
	return hipGetLastError();
}

hipError_t launchGroupSizes(GroupDesc* groups, int64_t nGroups, int64_t nMatches, hipStream_t s)
{
	if (nGroups == 0) return hipSuccess;
	hipLaunchKernelGGL(groupSizeKernel, gridFor(nGroups, 256), dim3(256), 0, s, groups, nGroups, nMatches);
	return hipGetLastError();
}

hipError_t launchDivStats(const flyehip_overlap_t* rows, const int64_t* qRowOff, const int32_t* curLen, int32_t nQueries,
						  int32_t* divCount, float* divOut, const int64_t* divOff, hipStream_t s)
{
	if (nQueries == 0) return hipSuccess;
	hipLaunchKernelGGL(divStatsKernel, gridFor(nQueries, 128), dim3(128), 0, s, rows, qRowOff, curLen, nQueries,
					   divCount, divOut, divOff);
	return hipGetLastError();
}

} // namespace fh
