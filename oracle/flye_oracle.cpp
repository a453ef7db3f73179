// flye_oracle.cpp - CPU restatement of Flye 2.9.6's overlap-detection hot path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker: the tests, the smoke
// entry and the CPU-baseline leg of the benchmark call it through liboracle.so
// (oracle/Makefile).  No product code links it.  It is written for clarity and
// exactness, not speed, and calls libstdc++'s real std::sort wherever Flye
// calls std::sort, so the tie order is Flye's by construction.
//
// What each function follows (reference paths are relative to the reference
// repository root):
//   encode / reverse complement / standard form
//       cuda/cuflye_cuda_kmer_encode_smoke.cu:69-117 (A=0 C=1 G=2 T=3, MSB first;
//       canonical = min(kmer, rc)); the flip flag is Flye's Kmer::standardForm
//       return value: true unless kmer < rc.
//   signed/internal id mapping
//       tools/replay_read_to_graph_source_pack.py:85-88 (+n -> 2(n-1), -n -> 2n-1).
//   candidate loop (collectCandidateMatchesCpu)
//       patches/flye/2.9.6/0002-cuflye-candidate-backend-seam.patch file lines 52-95
//       (repetitive -> filtered position; kmerFreq == 0 -> skip; drop the exact
//       self hit; emit KmerMatch(curPos, extPos, extId, numRepr)).
//   iterKmerPos orientation
//       patches/flye/2.9.6/0004-cuflye-cuda-adapter-shell.patch:74-85, 214-220 and
//       tests/fixtures/kmer-encode-smoke-v0/index.tsv: the returned position is
//       where the query k-mer itself is spelled on the returned strand.
//   match sort          patch 0001:88-91 (std::sort by (extId, curPos)).
//   group chaining      tools/replay_overlap_chains.py:228-380 (prefilters, DP,
//                       backtrack, overlapTest, f32 divergence with glibc logf,
//                       primary selection, divergence filter) with the libstdc++
//                       std::sort points of tools/replay_read_to_graph_source_pack.py
//                       :396-397 (ext re-sort), 439-442 (score order) and 481
//                       (extOverlaps by score); onlyMaxExt keeps the first
//                       proposal of maximal score
//                       (cuda/cuflye_cuda_overlap_chain_replay.cu:1627-1631).
//   divStats windows    patches/flye/2.9.6/0023-...perf-gate.patch:245-269.
//   read-to-graph       patch 0039:72-77, 241-251 (filter, std::sort by curBegin);
//                       tools/replay_read_alignment_chains.py:64-139
//                       (chainReadAlignments, with std::sort for the score order).
//   VertexIndex         not in the reference (Flye's vertex_index.* is absent);
//                       the definition here is the one include/flyehip.h states.

#include "../include/flyehip.h"

#include <openssl/evp.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

namespace
{

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

uint64_t rcKmer(uint64_t x, int k)
{
	uint64_t r = 0;
	for (int i = 0; i < k; ++i)
	{
		r = (r << 2) | (~x & 3ULL);
		x >>= 2;
	}
	return r;
}

uint32_t toInternal(int64_t sid)
{
	return sid > 0 ? static_cast<uint32_t>(2 * (sid - 1)) : static_cast<uint32_t>(2 * (-sid) - 1);
}

int64_t toSigned(uint32_t iid)
{
	return (iid & 1U) ? -static_cast<int64_t>((iid + 1) / 2) : static_cast<int64_t>(iid / 2 + 1);
}

struct Store
{
	std::vector<std::string> fwd;   // read n at fwd[n - 1]
};

std::string revComp(const std::string& s)
{
	std::string r(s.size(), 'A');
	for (size_t i = 0; i < s.size(); ++i)
	{
		char c = s[s.size() - 1 - i];
		r[i] = c == 'A' ? 'T' : c == 'C' ? 'G' : c == 'G' ? 'C' : 'A';
	}
	return r;
}

struct Index
{
	int k = 0;
	int solidMin = 0;
	int repMax = 0;
	std::vector<uint64_t> kmers;      // solid, non-repetitive canonical k-mers, ascending
	std::vector<int64_t> start;       // bucket b = entries [start[b], start[b+1])
	std::vector<uint32_t> entryId;    // internal id of the stored strand
	std::vector<int32_t> entryPos;
	std::vector<uint64_t> repetitive; // ascending
	std::vector<int32_t> readLen;     // by read number - 1
	int64_t distinct = 0;
	std::vector<int64_t> hist;        // hist[c] = distinct k-mers with count c
	// hash lookup (Flye's VertexIndex is a hash table): open addressing,
	// linear probing, canonical k-mer -> bucket b, or kRepetitive
	std::vector<uint64_t> hashKey;
	std::vector<int64_t> hashVal;     // -1 = empty slot
	uint64_t hashMask = 0;
	static constexpr int64_t kRepetitive = -2;

	static uint64_t mix(uint64_t x)
	{
		x ^= x >> 33;
		x *= 0xff51afd7ed558ccdULL;
		x ^= x >> 33;
		return x;
	}
	void buildLookup()
	{
		const size_t n = kmers.size() + repetitive.size();
		size_t cap = 16;
		while (cap < 2 * n + 1) cap <<= 1;
		hashKey.assign(cap, 0);
		hashVal.assign(cap, -1);
		hashMask = cap - 1;
		auto put = [&](uint64_t key, int64_t v) {
			uint64_t h = mix(key) & hashMask;
			while (hashVal[h] != -1) h = (h + 1) & hashMask;
			hashKey[h] = key;
			hashVal[h] = v;
		};
		// repetitive first: a k-mer given as both is filtered (probes stop at
		// the first slot holding the key)
		for (uint64_t r : repetitive) put(r, kRepetitive);
		for (size_t b = 0; b < kmers.size(); ++b) put(kmers[b], static_cast<int64_t>(b));
	}
	// bucket of canonical k-mer c, kRepetitive, or -1 (absent)
	int64_t lookup(uint64_t c) const
	{
		uint64_t h = mix(c) & hashMask;
		while (hashVal[h] != -1)
		{
			if (hashKey[h] == c) return hashVal[h];
			h = (h + 1) & hashMask;
		}
		return -1;
	}
};

// Same window walk for every caller: calls f(pos, kmer, canonical, flip).
template <class F>
void forEachWindow(const std::string& s, int k, F f)
{
	if (static_cast<int64_t>(s.size()) < k) return;
	const uint64_t mask = (k == 32) ? ~0ULL : ((1ULL << (2 * k)) - 1);
	uint64_t x = 0;
	uint64_t r = 0;
	for (size_t i = 0; i < s.size(); ++i)
	{
		uint64_t b = static_cast<uint64_t>(baseCode(static_cast<uint8_t>(s[i])));
		x = ((x << 2) | b) & mask;
		r = (r >> 2) | ((3ULL - b) << (2 * (k - 1)));
		if (static_cast<int64_t>(i) + 1 >= k)
		{
			int32_t pos = static_cast<int32_t>(i + 1 - k);
			bool flip = !(x < r);
			f(pos, x, flip ? r : x, flip);
		}
	}
}

int autoSolidMin(const std::vector<int64_t>& hist)
{
	// First c >= 2 where the histogram stops falling: hist[c] <= hist[c + 1].
	for (size_t c = 2; c + 1 < hist.size(); ++c)
	{
		if (hist[c] <= hist[c + 1]) return static_cast<int>(c);
	}
	return 2;
}

int autoRepMax(const std::vector<int64_t>& hist, int solidMin, float rate)
{
	double total = 0.0;
	double distinct = 0.0;
	for (size_t c = static_cast<size_t>(solidMin); c < hist.size(); ++c)
	{
		total += static_cast<double>(c) * static_cast<double>(hist[c]);
		distinct += static_cast<double>(hist[c]);
	}
	double mean = total / (distinct + 1.0);
	int rep = static_cast<int>(std::floor(static_cast<double>(rate) * mean));
	return std::max(rep, solidMin);
}

// Minimizer positions of one stored read (flyehip_build_index_minimizers'
// definition): in every run of w consecutive k-mer positions (all of them
// when there are fewer), the position with the smallest splitmix64 mix of its
// canonical k-mer, the leftmost on equal values.
uint64_t mix64(uint64_t x)
{
	x ^= x >> 30;
	x *= 0xbf58476d1ce4e5b9ULL;
	x ^= x >> 27;
	x *= 0x94d049bb133111ebULL;
	x ^= x >> 31;
	return x;
}

std::vector<char> minimizerFlags(const std::string& s, int k, int w)
{
	std::vector<uint64_t> ord;
	forEachWindow(s, k, [&](int32_t, uint64_t, uint64_t c, bool) { ord.push_back(mix64(c)); });
	const int64_t nk = static_cast<int64_t>(ord.size());
	std::vector<char> flag(static_cast<size_t>(nk), 0);
	if (nk == 0) return flag;
	const int64_t last = nk >= w ? nk - w : 0;
	for (int64_t st = 0; st <= last; ++st)
	{
		const int64_t e = std::min<int64_t>(nk, st + w);
		int64_t best = st;
		for (int64_t p = st + 1; p < e; ++p)
			if (ord[static_cast<size_t>(p)] < ord[static_cast<size_t>(best)]) best = p;
		flag[static_cast<size_t>(best)] = 1;
	}
	return flag;
}

// VertexIndex::countKmers + buildIndex, or buildIndexMinimizers when
// minWindow > 0 (same counts and thresholds; buckets of minimizer positions,
// and a k-mer without one is left out of the index).
Index* buildIndex(const Store& st, int k, int solidMin, int repMax, float rate, int minWindow = 0)
{
	Index* ix = new Index();
	ix->k = k;
	std::vector<uint64_t> all;
	size_t total = 0;
	for (const auto& s : st.fwd)
		if (static_cast<int64_t>(s.size()) >= k) total += s.size() - k + 1;
	all.reserve(total);
	for (const auto& s : st.fwd)
		forEachWindow(s, k, [&](int32_t, uint64_t, uint64_t c, bool) { all.push_back(c); });
	std::sort(all.begin(), all.end());

	std::vector<std::pair<uint64_t, int64_t>> counts;
	for (size_t i = 0; i < all.size();)
	{
		size_t j = i;
		while (j < all.size() && all[j] == all[i]) ++j;
		counts.emplace_back(all[i], static_cast<int64_t>(j - i));
		i = j;
	}
	std::vector<uint64_t>().swap(all);
	ix->distinct = static_cast<int64_t>(counts.size());
	int64_t maxCount = 0;
	for (auto& c : counts) maxCount = std::max(maxCount, c.second);
	ix->hist.assign(static_cast<size_t>(maxCount) + 2, 0);
	for (auto& c : counts) ix->hist[static_cast<size_t>(c.second)]++;

	ix->solidMin = solidMin > 0 ? solidMin : autoSolidMin(ix->hist);
	ix->repMax = repMax > 0 ? repMax : autoRepMax(ix->hist, ix->solidMin, rate);

	std::vector<std::vector<char>> flags;
	std::vector<std::pair<uint64_t, int64_t>> minCounts;  // k-mer, minimizer positions
	if (minWindow > 0)
	{
		std::vector<uint64_t> sel;
		for (const auto& s : st.fwd)
		{
			flags.push_back(minimizerFlags(s, k, minWindow));
			const std::vector<char>& f = flags.back();
			forEachWindow(s, k, [&](int32_t pos, uint64_t, uint64_t c, bool) {
				if (f[static_cast<size_t>(pos)]) sel.push_back(c);
			});
		}
		std::sort(sel.begin(), sel.end());
		for (size_t i = 0; i < sel.size();)
		{
			size_t j = i;
			while (j < sel.size() && sel[j] == sel[i]) ++j;
			minCounts.emplace_back(sel[i], static_cast<int64_t>(j - i));
			i = j;
		}
	}
	std::vector<int64_t> bucketSize;
	for (auto& c : counts)
	{
		if (c.second < ix->solidMin) continue;
		if (c.second > ix->repMax)
		{
			ix->repetitive.push_back(c.first);
			continue;
		}
		int64_t size = c.second;
		if (minWindow > 0)
		{
			auto it = std::lower_bound(minCounts.begin(), minCounts.end(), std::make_pair(c.first, int64_t{0}));
			size = (it != minCounts.end() && it->first == c.first) ? it->second : 0;
			if (size == 0) continue;
		}
		ix->kmers.push_back(c.first);
		bucketSize.push_back(size);
	}
	ix->start.assign(ix->kmers.size() + 1, 0);
	for (size_t b = 0; b < bucketSize.size(); ++b) ix->start[b + 1] = ix->start[b] + bucketSize[b];
	ix->entryId.assign(static_cast<size_t>(ix->start.back()), 0);
	ix->entryPos.assign(static_cast<size_t>(ix->start.back()), 0);
	std::vector<int64_t> cursor(ix->start.begin(), ix->start.end() - 1);
	for (size_t n = 0; n < st.fwd.size(); ++n)
	{
		const std::string& s = st.fwd[n];
		ix->readLen.push_back(static_cast<int32_t>(s.size()));
		const std::vector<char>* f = minWindow > 0 ? &flags[n] : nullptr;
		forEachWindow(s, k, [&](int32_t pos, uint64_t, uint64_t c, bool flip) {
			if (f && !(*f)[static_cast<size_t>(pos)]) return;
			auto it = std::lower_bound(ix->kmers.begin(), ix->kmers.end(), c);
			if (it == ix->kmers.end() || *it != c) return;
			size_t b = static_cast<size_t>(it - ix->kmers.begin());
			int64_t slot = cursor[b]++;
			uint32_t plusId = static_cast<uint32_t>(2 * n);
			if (flip)
			{
				ix->entryId[slot] = plusId ^ 1U;
				ix->entryPos[slot] = static_cast<int32_t>(s.size()) - pos - k;
			}
			else
			{
				ix->entryId[slot] = plusId;
				ix->entryPos[slot] = pos;
			}
		});
	}
	ix->buildLookup();
	return ix;
}

struct Match
{
	int32_t curPos;
	int32_t extPos;
	uint32_t extId;
	uint64_t kmer;
};

void collect(const Index& ix, uint32_t qInternal, const std::string& q,
			 std::vector<Match>& out, std::vector<int32_t>& filtered)
{
	const int k = ix.k;
	forEachWindow(q, k, [&](int32_t pos, uint64_t x, uint64_t c, bool flip) {
		const int64_t hit = ix.lookup(c);
		if (hit == Index::kRepetitive)
		{
			filtered.push_back(pos);
			return;
		}
		if (hit < 0) return;
		const size_t b = static_cast<size_t>(hit);
		for (int64_t e = ix.start[b]; e < ix.start[b + 1]; ++e)
		{
			uint32_t id = ix.entryId[e];
			int32_t p = ix.entryPos[e];
			if (flip)
			{
				id ^= 1U;
				p = ix.readLen[id >> 1] - p - k;
			}
			if (id == qInternal && p == pos) continue;
			out.push_back({pos, p, id, x});
		}
	});
}

struct Params
{
	flyehip_overlap_params_t p;
};

struct Ovlp
{
	int64_t curId, extId;
	int32_t curBegin, curEnd, curLen, extBegin, extEnd, extLen, score;
	float div;
	int32_t chainLength;
	int32_t curRange() const { return curEnd - curBegin; }
	int32_t extRange() const { return extEnd - extBegin; }
	bool containedBy(const Ovlp& o) const
	{
		return curId == o.curId && extId == o.extId && o.curBegin <= curBegin &&
			   curEnd <= o.curEnd && o.extBegin <= extBegin && extEnd <= o.extEnd;
	}
};

bool overlapTest(const Ovlp& o, const flyehip_overlap_params_t& p)
{
	if (o.curRange() < p.minimum_overlap || o.extRange() < p.minimum_overlap) return false;
	int32_t lengthDiff = std::abs(o.curRange() - o.extRange());
	if (static_cast<double>(lengthDiff) > 0.5 * std::min(o.curRange(), o.extRange())) return false;
	if (o.curId == o.extId)
	{
		int32_t inter = std::min(o.curEnd, o.extEnd) - std::max(o.curBegin, o.extBegin);
		if (static_cast<double>(inter) > o.curRange() / 2.0) return false;
	}
	if (o.curId == -o.extId)
	{
		int32_t inter = std::min(o.curEnd, o.extLen - o.extBegin) -
						std::max(o.curBegin, o.extLen - o.extEnd);
		if (static_cast<double>(inter) > o.curRange() / 2.0) return false;
	}
	if (!p.force_local && p.check_overhang)
	{
		int32_t oh = std::max(std::min(o.curBegin, o.extBegin),
							  std::min(o.curLen - o.curEnd, o.extLen - o.extEnd));
		if (oh > p.maximum_overhang) return false;
	}
	return true;
}

struct CE
{
	int32_t cur;
	int32_t ext;
};

// The chaining DP of one group (tools/replay_overlap_chains.py:280-313) over
// `hits` in their final order.  earlyBreak = false drops the break after an
// improvement on an exact diagonal step shorter than k: only the tests use
// that form, to show which inputs the rule changes.
void chainDp(const std::vector<CE>& hits, const flyehip_overlap_params_t& p, bool extSorted,
			 bool earlyBreak, std::vector<int32_t>& score, std::vector<int32_t>& bt)
{
	const int32_t n = static_cast<int32_t>(hits.size());
	const int32_t k = p.kmer_size;
	score.assign(static_cast<size_t>(n), 0);
	bt.assign(static_cast<size_t>(n), -1);
	for (int32_t i = 1; i < n; ++i)
	{
		int32_t maxScore = 0;
		int32_t maxId = 0;
		const int32_t curNext = hits[i].cur;
		const int32_t extNext = hits[i].ext;
		for (int32_t j = i - 1; j >= 0; --j)
		{
			const int32_t curPrev = hits[j].cur;
			const int32_t extPrev = hits[j].ext;
			const int32_t dc = curNext - curPrev;
			const int32_t de = extNext - extPrev;
			const int32_t jumpDiv = std::abs(dc - de);
			if (0 < dc && dc < p.maximum_jump && 0 < de && de < p.maximum_jump &&
				jumpDiv <= p.max_jump_gap)
			{
				const int32_t matchScore = std::min(std::min(dc, de), k);
				const float pen = jumpDiv > p.gap_jump_threshold ? p.large_gap_penalty
																 : p.small_gap_penalty;
				const int32_t gapCost = static_cast<int32_t>(pen * static_cast<float>(jumpDiv));
				const int32_t next = score[j] + matchScore - gapCost;
				if (next > maxScore)
				{
					maxScore = next;
					maxId = j;
					if (earlyBreak && jumpDiv == 0 && dc < k) break;
				}
			}
			if (extSorted && de > p.maximum_jump) break;
			if (!extSorted && dc > p.maximum_jump) break;
		}
		score[i] = std::max(maxScore, k);
		if (maxScore > k) bt[i] = maxId;
	}
}

// One (query, target) group: tools/replay_overlap_chains.py:228-380.
// `hits` is the group in (extId, curPos) order; `filtered` ascending.
void chainGroup(std::vector<CE>& hits, int64_t curId, int32_t curLen, int64_t extId,
				int32_t extLen, const std::vector<int32_t>& filtered,
				const flyehip_overlap_params_t& p, std::vector<Ovlp>& detected)
{
	const int32_t n = static_cast<int32_t>(hits.size());
	const int32_t k = p.kmer_size;
	int32_t unique = 0;
	int32_t prev = 0;
	for (const auto& h : hits)
	{
		if (h.cur != prev)
		{
			++unique;
			prev = h.cur;
		}
	}
	if (static_cast<float>(unique) <
		p.min_kmer_survival_rate * static_cast<float>(p.minimum_overlap)) return;

	int32_t minCur = hits.front().cur, maxCur = hits.back().cur;
	int32_t minExt = hits.front().ext, maxExt = hits.front().ext;
	for (const auto& h : hits)
	{
		minExt = std::min(minExt, h.ext);
		maxExt = std::max(maxExt, h.ext);
	}
	if (maxCur - minCur < p.minimum_overlap || maxExt - minExt < p.minimum_overlap) return;
	if (p.check_overhang && !p.force_local)
	{
		if (std::min(minCur, minExt) > p.maximum_overhang) return;
		if (std::min(curLen - maxCur, extLen - maxExt) > p.maximum_overhang) return;
	}

	const bool extSorted = extLen > curLen;
	if (extSorted)
	{
		std::sort(hits.begin(), hits.end(),
				  [](const CE& a, const CE& b) { return a.ext < b.ext; });
	}

	std::vector<int32_t> score, bt;
	chainDp(hits, p, extSorted, true, score, bt);

	std::vector<int32_t> order(static_cast<size_t>(n));
	for (int32_t i = 0; i < n; ++i) order[i] = i;
	std::sort(order.begin(), order.end(),
			  [&score](int32_t a, int32_t b) { return score[a] > score[b]; });

	std::vector<Ovlp> ext;
	for (int32_t start : order)
	{
		if (bt[start] == -1) continue;
		int32_t last = start, first = 0, len = 0, pos = start;
		while (pos != -1)
		{
			first = pos;
			++len;
			int32_t np = bt[pos];
			bt[pos] = -1;
			pos = np;
		}
		Ovlp o{};
		o.curId = curId;
		o.extId = extId;
		o.curBegin = hits[first].cur;
		o.curEnd = hits[last].cur + k - 1;
		o.curLen = curLen;
		o.extBegin = hits[first].ext;
		o.extEnd = hits[last].ext + k - 1;
		o.extLen = extLen;
		o.score = score[last] - score[first] + k - 1;
		o.chainLength = len;
		if (!overlapTest(o, p)) continue;

		auto lo = std::lower_bound(filtered.begin(), filtered.end(), o.curBegin);
		auto hi = std::upper_bound(filtered.begin(), filtered.end(), o.curEnd);
		int32_t filteredCount = static_cast<int32_t>(hi - lo);
		float normLen = static_cast<float>(std::max(o.curRange(), o.extRange()) - filteredCount);
		if (!(normLen > 0.0f)) continue;
		float rate = static_cast<float>(len) * p.sample_rate / normLen;
		rate = std::min(rate, 1.0f);
		if (!(rate > 0.0f)) continue;
		float inv = 1.0f / rate;
		float lg = logf(inv);
		o.div = lg / static_cast<float>(k);
		ext.push_back(o);
	}

	if (p.only_max_ext)
	{
		int best = -1;
		for (size_t i = 0; i < ext.size(); ++i)
			if (best < 0 || ext[i].score > ext[static_cast<size_t>(best)].score) best = static_cast<int>(i);
		if (best >= 0 && ext[static_cast<size_t>(best)].div < p.max_divergence)
			detected.push_back(ext[static_cast<size_t>(best)]);
		return;
	}
	std::sort(ext.begin(), ext.end(), [](const Ovlp& a, const Ovlp& b) { return a.score > b.score; });
	std::vector<Ovlp> primary;
	for (const auto& o : ext)
	{
		bool contained = false;
		for (const auto& pr : primary)
		{
			if (o.containedBy(pr) && pr.score > o.score)
			{
				contained = true;
				break;
			}
		}
		if (!contained) primary.push_back(o);
	}
	for (const auto& o : primary)
		if (o.div < p.max_divergence) detected.push_back(o);
}

// getSeqOverlaps for one query record, from the candidate list on.
void overlapsFromMatches(std::vector<Match>& matches, int64_t curId, int32_t curLen,
						 const std::vector<int32_t>& filtered, const std::vector<int32_t>& extLenOf,
						 const flyehip_overlap_params_t& p, std::vector<Ovlp>& detected)
{
	std::sort(matches.begin(), matches.end(), [](const Match& a, const Match& b) {
		return a.extId != b.extId ? a.extId < b.extId : a.curPos < b.curPos;
	});
	std::vector<CE> hits;
	size_t b = 0;
	while (b < matches.size())
	{
		size_t e = b + 1;
		while (e < matches.size() && matches[e].extId == matches[b].extId) ++e;
		if (p.max_overlaps != 0 && detected.size() >= static_cast<size_t>(p.max_overlaps)) break;
		hits.clear();
		for (size_t i = b; i < e; ++i) hits.push_back({matches[i].curPos, matches[i].extPos});
		uint32_t eid = matches[b].extId;
		chainGroup(hits, curId, curLen, toSigned(eid), extLenOf[eid >> 1], filtered, p, detected);
		b = e;
	}
}

void divStatsOf(const std::vector<Ovlp>& ovs, int32_t curLen, std::vector<float>& out)
{
	const int STAT_WND = 10000;
	std::vector<const Ovlp*> win(static_cast<size_t>(curLen / STAT_WND + 1), nullptr);
	for (const auto& o : ovs)
	{
		if (o.curBegin < 0) continue;
		size_t w = static_cast<size_t>(o.curBegin / STAT_WND);
		if (w >= win.size()) continue;
		int32_t have = win[w] ? win[w]->curRange() : 0;
		if (o.curRange() > have) win[w] = &o;
	}
	for (auto* o : win)
		if (o && o->curRange() > 0) out.push_back(o->div);
}

void toRow(const Ovlp& o, flyehip_overlap_t& r)
{
	std::memset(&r, 0, sizeof(r));
	r.cur_id = o.curId;
	r.ext_id = o.extId;
	r.cur_begin = o.curBegin;
	r.cur_end = o.curEnd;
	r.cur_len = o.curLen;
	r.ext_begin = o.extBegin;
	r.ext_end = o.extEnd;
	r.ext_len = o.extLen;
	r.score = o.score;
	r.seq_divergence = o.div;
	r.chain_length = o.chainLength;
}

struct Oracle
{
	Store targets;
	Index* index = nullptr;
	~Oracle() { delete index; }
};

template <class F>
void parallelFor(int32_t n, int threads, F f)
{
	if (threads <= 1 || n <= 1)
	{
		for (int32_t i = 0; i < n; ++i) f(i);
		return;
	}
	std::atomic<int32_t> next(0);
	std::vector<std::thread> pool;
	for (int t = 0; t < threads; ++t)
	{
		pool.emplace_back([&]() {
			for (;;)
			{
				int32_t i = next.fetch_add(1);
				if (i >= n) break;
				f(i);
			}
		});
	}
	for (auto& th : pool) th.join();
}

std::string querySeq(const Oracle& o, int64_t sid)
{
	int64_t n = sid > 0 ? sid : -sid;
	const std::string& s = o.targets.fwd[static_cast<size_t>(n - 1)];
	return sid > 0 ? s : revComp(s);
}

// chainReadAlignments (tools/replay_read_alignment_chains.py:64-139) over the
// kept, curBegin-sorted alignments of one read.  Chains are index lists.
struct RChain
{
	std::vector<int32_t> idx;
	int32_t score;
};

void chainRead(const std::vector<flyehip_edge_alignment_t>& a, const flyehip_chain_params_t& p,
			   std::vector<RChain>& accepted)
{
	std::vector<RChain> active, frozen;
	for (int32_t i = 0; i < static_cast<int32_t>(a.size()); ++i)
	{
		const auto& cur = a[static_cast<size_t>(i)];
		int32_t maxScore = 0;
		int maxChain = -1;
		int32_t numOutdated = 0;
		bool canExtend = cur.edge_begin < p.maximum_jump;
		bool canBeExtended = cur.edge_len - cur.edge_end < p.maximum_jump;
		if (canExtend)
		{
			for (size_t c = 0; c < active.size(); ++c)
			{
				const auto& prev = a[static_cast<size_t>(active[c].idx.back())];
				int32_t readDiff = cur.read_begin - prev.read_end;
				int32_t graphLeft = cur.edge_begin;
				int32_t graphRight = prev.edge_len - prev.edge_end;
				bool connected = prev.edge_right_node == cur.edge_left_node;
				if (connected && p.maximum_jump > readDiff && readDiff > -p.max_read_overlap &&
					graphLeft + graphRight < p.maximum_jump)
				{
					int32_t jumpDiv = std::abs(readDiff - (graphLeft + graphRight));
					int32_t gapCost = jumpDiv > 100 ? jumpDiv / 50 : 0;
					int32_t sc = active[c].score + cur.score - gapCost;
					if (sc > maxScore)
					{
						maxScore = sc;
						maxChain = static_cast<int>(c);
					}
				}
				if (readDiff > p.maximum_jump) ++numOutdated;
			}
		}
		if (maxChain >= 0)
		{
			RChain nc;
			nc.idx = active[static_cast<size_t>(maxChain)].idx;
			nc.idx.push_back(i);
			nc.score = maxScore;
			active.push_back(nc);
		}
		else if (canBeExtended)
		{
			active.push_back({{i}, cur.score});
		}
		else
		{
			frozen.push_back({{i}, cur.score});
		}
		if (numOutdated > static_cast<int32_t>(active.size()) / 2)
		{
			std::vector<RChain> keep;
			for (auto& c : active)
			{
				const auto& prev = a[static_cast<size_t>(c.idx.back())];
				if (cur.read_begin - prev.read_end > p.maximum_jump) frozen.push_back(c);
				else keep.push_back(c);
			}
			active.swap(keep);
		}
	}
	for (auto& c : frozen) active.push_back(c);
	std::sort(active.begin(), active.end(),
			  [](const RChain& x, const RChain& y) { return x.score > y.score; });
	for (const auto& c : active)
	{
		const auto& first = a[static_cast<size_t>(c.idx.front())];
		const auto& last = a[static_cast<size_t>(c.idx.back())];
		int32_t alnLen = last.read_end - first.read_begin;
		if (alnLen < p.minimum_overlap) continue;
		bool overlapsExisting = false;
		for (const auto& ex : accepted)
		{
			const auto& ef = a[static_cast<size_t>(ex.idx.front())];
			const auto& el = a[static_cast<size_t>(ex.idx.back())];
			int32_t rate = std::min(last.read_end, el.read_end) -
						   std::max(first.read_begin, ef.read_begin);
			if (rate > p.max_separation)
			{
				overlapsExisting = true;
				break;
			}
		}
		if (!overlapsExisting) accepted.push_back(c);
	}
}

} // namespace

extern "C" {

// ---- k-mer helpers (pinned by tests/fixtures/kmer-*-smoke-v0) -------------
uint64_t orc_encode_kmer(const char* s, int32_t k)
{
	uint64_t x = 0;
	for (int i = 0; i < k; ++i) x = (x << 2) | static_cast<uint64_t>(std::max(0, baseCode(static_cast<uint8_t>(s[i]))));
	return x;
}
uint64_t orc_reverse_complement(uint64_t x, int32_t k) { return rcKmer(x, k); }
uint64_t orc_standard_form(uint64_t x, int32_t k, int32_t* flipped)
{
	uint64_t r = rcKmer(x, k);
	bool flip = !(x < r);
	if (flipped) *flipped = flip ? 1 : 0;
	return flip ? r : x;
}
uint32_t orc_signed_to_internal(int64_t sid) { return toInternal(sid); }
int64_t orc_internal_to_signed(uint32_t iid) { return toSigned(iid); }

// ---- reads + index -------------------------------------------------------
void* orc_create(const uint8_t* bases, const int64_t* offsets, int32_t n_reads)
{
	Oracle* o = new Oracle();
	for (int32_t i = 0; i < n_reads; ++i)
		o->targets.fwd.emplace_back(reinterpret_cast<const char*>(bases + offsets[i]),
									static_cast<size_t>(offsets[i + 1] - offsets[i]));
	return o;
}

void orc_destroy(void* h) { delete static_cast<Oracle*>(h); }

// out[0..7] = solid_min, rep_max, distinct, solid, repetitive, positions, hist_len, 0
int orc_build_index_minimizers(void* h, int32_t k, int32_t window, int32_t solid_min, int32_t rep_max, float rate,
							   int64_t* out);
int orc_build_index(void* h, int32_t k, int32_t solid_min, int32_t rep_max, float rate, int64_t* out)
{
	return orc_build_index_minimizers(h, k, 0, solid_min, rep_max, rate, out);
}

// window 0: every position; else minimizer positions only
int orc_build_index_minimizers(void* h, int32_t k, int32_t window, int32_t solid_min, int32_t rep_max, float rate,
							   int64_t* out)
{
	Oracle* o = static_cast<Oracle*>(h);
	delete o->index;
	o->index = buildIndex(o->targets, k, solid_min, rep_max, rate, window);
	const Index& ix = *o->index;
	out[0] = ix.solidMin;
	out[1] = ix.repMax;
	out[2] = ix.distinct;
	out[3] = static_cast<int64_t>(ix.kmers.size());
	out[4] = static_cast<int64_t>(ix.repetitive.size());
	out[5] = ix.start.empty() ? 0 : ix.start.back();
	out[6] = static_cast<int64_t>(ix.hist.size());
	out[7] = 0;
	return 0;
}

// Minimizer flags of one sequence (one byte per k-mer position).
int64_t orc_minimizer_flags(const uint8_t* bases, int64_t len, int32_t k, int32_t window, uint8_t* out)
{
	const std::string s(reinterpret_cast<const char*>(bases), static_cast<size_t>(len));
	const std::vector<char> f = minimizerFlags(s, k, window);
	for (size_t i = 0; i < f.size(); ++i) out[i] = static_cast<uint8_t>(f[i]);
	return static_cast<int64_t>(f.size());
}

// An index given as tables (flyehip_load_index's arguments): canonical
// k-mers with their buckets (entries in the given order: signed id and the
// position where the canonical k-mer is spelled) and the repetitive k-mers.
int orc_load_index(void* h, int32_t k, const uint64_t* kmers, int64_t n_kmers, const int64_t* offsets,
				   const int64_t* ids, const int32_t* pos, const uint64_t* rep, int64_t n_rep)
{
	Oracle* o = static_cast<Oracle*>(h);
	delete o->index;
	Index* ix = new Index();
	ix->k = k;
	std::vector<int64_t> order(static_cast<size_t>(n_kmers));
	for (int64_t i = 0; i < n_kmers; ++i) order[static_cast<size_t>(i)] = i;
	std::sort(order.begin(), order.end(), [kmers](int64_t a, int64_t b) { return kmers[a] < kmers[b]; });
	ix->start.push_back(0);
	for (int64_t i : order)
	{
		ix->kmers.push_back(kmers[i]);
		for (int64_t e = offsets[i]; e < offsets[i + 1]; ++e)
		{
			ix->entryId.push_back(toInternal(ids[e]));
			ix->entryPos.push_back(pos[e]);
		}
		ix->start.push_back(static_cast<int64_t>(ix->entryId.size()));
	}
	ix->repetitive.assign(rep, rep + n_rep);
	std::sort(ix->repetitive.begin(), ix->repetitive.end());
	for (const auto& s : o->targets.fwd) ix->readLen.push_back(static_cast<int32_t>(s.size()));
	ix->distinct = n_kmers + n_rep;
	ix->buildLookup();
	o->index = ix;
	return 0;
}

int64_t orc_histogram(void* h, int64_t* out, int64_t cap)
{
	const Index& ix = *static_cast<Oracle*>(h)->index;
	int64_t n = std::min<int64_t>(cap, static_cast<int64_t>(ix.hist.size()));
	for (int64_t i = 0; i < n; ++i) out[i] = ix.hist[static_cast<size_t>(i)];
	return static_cast<int64_t>(ix.hist.size());
}

// Candidates of one query record (a resident read strand, or an external
// sequence when seq != NULL), in emission order (sorted = 0) or after the
// std::sort by (extId, curPos) (sorted = 1).  Returns the match count; writes
// at most cap rows.  Filtered positions go to filt (cap fcap, count *nfilt).
int64_t orc_candidates(void* h, int64_t query_id, const uint8_t* seq, int32_t len, int32_t sorted,
					   flyehip_candidate_t* out, int64_t cap, int32_t* filt, int64_t fcap,
					   int64_t* nfilt)
{
	Oracle* o = static_cast<Oracle*>(h);
	std::string q = seq ? std::string(reinterpret_cast<const char*>(seq), static_cast<size_t>(len))
						: querySeq(*o, query_id);
	std::vector<Match> m;
	std::vector<int32_t> f;
	collect(*o->index, toInternal(query_id), q, m, f);
	if (sorted)
	{
		std::sort(m.begin(), m.end(), [](const Match& a, const Match& b) {
			return a.extId != b.extId ? a.extId < b.extId : a.curPos < b.curPos;
		});
	}
	for (size_t i = 0; i < m.size() && static_cast<int64_t>(i) < cap; ++i)
	{
		out[i].query_id = query_id;
		out[i].target_id = toSigned(m[i].extId);
		out[i].kmer = m[i].kmer;
		out[i].query_pos = m[i].curPos;
		out[i].target_pos = m[i].extPos;
	}
	for (size_t i = 0; i < f.size() && static_cast<int64_t>(i) < fcap; ++i) filt[i] = f[i];
	if (nfilt) *nfilt = static_cast<int64_t>(f.size());
	return static_cast<int64_t>(m.size());
}

// getSeqOverlaps over many query records with `threads` workers.  Query i is a
// resident strand (seq_offsets == NULL) or bases[seq_offsets[i]..[i+1]).
// rows: cap total; row_offsets n+1; div stats likewise (may be NULL).
int64_t orc_overlaps(void* h, const int64_t* query_ids, int32_t n, const uint8_t* bases,
					 const int64_t* seq_offsets, const flyehip_overlap_params_t* params,
					 int32_t threads, flyehip_overlap_t* rows, int64_t cap, int64_t* row_offsets,
					 float* divs, int64_t dcap, int64_t* div_offsets)
{
	Oracle* o = static_cast<Oracle*>(h);
	std::vector<std::vector<Ovlp>> per(static_cast<size_t>(n));
	std::vector<std::vector<float>> perDiv(static_cast<size_t>(n));
	parallelFor(n, threads, [&](int32_t i) {
		std::string q = seq_offsets
							? std::string(reinterpret_cast<const char*>(bases + seq_offsets[i]),
										  static_cast<size_t>(seq_offsets[i + 1] - seq_offsets[i]))
							: querySeq(*o, query_ids[i]);
		std::vector<Match> m;
		std::vector<int32_t> f;
		collect(*o->index, toInternal(query_ids[i]), q, m, f);
		overlapsFromMatches(m, query_ids[i], static_cast<int32_t>(q.size()), f, o->index->readLen,
							*params, per[static_cast<size_t>(i)]);
		divStatsOf(per[static_cast<size_t>(i)], static_cast<int32_t>(q.size()),
				   perDiv[static_cast<size_t>(i)]);
	});
	int64_t total = 0, dtotal = 0;
	row_offsets[0] = 0;
	if (div_offsets) div_offsets[0] = 0;
	for (int32_t i = 0; i < n; ++i)
	{
		for (const auto& ov : per[static_cast<size_t>(i)])
		{
			if (total < cap) toRow(ov, rows[total]);
			++total;
		}
		row_offsets[i + 1] = total;
		for (float d : perDiv[static_cast<size_t>(i)])
		{
			if (divs && dtotal < dcap) divs[dtotal] = d;
			++dtotal;
		}
		if (div_offsets) div_offsets[i + 1] = dtotal;
	}
	return total;
}

// Chaining of given hit lists (hits already in group order).
int64_t orc_chain_lists(const int32_t* cur, const int32_t* ext, const int64_t* offs, int32_t n,
						const int64_t* cur_ids, const int64_t* ext_ids, const int32_t* cur_lens,
						const int32_t* ext_lens, const flyehip_overlap_params_t* params,
						int32_t threads, flyehip_overlap_t* rows, int64_t cap, int64_t* row_offsets)
{
	std::vector<std::vector<Ovlp>> per(static_cast<size_t>(n));
	std::vector<int32_t> noFiltered;
	parallelFor(n, threads, [&](int32_t i) {
		std::vector<CE> hits;
		for (int64_t j = offs[i]; j < offs[i + 1]; ++j) hits.push_back({cur[j], ext[j]});
		if (hits.empty()) return;
		chainGroup(hits, cur_ids[i], cur_lens[i], ext_ids[i], ext_lens[i], noFiltered, *params,
				   per[static_cast<size_t>(i)]);
	});
	int64_t total = 0;
	row_offsets[0] = 0;
	for (int32_t i = 0; i < n; ++i)
	{
		for (const auto& ov : per[static_cast<size_t>(i)])
		{
			if (total < cap) toRow(ov, rows[total]);
			++total;
		}
		row_offsets[i + 1] = total;
	}
	return total;
}

// Read-to-graph chaining for a batch of reads; mirrors flyehip_read_align_chains.
int64_t orc_read_align_chains(const flyehip_edge_alignment_t* alns, const int64_t* offs,
							  int32_t n_reads, int32_t long_edge, int32_t big_aln,
							  const flyehip_chain_params_t* params, int32_t* sorted_index,
							  int64_t* kept_offsets, flyehip_chain_segment_t* segs, int64_t cap,
							  int64_t* seg_offsets)
{
	int64_t total = 0;
	int64_t kept = 0;
	kept_offsets[0] = 0;
	seg_offsets[0] = 0;
	for (int32_t r = 0; r < n_reads; ++r)
	{
		std::vector<int32_t> idx;
		for (int64_t j = offs[r]; j < offs[r + 1]; ++j)
		{
			const auto& a = alns[j];
			int32_t curRange = a.read_end - a.read_begin;
			int32_t extRange = a.edge_end - a.edge_begin;
			if (a.edge_len < long_edge || std::min(curRange, extRange) > big_aln)
				idx.push_back(static_cast<int32_t>(j - offs[r]));
		}
		const flyehip_edge_alignment_t* base = alns + offs[r];
		std::sort(idx.begin(), idx.end(), [base](int32_t x, int32_t y) {
			return base[x].read_begin < base[y].read_begin;
		});
		std::vector<flyehip_edge_alignment_t> a;
		for (int32_t x : idx)
		{
			sorted_index[kept++] = x;
			a.push_back(base[x]);
		}
		kept_offsets[r + 1] = kept;
		std::vector<RChain> acc;
		chainRead(a, *params, acc);
		for (size_t c = 0; c < acc.size(); ++c)
		{
			for (size_t s = 0; s < acc[c].idx.size(); ++s)
			{
				if (total < cap)
				{
					segs[total].read_slot = r;
					segs[total].chain_id = static_cast<int32_t>(c);
					segs[total].segment_id = static_cast<int32_t>(s);
					segs[total].input_index = acc[c].idx[s];
					segs[total].chain_score = acc[c].score;
					segs[total].reserved = 0;
				}
				++total;
			}
		}
		seg_offsets[r + 1] = total;
	}
	return total;
}

// candidate-record-v1 canonical SHA-256 of the query records' candidates
// (tools/validate_candidate_dump.py:76-134: records as decimal TSV lines,
// sorted as numeric tuples, the text hashed).  Each query's records are
// sorted whole here; query records are taken in ascending id order and
// chunks of them are collected on `threads` workers, hashed in order.
int orc_candidates_canonical_sha256(void* h, const int64_t* query_ids, int32_t n, int32_t threads,
									uint8_t* out, int64_t* records)
{
	Oracle* o = static_cast<Oracle*>(h);
	std::vector<int64_t> ids(query_ids, query_ids + n);
	std::sort(ids.begin(), ids.end());
	EVP_MD_CTX* md = EVP_MD_CTX_new();
	if (!md || EVP_DigestInit_ex(md, EVP_sha256(), nullptr) != 1) return -1;
	int64_t total = 0;
	const int32_t chunk = 64;
	std::vector<std::string> text(static_cast<size_t>(chunk));
	std::vector<int64_t> counts(static_cast<size_t>(chunk));
	for (int32_t c0 = 0; c0 < n; c0 += chunk)
	{
		const int32_t m = std::min(chunk, n - c0);
		parallelFor(m, threads, [&](int32_t i) {
			const int64_t qid = ids[static_cast<size_t>(c0 + i)];
			std::string q = querySeq(*o, qid);
			std::vector<Match> ms;
			std::vector<int32_t> f;
			collect(*o->index, toInternal(qid), q, ms, f);
			struct Rec
			{
				int32_t qpos;
				uint64_t kmer;
				int64_t tid;
				int32_t tpos;
			};
			std::vector<Rec> rs;
			rs.reserve(ms.size());
			for (const auto& m2 : ms) rs.push_back({m2.curPos, m2.kmer, toSigned(m2.extId), m2.extPos});
			std::sort(rs.begin(), rs.end(), [](const Rec& a, const Rec& b) {
				if (a.qpos != b.qpos) return a.qpos < b.qpos;
				if (a.kmer != b.kmer) return a.kmer < b.kmer;
				if (a.tid != b.tid) return a.tid < b.tid;
				return a.tpos < b.tpos;
			});
			std::string& s = text[static_cast<size_t>(i)];
			s.clear();
			for (const auto& r : rs)
			{
				s += std::to_string(qid);
				s += '\t';
				s += std::to_string(r.qpos);
				s += '\t';
				s += std::to_string(r.kmer);
				s += '\t';
				s += std::to_string(r.tid);
				s += '\t';
				s += std::to_string(r.tpos);
				s += r.tid > 0 ? "\t+\n" : "\t-\n";
			}
			counts[static_cast<size_t>(i)] = static_cast<int64_t>(rs.size());
		});
		for (int32_t i = 0; i < m; ++i)
		{
			const std::string& s = text[static_cast<size_t>(i)];
			if (!s.empty()) EVP_DigestUpdate(md, s.data(), s.size());
			total += counts[static_cast<size_t>(i)];
		}
	}
	unsigned int len = 0;
	const int ok = EVP_DigestFinal_ex(md, out, &len);
	EVP_MD_CTX_free(md);
	if (records) *records = total;
	return ok == 1 && len == 32 ? 0 : -1;
}

// Real libstdc++ std::sort of 64-bit keys, returning the permutation: the
// reference the device's exact sort emulation is checked against.
void orc_std_sort_perm(const uint64_t* keys, int64_t n, int32_t* perm)
{
	std::vector<int32_t> idx(static_cast<size_t>(n));
	for (int64_t i = 0; i < n; ++i) idx[static_cast<size_t>(i)] = static_cast<int32_t>(i);
	std::sort(idx.begin(), idx.end(), [keys](int32_t a, int32_t b) { return keys[a] < keys[b]; });
	for (int64_t i = 0; i < n; ++i) perm[i] = idx[static_cast<size_t>(i)];
}

// glibc logf, the function Flye's divergence calls.
float orc_logf(float x) { return logf(x); }

// glibc logf over an array (the exhaustive comparison of the device's logf)
void orc_logf_array(const float* x, int64_t n, float* out)
{
	for (int64_t i = 0; i < n; ++i) out[i] = logf(x[i]);
}

// ---- pieces of one group, exported for the tests that pin them ------------

// The DP table of one group over hits in their final order (no sort).
void orc_chain_dp(const int32_t* cur, const int32_t* ext, int32_t n,
				  const flyehip_overlap_params_t* params, int32_t ext_sorted, int32_t early_break,
				  int32_t* score, int32_t* bt)
{
	std::vector<CE> hits(static_cast<size_t>(n));
	for (int32_t i = 0; i < n; ++i) hits[static_cast<size_t>(i)] = {cur[i], ext[i]};
	std::vector<int32_t> s, b;
	chainDp(hits, *params, ext_sorted != 0, early_break != 0, s, b);
	for (int32_t i = 0; i < n; ++i)
	{
		score[i] = s[static_cast<size_t>(i)];
		bt[i] = b[static_cast<size_t>(i)];
	}
}

// One (query, target) group with the query's filtered positions (ascending):
// chainGroup as getSeqOverlaps runs it.  Returns the row count; writes at
// most cap rows.
int64_t orc_chain_group(const int32_t* cur, const int32_t* ext, int32_t n, int64_t cur_id,
						int32_t cur_len, int64_t ext_id, int32_t ext_len, const int32_t* filtered,
						int32_t n_filtered, const flyehip_overlap_params_t* params,
						flyehip_overlap_t* rows, int64_t cap)
{
	if (n <= 0) return 0;
	std::vector<CE> hits(static_cast<size_t>(n));
	for (int32_t i = 0; i < n; ++i) hits[static_cast<size_t>(i)] = {cur[i], ext[i]};
	std::vector<int32_t> filt(filtered, filtered + n_filtered);
	std::vector<Ovlp> det;
	chainGroup(hits, cur_id, cur_len, ext_id, ext_len, filt, *params, det);
	for (size_t i = 0; i < det.size() && static_cast<int64_t>(i) < cap; ++i) toRow(det[i], rows[i]);
	return static_cast<int64_t>(det.size());
}

// overlapTest of one proposal (cur/ext ids, ranges and lengths from the row).
int32_t orc_overlap_test(const flyehip_overlap_t* row, const flyehip_overlap_params_t* params)
{
	Ovlp o{};
	o.curId = row->cur_id;
	o.extId = row->ext_id;
	o.curBegin = row->cur_begin;
	o.curEnd = row->cur_end;
	o.curLen = row->cur_len;
	o.extBegin = row->ext_begin;
	o.extEnd = row->ext_end;
	o.extLen = row->ext_len;
	return overlapTest(o, *params) ? 1 : 0;
}

// divStats windows of one query's detected overlaps, in row order.
int64_t orc_div_stats(const flyehip_overlap_t* rows, int64_t n, int32_t cur_len, float* out,
					  int64_t cap)
{
	std::vector<Ovlp> ovs(static_cast<size_t>(n));
	for (int64_t i = 0; i < n; ++i)
	{
		const flyehip_overlap_t& r = rows[i];
		Ovlp& o = ovs[static_cast<size_t>(i)];
		o.curId = r.cur_id;
		o.extId = r.ext_id;
		o.curBegin = r.cur_begin;
		o.curEnd = r.cur_end;
		o.curLen = r.cur_len;
		o.extBegin = r.ext_begin;
		o.extEnd = r.ext_end;
		o.extLen = r.ext_len;
		o.score = r.score;
		o.div = r.seq_divergence;
		o.chainLength = r.chain_length;
	}
	std::vector<float> d;
	divStatsOf(ovs, cur_len, d);
	for (size_t i = 0; i < d.size() && static_cast<int64_t>(i) < cap; ++i) out[i] = d[i];
	return static_cast<int64_t>(d.size());
}

// The index tables after orc_build_index: counts[0..2] = indexed k-mers,
// bucket entries, repetitive k-mers; each array is written when its pointer is
// not NULL (kmers ascending, start = bucket offsets, entry ids / positions in
// bucket order, repetitive k-mers ascending).
void orc_index_tables(void* h, int64_t* counts, uint64_t* kmers, int64_t* start, uint32_t* ids,
					  int32_t* pos, uint64_t* repetitive)
{
	const Index& ix = *static_cast<Oracle*>(h)->index;
	counts[0] = static_cast<int64_t>(ix.kmers.size());
	counts[1] = ix.start.empty() ? 0 : ix.start.back();
	counts[2] = static_cast<int64_t>(ix.repetitive.size());
	if (kmers) std::copy(ix.kmers.begin(), ix.kmers.end(), kmers);
	if (start) std::copy(ix.start.begin(), ix.start.end(), start);
	if (ids) std::copy(ix.entryId.begin(), ix.entryId.end(), ids);
	if (pos) std::copy(ix.entryPos.begin(), ix.entryPos.end(), pos);
	if (repetitive) std::copy(ix.repetitive.begin(), ix.repetitive.end(), repetitive);
}

} // extern "C"
