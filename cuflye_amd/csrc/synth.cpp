// synth.cpp - seeded synthetic genomes, reads, hit lists and repeat graphs
// for the benchmark workloads and the parity tests (host only, no HIP).
//
// Every generator is a pure function of its parameters and seed: it uses
// xoshiro256** seeded through splitmix64 (Blackman & Vigna) and glibc's
// log/exp/cos/sqrt for the normal and log-normal draws, so one image gives
// the same bytes on every machine.  Sequences are ASCII A/C/G/T.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

namespace
{

struct Rng
{
	uint64_t s[4];
	explicit Rng(uint64_t seed)
	{
		uint64_t z = seed;
		for (int i = 0; i < 4; ++i)
		{
			z += 0x9e3779b97f4a7c15ULL;
			uint64_t x = z;
			x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
			x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
			s[i] = x ^ (x >> 31);
		}
	}
	static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
	uint64_t next()
	{
		uint64_t r = rotl(s[1] * 5, 7) * 9;
		uint64_t t = s[1] << 17;
		s[2] ^= s[0];
		s[3] ^= s[1];
		s[1] ^= s[2];
		s[0] ^= s[3];
		s[2] ^= t;
		s[3] = rotl(s[3], 45);
		return r;
	}
	double uniform() { return static_cast<double>(next() >> 11) * (1.0 / 9007199254740992.0); }
	uint64_t below(uint64_t n) { return n ? next() % n : 0; }
	double normal()
	{
		double u1 = uniform();
		double u2 = uniform();
		if (u1 < 1e-300) u1 = 1e-300;
		return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
	}
};

const char kBases[4] = {'A', 'C', 'G', 'T'};

char randomBase(Rng& r, double gc)
{
	double u = r.uniform();
	if (u < gc) return (r.next() & 1) ? 'C' : 'G';
	return (r.next() & 1) ? 'A' : 'T';
}

char complement(char c)
{
	return c == 'A' ? 'T' : c == 'C' ? 'G' : c == 'G' ? 'C' : 'A';
}

std::string revComp(const std::string& s)
{
	std::string r(s.size(), 'A');
	for (size_t i = 0; i < s.size(); ++i) r[i] = complement(s[s.size() - 1 - i]);
	return r;
}

// Substitution-only copy of `unit` at divergence d.
std::string mutateCopy(Rng& r, const std::string& unit, double d, double gc)
{
	std::string c = unit;
	for (auto& b : c)
	{
		if (r.uniform() < d)
		{
			char nb = b;
			while (nb == b) nb = randomBase(r, gc);
			b = nb;
		}
	}
	return c;
}

struct Dataset
{
	std::string genome;
	std::vector<std::string> reads;
	std::vector<int64_t> readStart;   // genome coordinate of the template
	std::vector<int32_t> readStrand;  // 1 forward, 0 reverse
	// repeat graph (config 4): edge sequences and node ids
	std::vector<std::string> edgeSeqs;
	std::vector<int32_t> edgeLeft, edgeRight, edgeIsRepeat;
	// hit lists (config 3)
	std::vector<int32_t> cur, ext;
	std::vector<int64_t> listOffsets;
	std::vector<int32_t> curLen, extLen;
	std::vector<int64_t> curId, extId;
	// flat views
	std::string flat;
	std::vector<int64_t> offsets;
};

void flatten(Dataset& d, const std::vector<std::string>& seqs)
{
	d.flat.clear();
	d.offsets.assign(1, 0);
	size_t total = 0;
	for (auto& s : seqs) total += s.size();
	d.flat.reserve(total);
	for (auto& s : seqs)
	{
		d.flat += s;
		d.offsets.push_back(static_cast<int64_t>(d.flat.size()));
	}
}

// Overwrites repeat families into the genome until `frac` of its bases are
// repeat copies.  Each family: random unit of [lenMin, lenMax], [cMin, cMax]
// copies, each at a divergence drawn from [dMin, dMax]; placements avoid
// earlier copies.  Returns the placed intervals (start, length, family).
struct Placed
{
	int64_t start;
	int64_t len;
	int32_t family;
};

std::vector<Placed> overwriteRepeats(Rng& r, std::string& g, double frac, int lenMin,
									 int lenMax, int cMin, int cMax, double dMin, double dMax,
									 double gc, std::vector<std::string>* units)
{
	std::vector<Placed> placed;
	int64_t target = static_cast<int64_t>(frac * static_cast<double>(g.size()));
	int64_t have = 0;
	int32_t family = 0;
	int guard = 0;
	while (have < target && guard < 1000000)
	{
		int len = lenMin + static_cast<int>(r.below(static_cast<uint64_t>(lenMax - lenMin + 1)));
		int copies = cMin + static_cast<int>(r.below(static_cast<uint64_t>(cMax - cMin + 1)));
		std::string unit(static_cast<size_t>(len), 'A');
		for (auto& b : unit) b = randomBase(r, gc);
		if (units) units->push_back(unit);
		for (int c = 0; c < copies && have < target; ++c)
		{
			for (int attempt = 0; attempt < 64; ++attempt)
			{
				++guard;
				if (static_cast<int64_t>(g.size()) <= len) break;
				int64_t s = static_cast<int64_t>(r.below(g.size() - static_cast<uint64_t>(len)));
				bool clash = false;
				for (auto& p : placed)
				{
					if (s < p.start + p.len + 200 && p.start < s + len + 200)
					{
						clash = true;
						break;
					}
				}
				if (clash) continue;
				double d = dMin + (dMax - dMin) * r.uniform();
				std::string copy = mutateCopy(r, unit, d, gc);
				if (r.next() & 1) copy = revComp(copy);
				g.replace(static_cast<size_t>(s), static_cast<size_t>(len), copy);
				placed.push_back({s, len, family});
				have += len;
				break;
			}
		}
		++family;
	}
	std::sort(placed.begin(), placed.end(), [](const Placed& a, const Placed& b) { return a.start < b.start; });
	return placed;
}

// Error model: per template base, with probability err an error whose kind is
// drawn by weights (ins, del, sub); inside a homopolymer run (base equal to
// the previous one) the deletion weight is multiplied by hpDelBoost.
std::string addErrors(Rng& r, const std::string& t, double err, double wIns, double wDel,
					  double wSub, double hpDelBoost)
{
	std::string out;
	out.reserve(t.size() + t.size() / 16);
	for (size_t i = 0; i < t.size(); ++i)
	{
		char b = t[i];
		if (r.uniform() >= err)
		{
			out.push_back(b);
			continue;
		}
		double wd = wDel * ((i > 0 && t[i - 1] == b) ? hpDelBoost : 1.0);
		double u = r.uniform() * (wIns + wd + wSub);
		if (u < wIns)
		{
			out.push_back(kBases[r.below(4)]);
			out.push_back(b);
		}
		else if (u < wIns + wd)
		{
			// deleted
		}
		else
		{
			char nb = b;
			while (nb == b) nb = kBases[r.below(4)];
			out.push_back(nb);
		}
	}
	return out;
}

void sampleReads(Rng& r, Dataset& d, double coverage, double mu, double sigma, int minLen,
				 int maxLen, double err, double wIns, double wDel, double wSub, double hpBoost)
{
	const std::string& g = d.genome;
	const int64_t glen = static_cast<int64_t>(g.size());
	int64_t want = static_cast<int64_t>(coverage * static_cast<double>(glen));
	int64_t have = 0;
	while (have < want)
	{
		int len = 0;
		for (int tries = 0; tries < 1000; ++tries)
		{
			double v = std::exp(mu + sigma * r.normal());
			if (v >= minLen && v <= maxLen)
			{
				len = static_cast<int>(v);
				break;
			}
		}
		if (len <= 0) len = minLen;
		if (len > glen) len = static_cast<int>(glen);
		int64_t s = static_cast<int64_t>(r.below(static_cast<uint64_t>(glen - len + 1)));
		std::string t = g.substr(static_cast<size_t>(s), static_cast<size_t>(len));
		int strand = (r.next() & 1) ? 1 : 0;
		if (!strand) t = revComp(t);
		std::string read = err > 0 ? addErrors(r, t, err, wIns, wDel, wSub, hpBoost) : t;
		d.reads.push_back(read);
		d.readStart.push_back(s);
		d.readStrand.push_back(strand);
		have += len;
	}
}

} // namespace

extern "C" {

void syn_destroy(void* h) { delete static_cast<Dataset*>(h); }

// Reads (or edge sequences, `which` = 1; genome, `which` = 2) as one flat
// buffer plus offsets.
int64_t syn_bases_size(void* h, int32_t which)
{
	Dataset* d = static_cast<Dataset*>(h);
	if (which == 2) return static_cast<int64_t>(d->genome.size());
	flatten(*d, which == 1 ? d->edgeSeqs : d->reads);
	return static_cast<int64_t>(d->flat.size());
}
int32_t syn_count(void* h, int32_t which)
{
	Dataset* d = static_cast<Dataset*>(h);
	if (which == 1) return static_cast<int32_t>(d->edgeSeqs.size());
	if (which == 3) return static_cast<int32_t>(d->curLen.size());
	return static_cast<int32_t>(d->reads.size());
}
// Copies the buffer flattened by the last syn_bases_size(h, which) call.
void syn_copy_bases(void* h, int32_t which, uint8_t* bases, int64_t* offsets)
{
	Dataset* d = static_cast<Dataset*>(h);
	if (which == 2)
	{
		std::memcpy(bases, d->genome.data(), d->genome.size());
		return;
	}
	std::memcpy(bases, d->flat.data(), d->flat.size());
	std::memcpy(offsets, d->offsets.data(), d->offsets.size() * sizeof(int64_t));
}
void syn_copy_read_truth(void* h, int64_t* start, int32_t* strand)
{
	Dataset* d = static_cast<Dataset*>(h);
	std::memcpy(start, d->readStart.data(), d->readStart.size() * sizeof(int64_t));
	std::memcpy(strand, d->readStrand.data(), d->readStrand.size() * sizeof(int32_t));
}
void syn_copy_edges(void* h, int32_t* left, int32_t* right, int32_t* is_repeat)
{
	Dataset* d = static_cast<Dataset*>(h);
	size_t n = d->edgeSeqs.size();
	std::memcpy(left, d->edgeLeft.data(), n * sizeof(int32_t));
	std::memcpy(right, d->edgeRight.data(), n * sizeof(int32_t));
	std::memcpy(is_repeat, d->edgeIsRepeat.data(), n * sizeof(int32_t));
}
int64_t syn_hits_size(void* h) { return static_cast<int64_t>(static_cast<Dataset*>(h)->cur.size()); }
void syn_copy_lists(void* h, int32_t* cur, int32_t* ext, int64_t* offs, int64_t* cur_id,
					int64_t* ext_id, int32_t* cur_len, int32_t* ext_len)
{
	Dataset* d = static_cast<Dataset*>(h);
	size_t n = d->curLen.size();
	std::memcpy(cur, d->cur.data(), d->cur.size() * sizeof(int32_t));
	std::memcpy(ext, d->ext.data(), d->ext.size() * sizeof(int32_t));
	std::memcpy(offs, d->listOffsets.data(), (n + 1) * sizeof(int64_t));
	std::memcpy(cur_id, d->curId.data(), n * sizeof(int64_t));
	std::memcpy(ext_id, d->extId.data(), n * sizeof(int64_t));
	std::memcpy(cur_len, d->curLen.data(), n * sizeof(int32_t));
	std::memcpy(ext_len, d->extLen.data(), n * sizeof(int32_t));
}

// Read-set workload: a random genome of genome_len bases at GC fraction gc,
// repeat families per overwriteRepeats (rep_frac of the genome) or, when
// insert_copies > 0, that many exact copies of one insert_len unit inserted
// (the genome grows); then reads to `coverage` with log-normal lengths
// exp(N(mu, sigma)) kept in [min_len, max_len] and the error model above.
void* syn_read_set(int64_t genome_len, double gc, int32_t insert_copies, int32_t insert_len,
				   double rep_frac, int32_t rep_len_min, int32_t rep_len_max, int32_t copies_min,
				   int32_t copies_max, double div_min, double div_max, double coverage,
				   double mu, double sigma, int32_t min_len, int32_t max_len, double err,
				   double w_ins, double w_del, double w_sub, double hp_boost, uint64_t seed)
{
	Dataset* d = new Dataset();
	Rng r(seed);
	d->genome.resize(static_cast<size_t>(genome_len));
	for (auto& b : d->genome) b = randomBase(r, gc);
	if (insert_copies > 0)
	{
		std::string unit(static_cast<size_t>(insert_len), 'A');
		for (auto& b : unit) b = randomBase(r, gc);
		std::vector<int64_t> at;
		for (int i = 0; i < insert_copies; ++i)
			at.push_back(static_cast<int64_t>(r.below(static_cast<uint64_t>(genome_len))));
		std::sort(at.begin(), at.end());
		std::string g;
		int64_t prev = 0;
		for (int64_t p : at)
		{
			g.append(d->genome, static_cast<size_t>(prev), static_cast<size_t>(p - prev));
			g += unit;
			prev = p;
		}
		g.append(d->genome, static_cast<size_t>(prev), std::string::npos);
		d->genome.swap(g);
	}
	if (rep_frac > 0)
	{
		overwriteRepeats(r, d->genome, rep_frac, rep_len_min, rep_len_max, copies_min, copies_max,
						 div_min, div_max, gc, nullptr);
	}
	sampleReads(r, *d, coverage, mu, sigma, min_len, max_len, err, w_ins, w_del, w_sub, hp_boost);
	return d;
}

// Reads sampled from an existing genome (serving batches over a resident
// index): n_reads reads, same length and error model as syn_read_set.
void* syn_reads_from(const uint8_t* genome, int64_t genome_len, int32_t n_reads, double mu,
					 double sigma, int32_t min_len, int32_t max_len, double err, double w_ins,
					 double w_del, double w_sub, double hp_boost, uint64_t seed)
{
	Dataset* d = new Dataset();
	Rng r(seed);
	d->genome.assign(reinterpret_cast<const char*>(genome), static_cast<size_t>(genome_len));
	// coverage chosen so the loop stops after n_reads reads
	const int64_t glen = genome_len;
	for (int32_t i = 0; i < n_reads; ++i)
	{
		int len = 0;
		for (int tries = 0; tries < 1000; ++tries)
		{
			double v = std::exp(mu + sigma * r.normal());
			if (v >= min_len && v <= max_len)
			{
				len = static_cast<int>(v);
				break;
			}
		}
		if (len <= 0) len = min_len;
		if (len > glen) len = static_cast<int>(glen);
		int64_t s = static_cast<int64_t>(r.below(static_cast<uint64_t>(glen - len + 1)));
		std::string t = d->genome.substr(static_cast<size_t>(s), static_cast<size_t>(len));
		int strand = (r.next() & 1) ? 1 : 0;
		if (!strand) t = revComp(t);
		d->reads.push_back(err > 0 ? addErrors(r, t, err, w_ins, w_del, w_sub, hp_boost) : t);
		d->readStart.push_back(s);
		d->readStrand.push_back(strand);
	}
	std::string().swap(d->genome);
	return d;
}

// Chaining-DP stress lists: n_lists lists with Pareto(alpha, xm) sizes capped
// at max_hits.  70% of a list's hits walk one diagonal (cur gaps uniform in
// [1, 2*gap_mean], ext = cur + offset + a random walk of sd drift_sd per
// sqrt(bp)), 30% are uniform spurious hits.  Lists are returned in Flye's
// group order: curPos ascending, equal curPos in generation order.
void* syn_hit_lists(int32_t n_lists, double alpha, int32_t xm, int32_t max_hits, int32_t gap_mean,
					double drift_sd, double spurious, uint64_t seed)
{
	Dataset* d = new Dataset();
	Rng r(seed);
	d->listOffsets.push_back(0);
	std::vector<std::pair<int32_t, int32_t>> hits;
	for (int32_t l = 0; l < n_lists; ++l)
	{
		double u = r.uniform();
		if (u < 1e-12) u = 1e-12;
		double sz = static_cast<double>(xm) / std::pow(u, 1.0 / alpha);
		int32_t n = static_cast<int32_t>(std::min<double>(sz, max_hits));
		if (n < xm) n = xm;
		int32_t nTrue = static_cast<int32_t>(std::lround(n * (1.0 - spurious)));
		hits.clear();
		int64_t span = static_cast<int64_t>(nTrue) * gap_mean + 2000;
		int32_t curLen = static_cast<int32_t>(std::min<int64_t>(span + static_cast<int64_t>(r.below(4000)), 2000000000));
		int32_t extLen = static_cast<int32_t>(std::min<int64_t>(span + static_cast<int64_t>(r.below(4000)), 2000000000));
		if (r.next() & 1) extLen = curLen + 1 + static_cast<int32_t>(r.below(3000));
		int32_t c = static_cast<int32_t>(r.below(1000));
		int32_t off = static_cast<int32_t>(r.below(1000));
		double drift = 0.0;
		for (int32_t i = 0; i < nTrue; ++i)
		{
			int32_t gap = 1 + static_cast<int32_t>(r.below(static_cast<uint64_t>(2 * gap_mean)));
			c += gap;
			drift += drift_sd * std::sqrt(static_cast<double>(gap)) * r.normal();
			int32_t e = c + off + static_cast<int32_t>(std::lround(drift));
			if (c >= curLen - 20 || e < 0 || e >= extLen - 20) continue;
			hits.emplace_back(c, e);
		}
		for (int32_t i = nTrue; i < n; ++i)
		{
			int32_t cc = static_cast<int32_t>(r.below(static_cast<uint64_t>(curLen - 20)));
			int32_t ee = static_cast<int32_t>(r.below(static_cast<uint64_t>(extLen - 20)));
			hits.emplace_back(cc, ee);
		}
		std::stable_sort(hits.begin(), hits.end(),
						 [](const std::pair<int32_t, int32_t>& a, const std::pair<int32_t, int32_t>& b) {
							 return a.first < b.first;
						 });
		for (auto& h : hits)
		{
			d->cur.push_back(h.first);
			d->ext.push_back(h.second);
		}
		d->listOffsets.push_back(static_cast<int64_t>(d->cur.size()));
		d->curLen.push_back(curLen);
		d->extLen.push_back(extLen);
		int64_t cid = 1 + 2 * static_cast<int64_t>(l);
		d->curId.push_back(cid);
		d->extId.push_back((r.below(50) == 0) ? cid : -(cid + 1));
	}
	return d;
}

// Repeat graph over a random genome: repeat families are placed as in
// syn_read_set; the genome is cut at every copy boundary.  All copies of one
// family collapse into one repeat edge carrying the family unit.  Each unique
// stretch between copies is one edge, or (seg_alpha > 0) a path of edges
// whose lengths are drawn from a Pareto law (minimum seg_min, shape
// seg_alpha, at most seg_max; a remainder below seg_min joins the last
// piece), standing for the junctions of the many short repeats and bubbles
// a real repeat graph has.  Node ids: the junction before and after each copy
// of family f are nodes 2f and 2f+1 (shared by all copies); every other
// junction gets a node of its own.  Edge sequences are the forward strands;
// edge i has signed id i + 1 and its complement -(i + 1).
void* syn_repeat_graph(int64_t genome_len, double gc, double rep_frac, int32_t rep_len_min,
					   int32_t rep_len_max, int32_t copies_min, int32_t copies_max, double div_min,
					   double div_max, double coverage, double mu, double sigma, int32_t min_len,
					   int32_t max_len, double err, double w_ins, double w_del, double w_sub,
					   double hp_boost, int32_t seg_min, int32_t seg_max, double seg_alpha, uint64_t seed)
{
	Dataset* d = new Dataset();
	Rng r(seed);
	d->genome.resize(static_cast<size_t>(genome_len));
	for (auto& b : d->genome) b = randomBase(r, gc);
	std::vector<std::string> units;
	// copies are placed forward-only here so a family's unit is its copies' consensus
	std::vector<Placed> placed;
	{
		int64_t target = static_cast<int64_t>(rep_frac * static_cast<double>(genome_len));
		int64_t have = 0;
		int32_t family = 0;
		while (have < target)
		{
			int len = rep_len_min + static_cast<int>(r.below(static_cast<uint64_t>(rep_len_max - rep_len_min + 1)));
			int copies = copies_min + static_cast<int>(r.below(static_cast<uint64_t>(copies_max - copies_min + 1)));
			std::string unit(static_cast<size_t>(len), 'A');
			for (auto& b : unit) b = randomBase(r, gc);
			units.push_back(unit);
			for (int c = 0; c < copies; ++c)
			{
				for (int attempt = 0; attempt < 64; ++attempt)
				{
					int64_t s = 1000 + static_cast<int64_t>(r.below(static_cast<uint64_t>(genome_len - len - 2000)));
					bool clash = false;
					for (auto& p : placed)
						if (s < p.start + p.len + 1000 && p.start < s + len + 1000) { clash = true; break; }
					if (clash) continue;
					double dv = div_min + (div_max - div_min) * r.uniform();
					d->genome.replace(static_cast<size_t>(s), static_cast<size_t>(len), mutateCopy(r, unit, dv, gc));
					placed.push_back({s, len, family});
					have += len;
					break;
				}
			}
			++family;
		}
		std::sort(placed.begin(), placed.end(), [](const Placed& a, const Placed& b) { return a.start < b.start; });
	}
	int32_t nextNode = 2 * static_cast<int32_t>(units.size());
	int32_t prevNode = nextNode++;
	// one unique stretch [a, b) from node `from` to node `to`, cut into a path
	auto addStretch = [&](int64_t a, int64_t b, int32_t from, int32_t to) {
		std::vector<int64_t> cuts{a};
		if (seg_alpha > 0.0)
		{
			int64_t x = a;
			for (;;)
			{
				double u = r.uniform();
				if (u < 1e-12) u = 1e-12;
				double len = static_cast<double>(seg_min) * std::pow(u, -1.0 / seg_alpha);
				if (len > static_cast<double>(seg_max)) len = static_cast<double>(seg_max);
				const int64_t next = x + static_cast<int64_t>(len);
				if (next + seg_min > b) break;
				cuts.push_back(next);
				x = next;
			}
		}
		cuts.push_back(b);
		for (size_t i = 0; i + 1 < cuts.size(); ++i)
		{
			d->edgeSeqs.push_back(d->genome.substr(static_cast<size_t>(cuts[i]), static_cast<size_t>(cuts[i + 1] - cuts[i])));
			d->edgeLeft.push_back(i == 0 ? from : nextNode - 1);
			const bool lastPiece = i + 2 == cuts.size();
			d->edgeRight.push_back(lastPiece ? to : nextNode++);
			d->edgeIsRepeat.push_back(0);
		}
	};
	int64_t cursor = 0;
	for (auto& p : placed)
	{
		int32_t inNode = 2 * p.family;
		int32_t outNode = 2 * p.family + 1;
		if (p.start > cursor) addStretch(cursor, p.start, prevNode, inNode);
		prevNode = outNode;
		cursor = p.start + p.len;
	}
	if (cursor < genome_len)
	{
		const int32_t endNode = nextNode++;
		addStretch(cursor, genome_len, prevNode, endNode);
	}
	for (size_t f = 0; f < units.size(); ++f)
	{
		d->edgeSeqs.push_back(units[f]);
		d->edgeLeft.push_back(static_cast<int32_t>(2 * f));
		d->edgeRight.push_back(static_cast<int32_t>(2 * f + 1));
		d->edgeIsRepeat.push_back(1);
	}
	sampleReads(r, *d, coverage, mu, sigma, min_len, max_len, err, w_ins, w_del, w_sub, hp_boost);
	return d;
}

} // extern "C"
