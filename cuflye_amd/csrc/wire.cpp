// wire.cpp - streaming canonical SHA-256 of candidate-record-v1 (see wire.hpp).
#include "wire.hpp"

#include <openssl/evp.h>

#include <algorithm>
#include <thread>
#include <vector>

namespace fh
{

struct CanonSha256::Impl
{
	EVP_MD_CTX* md = nullptr;
};

namespace
{

// Decimal digits of v into p (no terminator); returns the end.
char* putU64(char* p, uint64_t v)
{
	char tmp[20];
	int n = 0;
	do
	{
		tmp[n++] = static_cast<char>('0' + v % 10);
		v /= 10;
	} while (v);
	while (n) *p++ = tmp[--n];
	return p;
}

char* putI64(char* p, int64_t v)
{
	if (v < 0)
	{
		*p++ = '-';
		return putU64(p, static_cast<uint64_t>(0) - static_cast<uint64_t>(v));
	}
	return putU64(p, static_cast<uint64_t>(v));
}

// Sorts each window's run by (target_id, target_pos) and formats the records.
void formatQueries(flyehip_candidate_t* recs, const int64_t* off, int32_t q0, int32_t q1, std::string& out)
{
	out.clear();
	out.reserve(static_cast<size_t>(off[q1] - off[q0]) * 40);
	char line[128];
	for (int32_t q = q0; q < q1; ++q)
	{
		flyehip_candidate_t* b = recs + off[q];
		flyehip_candidate_t* e = recs + off[q + 1];
		for (flyehip_candidate_t* w = b; w < e;)
		{
			flyehip_candidate_t* we = w + 1;
			while (we < e && we->query_pos == w->query_pos) ++we;
			if (we - w > 1)
				std::sort(w, we, [](const flyehip_candidate_t& x, const flyehip_candidate_t& y) {
					return x.target_id != y.target_id ? x.target_id < y.target_id : x.target_pos < y.target_pos;
				});
			for (const flyehip_candidate_t* r = w; r < we; ++r)
			{
				char* p = line;
				p = putI64(p, r->query_id);
				*p++ = '\t';
				p = putI64(p, r->query_pos);
				*p++ = '\t';
				p = putU64(p, r->kmer);
				*p++ = '\t';
				p = putI64(p, r->target_id);
				*p++ = '\t';
				p = putI64(p, r->target_pos);
				*p++ = '\t';
				*p++ = r->target_id > 0 ? '+' : '-';
				*p++ = '\n';
				out.append(line, static_cast<size_t>(p - line));
			}
			w = we;
		}
	}
}

} // namespace

CanonSha256::CanonSha256() : impl_(new Impl())
{
	impl_->md = EVP_MD_CTX_new();
	if (!impl_->md || EVP_DigestInit_ex(impl_->md, EVP_sha256(), nullptr) != 1) err_ = "SHA-256 init failed";
}

CanonSha256::~CanonSha256()
{
	if (impl_->md) EVP_MD_CTX_free(impl_->md);
}

void CanonSha256::addQueries(flyehip_candidate_t* recs, const int64_t* off, int32_t nQueries, int threads)
{
	if (!err_.empty() || nQueries <= 0) return;
	threads = std::max(1, std::min(threads, nQueries));
	// contiguous query ranges of about equal record counts, one per thread
	std::vector<int32_t> cut(static_cast<size_t>(threads) + 1, nQueries);
	cut[0] = 0;
	const int64_t total = off[nQueries] - off[0];
	for (int t = 1; t < threads; ++t)
	{
		const int64_t target = off[0] + total * t / threads;
		cut[static_cast<size_t>(t)] = static_cast<int32_t>(std::lower_bound(off, off + nQueries, target) - off);
	}
	for (int t = 1; t <= threads; ++t)
		cut[static_cast<size_t>(t)] = std::max(cut[static_cast<size_t>(t)], cut[static_cast<size_t>(t) - 1]);
	std::vector<std::string> text(static_cast<size_t>(threads));
	std::vector<std::thread> th;
	for (int t = 1; t < threads; ++t)
		th.emplace_back(formatQueries, recs, off, cut[static_cast<size_t>(t)], cut[static_cast<size_t>(t) + 1],
						std::ref(text[static_cast<size_t>(t)]));
	formatQueries(recs, off, cut[0], cut[1], text[0]);
	for (auto& x : th) x.join();
	for (const std::string& s : text)
		if (!s.empty() && EVP_DigestUpdate(impl_->md, s.data(), s.size()) != 1) err_ = "SHA-256 update failed";
	records_ += total;
}

bool CanonSha256::finish(uint8_t* out)
{
	if (!err_.empty()) return false;
	unsigned int len = 0;
	if (EVP_DigestFinal_ex(impl_->md, out, &len) != 1 || len != 32)
	{
		err_ = "SHA-256 final failed";
		return false;
	}
	return true;
}

} // namespace fh
