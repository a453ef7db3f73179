// Checks the std::sort emulation of cuflye_amd/csrc/exsort.hpp against the
// host libstdc++ std::sort: same permutation (payload order), on random keys
// with many ties, sorted / reversed / organ-pipe inputs, and inputs built by
// McIlroy's "killer adversary" (antiqsort) that drive introsort into its
// heapsort fallback.  Also checks the rank-formulated partition against the
// scanning partition.  Exit code 0 = all equal.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <random>
#include <vector>

#include "../../cuflye_amd/csrc/exsort.hpp"

static int failures = 0;

static std::vector<uint32_t> refPerm(const std::vector<uint64_t>& keys)
{
	std::vector<uint32_t> idx(keys.size());
	for (size_t i = 0; i < idx.size(); ++i) idx[i] = static_cast<uint32_t>(i);
	std::sort(idx.begin(), idx.end(), [&](uint32_t a, uint32_t b) { return keys[a] < keys[b]; });
	return idx;
}

static void check32(const std::vector<uint64_t>& keys, const char* what)
{
	std::vector<uint32_t> k(keys.size());
	for (size_t i = 0; i < k.size(); ++i) k[i] = static_cast<uint32_t>(keys[i]);
	std::vector<uint32_t> v(keys.size());
	for (size_t i = 0; i < v.size(); ++i) v[i] = static_cast<uint32_t>(i);
	exsort::stdSort(exsort::Span<uint32_t>{k.data(), v.data()}, 0, static_cast<int>(k.size()));
	std::vector<uint64_t> trunc(keys.size());
	for (size_t i = 0; i < k.size(); ++i) trunc[i] = static_cast<uint32_t>(keys[i]);
	if (v != refPerm(trunc))
	{
		if (failures < 10) std::printf("MISMATCH32 %s n=%zu\n", what, keys.size());
		++failures;
	}
}

static void check(const std::vector<uint64_t>& keys, const char* what)
{
	check32(keys, what);
	std::vector<uint64_t> k = keys;
	std::vector<uint32_t> v(keys.size());
	for (size_t i = 0; i < v.size(); ++i) v[i] = static_cast<uint32_t>(i);
	exsort::stdSort(exsort::Span<uint64_t>{k.data(), v.data()}, 0, static_cast<int>(k.size()));
	if (v != refPerm(keys))
	{
		if (failures < 10) std::printf("MISMATCH %s n=%zu\n", what, keys.size());
		++failures;
	}
}

// Packed heap sort vs libstdc++'s partial_sort(first, last, last) (make_heap +
// sort_heap) on 32-bit keys with many ties.
static void checkHeapPacked(const std::vector<uint64_t>& keys)
{
	const size_t n = keys.size();
	std::vector<uint64_t> a(n);
	for (size_t i = 0; i < n; ++i) a[i] = (static_cast<uint64_t>(static_cast<uint32_t>(keys[i])) << 32) | i;
	std::vector<uint64_t> b = a;
	exsort::heapSortPacked(a.data(), static_cast<int>(n));
	exsort::heapSortPipelinedModel(b.data(), static_cast<int>(n));
	if (a != b)
	{
		if (failures < 10) std::printf("MISMATCH pipelined heap n=%zu\n", n);
		++failures;
	}
	std::vector<uint32_t> idx(n);
	for (size_t i = 0; i < n; ++i) idx[i] = static_cast<uint32_t>(i);
	std::partial_sort(idx.begin(), idx.end(), idx.end(), [&](uint32_t x, uint32_t y) {
		return static_cast<uint32_t>(keys[x]) < static_cast<uint32_t>(keys[y]);
	});
	for (size_t i = 0; i < n; ++i)
		if (static_cast<uint32_t>(a[i]) != idx[i])
		{
			if (failures < 10) std::printf("MISMATCH heapPacked n=%zu at %zu\n", n, i);
			++failures;
			return;
		}
}

// Rank partition vs scanning partition on the same median-moved array.
static void checkPartition(const std::vector<uint64_t>& keys)
{
	const int n = static_cast<int>(keys.size());
	if (n <= 16) return;
	std::vector<uint64_t> k1 = keys, k2 = keys;
	std::vector<uint32_t> v1(n), v2(n);
	for (int i = 0; i < n; ++i) v1[i] = v2[i] = static_cast<uint32_t>(i);
	exsort::Span<uint64_t> s1{k1.data(), v1.data()}, s2{k2.data(), v2.data()};
	int c1 = exsort::partitionPivot(s1, 0, n);
	exsort::medianToFirst(s2, 0, 1, n / 2, n - 1);
	std::vector<int> pl(n), pr(n);
	int c2 = exsort::partitionByRanks(s2, 0, n, pl.data(), pr.data());
	if (c1 != c2 || v1 != v2)
	{
		if (failures < 10) std::printf("PARTITION MISMATCH n=%d cut %d vs %d\n", n, c1, c2);
		++failures;
	}
}

// McIlroy, "A Killer Adversary for Quicksort" (1999).
struct Adversary
{
	std::vector<int> val;
	int gas, nsolid = 0, candidate = 0;
	explicit Adversary(int n) : val(n), gas(n) { std::fill(val.begin(), val.end(), gas); }
	bool less(int x, int y)
	{
		if (val[x] == gas && val[y] == gas)
		{
			if (x == candidate) val[x] = nsolid++;
			else val[y] = nsolid++;
		}
		if (val[x] == gas) candidate = x;
		else if (val[y] == gas) candidate = y;
		return val[x] < val[y];
	}
};

static std::vector<uint64_t> killer(int n)
{
	Adversary adv(n);
	std::vector<int> idx(n);
	for (int i = 0; i < n; ++i) idx[i] = i;
	std::sort(idx.begin(), idx.end(), [&](int a, int b) { return adv.less(a, b); });
	std::vector<uint64_t> keys(n);
	for (int i = 0; i < n; ++i) keys[i] = static_cast<uint64_t>(adv.val[i] == adv.gas ? n : adv.val[i]);
	return keys;
}

int main()
{
	std::mt19937_64 rng(12345);
	int cases = 0;
	for (int n : {0, 1, 2, 3, 15, 16, 17, 18, 31, 32, 33, 64, 100, 257, 1000, 4096, 10007, 65536})
	{
		for (int distinct : {1, 2, 3, 7, 50, 1000000})
		{
			for (int rep = 0; rep < 3; ++rep)
			{
				std::vector<uint64_t> keys(n);
				for (auto& x : keys) x = rng() % static_cast<uint64_t>(distinct);
				check(keys, "random");
				checkPartition(keys);
				checkHeapPacked(keys);
				++cases;
			}
		}
		std::vector<uint64_t> asc(n), desc(n), pipe(n);
		for (int i = 0; i < n; ++i)
		{
			asc[i] = i / 3;
			desc[i] = (n - i) / 2;
			pipe[i] = std::min(i, n - 1 - i);
		}
		check(asc, "ascending");
		check(desc, "descending");
		check(pipe, "organ-pipe");
		cases += 3;
	}
	for (int n : {100, 1000, 5000, 20000})
	{
		check(killer(n), "killer");
		++cases;
	}
	// structured like the match sort: curPos-major input sorted by (extId, curPos)
	for (int rep = 0; rep < 20; ++rep)
	{
		const int n = 2000 + static_cast<int>(rng() % 30000);
		std::vector<uint64_t> keys(n);
		int pos = 0;
		for (int i = 0; i < n; ++i)
		{
			if (rng() % 4 == 0) pos += 1 + static_cast<int>(rng() % 5);
			uint64_t ext = rng() % 40;
			keys[i] = (ext << 32) | static_cast<uint32_t>(pos);
		}
		check(keys, "match-like");
		checkPartition(keys);
		++cases;
	}
	std::printf("exsort cases=%d failures=%d\n", cases, failures);
	return failures ? 1 : 0;
}
