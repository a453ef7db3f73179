// gen_logf_table.cpp - build-time generator of the glibc-logf agreement table.
//
// Flye computes seq_divergence = logf(1 / matchRate) / k in float with glibc's
// logf.  The device evaluates log in double and rounds to float, which equals
// the correctly rounded logf; glibc's logf is not correctly rounded (its
// documented bound is about 0.82 ulp), so the two differ on a small set of
// arguments.  This program calls the host's glibc logf on every float in
// [1, 2^25) - every value 1 / matchRate can take for records shorter than
// 32 Mb - and writes the (argument bits, glibc result bits) pairs where the
// rounded double log disagrees, plus the arguments whose double log lies
// within 16 double ulps of a float rounding boundary (so a last-bit difference
// between two double log implementations cannot change the float result).
// Output: little-endian uint32 pairs sorted by argument, binary.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

int main(int argc, char** argv)
{
	if (argc != 2)
	{
		std::fprintf(stderr, "usage: gen_logf_table OUT\n");
		return 2;
	}
	const float lo = 1.0f, hi = 33554432.0f;
	uint32_t ulo, uhi;
	std::memcpy(&ulo, &lo, 4);
	std::memcpy(&uhi, &hi, 4);
	std::vector<uint32_t> out;
	for (uint32_t u = ulo; u < uhi; ++u)
	{
		float x;
		std::memcpy(&x, &u, 4);
		const float g = logf(x);
		const double d = std::log(static_cast<double>(x));
		const float r = static_cast<float>(d);
		bool keep = std::memcmp(&g, &r, 4) != 0;
		if (!keep)
		{
			const double rd = r;
			const float nb = std::nextafter(r, d > rd ? INFINITY : -INFINITY);
			const double mid = (rd + static_cast<double>(nb)) / 2.0;
			const double ulp = std::nextafter(d, INFINITY) - d;
			keep = std::fabs(d - mid) < 16.0 * ulp;
		}
		if (keep)
		{
			uint32_t gb;
			std::memcpy(&gb, &g, 4);
			out.push_back(u);
			out.push_back(gb);
		}
	}
	FILE* f = std::fopen(argv[1], "wb");
	if (!f) return 1;
	std::fwrite(out.data(), 4, out.size(), f);
	std::fclose(f);
	std::printf("logf table: %zu entries\n", out.size() / 2);
	return 0;
}
