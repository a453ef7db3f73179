// ktimer.hpp - GPU time of kernel families from HIP events recorded on the
// stream each launch runs on (host code).  A lane records begin / end pairs
// while it enqueues a batch and sums them once the batch has synchronised.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/flyehip.h"

namespace fh
{

struct KTimer
{
	static constexpr int kMax = 2048;
	hipEvent_t ev[2 * kMax] = {};
	int family[kMax] = {};
	double bytes[kMax] = {};
	int n = 0;

	hipError_t create()
	{
		for (auto& e : ev)
		{
			hipError_t r = hipEventCreate(&e);
			if (r != hipSuccess) return r;
		}
		return hipSuccess;
	}
	void destroy()
	{
		for (auto& e : ev)
			if (e) (void)hipEventDestroy(e);
	}
	// Starts an interval of `fam` on stream s; returns its slot (-1 when full).
	int begin(int fam, hipStream_t s, double algBytes = 0.0)
	{
		if (n >= kMax || hipEventRecord(ev[2 * n], s) != hipSuccess) return -1;
		family[n] = fam;
		bytes[n] = algBytes;
		return n++;
	}
	void addBytes(int slot, double moreBytes)
	{
		if (slot >= 0) bytes[slot] += moreBytes;
	}
	void end(int slot, hipStream_t s, double moreBytes = 0.0)
	{
		if (slot < 0) return;
		bytes[slot] += moreBytes;
		if (hipEventRecord(ev[2 * slot + 1], s) != hipSuccess) family[slot] = -1;
	}
	// After every stream the intervals ran on has synchronised.
	void collect(flyehip_stats_t& st)
	{
		for (int i = 0; i < n; ++i)
		{
			if (family[i] < 0) continue;
			float ms = 0.0f;
			if (hipEventElapsedTime(&ms, ev[2 * i], ev[2 * i + 1]) != hipSuccess) continue;
			st.kernel_ms[family[i]] += ms;
			st.kernel_launches[family[i]] += 1;
			st.kernel_bytes[family[i]] += bytes[i];
		}
		n = 0;
	}
};

} // namespace fh
