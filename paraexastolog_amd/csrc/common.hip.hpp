// Device-side helpers shared by the kernels (gfx950: 64-lane wavefronts).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace esl {

constexpr int wave = 64;

__device__ __forceinline__ int lane_id() { return threadIdx.x & (wave - 1); }

// Inclusive prefix sum across the 64 lanes of a wavefront.
template <typename T>
__device__ __forceinline__ T wave_incl_scan(T v)
{
#pragma unroll
	for (int d = 1; d < wave; d <<= 1)
	{
		T u = __shfl_up(v, d, wave);
		if (lane_id() >= d)
			v += u;
	}
	return v;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v)
{
#pragma unroll
	for (int d = wave / 2; d > 0; d >>= 1)
		v += __shfl_xor(v, d, wave);
	return v;
}

template <typename T>
__device__ __forceinline__ T wave_max(T v)
{
#pragma unroll
	for (int d = wave / 2; d > 0; d >>= 1)
	{
		T u = __shfl_xor(v, d, wave);
		v = u > v ? u : v;
	}
	return v;
}

// Exclusive prefix sum over a workgroup of BLOCK threads; also returns the total.
// `lds` must hold BLOCK / 64 + 1 elements.
template <int BLOCK, typename T>
__device__ __forceinline__ T block_excl_scan(T v, T* lds, T& total)
{
	constexpr int nw = BLOCK / wave;
	const int w = threadIdx.x / wave;
	T inc = wave_incl_scan(v);
	if (lane_id() == wave - 1)
		lds[w] = inc;
	__syncthreads();
	if (threadIdx.x == 0)
	{
		T acc = 0;
		for (int i = 0; i < nw; ++i)
		{
			T t = lds[i];
			lds[i] = acc;
			acc += t;
		}
		lds[nw] = acc;
	}
	__syncthreads();
	T res = lds[w] + inc - v;
	total = lds[nw];
	__syncthreads();
	return res;
}

template <int BLOCK, typename T>
__device__ __forceinline__ T block_sum(T v, T* lds)
{
	constexpr int nw = BLOCK / wave;
	const int w = threadIdx.x / wave;
	T s = wave_sum(v);
	if (lane_id() == 0)
		lds[w] = s;
	__syncthreads();
	T r = 0;
	if (threadIdx.x == 0)
	{
		for (int i = 0; i < nw; ++i)
			r += lds[i];
		lds[nw] = r;
	}
	__syncthreads();
	r = lds[nw];
	__syncthreads();
	return r;
}

inline int ceil_div(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }

} // namespace esl
