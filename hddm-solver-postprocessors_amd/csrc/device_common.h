// Device helpers shared by the forward and adjoint kernels (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "layout.h"

namespace hddm_lb {

// Loads through the constant address space: a wave-uniform address becomes a
// scalar (SMEM) load, which keeps node records and surplus rows out of VGPRs.
template <typename T>
__device__ __forceinline__ T ldc(const T* p)
{
	return *(const __attribute__((address_space(4))) T*)p;
}

// A 16-byte op through one scalar load.
__device__ __forceinline__ Op ldc_op(const Op* p)
{
	typedef uint32_t u4v __attribute__((ext_vector_type(4)));
	const u4v v = *(const __attribute__((address_space(4))) u4v*)p;
	Op op;
	op.dim = v[0]; op.ctrl = v[1]; op.ihi = v[2]; op.njhi = v[3];
	return op;
}

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ uint64_t uniform64(uint64_t v)
{
	uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
	uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
	return ((uint64_t)hi << 32) | lo;
}

// True if any lane of the wave has pred set.
__device__ __forceinline__ bool wave_any(bool pred)
{
	return __builtin_amdgcn_ballot_w64(pred) != 0;
}

// Top `bits` bits of the interleaved dyadic expansion of x: bit order
// (level 1, dim 0), (level 1, dim 1), ..., (level 2, dim 0), ...
// Bit (l, d) is the l-th binary digit of x_d, i.e. which half, quarter, ...
// of [0, 1] holds x_d. Points adjacent in this order share supports.
// Each coordinate is read once: with L the deepest level used, the digit of
// level l is bit L - l of floor(x_d 2^L) (scaling by 2^L is exact, and
// floor(floor(y) / 2^k) = floor(y / 2^k) for y >= 0).
__device__ __forceinline__ uint64_t morton_key(const double* x, int dim, int bits)
{
	if (bits <= 0 || dim <= 0) return 0;
	int L = (bits + dim - 1) / dim;
	if (L > 30) L = 30;
	const int nd = dim < bits ? dim : bits;
	const double scale = (double)(1u << L);
	uint64_t key = 0;
	for (int d = 0; d < nd; d++)
	{
		double v = x[d];
		v = v < 0.0 ? 0.0 : (v >= 1.0 ? 0.9999999999999999 : v);
		const uint32_t q = (uint32_t)(v * scale);
		// bit (lev, d) sits at position bits - 1 - ((lev - 1) dim + d)
		for (int lev = 1, pos = d; lev <= L && pos < bits; lev++, pos += dim)
			key |= (uint64_t)((q >> (L - lev)) & 1u) << (bits - 1 - pos);
	}
	return key;
}

} // namespace hddm_lb
