// MFMA operand helpers shared by the forward sweeps (gfx950 only): f16 / bf16
// packing and the surplus operands (the MFMA A side) of one node chunk.
#pragma once

#include <type_traits>

#include "hddm_lb.h"
#include "device_common.h"

namespace hddm_lb {

typedef _Float16 h2_t __attribute__((ext_vector_type(2)));
typedef _Float16 h8_t __attribute__((ext_vector_type(8)));
typedef __bf16 b2_t __attribute__((ext_vector_type(2)));
typedef __bf16 b8_t __attribute__((ext_vector_type(8)));
typedef float f16_t __attribute__((ext_vector_type(16)));


__device__ __forceinline__ uint32_t pack_h2(float a, float b)
{
	h2_t v;
	v[0] = (_Float16)a;
	v[1] = (_Float16)b;
	return __builtin_bit_cast(uint32_t, v);
}

__device__ __forceinline__ uint32_t pack_b2(float a, float b)
{
	b2_t v;
	v[0] = (__bf16)a;
	v[1] = (__bf16)b;
	return __builtin_bit_cast(uint32_t, v);
}

__device__ __forceinline__ float lo_half_f16(uint32_t u) { return (float)__builtin_bit_cast(h2_t, u)[0]; }
__device__ __forceinline__ float hi_half_f16(uint32_t u) { return (float)__builtin_bit_cast(h2_t, u)[1]; }
__device__ __forceinline__ float lo_half_bf16(uint32_t u) { return __builtin_bit_cast(float, u << 16); }
__device__ __forceinline__ float hi_half_bf16(uint32_t u) { return __builtin_bit_cast(float, u & 0xffff0000u); }

template <typename V>
__device__ __forceinline__ V as_vec8(uint32_t a, uint32_t b, uint32_t c, uint32_t d)
{
	u4_t u = { a, b, c, d };
	return __builtin_bit_cast(V, u);
}

// S^T operands (the MFMA A side) of one chunk (16 nodes; 32 for F32, two
// consecutive 16-node surplus chunks), loaded ahead of use.
template <int PREC, int DT>
struct SOperands
{
	typedef typename std::conditional<PREC == HDDM_LB_F16X2, h8_t,
		typename std::conditional<PREC == HDDM_LB_BF16X3, b8_t, float>::type>::type V;
	static constexpr int kTerms = PREC == HDDM_LB_F16X2 ? 2 : (PREC == HDDM_LB_BF16X3 ? 3 : 16);
	V A[DT][kTerms];

	__device__ __forceinline__ void load(const StateDesc* sd, int chunk, int DP, int dt0, int ndt, int lane)
	{
		const int row = lane & 31;
		const int khalf = lane >> 5;
#pragma unroll
		for (int dt = 0; dt < DT; dt++)
			if (dt < ndt)
			{
				const int dr = (dt0 + dt) * 32 + row;
				if constexpr (PREC == HDDM_LB_F16X2)
				{
					const uint16_t* sc = ldc(&sd->s16) + (size_t)chunk * 2 * DP * kChunk;
					A[dt][0] = *(const h8_t*)(sc + ((size_t)dr) * kChunk + khalf * 8);
					A[dt][1] = *(const h8_t*)(sc + ((size_t)DP + dr) * kChunk + khalf * 8);
				}
				else if constexpr (PREC == HDDM_LB_BF16X3)
				{
					const uint16_t* sc = ldc(&sd->sbf) + (size_t)chunk * 3 * DP * kChunk;
#pragma unroll
					for (int term = 0; term < 3; term++)
						A[dt][term] = *(const b8_t*)(sc + ((size_t)term * DP + dr) * kChunk + khalf * 8);
				}
				else
				{
					// K-step q (nodes 2q, 2q+1 of the 32) is step q % 8 of
					// surplus chunk 2 chunk + q / 8; past the last chunk: 0
					const int nch = ldc(&sd->nchunks);
#pragma unroll
					for (int h = 0; h < 2; h++)
					{
						const int sc = 2 * chunk + h;
						const float* ap = ldc(&sd->s32) + (size_t)(sc < nch ? sc : 0) * DP * kChunk
							+ ((size_t)dr * 2 + khalf) * 8;
#pragma unroll
						for (int q = 0; q < 8; q++) A[dt][8 * h + q] = sc < nch ? ap[q] : 0.0f;
					}
				}
			}
	}
};

// Split-precision MFMA contraction of one 16-node chunk (F16X2, BF16X3):
// acc[t][dt] (points 32t.., dofs 32(dt0+dt)..) += S^T[dofs x 16 nodes] *
// B^T[16 nodes x 32 points]. The F32 path contracts in the sweep kernel.
// Basis values are scaled by this power of two before the f16 hi/lo split, so
// that the low halves of basis values down to 2^-12 stay normal f16 numbers.
constexpr float kBScale = 4096.0f;

} // namespace hddm_lb
