// Device pieces shared by the forward and adjoint sweeps (gfx950 only):
// work-item decoding, coordinate banks and the LDS tiles of the op stream.
#pragma once

#include "device_common.h"

namespace hddm_lb {

typedef uint32_t u4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ double f64_from_high(uint32_t hi)
{
	return __builtin_bit_cast(double, (uint64_t)hi << 32);
}

// *p += v on an LDS row owned by the calling wave, as an LDS atomic add
// whose result is not returned (no wait for it): a wave's LDS operations
// complete in issue order, so later reads see the sum and the additions
// happen in program order, as with a read-modify-write.
__device__ __forceinline__ void lds_add(double* p, double v)
{
	(void)__hip_atomic_fetch_add((__attribute__((address_space(3))) double*)p, v, __ATOMIC_RELAXED,
		__HIP_MEMORY_SCOPE_WORKGROUP);
}

// One work item: G neighbouring 64-point blocks of one state (one block per
// wave of the workgroup) against one node segment of that state.
struct ItemInfo
{
	int s, seg;
	int64_t grp, start, count;
};

// Decodes a work-item number; items are laid out state by state in cost
// order, each state's items segment-major over its groups of G blocks.
// A is any argument block with states, plans, cost_order and nstates.
template <typename A>
__device__ __forceinline__ void decode_item(const A& a, int64_t item, int G, ItemInfo& it)
{
	int s = 0;
	int64_t base = 0, ngrp = 1;
	for (int i = 0; i < a.nstates; i++)
	{
		const int si = ldc(&a.cost_order[i]);
		const int64_t ib = ldc(&a.plans[si].item_base);
		const int64_t nb = ldc(&a.plans[si].nblk);
		const int64_t ng = (nb + G - 1) / G;
		const int ns = ldc(&a.states[si].nseg);
		if (item < ib + ng * ns) { s = si; base = ib; ngrp = ng; break; }
	}
	const int64_t local = item - base;
	it.s = s;
	it.seg = (int)(local / ngrp);
	it.grp = local - (int64_t)it.seg * ngrp;
	it.start = ldc(&a.plans[s].start);
	it.count = ldc(&a.plans[s].count);
}

// Coordinates of the lane's point are held in XB banks of 16 doubles
// (separate arrays), indexed with a wave-uniform dim (s_set_gpr_idx).
// SELECT chooses how more than one bank is read (see below): measured on a
// D=20 workload, selecting values is faster where registers are left over
// (FP64 paths), a selected pointer where they are not (MFMA paths).
template <int XB, bool SELECT>
__device__ __forceinline__ double xget(const double (&b0)[16], const double (&b1)[16], const double (&b2)[16],
	const double (&b3)[16], uint32_t dim)
{
	const uint32_t idx = dim & 15u;
	if constexpr (XB == 1) return b0[idx];
	else if constexpr (SELECT)
	{
		// every bank is read with the same index and the value selected: no
		// pointer into the banks is formed, so they stay in registers (more
		// VGPRs, no scratch)
		const uint32_t bank = dim >> 4;
		const double v0 = b0[idx], v1 = b1[idx];
		if constexpr (XB == 2) return bank == 0u ? v0 : v1;
		else
		{
			const double v2 = b2[idx], v3 = b3[idx];
			const double lo = bank == 0u ? v0 : v1, hi = bank == 2u ? v2 : v3;
			return bank < 2u ? lo : hi;
		}
	}
	else
	{
		// one bank read through a selected pointer: the compiler keeps the
		// banks beyond the first in scratch, which costs fewer VGPRs
		const uint32_t bank = dim >> 4;
		if (bank == 0) return b0[idx];
		if constexpr (XB == 2) return b1[idx];
		else
		{
			if (bank == 1) return b1[idx];
			if (bank == 2) return b2[idx];
			return b3[idx];
		}
	}
}

// Ops per LDS tile of the op stream; the workgroup double-buffers tiles.
constexpr int kTileOps = 512;

// Workgroup LDS: [item slot 16 B][ops 2 x kTileOps x 16 B][skip 2 x kTileOps x 4 B]
// then per wave the path stack [PM][64] doubles and the chunk buffer [16][64] BT.
constexpr size_t kLdsHeader = 16;
constexpr size_t kLdsTiles = 2 * (size_t)kTileOps * (sizeof(Op) + sizeof(uint32_t));

// Issues the copy of ops [tb, tb + kTileOps) and their skips into one tile
// buffer, spread over the G waves of the workgroup (one 64-op row per wave
// and instruction). Rows past the end of the stream read its sentinel.
__device__ __forceinline__ void load_tile(const Op* gops, const uint32_t* gskip, int nops, int tb,
	Op* tops, uint32_t* tskip, int wave, int G, int lane)
{
	for (int r = wave; r < kTileOps / kWave; r += G)
	{
		int q = tb + r * kWave + lane;
		q = q < nops ? q : nops;
		__builtin_amdgcn_global_load_lds(gops + q, tops + r * kWave, 16, 0, 0);
		__builtin_amdgcn_global_load_lds(gskip + q, tskip + r * kWave, 4, 0, 0);
	}
}

} // namespace hddm_lb
