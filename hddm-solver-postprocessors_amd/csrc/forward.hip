// Forward kernels of the LinearBasis interpolation engine (gfx950).
//
// Math (LinearBasis/gold/src/InterpolateArray.cpp:8-43, LinearBasis/include/LinearBasis.h:51):
//   y[p][c] = sum_k B_k(x_p) S[k][c],  B_k(x) = prod_d max(0, 1 - |x_d i_kd - j_kd|),
//   B_k := 0 as soon as one factor is <= 0.
//
// Pipeline of one forward call (all on the caller's stream, no host sync, so
// the whole sequence can be captured into a hipGraph):
//   1. fill_u32      zero the bucket histogram (a kernel, never a memset)
//   2. bucket_key    per point: bucket = (state, top B bits of its dyadic
//                    interleaved expansion); histogram by atomics
//   3. scan_plan     one workgroup: exclusive scan of the histogram, per-state
//                    point ranges and the work-item plan (items = state x
//                    group of G 64-point blocks x node segment, states in
//                    cost order)
//   4. scatter       counting-sort scatter: perm[sorted position] = point
//   5. sweep         persistent workgroups of G waves pop items; one point per
//                    lane, one block per wave. Each wave reads the segment's
//                    op stream through scalar loads (or the workgroup streams
//                    it through double-buffered LDS tiles), walks the prefix
//                    trie of nodes, pruning every subtree whose prefix product
//                    is zero on all 64 lanes, and contracts the surviving
//                    nodes against the surplus (FP64 VALU on 8-node chunks,
//                    MFMA on 32-node F32 or 16-node split-operand chunks)
//   6. reduce        sums node-segment partials of split states (fixed order)
#include "sweep_common.h"
#include <type_traits>
#include "kernels.h"
#include "walk.h"
#include "contract.h"

namespace hddm_lb {


__global__ void fill_u32_kernel(uint32_t* p, int64_t n, uint32_t v)
{
	for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
		p[e] = v;
}

// Bucket bits never exceed 16, so a key reads at most the first 16
// coordinates of a point.
constexpr int kKeyDims = 16;
constexpr int kKeyBlock = 256;

// One point per thread. The workgroup first stages the coordinates its keys
// read (the first min(dim, bbits) of each of its 256 points) through LDS with
// coalesced loads: read lane by lane, rows of 128 bytes are re-fetched from
// HBM once per coordinate when the many resident waves evict them from L2.
__global__ void __launch_bounds__(kKeyBlock) bucket_key_kernel(int64_t n, const double* __restrict__ x, int64_t ldx,
	const int32_t* __restrict__ sid, const uint8_t* __restrict__ state_ok, int nstates,
	int dim, int bbits, uint32_t* __restrict__ bkey, uint32_t* __restrict__ hist)
{
	__shared__ double rows[kKeyBlock * (kKeyDims + 1)];
	const int tid = threadIdx.x;
	const int nd = bbits < dim ? bbits : dim;
	for (int64_t base = (int64_t)blockIdx.x * kKeyBlock; base < n; base += (int64_t)gridDim.x * kKeyBlock)
	{
		const int npt = n - base < kKeyBlock ? (int)(n - base) : kKeyBlock;
		for (int k = tid; k < npt * nd; k += kKeyBlock)
		{
			const int pt = k / nd, d = k - pt * nd;
			rows[pt * (kKeyDims + 1) + d] = x[(base + pt) * ldx + d];
		}
		__syncthreads();
		if (tid < npt)
		{
			const int64_t p = base + tid;
			int s = sid ? sid[p] : 0;
			if (s < 0 || s >= nstates || !state_ok[s]) s = nstates;  // never planned: row left untouched
			uint32_t key = ((uint32_t)s << bbits);
			if (bbits) key |= (uint32_t)morton_key(rows + tid * (kKeyDims + 1), dim, bbits);
			bkey[p] = key;
			atomicAdd(&hist[key], 1u);
		}
		__syncthreads();
	}
}

// One workgroup of 1024 threads.
__global__ void __launch_bounds__(1024) scan_plan_kernel(const uint32_t* __restrict__ hist,
	uint32_t* __restrict__ cursor, int64_t nbuckets, int bbits, int nstates,
	const StateDesc* __restrict__ states, const int32_t* __restrict__ cost_order,
	StatePlan* __restrict__ plans, PlanHeader* __restrict__ hdr, int group, int dxw)
{
	// exclusive scan of the histogram into the cursors, 1024 buckets per
	// step with coalesced reads: a wave-level scan, then the wave totals
	__shared__ uint32_t wsum[16];
	const int t = threadIdx.x;
	const int lane = t & 63, w = t >> 6;
	uint32_t carry = 0;
	for (int64_t base = 0; base < nbuckets; base += 1024)
	{
		const int64_t b = base + t;
		const uint32_t v = b < nbuckets ? hist[b] : 0u;
		uint32_t x = v;
#pragma unroll
		for (int off = 1; off < 64; off <<= 1)
		{
			const uint32_t y = __shfl_up(x, off, 64);
			if (lane >= off) x += y;
		}
		if (lane == 63) wsum[w] = x;
		__syncthreads();
		uint32_t wpre = 0, total = 0;
#pragma unroll
		for (int i = 0; i < 16; i++)
		{
			const uint32_t ws = wsum[i];
			wpre += i < w ? ws : 0u;
			total += ws;
		}
		if (b < nbuckets) cursor[b] = carry + wpre + x - v;
		carry += total;
		__syncthreads();
	}
	__threadfence_block();
	__syncthreads();

	for (int s = t; s < nstates; s += 1024)
	{
		int64_t lo = (int64_t)cursor[(int64_t)s << bbits];
		int64_t hi = (int64_t)cursor[(int64_t)(s + 1) << bbits];
		plans[s].start = lo;
		plans[s].count = hi - lo;
		plans[s].nblk = (hi - lo + kWave - 1) / kWave;
	}
	__syncthreads();
	if (t == 0)
	{
		int64_t item = 0, pbase = 0, xbase = 0;
		for (int i = 0; i < nstates; i++)
		{
			int s = cost_order[i];
			int nseg = states[s].nseg;
			plans[s].item_base = item;
			item += (plans[s].nblk + group - 1) / group * nseg;
			plans[s].partial_base = pbase;
			plans[s].dx_base = xbase;
			if (nseg > 1)
			{
				pbase += (int64_t)nseg * plans[s].count * states[s].dofs_pad;
				xbase += (int64_t)nseg * plans[s].count * dxw;
			}
		}
		hdr->nitems = item;
		hdr->counter = 0;
		hdr->partial_elems = pbase;
	}
}

__global__ void scatter_kernel(int64_t n, const uint32_t* __restrict__ bkey,
	uint32_t* __restrict__ cursor, uint32_t* __restrict__ perm)
{
	for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x)
	{
		uint32_t pos = atomicAdd(&cursor[bkey[p]], 1u);
		perm[pos] = (uint32_t)p;
	}
}

namespace {


template <int PREC, int DT>
__device__ __forceinline__ void contract_chunk(const SOperands<PREC, DT>& so, int ndt,
	const double (&b)[kChunk], f16_t (&acc)[2][DT])
{
	if constexpr (PREC == HDDM_LB_F16X2)
	{
		// B is scaled by 2^12 before the split so that the low halves of
		// basis values down to 2^-12 stay normal f16 numbers
		uint32_t hi[8], lo[8];
#pragma unroll
		for (int m = 0; m < 8; m++)
		{
			// split in FP32 (one rounding from FP64; the f16 pair keeps 22 bits)
			const float f0 = (float)b[2 * m] * kBScale, f1 = (float)b[2 * m + 1] * kBScale;
			hi[m] = pack_h2(f0, f1);
			lo[m] = pack_h2(f0 - lo_half_f16(hi[m]), f1 - hi_half_f16(hi[m]));
		}
		// Lanes 0-31 own points 0-31, lanes 32-63 points 32-63; each holds
		// nodes 0-15 of its point. The B operand of points 0-31 needs nodes
		// 0-7 in lanes 0-31 and nodes 8-15 in lanes 32-63: one half-swap.
#pragma unroll
		for (int m = 0; m < 4; m++)
		{
			auto rh = __builtin_amdgcn_permlane32_swap(hi[m], hi[m + 4], false, false);
			hi[m] = rh[0]; hi[m + 4] = rh[1];
			auto rl = __builtin_amdgcn_permlane32_swap(lo[m], lo[m + 4], false, false);
			lo[m] = rl[0]; lo[m + 4] = rl[1];
		}
		const h8_t Bh0 = as_vec8<h8_t>(hi[0], hi[1], hi[2], hi[3]);
		const h8_t Bh1 = as_vec8<h8_t>(hi[4], hi[5], hi[6], hi[7]);
		const h8_t Bl0 = as_vec8<h8_t>(lo[0], lo[1], lo[2], lo[3]);
		const h8_t Bl1 = as_vec8<h8_t>(lo[4], lo[5], lo[6], lo[7]);
#pragma unroll
		for (int dt = 0; dt < DT; dt++)
			if (dt < ndt)
			{
				const h8_t Ah = so.A[dt][0], Al = so.A[dt][1];
				acc[0][dt] = __builtin_amdgcn_mfma_f32_32x32x16_f16(Al, Bh0, acc[0][dt], 0, 0, 0);
				acc[0][dt] = __builtin_amdgcn_mfma_f32_32x32x16_f16(Ah, Bl0, acc[0][dt], 0, 0, 0);
				acc[0][dt] = __builtin_amdgcn_mfma_f32_32x32x16_f16(Ah, Bh0, acc[0][dt], 0, 0, 0);
				acc[1][dt] = __builtin_amdgcn_mfma_f32_32x32x16_f16(Al, Bh1, acc[1][dt], 0, 0, 0);
				acc[1][dt] = __builtin_amdgcn_mfma_f32_32x32x16_f16(Ah, Bl1, acc[1][dt], 0, 0, 0);
				acc[1][dt] = __builtin_amdgcn_mfma_f32_32x32x16_f16(Ah, Bh1, acc[1][dt], 0, 0, 0);
			}
	}
	else if constexpr (PREC == HDDM_LB_BF16X3)
	{
		uint32_t t1[8], t2[8], t3[8];
#pragma unroll
		for (int m = 0; m < 8; m++)
		{
			// split in FP32: three bf16 terms carry the 24 bits of the FP32 value
			const float f0 = (float)b[2 * m], f1 = (float)b[2 * m + 1];
			t1[m] = pack_b2(f0, f1);
			const float r0 = f0 - lo_half_bf16(t1[m]), r1 = f1 - hi_half_bf16(t1[m]);
			t2[m] = pack_b2(r0, r1);
			t3[m] = pack_b2(r0 - lo_half_bf16(t2[m]), r1 - hi_half_bf16(t2[m]));
		}
#pragma unroll
		for (int m = 0; m < 4; m++)
		{
			auto r1 = __builtin_amdgcn_permlane32_swap(t1[m], t1[m + 4], false, false);
			t1[m] = r1[0]; t1[m + 4] = r1[1];
			auto r2 = __builtin_amdgcn_permlane32_swap(t2[m], t2[m + 4], false, false);
			t2[m] = r2[0]; t2[m + 4] = r2[1];
			auto r3 = __builtin_amdgcn_permlane32_swap(t3[m], t3[m + 4], false, false);
			t3[m] = r3[0]; t3[m + 4] = r3[1];
		}
		b8_t B[2][3];
		B[0][0] = as_vec8<b8_t>(t1[0], t1[1], t1[2], t1[3]);
		B[1][0] = as_vec8<b8_t>(t1[4], t1[5], t1[6], t1[7]);
		B[0][1] = as_vec8<b8_t>(t2[0], t2[1], t2[2], t2[3]);
		B[1][1] = as_vec8<b8_t>(t2[4], t2[5], t2[6], t2[7]);
		B[0][2] = as_vec8<b8_t>(t3[0], t3[1], t3[2], t3[3]);
		B[1][2] = as_vec8<b8_t>(t3[4], t3[5], t3[6], t3[7]);
#pragma unroll
		for (int dt = 0; dt < DT; dt++)
			if (dt < ndt)
			{
#pragma unroll
				for (int t = 0; t < 2; t++)
				{
					// smallest products first
					acc[t][dt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(so.A[dt][2], B[t][0], acc[t][dt], 0, 0, 0);
					acc[t][dt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(so.A[dt][1], B[t][1], acc[t][dt], 0, 0, 0);
					acc[t][dt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(so.A[dt][0], B[t][2], acc[t][dt], 0, 0, 0);
					acc[t][dt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(so.A[dt][1], B[t][0], acc[t][dt], 0, 0, 0);
					acc[t][dt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(so.A[dt][0], B[t][1], acc[t][dt], 0, 0, 0);
					acc[t][dt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(so.A[dt][0], B[t][0], acc[t][dt], 0, 0, 0);
				}
			}
	}
}

} // namespace


// Node chunk of the walk's basis buffer: 16 nodes (the MFMA K extent) for
// the split MFMA paths; 32 for F32, whose buffer holds floats (half as many
// contractions, each folding into FP64 once); 8 for the FP64 path, which
// contracts node by node and whose smaller buffer leaves LDS for more
// resident waves.
template <int PREC>
constexpr int chunk_bits() { return PREC == HDDM_LB_FP64 ? 3 : (PREC == HDDM_LB_F32 ? 5 : 4); }

// Bytes of one basis value in the chunk buffer.
template <int PREC>
constexpr int chunk_value_bytes() { return PREC == HDDM_LB_F32 ? 4 : 8; }

// Per-wave LDS stage: the path stack [PM][64] doubles and the chunk buffer [2^CB][64].
template <int PREC, int PM>
constexpr size_t wave_stage_bytes()
{
	return (size_t)PM * kWave * 8 + ((size_t)1 << chunk_bits<PREC>()) * kWave * chunk_value_bytes<PREC>();
}

template <int PREC, int DT, int XB, int PM, int WALK>
__global__ void __launch_bounds__(256)
	__attribute__((amdgpu_waves_per_eu(PREC != HDDM_LB_FP64 && DT == 1 ? (XB == 1 ? 3 : (XB == 2 ? 2 : 1)) : 1)))
	sweep_kernel(SweepArgs a)
{
	extern __shared__ __attribute__((aligned(16))) double lds[];
	const int lane = threadIdx.x & 63;
	// wave-uniform values the compiler cannot see as uniform by itself
	const int wave = uniform((int)(threadIdx.x >> 6));
	// blocks per item: the workgroup's waves with LDS tiles, one otherwise
	const int G = WALK == kWalkLds ? uniform((int)(blockDim.x >> 6)) : 1;
	// basis values of the chunk: FP64 until the operand split, F32 for the F32 path
	typedef typename std::conditional<PREC == HDDM_LB_F32, float, double>::type BT;
	char* const lbase = (char*)lds;
	int64_t* const item_slot = (int64_t*)lbase;
	Op* const tile_ops = (Op*)(lbase + kLdsHeader);
	uint32_t* const tile_skip = (uint32_t*)(lbase + kLdsHeader + 2 * (size_t)kTileOps * sizeof(Op));
	const size_t stage0 = WALK == kWalkLds ? kLdsHeader + kLdsTiles : 0;
	constexpr int CB = chunk_bits<PREC>();
	double* const Pl = (double*)(lbase + stage0 + (size_t)wave * wave_stage_bytes<PREC, PM>());
	BT* const Bl = (BT*)(Pl + PM * kWave);
	const int64_t nitems = ldc(&a.hdr->nitems);

	for (;;)
	{
		int64_t item;
		if constexpr (WALK == kWalkLds)
		{
			if (threadIdx.x == 0) *item_slot = (int64_t)atomicAdd(&a.hdr->counter, 1ull);
			__syncthreads();
			item = uniform64((uint64_t)*item_slot);
		}
		else
		{
			uint64_t v = 0;
			if (lane == 0) v = atomicAdd(&a.hdr->counter, 1ull);
			item = (int64_t)uniform64(v);
		}
		if (item >= nitems) break;
		ItemInfo it;
		decode_item(a, item, G, it);
		const StateDesc* sd = a.states + it.s;
		const int dofs = ldc(&sd->dofs);
		const int DP = ldc(&sd->dofs_pad);
		const int nseg = ldc(&sd->nseg);
		const int nops_s = ldc(&sd->nops);
		const Op* const gops = ldc(&sd->ops);
		const uint32_t* const gskip = ldc(&sd->skip);
		const int32_t* const seg_op = ldc(&sd->seg_op);
		const int q0 = ldc(seg_op + it.seg), q1 = ldc(seg_op + it.seg + 1);
		const int dt0 = a.dt_base;
		const int ndt = min(DT, (DP >> 5) - dt0);

		// this wave's block; with LDS tiles, waves past the state's last block only help load tiles
		const int64_t blk = it.grp * G + (WALK == kWalkLds ? wave : 0);
		const int64_t rem = it.count - blk * kWave;
		const int np = ndt <= 0 ? 0 : (rem <= 0 ? 0 : (rem < kWave ? (int)rem : kWave));
		if (WALK == kWalkSmem && np == 0) continue;
		const int myp = lane < np ? lane : (np > 0 ? np - 1 : 0);
		const uint32_t orig = np > 0 ? a.perm[it.start + blk * kWave + myp] : 0u;
		const double* xrow = a.x + (int64_t)orig * a.ldx;
		double xb0[16], xb1[16], xb2[16], xb3[16];
#pragma unroll
		for (int d = 0; d < 16; d++)
		{
			xb0[d] = np > 0 ? xrow[d < a.dim ? d : 0] : 0.0;
			xb1[d] = XB > 1 && np > 0 ? xrow[16 + d < a.dim ? 16 + d : 0] : 0.0;
			xb2[d] = XB > 2 && np > 0 ? xrow[32 + d < a.dim ? 32 + d : 0] : 0.0;
			xb3[d] = XB > 2 && np > 0 ? xrow[48 + d < a.dim ? 48 + d : 0] : 0.0;
		}
#pragma unroll
		for (int k = 0; k < (1 << CB); k++) Bl[k * kWave + lane] = (BT)0;

		Walker<XB, BT, WALK, PM, PREC == HDDM_LB_FP64, CB> w;
		w.init(q0, np > 0 ? q1 : q0, Pl, Bl, lane, gops, gskip);
		uint32_t stat_chunks = 0, stat_mfma = 0;
		const bool split = nseg > 1;
		double* const part = a.partial + ldc(&a.plans[it.s].partial_base);
		// one pass over [q0, q1) for scalar-load walks; tiles of kTileOps otherwise
		const int tstep = WALK == kWalkLds ? kTileOps : q1 - q0;

		if constexpr (WALK == kWalkLds)
		{
			load_tile(gops, gskip, nops_s, q0, tile_ops, tile_skip, wave, G, lane);
			__syncthreads();
		}

		if constexpr (PREC == HDDM_LB_FP64)
		{
			const double* s64 = ldc(&sd->s64) + 32 * dt0;
			double acc[32];
#pragma unroll
			for (int c = 0; c < 32; c++) acc[c] = 0.0;
			// FP64 contraction of the finished chunk, positions in ascending order
			auto contract = [&]() {
				if (a.stats) stat_chunks++;
				for (uint32_t m = w.mask; m; m &= m - 1)
				{
					const int kk = __builtin_ctz(m);
					const double b = Bl[kk * kWave + lane];
					const double* srow = s64 + (size_t)((w.chunk << CB) + kk) * DP;
#pragma unroll
					for (int c = 0; c < 32; c++) acc[c] = fma(b, ldc(srow + c), acc[c]);
				}
			};
			int k = 0;
			for (int tb = q0; tb < q1; tb += tstep, k ^= 1)
			{
				if (WALK == kWalkLds && tb + kTileOps < q1)
					load_tile(gops, gskip, nops_s, tb + kTileOps, tile_ops + (k ^ 1) * kTileOps,
						tile_skip + (k ^ 1) * kTileOps, wave, G, lane);
				const int tend = min(tb + tstep, w.q1);
				for (;;)
				{
					int node = 0;
					double val = 0.0;
					if (!w.run(tile_ops + k * kTileOps, tile_skip + k * kTileOps, tb, tend, lane,
							xb0, xb1, xb2, xb3, node, val)) break;
					if (w.chunk >= 0) contract();
					w.start_chunk(node, val, lane);
				}
				if constexpr (WALK == kWalkLds) __syncthreads();
			}
			if (w.chunk >= 0) contract();
			if (lane < np)
			{
				double* out = split ? part + ((int64_t)it.seg * it.count + blk * kWave + lane) * DP
					: a.y + (int64_t)orig * a.ldy;
				out += 32 * dt0;
				const int dlim = dofs - 32 * dt0;
#pragma unroll
				for (int c = 0; c < 32; c++)
					if (c < dlim) out[c] = acc[c];
			}
		}
		else
		{
			// Each chunk's MFMA sum starts from zero and is folded into the
			// FP64 accumulators when the chunk is contracted.
			double acc64[2][DT][16];
#pragma unroll
			for (int t = 0; t < 2; t++)
#pragma unroll
				for (int dt = 0; dt < DT; dt++)
#pragma unroll
					for (int r = 0; r < 16; r++) acc64[t][dt][r] = 0.0;
			SOperands<PREC, DT> so;
			auto contract = [&]() {
				if (a.stats)
				{
					// counted in a pass of its own (statistics mode only) so
					// that the contraction below stays as it runs otherwise
					stat_chunks++;
					if constexpr (PREC == HDDM_LB_F32)
					{
#pragma unroll 1
						for (int q = 0; q < 16; q++)
							if ((w.mask >> (2 * q)) & 3u)
							{
								auto r = __builtin_amdgcn_permlane32_swap(
									__builtin_bit_cast(uint32_t, Bl[(2 * q) * kWave + lane]),
									__builtin_bit_cast(uint32_t, Bl[(2 * q + 1) * kWave + lane]), false, false);
								const uint32_t n0 = __builtin_amdgcn_ballot_w64(__builtin_bit_cast(float, (uint32_t)r[0]) != 0.0f) != 0;
								const uint32_t n1 = __builtin_amdgcn_ballot_w64(__builtin_bit_cast(float, (uint32_t)r[1]) != 0.0f) != 0;
								stat_mfma += (n0 + n1) * (uint32_t)ndt;
							}
					}
					else
						stat_mfma += (PREC == HDDM_LB_F16X2 ? 6u : 12u) * (uint32_t)ndt;
				}
				f16_t acc[2][DT];
#pragma unroll
				for (int t = 0; t < 2; t++)
#pragma unroll
					for (int dt = 0; dt < DT; dt++)
#pragma unroll
						for (int r = 0; r < 16; r++) acc[t][dt][r] = 0.0f;
				if constexpr (PREC == HDDM_LB_F32)
				{
					// K = 2 nodes per MFMA step, 16 steps over the 32-node
					// chunk: steps whose two positions received no node are
					// skipped (wave-uniform mask)
#pragma unroll
					for (int q = 0; q < 16; q++)
						if ((w.mask >> (2 * q)) & 3u)
						{
							const float b0 = Bl[(2 * q) * kWave + lane], b1 = Bl[(2 * q + 1) * kWave + lane];
							Bl[(2 * q) * kWave + lane] = 0.0f;
							Bl[(2 * q + 1) * kWave + lane] = 0.0f;
							auto r = __builtin_amdgcn_permlane32_swap(__builtin_bit_cast(uint32_t, b0),
								__builtin_bit_cast(uint32_t, b1), false, false);
							const float B0 = __builtin_bit_cast(float, (uint32_t)r[0]);
							const float B1 = __builtin_bit_cast(float, (uint32_t)r[1]);
							// a 32-point half where both nodes are zero adds nothing
							const bool h0 = __builtin_amdgcn_ballot_w64(B0 != 0.0f) != 0;
							const bool h1 = __builtin_amdgcn_ballot_w64(B1 != 0.0f) != 0;
#pragma unroll
							for (int dt = 0; dt < DT; dt++)
								if (dt < ndt)
								{
									if (h0) acc[0][dt] = __builtin_amdgcn_mfma_f32_32x32x2f32(so.A[dt][q], B0, acc[0][dt], 0, 0, 0);
									if (h1) acc[1][dt] = __builtin_amdgcn_mfma_f32_32x32x2f32(so.A[dt][q], B1, acc[1][dt], 0, 0, 0);
								}
						}
				}
				else
				{
					double b[kChunk];
#pragma unroll
					for (int kk = 0; kk < kChunk; kk++) { b[kk] = Bl[kk * kWave + lane]; Bl[kk * kWave + lane] = 0.0; }
					contract_chunk<PREC, DT>(so, ndt, b, acc);
				}
#pragma unroll
				for (int t = 0; t < 2; t++)
#pragma unroll
					for (int dt = 0; dt < DT; dt++)
#pragma unroll
						for (int r = 0; r < 16; r++) acc64[t][dt][r] += (double)acc[t][dt][r];
			};
			int k = 0;
			for (int tb = q0; tb < q1; tb += tstep, k ^= 1)
			{
				if (WALK == kWalkLds && tb + kTileOps < q1)
					load_tile(gops, gskip, nops_s, tb + kTileOps, tile_ops + (k ^ 1) * kTileOps,
						tile_skip + (k ^ 1) * kTileOps, wave, G, lane);
				const int tend = min(tb + tstep, w.q1);
				for (;;)
				{
					int node = 0;
					double val = 0.0;
					if (!w.run(tile_ops + k * kTileOps, tile_skip + k * kTileOps, tb, tend, lane,
							xb0, xb1, xb2, xb3, node, val)) break;
					if (w.chunk >= 0) contract();
					w.start_chunk(node, val, lane);
					// issue the surplus operands of the new chunk now; they are
					// consumed after the walk has finished the chunk
					so.load(sd, w.chunk, DP, dt0, ndt, lane);
				}
				if constexpr (WALK == kWalkLds) __syncthreads();
			}
			if (w.chunk >= 0) contract();

			// C layout of v_mfma_f32_32x32x*: register r of lane l holds
			// row 8*(r/4) + 4*(l/32) + r%4 (the dof) and column l%32 (the point).
			const double* iscale = ldc(&sd->iscale);
			const int row = lane & 31, khalf = lane >> 5;
#pragma unroll
			for (int t = 0; t < 2; t++)
			{
				const int qp = 32 * t + row;
				const uint32_t oq = (uint32_t)__shfl((int)orig, qp, 64);
				if (qp < np)
				{
					double* out = split ? part + ((int64_t)it.seg * it.count + blk * kWave + qp) * DP
						: a.y + (int64_t)oq * a.ldy;
#pragma unroll
					for (int dt = 0; dt < DT; dt++)
						if (dt < ndt)
						{
#pragma unroll
							for (int r = 0; r < 16; r++)
							{
								const int dof = (dt0 + dt) * 32 + 8 * (r >> 2) + 4 * khalf + (r & 3);
								if (dof < dofs)
								{
									double v = acc64[t][dt][r];
									if constexpr (PREC == HDDM_LB_F16X2) v *= ldc(iscale + dof) * (1.0 / kBScale);
									out[dof] = v;
								}
							}
						}
				}
			}
		}
		if (a.stats && lane == 0 && np > 0)
		{
			atomicAdd(&a.stats[0], (unsigned long long)w.nops);
			atomicAdd(&a.stats[1], (unsigned long long)w.nprunes);
			atomicAdd(&a.stats[2], (unsigned long long)stat_chunks);
			atomicAdd(&a.stats[3], 1ull);
			atomicAdd(&a.stats[4], (unsigned long long)stat_mfma);
		}
	}
}

size_t sweep_lds_bytes(int prec, int waves, int depth, int walk)
{
	size_t stage;
	switch (prec)
	{
	case HDDM_LB_FP64: stage = depth < 8 ? wave_stage_bytes<HDDM_LB_FP64, 8>() : wave_stage_bytes<HDDM_LB_FP64, 16>(); break;
	case HDDM_LB_F32: stage = depth < 8 ? wave_stage_bytes<HDDM_LB_F32, 8>() : wave_stage_bytes<HDDM_LB_F32, 16>(); break;
	default: stage = depth < 8 ? wave_stage_bytes<HDDM_LB_F16X2, 8>() : wave_stage_bytes<HDDM_LB_F16X2, 16>(); break;
	}
	return (walk == kWalkLds ? kLdsHeader + kLdsTiles : 0) + (size_t)waves * stage;
}

__global__ void reset_counter_kernel(PlanHeader* hdr)
{
	if (threadIdx.x == 0) hdr->counter = 0;
}

// Sums the node-segment partials of every point in segment order. Block
// (32, 8): threadIdx.x walks the dofs, threadIdx.y picks one of 8 points;
// blocks stride over the state's points. No 64-bit division per element.
__global__ void __launch_bounds__(256) reduce_partials_kernel(const StateDesc* __restrict__ states,
	const StatePlan* __restrict__ plans, const uint32_t* __restrict__ perm, const double* __restrict__ partial,
	double* __restrict__ y, int64_t ldy)
{
	const int s = blockIdx.y;
	const int nseg = states[s].nseg;
	if (nseg <= 1) return;
	const int dofs = states[s].dofs, DP = states[s].dofs_pad;
	const int64_t count = plans[s].count, start = plans[s].start;
	const double* base = partial + plans[s].partial_base;
	const int64_t gstride = count * DP;
	for (int64_t lp = (int64_t)blockIdx.x * 8 + threadIdx.y; lp < count; lp += (int64_t)gridDim.x * 8)
	{
		const uint32_t orow = perm[start + lp];
		const double* src = base + lp * DP;
		for (int c = threadIdx.x; c < dofs; c += 32)
		{
			double sum = 0.0;
			for (int g = 0; g < nseg; g++) sum += src[g * gstride + c];
			y[(int64_t)orow * ldy + c] = sum;
		}
	}
}

// ---- launchers ----

template <int PREC, int DT, int XB, int PM>
static const void* sweep_ptr(int walk)
{
	return walk == kWalkLds ? (const void*)&sweep_kernel<PREC, DT, XB, PM, kWalkLds>
		: (const void*)&sweep_kernel<PREC, DT, XB, PM, kWalkSmem>;
}

template <int PREC, int DT, int XB>
static const void* sweep_pm(int depth, int walk)
{
	return depth < 8 ? sweep_ptr<PREC, DT, XB, 8>(walk) : sweep_ptr<PREC, DT, XB, 16>(walk);
}

template <int PREC, int DT>
static const void* sweep_xb(int xbanks, int depth, int walk)
{
	switch (xbanks)
	{
	case 1: return sweep_pm<PREC, DT, 1>(depth, walk);
	case 2: return sweep_pm<PREC, DT, 2>(depth, walk);
	case 3: case 4: return sweep_pm<PREC, DT, 4>(depth, walk);
	default: return nullptr;
	}
}

const void* sweep_kernel_ptr(int prec, int dt, int xbanks, int depth, int walk)
{
	switch (prec)
	{
	case HDDM_LB_FP64: return sweep_xb<HDDM_LB_FP64, 1>(xbanks, depth, walk);
	case HDDM_LB_F16X2: return dt == 1 ? sweep_xb<HDDM_LB_F16X2, 1>(xbanks, depth, walk) : sweep_xb<HDDM_LB_F16X2, 2>(xbanks, depth, walk);
	case HDDM_LB_BF16X3: return dt == 1 ? sweep_xb<HDDM_LB_BF16X3, 1>(xbanks, depth, walk) : sweep_xb<HDDM_LB_BF16X3, 2>(xbanks, depth, walk);
	case HDDM_LB_F32: return dt == 1 ? sweep_xb<HDDM_LB_F32, 1>(xbanks, depth, walk) : sweep_xb<HDDM_LB_F32, 2>(xbanks, depth, walk);
	}
	return nullptr;
}

int xbanks_for(int dim)
{
	if (dim <= 16) return 1;
	if (dim <= 32) return 2;
	if (dim <= 64) return 4;
	return 0;
}

hipError_t launch_fill_u32(uint32_t* p, int64_t n, uint32_t v, hipStream_t st)
{
	int blocks = (int)((n + 255) / 256);
	if (blocks > 4096) blocks = 4096;
	if (blocks < 1) blocks = 1;
	hipLaunchKernelGGL(fill_u32_kernel, dim3(blocks), dim3(256), 0, st, p, n, v);
	return hipGetLastError();
}

hipError_t launch_bucket_key(int64_t n, const double* x, int64_t ldx, const int32_t* sid,
	const uint8_t* state_ok, int nstates, int dim, int bbits, uint32_t* bkey, uint32_t* hist, hipStream_t st)
{
	if (bbits > kKeyDims) return hipErrorInvalidValue;
	int blocks = (int)((n + kKeyBlock - 1) / kKeyBlock);
	if (blocks > 8192) blocks = 8192;
	if (blocks < 1) blocks = 1;
	hipLaunchKernelGGL(bucket_key_kernel, dim3(blocks), dim3(kKeyBlock), 0, st, n, x, ldx, sid, state_ok, nstates, dim, bbits, bkey, hist);
	return hipGetLastError();
}

hipError_t launch_scan_plan(const uint32_t* hist, uint32_t* cursor, int64_t nbuckets, int bbits, int nstates,
	const StateDesc* states, const int32_t* cost_order, StatePlan* plans, PlanHeader* hdr, int group, int dxw,
	hipStream_t st)
{
	hipLaunchKernelGGL(scan_plan_kernel, dim3(1), dim3(1024), 0, st, hist, cursor, nbuckets, bbits, nstates,
		states, cost_order, plans, hdr, group, dxw);
	return hipGetLastError();
}

hipError_t launch_scatter(int64_t n, const uint32_t* bkey, uint32_t* cursor, uint32_t* perm, hipStream_t st)
{
	int blocks = (int)((n + 255) / 256);
	if (blocks > 8192) blocks = 8192;
	if (blocks < 1) blocks = 1;
	hipLaunchKernelGGL(scatter_kernel, dim3(blocks), dim3(256), 0, st, n, bkey, cursor, perm);
	return hipGetLastError();
}

hipError_t launch_reset_counter(PlanHeader* hdr, hipStream_t st)
{
	hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(64), 0, st, hdr);
	return hipGetLastError();
}

hipError_t launch_sweep(const void* kern, int grid, int waves, size_t lds, const SweepArgs& args, hipStream_t st)
{
	void* params[] = { (void*)&args };
	return hipLaunchKernel(kern, dim3(grid), dim3(waves * kWave), params, lds, st);
}

hipError_t launch_reduce(const StateDesc* states, const StatePlan* plans, const uint32_t* perm,
	const double* partial, double* y, int64_t ldy, int nstates, hipStream_t st)
{
	hipLaunchKernelGGL(reduce_partials_kernel, dim3(2048, nstates), dim3(32, 8), 0, st, states, plans, perm, partial, y, ldy);
	return hipGetLastError();
}

} // namespace hddm_lb
