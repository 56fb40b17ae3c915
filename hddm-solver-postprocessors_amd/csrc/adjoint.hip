// Adjoint of the LinearBasis forward (gfx950). No counterpart in the reference.
//
// For y_p = sum_k B_k(x_p) S_k and an upstream gradient g_p = dL/dy_p:
//   dS_k   = sum_p B_k(x_p) g_p                                     (dsurplus)
//   dx_p,d = sum_k w_pk dB_k/dx_d,  w_pk = g_p . S_k                 (dx)
// with the piecewise-constant hat derivative d/dx max(0, 1 - |x i - j|) =
// -i sign(x i - j) inside the support and 0 outside (and 0 at the peak).
// The basis is recomputed by the same prefix-trie walk as the forward sweep
// (independent waves, one 64-point block per wave, the op stream through
// scalar loads). Two kernels, each with only the state it needs:
// * dx: along the walk each trie edge e at depth t keeps
//     R_e = sum over the surviving nodes k below e of w_k B_k,
//   and when the walk leaves e (a sibling or an ancestor's sibling comes
//   next) it adds R_e * phi'_e / phi_e to dx[dim(e)] and R_e to its parent
//   edge: dB_k/dx_d = B_k phi'_d / phi_d for the one factor of k in dim d, so
//   every node's contribution reaches each of its factors exactly once. w_pk
//   is an FP64 dot product against the node's surplus row at emission.
// * dS: the forward's walk emits basis values into 16-node chunks; each
//   finished chunk is contracted with v_mfma_f64_16x16x4f64 (dofs x nodes,
//   K = 4 points per step) and added to the output with FP64 atomics.
#include "sweep_common.h"
#include "walk.h"
#include "kernels.h"
#include "context.h"
#include "errors.h"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstring>

namespace hddm_lb {

typedef double d4_t __attribute__((ext_vector_type(4)));



// Per-wave LDS of the dx walk: P, R [PM][64]; dx [dim rounded to even][64]
// (doubles); per depth of the path the edge's op words {dim, i, -j, 0}
// (uint32 x 4), from which phi'/phi is recomputed when the edge is left.
__host__ __device__ constexpr int adj_dx_rows(int dim) { return (dim + 1) & ~1; }
__host__ __device__ constexpr size_t adj_lds_bytes(int pm, int dim)
{
	return (size_t)(2 * pm + adj_dx_rows(dim)) * kWave * 8 + (size_t)pm * 16;
}

// Per-wave LDS of the dS-only sweep: path stack [PM][64] and chunk basis [16][64].
template <int PM>
constexpr size_t ds_stage_bytes() { return (size_t)(PM + kChunk) * kWave * 8; }

namespace {

// Copies of dS for the per-XCD accumulation (MI355X has 8 XCDs).
constexpr int kXcdCopies = 8;
// Largest per-XCD dS scratch, in elements of all copies together; beyond it
// dS is accumulated in place.
constexpr int64_t kXcdCopiesMaxElems = (int64_t)1 << 27;

// Owned units of the wave-specialized adjoint: the dS copies they may use,
// all together at most this many bytes (and an eighth of device memory).
constexpr int kOwnMaxCopies = 32;
constexpr int64_t kOwnMaxBytes = (int64_t)16 << 30;

// out[i] = sum of the ncopies copies of element i, in copy order.
__global__ void sum_ds_copies_kernel(const double* __restrict__ part, int64_t n, int ncopies, double* __restrict__ out)
{
	for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
	{
		double v = 0.0;
		for (int c = 0; c < ncopies; c++) v += part[c * n + i];
		out[i] = v;
	}
}

// A operands of the dS contraction for one 64-point block: the upstream
// gradient in the v_mfma_f64_16x16x4f64 A layout, A[i][k] from lane i + 16k
// (i: dof of the 16-dof tile, k: point 4s + k of step s). Loaded once per
// block; lanes past the block hold 0.
struct GradTiles
{
	double v[2][16];

	__device__ __forceinline__ void load(const AdjointArgs& a, uint32_t orig, int np, int dofs, int lane)
	{
		const int col = lane & 15, kq = lane >> 4;
#pragma unroll
		for (int tile = 0; tile < 2; tile++)
		{
			const int dofA = (a.dt_base * 2 + tile) * 16 + col;
#pragma unroll
			for (int s = 0; s < 16; s++)
			{
				const int pt = 4 * s + kq;
				const uint32_t og = (uint32_t)__shfl((int)orig, pt, 64);
				v[tile][s] = pt < np && dofA < dofs ? a.g[(int64_t)og * a.ldg + dofA] : 0.0;
			}
		}
	}
};

// dS of one finished chunk: D[dof][node] += sum_pt g[pt][dof] B[node][pt],
// 16 dofs x 16 nodes per v_mfma_f64_16x16x4f64 tile, 16 steps of 4 points.
// Operand layouts: A[i][k] from lane i + 16k, B[k][j] from lane j + 16k;
// D[i][j] in lane j + 16 (i % 4), element i / 4.
__device__ __forceinline__ void contract_ds(const AdjointArgs& a, const StateDesc* sd, const double* Bc,
	uint32_t mask, int chunk, int lane, const GradTiles& gt, int dofs)
{
	// dS is added to the copy of the XCD the wave runs on (atomics stay in
	// that XCD's L2); the copies are summed in a fixed order afterwards
	double* ds = a.ds + ldc(&sd->ds_offset);
	if (a.ds_xcd > 0)
	{
		uint32_t xcc;
		asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(xcc));
		ds += (int64_t)(xcc & (kXcdCopies - 1)) * a.ds_xcd;
	}
	const int32_t* order = ldc(&sd->order);
	const int col = lane & 15, kq = lane >> 4;
	const int node = chunk * kChunk + col;
	const bool live = (mask >> col) & 1u;
	const int orow = live ? order[node] : 0;
	// the two dof tiles' chains are interleaved (independent accumulators)
	d4_t acc[2] = { { 0.0, 0.0, 0.0, 0.0 }, { 0.0, 0.0, 0.0, 0.0 } };
#pragma unroll
	for (int s = 0; s < 16; s++)
	{
		const double bv = Bc[col * kWave + 4 * s + kq];
		acc[0] = __builtin_amdgcn_mfma_f64_16x16x4f64(gt.v[0][s], bv, acc[0], 0, 0, 0);
		acc[1] = __builtin_amdgcn_mfma_f64_16x16x4f64(gt.v[1][s], bv, acc[1], 0, 0, 0);
	}
	if (live)
	{
#pragma unroll
		for (int tile = 0; tile < 2; tile++)
#pragma unroll
			for (int r = 0; r < 4; r++)
			{
				const int dof = (a.dt_base * 2 + tile) * 16 + kq + 4 * r;
				if (dof < dofs) unsafeAtomicAdd(ds + (int64_t)orow * dofs + dof, acc[tile][r]);
			}
	}
}

} // namespace

// Leaves the path edge at depth tt with sum r: dx[dim] += r * phi'/phi and
// r goes to the parent edge. phi'/phi = -i sign(x i - j) / phi inside the
// support (0 outside and at the peak), from the hardware reciprocal refined
// by two Newton steps (error ~1 ulp) instead of a full FP64 division.
template <int XB>
__device__ __forceinline__ void leave_edge(const u4_t& e, double r, const double (&xb0)[16], const double (&xb1)[16],
	const double (&xb2)[16], const double (&xb3)[16], double* DX, double* R, int tt, int lane)
{
	const uint32_t dd = __builtin_amdgcn_readfirstlane(e[0]);
	const double iv = f64_from_high(__builtin_amdgcn_readfirstlane(e[1]));
	const double u = fma(xget<XB, true>(xb0, xb1, xb2, xb3, dd), iv, f64_from_high(__builtin_amdgcn_readfirstlane(e[2])));
	const double phi = 1.0 - fabs(u);
	const double ph = phi > 0.0 ? phi : 1.0;
	double rc = __builtin_amdgcn_rcp(ph);
	rc = fma(rc, fma(-ph, rc, 1.0), rc);
	rc = fma(rc, fma(-ph, rc, 1.0), rc);
	const double qv = phi > 0.0 ? (u > 0.0 ? -iv : (u < 0.0 ? iv : 0.0)) * rc : 0.0;
	lds_add(DX + dd * kWave + lane, r * qv);
	if (tt > 0) lds_add(R + (tt - 1) * kWave + lane, r);
}

// dx along the edge-sum walk.
template <int XB, int PM>
__global__ void __launch_bounds__(256) adjoint_kernel(AdjointArgs a)
{
	// Waves are independent: each pops its own items (one 64-point block
	// against one node segment) and reads the op stream through scalar loads.
	extern __shared__ __attribute__((aligned(16))) double lds[];
	const int lane = threadIdx.x & 63;
	const int wave = uniform((int)(threadIdx.x >> 6));
	constexpr int DXR = 16 * XB;
	char* const lbase = (char*)lds;
	double* const P = (double*)(lbase + (size_t)wave * adj_lds_bytes(PM, a.dim));
	double* const R = P + PM * kWave;
	double* const DX = R + PM * kWave;
	u4_t* const edges = (u4_t*)(DX + adj_dx_rows(a.dim) * kWave);
	const int64_t nitems = ldc(&a.hdr->nitems);

	for (;;)
	{
		uint64_t iv0 = 0;
		if (lane == 0) iv0 = atomicAdd(&a.hdr->counter, 1ull);
		const int64_t item = (int64_t)uniform64(iv0);
		if (item >= nitems) break;
		ItemInfo it;
		decode_item(a, item, 1, it);
		const StateDesc* sd = a.states + it.s;
		const int dofs = ldc(&sd->dofs);
		const int DP = ldc(&sd->dofs_pad);
		const Op* const gops = ldc(&sd->ops);
		const int32_t* const seg_op = ldc(&sd->seg_op);
		const int q0 = ldc(seg_op + it.seg), q1 = ldc(seg_op + it.seg + 1);
		const int dt0 = a.dt_base;
		const bool active = 32 * dt0 < dofs;

		const int64_t blk = it.grp;
		const int64_t rem = it.count - blk * kWave;
		const int np = !active ? 0 : (rem <= 0 ? 0 : (rem < kWave ? (int)rem : kWave));
		if (np == 0) continue;
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
		// the lane's upstream gradient on this dof slab (0 for lanes past the block)
		double g[32];
		const double* grow = a.g + (int64_t)orig * a.ldg + 32 * dt0;
#pragma unroll
		for (int c = 0; c < 32; c++) g[c] = lane < np && 32 * dt0 + c < dofs ? grow[c] : 0.0;
#pragma unroll
		for (int d = 0; d < DXR; d++)
			if (d < a.dim) DX[d * kWave + lane] = 0.0;
		P[lane] = 1.0;

		const double* const s64 = ldc(&sd->s64) + 32 * dt0;
		int q = q0;
		int top = -1;          // deepest live edge of the path
		double cur = 1.0, base_prev = 1.0;
		// An op without children is left by the very next op, so the edge
		// sum and phi'/phi of such a top edge stay in registers; edges of
		// ops with children go through the LDS path (R, op words).
		bool top_leaf = false;
		double r_top = 0.0, q_top = 0.0;
		uint32_t dim_top = 0;
		Op op = ldc_op(gops + q);
		while (q < q1)
		{
			// the stream ends with kOpPad sentinels, so q + 1 is in bounds;
			// the load is issued first (scheduling fence) and stays in flight
			const Op nop = ldc_op(gops + q + 1);
			__builtin_amdgcn_sched_barrier(0);
			const uint32_t dimw = op.dim, ctrl = op.ctrl;
			const uint32_t dim = op_dim(dimw);
			const int t = (int)(ctrl & 15u);
			// leave the edges at depths >= t
			int tt = top;
			if (top_leaf && tt >= t)
			{
				if (wave_any(r_top != 0.0))
				{
					lds_add(DX + dim_top * kWave + lane, r_top * q_top);
					if (tt > 0) lds_add(R + (tt - 1) * kWave + lane, r_top);
				}
				tt--;
			}
			for (; tt >= t; tt--)
			{
				const double r = R[tt * kWave + lane];
				if (wave_any(r != 0.0)) leave_edge<XB>(edges[tt], r, xb0, xb1, xb2, xb3, DX, R, tt, lane);
			}
			top = t;
			top_leaf = !(ctrl & kOpKids);
			// parent product from the op's relation to its stream predecessor
			const uint32_t rel = (ctrl >> kOpRelShift) & 3u;
			double bt;
			if (rel == 0u) bt = cur;
			else if (rel == 1u) bt = base_prev;
			else bt = P[t * kWave + lane];
			const double iv = f64_from_high(op.ihi);
			const double u = fma(xget<XB, true>(xb0, xb1, xb2, xb3, dim), iv, f64_from_high(op.njhi));
			const double phi = 1.0 - fabs(u);
			const double p = bt * fmax(phi, 0.0);
			if (top_leaf)
			{
				r_top = 0.0;
				dim_top = dim;
			}
			else
			{
				P[(t + 1) * kWave + lane] = p;
				R[t * kWave + lane] = 0.0;
				if (lane == 0) edges[t] = u4_t{ dim, op.ihi, op.njhi, 0u };
			}
			base_prev = bt;
			cur = p;
			if (!wave_any(p != 0.0))
			{
				// the edge's sum stays 0, so it is never left with a
				// contribution; a pruned leaf moves to the already loaded next op
				const int rs = (int)(dimw >> kOpSkipShift);
				q += rs;
				op = rs == 1 ? nop : ldc_op(gops + q);
				continue;
			}
			q++;
			op = nop;
			if (top_leaf)
			{
				// phi'/phi of the edge, for the next op that leaves it
				const double ph = phi > 0.0 ? phi : 1.0;
				double rc = __builtin_amdgcn_rcp(ph);
				rc = fma(rc, fma(-ph, rc, 1.0), rc);
				rc = fma(rc, fma(-ph, rc, 1.0), rc);
				q_top = phi > 0.0 ? (u > 0.0 ? -iv : (u < 0.0 ? iv : 0.0)) * rc : 0.0;
			}
			if (ctrl & kOpLast)
			{
				const int n = (int)(ctrl >> kOpNodeShift);
				const double* srow = s64 + (size_t)n * DP;
				double w = 0.0;
#pragma unroll
				for (int c = 0; c < 32; c++) w = fma(g[c], ldc(srow + c), w);
				if (top_leaf) r_top = w * p;
				else R[t * kWave + lane] = w * p;
			}
		}
		int tt = top;
		if (top_leaf && tt >= 0)
		{
			if (wave_any(r_top != 0.0))
			{
				lds_add(DX + dim_top * kWave + lane, r_top * q_top);
				if (tt > 0) lds_add(R + (tt - 1) * kWave + lane, r_top);
			}
			tt--;
		}
		for (; tt >= 0; tt--)
		{
			const double r = R[tt * kWave + lane];
			if (wave_any(r != 0.0)) leave_edge<XB>(edges[tt], r, xb0, xb1, xb2, xb3, DX, R, tt, lane);
		}
		// Each (segment, block) item belongs to one wave per launch and the
		// dof-slab launches are stream-ordered, so the first slab stores and
		// later slabs add without atomics; split states go through
		// per-segment partials summed in segment order afterwards.
		if (lane < np)
		{
			double* out = ldc(&sd->nseg) > 1
				? a.dxpart + ldc(&a.plans[it.s].dx_base) + ((int64_t)it.seg * it.count + blk * kWave + lane) * a.dim
				: a.dx + (int64_t)orig * a.lddx;
#pragma unroll
			for (int d = 0; d < DXR; d++)
				if (d < a.dim)
				{
					const double v = DX[d * kWave + lane];
					out[d] = dt0 == 0 ? v : out[d] + v;
				}
		}
	}
}

// dx of split states: sums the node-segment partials of every point in
// segment order. Block (64, 4): threadIdx.x walks the dims, threadIdx.y picks
// one of 4 points; blocks stride over the state's points.
__global__ void __launch_bounds__(256) reduce_dx_partials_kernel(const StateDesc* __restrict__ states,
	const StatePlan* __restrict__ plans, const uint32_t* __restrict__ perm, const double* __restrict__ part,
	int dim, double* __restrict__ dx, int64_t lddx)
{
	const int s = blockIdx.y;
	const int nseg = states[s].nseg;
	if (nseg <= 1) return;
	const int64_t count = plans[s].count, start = plans[s].start;
	const double* base = part + plans[s].dx_base;
	const int64_t gstride = count * dim;
	for (int64_t lp = (int64_t)blockIdx.x * 4 + threadIdx.y; lp < count; lp += (int64_t)gridDim.x * 4)
	{
		const uint32_t orow = perm[start + lp];
		const double* src = base + lp * dim;
		for (int d = threadIdx.x; d < dim; d += 64)
		{
			double sum = 0.0;
			for (int g = 0; g < nseg; g++) sum += src[g * gstride + d];
			dx[(int64_t)orow * lddx + d] = sum;
		}
	}
}

// dS alone: the forward's walk (nodes emitted into 16-node chunks of the
// wave's LDS stage) with the dS contraction per finished chunk; no dx state,
// so it needs far fewer registers and LDS than the edge-sum walk.
template <int XB, int PM>
__global__ void __launch_bounds__(256)
	__attribute__((amdgpu_waves_per_eu(XB == 1 ? 3 : (XB == 2 ? 2 : 1))))
	adjoint_ds_kernel(AdjointArgs a)
{
	extern __shared__ __attribute__((aligned(16))) double lds[];
	const int lane = threadIdx.x & 63;
	const int wave = uniform((int)(threadIdx.x >> 6));
	double* const Pl = (double*)((char*)lds + (size_t)wave * ds_stage_bytes<PM>());
	double* const Bl = Pl + PM * kWave;
	const int64_t nitems = ldc(&a.hdr->nitems);

	for (;;)
	{
		uint64_t iv0 = 0;
		if (lane == 0) iv0 = atomicAdd(&a.hdr->counter, 1ull);
		const int64_t item = (int64_t)uniform64(iv0);
		if (item >= nitems) break;
		ItemInfo it;
		decode_item(a, item, 1, it);
		const StateDesc* sd = a.states + it.s;
		const int dofs = ldc(&sd->dofs);
		const Op* const gops = ldc(&sd->ops);
		const uint32_t* const gskip = ldc(&sd->skip);
		const int32_t* const seg_op = ldc(&sd->seg_op);
		const int q0 = ldc(seg_op + it.seg), q1 = ldc(seg_op + it.seg + 1);
		if (32 * a.dt_base >= dofs) continue;
		const int64_t rem = it.count - it.grp * kWave;
		const int np = rem <= 0 ? 0 : (rem < kWave ? (int)rem : kWave);
		if (np == 0) continue;
		const int myp = lane < np ? lane : np - 1;
		const uint32_t orig = a.perm[it.start + it.grp * kWave + myp];
		const double* xrow = a.x + (int64_t)orig * a.ldx;
		double xb0[16], xb1[16], xb2[16], xb3[16];
#pragma unroll
		for (int d = 0; d < 16; d++)
		{
			xb0[d] = xrow[d < a.dim ? d : 0];
			xb1[d] = XB > 1 ? xrow[16 + d < a.dim ? 16 + d : 0] : 0.0;
			xb2[d] = XB > 2 ? xrow[32 + d < a.dim ? 32 + d : 0] : 0.0;
			xb3[d] = XB > 2 ? xrow[48 + d < a.dim ? 48 + d : 0] : 0.0;
		}
#pragma unroll
		for (int k = 0; k < kChunk; k++) Bl[k * kWave + lane] = 0.0;
		GradTiles gt;
		gt.load(a, orig, np, dofs, lane);

		Walker<XB, double, kWalkSmem, PM, true, 4> w;
		w.init(q0, q1, Pl, Bl, lane, gops, gskip);
		for (;;)
		{
			int node = 0;
			double val = 0.0;
			if (!w.run(nullptr, nullptr, q0, q1, lane, xb0, xb1, xb2, xb3, node, val)) break;
			if (w.chunk >= 0)
			{
				contract_ds(a, sd, Bl, w.mask, w.chunk, lane, gt, dofs);
#pragma unroll
				for (int k = 0; k < kChunk; k++) Bl[k * kWave + lane] = 0.0;
			}
			w.start_chunk(node, val, lane);
		}
		if (w.chunk >= 0)
		{
			contract_ds(a, sd, Bl, w.mask, w.chunk, lane, gt, dofs);
#pragma unroll
			for (int k = 0; k < kChunk; k++) Bl[k * kWave + lane] = 0.0;
		}
	}
}

enum AdjMode { kAdjDx = 0, kAdjDs = 1 };

template <int XB, int PM>
static const void* adjoint_ptr(int mode, int dim, size_t& per_wave)
{
	if (mode == kAdjDs) { per_wave = ds_stage_bytes<PM>(); return (const void*)&adjoint_ds_kernel<XB, PM>; }
	per_wave = adj_lds_bytes(PM, dim);
	return (const void*)&adjoint_kernel<XB, PM>;
}

static const void* adjoint_kernel_ptr(int dim, int depth, int mode, size_t& per_wave)
{
	const bool deep = depth >= 8;
	switch (xbanks_for(dim))
	{
	case 1: return deep ? adjoint_ptr<1, 16>(mode, dim, per_wave) : adjoint_ptr<1, 8>(mode, dim, per_wave);
	case 2: return deep ? adjoint_ptr<2, 16>(mode, dim, per_wave) : adjoint_ptr<2, 8>(mode, dim, per_wave);
	case 3: case 4: return deep ? adjoint_ptr<4, 16>(mode, dim, per_wave) : adjoint_ptr<4, 8>(mode, dim, per_wave);
	}
	return nullptr;
}

} // namespace hddm_lb

using namespace hddm_lb;

#define HIP_TRY(expr) do { hipError_t e_ = (expr); if (e_ != hipSuccess) return hip_fail(e_, #expr, __FILE__, __LINE__); } while (0)

// After the adjoint kernels: dx of split states summed over node segments
// (of the descriptor table the kernels used), dS summed over its copies
// (ncopies > 1), both in a fixed order.
static int adjoint_tail(hddm_lb_ctx* ctx, Scratch& sc, const StateDesc* tables, int max_nseg, double* dx,
	int64_t lddx, double* dsurplus, int64_t ds_total, int ncopies, int ns, hipStream_t st)
{
	if (dx && max_nseg > 1)
	{
		hipLaunchKernelGGL(reduce_dx_partials_kernel, dim3(1024, ns), dim3(64, 4), 0, st,
			tables, (const StatePlan*)sc.plans, (const uint32_t*)sc.perm,
			(const double*)sc.dxpart, ctx->dim, dx, lddx);
		HIP_TRY(hipGetLastError());
	}
	if (ncopies > 1)
	{
		int blocks = (int)std::min<int64_t>(4096, (ds_total + 255) / 256);
		hipLaunchKernelGGL(sum_ds_copies_kernel, dim3(std::max(blocks, 1)), dim3(256), 0, st,
			(const double*)sc.dspart, ds_total, ncopies, dsurplus);
		HIP_TRY(hipGetLastError());
	}
	return scratch_mark(ctx, st);
}

// dS copies K (a power of two) for the owned units of the wave-specialized
// adjoint: the fewest that give each workgroup of the grid about four units
// ((state, segment, copy) triples), within the byte limits; K = 1 adds to
// dsurplus itself. 0 (atomics) when the limits leave fewer than two units
// per workgroup.
static int owned_copies(const hddm_lb_ctx* ctx, int64_t ds_total, int grid)
{
	const int64_t segs = ctx->nseg_adj_total;
	if (!ctx->opt_adj_owned || segs <= 0 || ds_total <= 0 || grid <= 0) return 0;
	const int64_t cap = std::min<int64_t>(kOwnMaxBytes, (int64_t)(ctx->total_mem / 8));
	const int64_t copy_bytes = ds_total * (int64_t)sizeof(double);
	int K = 1;
	while (K < kOwnMaxCopies && segs * K < 4 * (int64_t)grid && 2 * K * copy_bytes <= cap) K *= 2;
	return segs * K >= 2 * (int64_t)grid ? K : 0;
}

__global__ void iota_u32_kernel(int64_t n, uint32_t* __restrict__ p)
{
	for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
		p[i] = (uint32_t)i;
}

template <typename T>
static int grow_buf(T*& p, size_t& cap, size_t need)
{
	if (need <= cap && p) return HDDM_LB_SUCCESS;
	if (p) HIP_TRY(hipDeviceSynchronize());
	dev_free(p);
	p = nullptr;
	cap = 0;
	int rc = 0;
	p = (T*)dev_alloc(sizeof(T) * std::max<size_t>(need, 1), rc);
	if (rc) return rc;
	cap = std::max<size_t>(need, 1);
	return HDDM_LB_SUCCESS;
}

// sc.perm = the points ordered by bucket key and, within a bucket, by index:
// a stable radix sort of the keys (the forward's atomic scatter orders a
// bucket by arrival). Which points share a 64-point block then does not vary
// between runs, so neither do the adjoint's per-block dS partials.
static int stable_order(Scratch& sc, int64_t npoints, int64_t nbuckets, hipStream_t st)
{
	int bits = 1;
	while (bits < 32 && ((int64_t)1 << bits) < nbuckets) bits++;
	const int n = (int)npoints;
	int rc;
	if ((rc = grow_buf(sc.bkey2, sc.cap_bkey2, (size_t)npoints))) return rc;
	if ((rc = grow_buf(sc.pidx, sc.cap_pidx, (size_t)npoints))) return rc;
	size_t tmp = 0;
	HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, sc.bkey, sc.bkey2, sc.pidx, sc.perm, n, 0, bits, st));
	if ((rc = grow_buf(sc.sort_tmp, sc.cap_sort_tmp, tmp))) return rc;
	const int blocks = (int)std::min<int64_t>(4096, (npoints + 255) / 256);
	hipLaunchKernelGGL(iota_u32_kernel, dim3(std::max(blocks, 1)), dim3(256), 0, st, npoints, sc.pidx);
	HIP_TRY(hipGetLastError());
	HIP_TRY(hipcub::DeviceRadixSort::SortPairs(sc.sort_tmp, tmp, sc.bkey, sc.bkey2, sc.pidx, sc.perm, n, 0, bits, st));
	return HDDM_LB_SUCCESS;
}

// Grows the dS copy scratch to `need` elements.
static int grow_dspart(Scratch& scr, size_t need)
{
	if (need <= scr.cap_dspart) return HDDM_LB_SUCCESS;
	if (scr.dspart) HIP_TRY(hipDeviceSynchronize());
	dev_free(scr.dspart);
	scr.dspart = nullptr;
	scr.cap_dspart = 0;
	int arc = 0;
	scr.dspart = (double*)dev_alloc(need * sizeof(double), arc);
	if (arc) return arc;
	scr.cap_dspart = need;
	return HDDM_LB_SUCCESS;
}

extern "C" int hddm_lb_adjoint(hddm_lb_ctx* ctx, int64_t npoints,
	const double* x, int64_t ldx, const int32_t* state_ids,
	const double* gbar, int64_t ldg, double* dx, int64_t lddx,
	double* dsurplus, int precision, void* stream)
{
	if (!ctx) return set_error(HDDM_LB_ERROR_INVALID_VALUE, "adjoint: ctx is NULL");
	if (precision != HDDM_LB_FP64 && precision != HDDM_LB_F32)
		return set_error(HDDM_LB_ERROR_UNSUPPORTED, "adjoint: precision %d is not available; use HDDM_LB_FP64 or HDDM_LB_F32",
			precision);
	if (npoints < 0 || npoints > ((int64_t)1 << 31) - 1)
		return set_error(HDDM_LB_ERROR_INVALID_VALUE, "adjoint: npoints %lld out of range", (long long)npoints);
	if (ctx->dim < 1) return set_error(HDDM_LB_ERROR_STATE, "adjoint: no state is loaded");
	int maxdofs = 0;
	for (auto& s : ctx->states) if (s && s->loaded) maxdofs = std::max(maxdofs, s->desc.dofs);
	if (npoints && (!x || !gbar)) return set_error(HDDM_LB_ERROR_INVALID_VALUE, "adjoint: x and gbar must not be NULL");
	if (ldx < ctx->dim) return set_error(HDDM_LB_ERROR_INVALID_VALUE, "adjoint: ldx %lld < dim %d", (long long)ldx, ctx->dim);
	if (ldg < maxdofs) return set_error(HDDM_LB_ERROR_INVALID_VALUE, "adjoint: ldg %lld < dofs %d", (long long)ldg, maxdofs);
	if (dx && lddx < ctx->dim) return set_error(HDDM_LB_ERROR_INVALID_VALUE, "adjoint: lddx %lld < dim %d", (long long)lddx, ctx->dim);
	DeviceGuard guard(ctx->device);
	hipStream_t st = (hipStream_t)stream;
	int rc = prepare_scratch(ctx, ctx->scratch, std::max<int64_t>(npoints, 1));
	if (rc) return rc;
	if ((rc = scratch_wait(ctx, st))) return rc;
	int64_t ds_total = 0;
	if ((rc = hddm_lb_surplus_offset(ctx, -1, nullptr, &ds_total))) return rc;

	// automatic: the wave-specialized kernel, which walks each trie once for
	// four dof tiles instead of once per tile, for states of two or more dof
	// tiles and for wide states (D > 32) in F32; the per-tile kernels for one
	// tile at D <= 32, where their edge-sum walk keeps up with the dx waves
	// (65,536 points, F32: D=24 / 12k nodes / 128 dofs 6.3 ms against 37.8
	// per-tile, D=16 / 60k / 128 dofs 20.6 against 57.5, D=32 / 40k / 64 dofs
	// 20.4 against 49.2, D=20 / 100k / 64 dofs 37.0 against 71.3; one tile:
	// D=20 / 100k / 32 dofs 36.7 against 35.8, D=16 / 60k / 32 dofs 18.7
	// against 14.4). The per-tile kernels compute in FP64 whatever the
	// precision asked for, which meets the F32 bound too.
	const bool wide = xbanks_for(ctx->dim) == 4;
	// nodes of more than 8 factors next to four coordinate banks exceed the
	// wave-specialized kernel's registers; its dx waves take each factor's
	// centre -j / i from the factor table, exact when i is a power of two
	const bool ws_fits = !(wide && ctx->fmax_all > 8) && ctx->centre_exact_all;
	if (ctx->opt_form == 2 && !ws_fits)
		return set_error(HDDM_LB_ERROR_UNSUPPORTED,
			"adjoint: the wave-specialized kernel takes nodes of at most 8 factors when dim > 32, "
			"and index_i values that are powers of two");
	const bool ws = ctx->opt_form == 2
		|| (ctx->opt_form == 0 && ws_fits
			&& (ctx->dt_needed >= 2 || (wide && precision == HDDM_LB_F32)));
	// the wave-specialized kernel's launch shape: one walker wave, four
	// consumer waves and two dx waves per workgroup, four dof tiles per
	// launch; the second dx wave works when its rows fit the LDS
	const void* ws_kern = nullptr;
	size_t ws_lds = 0;
	int ws_grid = 0, ws_ndx = 1;
	if (ws)
	{
		ws_kern = adjoint_ws_kernel_ptr(precision, xbanks_for(ctx->dim), ctx->fmax_all);
		if (!ws_kern) return set_error(HDDM_LB_ERROR_UNSUPPORTED, "adjoint: dim %d is not supported", ctx->dim);
		ws_ndx = adjoint_ws_lds_bytes(precision, ctx->fmax_all, ctx->dim, 2) <= ctx->max_lds ? 2 : 1;
		ws_lds = adjoint_ws_lds_bytes(precision, ctx->fmax_all, ctx->dim, ws_ndx);
		if (ws_lds > ctx->max_lds)
			return set_error(HDDM_LB_ERROR_UNSUPPORTED, "adjoint: LDS of one workgroup (%zu B) exceeds the CU", ws_lds);
		HIP_TRY(hipFuncSetAttribute(ws_kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)ws_lds));
		int per_cu = 0;
		HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, ws_kern, 7 * kWave, ws_lds));
		if (per_cu < 1) return set_error(HDDM_LB_ERROR_UNSUPPORTED, "adjoint: kernel does not fit a CU");
		ws_grid = per_cu * ctx->num_cus;
	}
	// the wave-specialized kernel splits states into finer node segments
	const StateDesc* tables = ws ? ctx->d_states_adj : ctx->d_states;
	const int max_nseg = ws ? ctx->max_nseg_adj : ctx->max_nseg;

	// dS is accumulated: clear it with a kernel (never a memset, so the
	// sequence stays capturable). The wave-specialized kernel adds through
	// owned units into K copies (K = 1: dsurplus itself); otherwise atomics
	// go to per-XCD copies when they fit the scratch limit, else to dsurplus.
	const int own = ws && npoints > 0 && dsurplus && ds_total ? owned_copies(ctx, ds_total, ws_grid) : 0;
	const bool xcd_copies = !own && npoints > 0 && dsurplus && ds_total && ds_total * kXcdCopies <= kXcdCopiesMaxElems;
	const int ncopies = own > 1 ? own : (xcd_copies ? kXcdCopies : 1);
	if (ncopies > 1)
	{
		Scratch& scr = ctx->scratch;
		const size_t need = (size_t)ds_total * ncopies;
		if ((rc = grow_dspart(scr, need))) return rc;
		HIP_TRY(launch_fill_u32((uint32_t*)scr.dspart, 2 * (int64_t)need, 0u, st));
	}
	else if (dsurplus && ds_total) HIP_TRY(launch_fill_u32((uint32_t*)dsurplus, 2 * ds_total, 0u, st));
	ctx->last_adj[0] = ws ? 2 : 1;
	ctx->last_adj[1] = own;
	ctx->last_adj[2] = ncopies;
	ctx->last_adj[3] = max_nseg;
	if (npoints == 0 || (!dx && !dsurplus)) return scratch_mark(ctx, st);
	// dx is stored, not accumulated: no clear, and columns past dim of a
	// padded dx stay as the caller left them
	if (dx && max_nseg > 1)
	{
		Scratch& scr = ctx->scratch;
		const size_t need = (size_t)npoints * (size_t)max_nseg * (size_t)ctx->dim;
		if (need > scr.cap_dxpart)
		{
			if (scr.dxpart) HIP_TRY(hipDeviceSynchronize());
			dev_free(scr.dxpart);
			scr.dxpart = nullptr;
			scr.cap_dxpart = 0;
			int arc = 0;
			scr.dxpart = (double*)dev_alloc(need * sizeof(double), arc);
			if (arc) return arc;
			scr.cap_dxpart = need;
		}
	}

	Scratch& sc = ctx->scratch;
	const int ns = ctx->tables_n;
	const int bb = bucket_bits(npoints, ns, ctx->opt_sort);
	const int64_t nb = (int64_t)(ns + 1) << bb;
	AdjointArgs a;
	std::memset(&a, 0, sizeof(a));
	a.states = tables; a.cost_order = ctx->d_cost_order; a.hdr = sc.hdr;
	a.perm = sc.perm; a.x = x; a.ldx = ldx; a.g = gbar; a.ldg = ldg; a.dx = dx; a.lddx = lddx;
	a.dxpart = sc.dxpart; a.plans = sc.plans;
	a.ds = ncopies > 1 ? sc.dspart : dsurplus; a.ds_xcd = ncopies > 1 ? ds_total : 0;
	a.ds_own = own;
	a.nstates = ns; a.dim = ctx->dim;
	a.status = ctx->d_status;

	if (ws && ctx->opt_stats)
	{
		if (!ctx->d_adj_stats)
		{
			int arc = 0;
			ctx->d_adj_stats = (unsigned long long*)dev_alloc(hddm_lb_ctx::kAdjStatCount * 8, arc);
			if (arc) return arc;
		}
		HIP_TRY(launch_fill_u32((uint32_t*)ctx->d_adj_stats, 2 * hddm_lb_ctx::kAdjStatCount, 0u, st));
		a.stats = ctx->d_adj_stats;
	}
	if (ws)
	{
		a.ndx = ws_ndx;
		HIP_TRY(launch_fill_u32(sc.hist, nb, 0u, st));
		HIP_TRY(launch_bucket_key(npoints, x, ldx, state_ids, ctx->d_state_ok, ns, ctx->dim, bb, sc.bkey, sc.hist, st));
		HIP_TRY(launch_scan_plan(sc.hist, sc.cursor, nb, bb, ns, tables, ctx->d_cost_order, sc.plans, sc.hdr, 1,
			ctx->dim, st));
		if ((rc = stable_order(sc, npoints, nb, st))) return rc;
		const int npass = (ctx->dt_needed + 3) / 4;
		auto& ring = ctx->tring[1];
		const int tslot = (int)(ring.calls % hddm_lb_ctx::kTimingRing);
		if (ctx->opt_timing) HIP_TRY(hipEventRecord(ring.e0[tslot], st));
		for (int ps = 0; ps < npass; ps++)
		{
			a.dt_base = 4 * ps;
			if (ps > 0) HIP_TRY(launch_reset_counter(sc.hdr, st));
			void* params[] = { (void*)&a };
			HIP_TRY(hipLaunchKernel(ws_kern, dim3(ws_grid), dim3(7 * kWave), params, ws_lds, st));
		}
		if (ctx->opt_timing)
		{
			HIP_TRY(hipEventRecord(ring.e1[tslot], st));
			ring.calls++;
		}
		return adjoint_tail(ctx, sc, tables, max_nseg, dx, lddx, dsurplus, ds_total, ncopies, ns, st);
	}

	// dx and dS together: the edge-sum walk for dx, then the forward's walk
	// for dS (two lighter kernels run faster than one that carries both)
	int modes[2] = { dx ? kAdjDx : kAdjDs, -1 };
	if (dx && dsurplus) modes[1] = kAdjDs;
	const void* kerns[2] = { nullptr, nullptr };
	int grids[2] = { 0, 0 }, groups[2] = { 0, 0 };
	size_t ldss[2] = { 0, 0 };
	for (int m = 0; m < 2 && modes[m] >= 0; m++)
	{
		size_t per_wave = 0;
		const void* kern = adjoint_kernel_ptr(ctx->dim, ctx->fmax_all, modes[m], per_wave);
		if (!kern) return set_error(HDDM_LB_ERROR_UNSUPPORTED, "adjoint: dim %d is not supported", ctx->dim);
		if (per_wave > ctx->max_lds)
			return set_error(HDDM_LB_ERROR_UNSUPPORTED, "adjoint: LDS of one wave (%zu B) exceeds the CU", per_wave);
		// waves are independent: take the workgroup size (1, 2 or 4 waves)
		// that keeps the most waves resident per CU (LDS rarely divides evenly)
		int bestG = 0, best_waves = 0, best_per_cu = 0;
		for (int G = 1; G <= 4; G *= 2)
		{
			if (ctx->opt_waves && G > ctx->opt_waves) break;
			const size_t lds = (size_t)G * per_wave;
			if (lds > ctx->max_lds) break;
			HIP_TRY(hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
			int per_cu = 0;
			HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, G * kWave, lds));
			if (per_cu * G >= best_waves && per_cu > 0) { bestG = G; best_waves = per_cu * G; best_per_cu = per_cu; }
		}
		if (bestG < 1) return set_error(HDDM_LB_ERROR_UNSUPPORTED, "adjoint: kernel does not fit a CU");
		const size_t lds = (size_t)bestG * per_wave;
		HIP_TRY(hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
		kerns[m] = kern; grids[m] = best_per_cu * ctx->num_cus; groups[m] = bestG; ldss[m] = lds;
	}

	HIP_TRY(launch_fill_u32(sc.hist, nb, 0u, st));
	HIP_TRY(launch_bucket_key(npoints, x, ldx, state_ids, ctx->d_state_ok, ns, ctx->dim, bb, sc.bkey, sc.hist, st));
	HIP_TRY(launch_scan_plan(sc.hist, sc.cursor, nb, bb, ns, tables, ctx->d_cost_order, sc.plans, sc.hdr, 1,
		ctx->dim, st));
	if ((rc = stable_order(sc, npoints, nb, st))) return rc;

	auto& ring = ctx->tring[1];
	const int tslot = (int)(ring.calls % hddm_lb_ctx::kTimingRing);
	if (ctx->opt_timing) HIP_TRY(hipEventRecord(ring.e0[tslot], st));
	bool first = true;
	for (int m = 0; m < 2 && kerns[m]; m++)
		for (int sl = 0; sl < ctx->dt_needed; sl++)
		{
			a.dt_base = sl;
			if (!first) HIP_TRY(launch_reset_counter(sc.hdr, st));
			first = false;
			void* params[] = { (void*)&a };
			HIP_TRY(hipLaunchKernel(kerns[m], dim3(grids[m]), dim3(groups[m] * kWave), params, ldss[m], st));
		}
	if (ctx->opt_timing)
	{
		HIP_TRY(hipEventRecord(ring.e1[tslot], st));
		ring.calls++;
	}
	return adjoint_tail(ctx, sc, tables, max_nseg, dx, lddx, dsurplus, ds_total, ncopies, ns, st);
}
