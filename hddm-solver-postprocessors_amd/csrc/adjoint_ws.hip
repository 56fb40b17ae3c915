// Wave-specialized adjoint (gfx950): one walker wave and NC = 4 consumer waves
// per workgroup. No counterpart in the reference.
//
// For y_p = sum_k B_k(x_p) S_k and an upstream gradient g_p:
//   dS_k  = sum_p B_k(x_p) g_p                    (dsurplus)
//   dx_pd = sum_k w_pk dB_k/dx_pd,  w_pk = g_p . S_k (dx)
// Work items are the forward's: one 64-point block of one state against one
// node segment. The walker wave runs two walks of the segment's trie, one
// node chunk apart:
//   A  the forward's assembly walk (walk.h): surviving nodes' basis values go
//      into chunks of an LDS ring (32 nodes in FP32, 16 in FP64) that are
//      published to the consumers;
//   B  the edge-sum walk of the dx kernel (adjoint.hip): every trie edge e
//      keeps R_e = sum of w_k B_k over the surviving nodes below it, and
//      leaving e adds R_e phi'_e / phi_e to dx[dim(e)] and R_e to its parent.
//      Walk B reaches the nodes of chunk c only after the consumers have
//      computed w for chunk c, which they do as soon as walk A publishes it.
// Both walks evaluate the same FP64 operations in the same order (fma,
// 1 - |u|, max, product, all-zero test), so they survive and prune the same
// nodes.
// Consumer c takes one 32-dof tile of the surplus (or, when a state has at
// most two tiles in the pass, one of the two products of a tile) and, per
// published chunk, on MFMA:
//   dS   D[node][dof] = sum over the block's points of B[point][node] g[point][dof],
//        added to one copy of dS per XCD with atomics (copies summed after);
//   w    W[point][node] = sum over the tile's dofs of g[point][dof] S[node][dof],
//        written to an LDS buffer of its own; the walker sums the tiles'
//        buffers in consumer order (so dx is bitwise repeatable).
// FP64: v_mfma_f64_16x16x4f64 throughout. F32: v_mfma_f32_32x32x2f32 with
// FP32 operands; dS partials of one chunk (64 points) are added in FP64, w is
// summed over tiles and carried in FP64 by walk B.
// Synchronization is the forward WS sweep's: an LDS message ring with
// release/acquire counters and bounded waits (sweep_ws.hip).
#include "sweep_common.h"
#include "kernels.h"
#include "walk.h"
#include "contract.h"

namespace hddm_lb {

namespace {

constexpr int kARing = 4;
constexpr int kASpinLimit = 1 << 22;
constexpr int kANC = 4;        // consumer waves
constexpr int kARow = kWave + 1;   // padded rows of the chunk and w buffers
enum : int { kAMsgItem = 1, kAMsgChunk = 2, kAMsgExit = 3 };
enum : int { kRoleDs = 1, kRoleW = 2 };

struct AMsg
{
	int type, a, b, np;
	int64_t start, lpos;
};

struct ASync
{
	int produced;
	int abort;
	int done[6];
};

template <int PREC> constexpr int adj_chunk_bits() { return PREC == HDDM_LB_FP64 ? 4 : 5; }
template <int PREC> using adj_t = typename std::conditional<PREC == HDDM_LB_FP64, double, float>::type;
// one chunk buffer (and one consumer's w buffer): nodes x padded row
template <int PREC> constexpr size_t adj_buf_bytes() { return ((size_t)1 << adj_chunk_bits<PREC>()) * kARow * sizeof(adj_t<PREC>); }

// LDS layout of the workgroup (bytes): kARing chunk slots, 2 x kANC w buffers,
// messages, counters, walk A's path stack [PM][64], walk B's P and R stacks
// [PM][64], its edge words [PM], dx rows [dim rounded up to even][64].
template <int PREC>
struct ALayout
{
	size_t slots, wbuf, msgs, sync, pa, pb, rb, edges, dx, total;
	__host__ __device__ ALayout(int pm, int dim)
	{
		slots = 0;
		wbuf = slots + kARing * adj_buf_bytes<PREC>();
		msgs = wbuf + 2 * kANC * adj_buf_bytes<PREC>();
		sync = msgs + kARing * sizeof(AMsg);
		pa = sync + 64;
		pb = pa + (size_t)pm * kWave * 8;
		rb = pb + (size_t)pm * kWave * 8;
		edges = rb + (size_t)pm * kWave * 8;
		dx = edges + (size_t)pm * 16;
		total = dx + (size_t)((dim + 1) & ~1) * kWave * 8;
	}
};

__device__ __forceinline__ int a_acquire(const int* p)
{
	return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ void a_release(int* p, int v)
{
	__hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ bool a_wait(const int* p, int target, ASync* sy, uint32_t* status)
{
	for (int spin = 0; spin < kASpinLimit; spin++)
	{
		if (uniform(a_acquire(p)) >= target) return true;
		if (uniform(a_acquire(&sy->abort))) return false;
		__builtin_amdgcn_s_sleep(2);
	}
	a_release(&sy->abort, 1);
	if ((threadIdx.x & 63) == 0) atomicOr(status, (uint32_t)HDDM_LB_STATUS_WAIT_LIMIT);
	return false;
}

// x[dim] of the lane's point, read from the coordinate banks pinned in
// v64.. (where the assembly walk keeps them) with GPR indexing; idx2 is
// twice the dim (the op's dim field). m0 is saved and restored.
template <int XB>
__device__ __forceinline__ double xpin(const d16_t& xv0, const d16_t& xv1, const d16_t& xv2, const d16_t& xv3,
	uint32_t idx2)
{
	double v;
	uint32_t sv;
#define HDDM_XPIN "s_mov_b32 %[sv], m0\n" "s_set_gpr_idx_on %[i], gpr_idx(SRC0)\n" "v_mov_b64 %[v], v[64:65]\n" \
	"s_set_gpr_idx_off\n" "s_mov_b32 m0, %[sv]\n"
	if constexpr (XB == 1)
		asm volatile(HDDM_XPIN : [v] "=&v"(v), [sv] "=&s"(sv) : [i] "s"(idx2), "{v[64:95]}"(xv0));
	else if constexpr (XB == 2)
		asm volatile(HDDM_XPIN : [v] "=&v"(v), [sv] "=&s"(sv) : [i] "s"(idx2), "{v[64:95]}"(xv0), "{v[96:127]}"(xv1));
	else
		asm volatile(HDDM_XPIN : [v] "=&v"(v), [sv] "=&s"(sv) : [i] "s"(idx2), "{v[64:95]}"(xv0), "{v[96:127]}"(xv1),
			"{v[128:159]}"(xv2), "{v[160:191]}"(xv3));
#undef HDDM_XPIN
	(void)xv1; (void)xv2; (void)xv3;
	return v;
}

// phi'/phi of a factor: -i sign(u) / phi inside the support (u = x i - j,
// phi = 1 - |u|), 0 outside and at the peak; the hardware reciprocal refined
// by two Newton steps.
__device__ __forceinline__ double dphi_over_phi(double u, double iv)
{
	const double phi = 1.0 - fabs(u);
	const double ph = phi > 0.0 ? phi : 1.0;
	double rc = __builtin_amdgcn_rcp(ph);
	rc = fma(rc, fma(-ph, rc, 1.0), rc);
	rc = fma(rc, fma(-ph, rc, 1.0), rc);
	return phi > 0.0 ? (u > 0.0 ? -iv : (u < 0.0 ? iv : 0.0)) * rc : 0.0;
}

// Walk B: the dx edge-sum walk, paused before emitting a node whose chunk's
// w is not computed yet.
template <int XB, int CB>
struct DxWalk
{
	double* P;       // [PM][64] prefix products
	double* R;       // [PM][64] edge sums
	double* DX;      // [dim][64]
	u4_t* edges;     // [PM] op words {2 dim, i, -j, 0} of the path's edges
	const Op* gops;
	Op nop;          // op q + 1, loaded one step ahead
	int q, q1, top;
	bool top_leaf;
	double r_top, q_top, cur, base_prev;
	uint32_t dim_top;
	Op op;
	bool pending;
	int pend_n, pend_t;
	double pend_p;

	__device__ __forceinline__ void init(int q0, int qend, const Op* ops, int lane)
	{
		gops = ops;
		q = q0;
		q1 = qend;
		top = -1;
		top_leaf = false;
		r_top = q_top = 0.0;
		cur = base_prev = 1.0;
		dim_top = 0;
		pending = false;
		pend_n = pend_t = 0;
		pend_p = 0.0;
		P[lane] = 1.0;
		op = ldc_op(gops + q);
		nop = ldc_op(gops + q + 1);
	}

	__device__ __forceinline__ void leave(const u4_t& e, double r, int tt, int lane,
		const d16_t& xv0, const d16_t& xv1, const d16_t& xv2, const d16_t& xv3)
	{
		const uint32_t d2 = __builtin_amdgcn_readfirstlane(e[0]);
		const double iv = f64_from_high(__builtin_amdgcn_readfirstlane(e[1]));
		const double u = fma(xpin<XB>(xv0, xv1, xv2, xv3, d2), iv, f64_from_high(__builtin_amdgcn_readfirstlane(e[2])));
		lds_add(DX + (d2 >> 1) * kWave + lane, r * dphi_over_phi(u, iv));
		if (tt > 0) lds_add(R + (tt - 1) * kWave + lane, r);
	}

	// Leaves every edge at depth >= t.
	__device__ __forceinline__ void leave_from(int t, int lane,
		const d16_t& xv0, const d16_t& xv1, const d16_t& xv2, const d16_t& xv3)
	{
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
			if (wave_any(r != 0.0)) leave(edges[tt], r, tt, lane, xv0, xv1, xv2, xv3);
		}
	}

	__device__ __forceinline__ void emit(double w, double p, int t, int lane)
	{
		if (top_leaf) r_top = w * p;
		else R[t * kWave + lane] = w * p;
	}

	// Walks until the segment ends (true) or until the next node to emit
	// lies in a chunk above `ready` (false, the emission kept pending).
	template <typename WF>
	__device__ __forceinline__ bool advance(int ready, WF wof, int lane,
		const d16_t& xv0, const d16_t& xv1, const d16_t& xv2, const d16_t& xv3)
	{
		if (pending)
		{
			if ((pend_n >> CB) > ready) return false;
			emit(wof(pend_n), pend_p, pend_t, lane);
			pending = false;
		}
		while (q < q1)
		{
			// the stream ends with kOpPad sentinels, so q + 2 is in bounds;
			// op q + 2 is loaded now and waited for a step later
			const Op nnop = ldc_op(gops + q + 2);
			__builtin_amdgcn_sched_barrier(0);
			const uint32_t dimw = op.dim, ctrl = op.ctrl;
			const uint32_t d2 = dimw & kOpDimMask;
			const int t = (int)(ctrl & 15u);
			leave_from(t, lane, xv0, xv1, xv2, xv3);
			top = t;
			top_leaf = !(ctrl & kOpKids);
			const uint32_t rel = (ctrl >> kOpRelShift) & 3u;
			double bt;
			if (rel == 0u) bt = cur;
			else if (rel == 1u) bt = base_prev;
			else bt = P[t * kWave + lane];
			const double iv = f64_from_high(op.ihi);
			const double u = fma(xpin<XB>(xv0, xv1, xv2, xv3, d2), iv, f64_from_high(op.njhi));
			const double p = bt * fmax(1.0 - fabs(u), 0.0);
			if (top_leaf)
			{
				r_top = 0.0;
				dim_top = d2 >> 1;
			}
			else
			{
				P[(t + 1) * kWave + lane] = p;
				R[t * kWave + lane] = 0.0;
				if (lane == 0) edges[t] = u4_t{ d2, op.ihi, op.njhi, 0u };
			}
			base_prev = bt;
			cur = p;
			if (!wave_any(p != 0.0))
			{
				const int rs = (int)(dimw >> kOpSkipShift);
				q += rs;
				if (rs == 1) { op = nop; nop = nnop; }
				else if (rs == 2) { op = nnop; nop = ldc_op(gops + q + 1); }
				else { op = ldc_op(gops + q); nop = ldc_op(gops + q + 1); }
				continue;
			}
			q++;
			op = nop;
			nop = nnop;
			if (top_leaf) q_top = dphi_over_phi(u, iv);
			if (ctrl & kOpLast)
			{
				const int n = (int)(ctrl >> kOpNodeShift);
				if ((n >> CB) > ready)
				{
					pending = true;
					pend_n = n;
					pend_t = t;
					pend_p = p;
					return false;
				}
				emit(wof(n), p, t, lane);
			}
		}
		return true;
	}
};

// Per-item gradient operands of one consumer (tile of 32 dofs) and its
// per-chunk products.
template <int PREC>
struct AdjCons;

// FP64, v_mfma_f64_16x16x4f64: A[i][k] from lane i + 16 k, B[k][j] from
// lane j + 16 k, D[i][j] in lane j + 16 (i % 4), element i / 4.
//   dS: D[node i][dof j] = sum_k B[node i][point 4 s + k] g[point][dof j]
//   w:  D[point i][node j] = sum_k g[point 16 pb + i][dof 4 s + k] S[node j][dof]
template <>
struct AdjCons<HDDM_LB_FP64>
{
	typedef double d4_t __attribute__((ext_vector_type(4)));
	double gd[2][16];   // g[point 4 s + lane / 16][dof 16 db + lane % 16]
	double gw[4][8];    // g[point 16 pb + lane % 16][dof 4 s + lane / 16]

	template <typename ORIG>
	__device__ __forceinline__ void load(const AdjointArgs& a, ORIG orig, int np, int tile, int dofs, int role, int lane)
	{
		const int col = lane & 15, kq = lane >> 4;
#pragma unroll
		for (int db = 0; db < 2; db++)
#pragma unroll
			for (int s = 0; s < 16; s++)
			{
				const int pt = 4 * s + kq, dof = 32 * tile + 16 * db + col;
				gd[db][s] = (role & kRoleDs) && pt < np && dof < dofs ? a.g[(int64_t)orig(pt) * a.ldg + dof] : 0.0;
			}
#pragma unroll
		for (int pb = 0; pb < 4; pb++)
#pragma unroll
			for (int s = 0; s < 8; s++)
			{
				const int pt = 16 * pb + col, dof = 32 * tile + 4 * s + kq;
				gw[pb][s] = (role & kRoleW) && pt < np && dof < dofs ? a.g[(int64_t)orig(pt) * a.ldg + dof] : 0.0;
			}
	}

	__device__ __forceinline__ void ds(const AdjointArgs& a, const StateDesc* sd, double* dsb, const double* Bc,
		int chunk, uint32_t mask, int tile, int dofs, int lane)
	{
		const int col = lane & 15, kq = lane >> 4;
		const bool on = (mask >> col) & 1u;
		d4_t acc[2] = { { 0.0, 0.0, 0.0, 0.0 }, { 0.0, 0.0, 0.0, 0.0 } };
#pragma unroll
		for (int s = 0; s < 16; s++)
		{
			const double bv = on ? Bc[col * kARow + 4 * s + kq] : 0.0;
			acc[0] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv, gd[0][s], acc[0], 0, 0, 0);
			acc[1] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv, gd[1][s], acc[1], 0, 0, 0);
		}
		const int32_t* order = ldc(&sd->order);
		(void)a;
#pragma unroll
		for (int e = 0; e < 4; e++)
		{
			const int nl = 4 * e + kq;
			if (!((mask >> nl) & 1u)) continue;
			double* row = dsb + (int64_t)ldc(order + chunk * 16 + nl) * dofs;
#pragma unroll
			for (int db = 0; db < 2; db++)
			{
				const int dof = 32 * tile + 16 * db + col;
				if (dof < dofs) unsafeAtomicAdd(row + dof, acc[db][e]);
			}
		}
	}

	__device__ __forceinline__ void w(const StateDesc* sd, double* Wb, int chunk, uint32_t mask, int tile, int DP, int lane)
	{
		const int col = lane & 15, kq = lane >> 4;
		const double* s64 = ldc(&sd->s64);
		const bool on = (mask >> col) & 1u;
		double sv[8];
		const double* srow = s64 + (size_t)(chunk * 16 + col) * DP + 32 * tile + kq;
#pragma unroll
		for (int s = 0; s < 8; s++) sv[s] = on ? srow[4 * s] : 0.0;
		d4_t acc[4];
#pragma unroll
		for (int pb = 0; pb < 4; pb++) acc[pb] = d4_t{ 0.0, 0.0, 0.0, 0.0 };
#pragma unroll
		for (int s = 0; s < 8; s++)
#pragma unroll
			for (int pb = 0; pb < 4; pb++)
				acc[pb] = __builtin_amdgcn_mfma_f64_16x16x4f64(gw[pb][s], sv[s], acc[pb], 0, 0, 0);
		// W[node][point]: node = lane % 16, point = 16 pb + 4 e + lane / 16
#pragma unroll
		for (int pb = 0; pb < 4; pb++)
#pragma unroll
			for (int e = 0; e < 4; e++) Wb[col * kARow + 16 * pb + 4 * e + kq] = acc[pb][e];
	}
};

// F32, v_mfma_f32_32x32x2f32: A[i][k] from lane i + 32 k, B[k][j] from lane
// j + 32 k, D[i][j] in register 4 (i / 8) + i % 4 of lane j + 32 ((i / 4) % 2).
//   dS: D[node i][dof j] = sum over 32 steps of B[node i][point 2 s + k] g[point][dof j]
//   w:  D[point 32 h + i][node j] = sum over 16 steps of g[point][dof 2 s + k] S[node j][dof]
template <>
struct AdjCons<HDDM_LB_F32>
{
	float gd[32];       // g[point 2 s + lane / 32][dof lane % 32]
	float gw[2][16];    // g[point 32 h + lane % 32][dof 2 s + lane / 32]

	template <typename ORIG>
	__device__ __forceinline__ void load(const AdjointArgs& a, ORIG orig, int np, int tile, int dofs, int role, int lane)
	{
		const int col = lane & 31, kh = lane >> 5;
#pragma unroll
		for (int s = 0; s < 32; s++)
		{
			const int pt = 2 * s + kh, dof = 32 * tile + col;
			gd[s] = (role & kRoleDs) && pt < np && dof < dofs ? (float)a.g[(int64_t)orig(pt) * a.ldg + dof] : 0.0f;
		}
#pragma unroll
		for (int h = 0; h < 2; h++)
#pragma unroll
			for (int s = 0; s < 16; s++)
			{
				const int pt = 32 * h + col, dof = 32 * tile + 2 * s + kh;
				gw[h][s] = (role & kRoleW) && pt < np && dof < dofs ? (float)a.g[(int64_t)orig(pt) * a.ldg + dof] : 0.0f;
			}
	}

	__device__ __forceinline__ void ds(const AdjointArgs& a, const StateDesc* sd, double* dsb, const float* Bc,
		int chunk, uint32_t mask, int tile, int dofs, int lane)
	{
		const int col = lane & 31, kh = lane >> 5;
		const bool on = (mask >> col) & 1u;
		f16_t acc;
#pragma unroll
		for (int r = 0; r < 16; r++) acc[r] = 0.0f;
#pragma unroll
		for (int s = 0; s < 32; s++)
		{
			const float bv = on ? Bc[col * kARow + 2 * s + kh] : 0.0f;
			acc = __builtin_amdgcn_mfma_f32_32x32x2f32(bv, gd[s], acc, 0, 0, 0);
		}
		const int32_t* order = ldc(&sd->order);
		(void)a;
		const int dof = 32 * tile + col;
#pragma unroll
		for (int r = 0; r < 16; r++)
		{
			const int nl = 8 * (r >> 2) + 4 * kh + (r & 3);
			if (!((mask >> nl) & 1u) || dof >= dofs) continue;
			unsafeAtomicAdd(dsb + (int64_t)ldc(order + chunk * 32 + nl) * dofs + dof, (double)acc[r]);
		}
	}

	__device__ __forceinline__ void w(const StateDesc* sd, float* Wb, int chunk, uint32_t mask, int tile, int DP, int lane)
	{
		const int col = lane & 31, kh = lane >> 5;
		const float* s32r = ldc(&sd->s32r);
		const bool on = (mask >> col) & 1u;
		float sv[16];
		const float* srow = s32r + (size_t)(chunk * 32 + col) * DP + 32 * tile + kh;
#pragma unroll
		for (int s = 0; s < 16; s++) sv[s] = on ? srow[2 * s] : 0.0f;
		f16_t acc[2];
#pragma unroll
		for (int h = 0; h < 2; h++)
		{
#pragma unroll
			for (int r = 0; r < 16; r++) acc[h][r] = 0.0f;
#pragma unroll
			for (int s = 0; s < 16; s++) acc[h] = __builtin_amdgcn_mfma_f32_32x32x2f32(gw[h][s], sv[s], acc[h], 0, 0, 0);
		}
		// W[node][point]: node = lane % 32, point = 32 h + 8 (r / 4) + 4 (lane / 32) + r % 4
#pragma unroll
		for (int h = 0; h < 2; h++)
#pragma unroll
			for (int r = 0; r < 16; r++) Wb[col * kARow + 32 * h + 8 * (r >> 2) + 4 * kh + (r & 3)] = acc[h][r];
	}
};

} // namespace

template <int PREC, int XB, int PM>
__global__ void __launch_bounds__(64 * (kANC + 1)) adjoint_ws_kernel(AdjointArgs a)
{
	extern __shared__ __attribute__((aligned(16))) double lds[];
	typedef adj_t<PREC> T;
	constexpr int CB = adj_chunk_bits<PREC>();
	constexpr size_t kBuf = adj_buf_bytes<PREC>();
	const ALayout<PREC> L(PM, a.dim);
	char* const lbase = (char*)lds;
	AMsg* const msgs = (AMsg*)(lbase + L.msgs);
	ASync* const sy = (ASync*)(lbase + L.sync);
	const int lane = threadIdx.x & 63;
	const int wave = uniform((int)(threadIdx.x >> 6));
	auto slot = [&](int n) { return (T*)(lbase + L.slots + (size_t)(n & (kARing - 1)) * kBuf); };
	auto wbuf = [&](int par, int c) { return (T*)(lbase + L.wbuf + ((size_t)(par & 1) * kANC + c) * kBuf); };

	if (threadIdx.x == 0)
	{
		sy->produced = 0;
		sy->abort = 0;
		for (int c = 0; c < 6; c++) sy->done[c] = 0;
	}
	__syncthreads();

	if (wave == 0)
	{
		// ---- walker ----
		const int64_t nitems = ldc(&a.hdr->nitems);
		Walker<XB, T, kWalkSmem, PM, PREC == HDDM_LB_FP64, CB, kARow> wa;
		DxWalk<XB, CB> wb;
		wb.P = (double*)(lbase + L.pb);
		wb.R = (double*)(lbase + L.rb);
		wb.edges = (u4_t*)(lbase + L.edges);
		wb.DX = (double*)(lbase + L.dx);
		int npub = 0, nchunks = 0;
		bool ok = true;
		// statistics mode: shader-clock cycles of walk A, walk B, waits for w
		// and waits for ring slots; items and chunks
		const bool st_on = a.stats != nullptr;
		uint64_t cyc_a = 0, cyc_b = 0, cyc_w = 0, cyc_s = 0, st_items = 0, st_chunks = 0;
		auto clk = [&]() -> uint64_t { return st_on ? __builtin_amdgcn_s_memtime() : 0; };
		auto acquire = [&]() -> bool {
			if (npub < kARing) return true;
			const uint64_t t0 = clk();
			bool r = true;
#pragma unroll 1
			for (int c = 0; c < kANC; c++)
				if (!a_wait(&sy->done[c], npub - kARing + 1, sy, a.status)) { r = false; break; }
			cyc_s += clk() - t0;
			return r;
		};
		auto publish = [&](int type, int a0, int b0, int np, int64_t start, int64_t lpos) {
			AMsg* m = msgs + (npub & (kARing - 1));
			if (lane == 0)
			{
				m->type = type; m->a = a0; m->b = b0; m->np = np;
				m->start = start; m->lpos = lpos;
			}
			a_release(&sy->produced, npub + 1);
			npub++;
		};
		// waits until every consumer has handled message m
		auto wait_msg = [&](int m) -> bool {
			const uint64_t t0 = clk();
			bool r = true;
#pragma unroll 1
			for (int c = 0; c < kANC; c++)
				if (!a_wait(&sy->done[c], m + 1, sy, a.status)) { r = false; break; }
			cyc_w += clk() - t0;
			return r;
		};
		for (;;)
		{
			uint64_t v = 0;
			if (lane == 0) v = atomicAdd(&a.hdr->counter, 1ull);
			const int64_t item = (int64_t)uniform64(v);
			if (item >= nitems) break;
			ItemInfo it;
			decode_item(a, item, 1, it);
			const StateDesc* sd = a.states + it.s;
			const int DP = ldc(&sd->dofs_pad);
			const int tpass = min(kANC, (DP >> 5) - a.dt_base);
			if (tpass <= 0) continue;
			const Op* const gops = ldc(&sd->ops);
			const uint32_t* const gskip = ldc(&sd->skip);
			const int32_t* const seg_op = ldc(&sd->seg_op);
			const int q0 = ldc(seg_op + it.seg), q1 = ldc(seg_op + it.seg + 1);
			const int64_t blk = it.grp;
			const int64_t rem = it.count - blk * kWave;
			const int np = rem <= 0 ? 0 : (rem < kWave ? (int)rem : kWave);
			if (np == 0) continue;
			const int myp = lane < np ? lane : np - 1;
			const uint32_t orig = a.perm[it.start + blk * kWave + myp];
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
			d16_t xv0, xv1, xv2, xv3;
#pragma unroll
			for (int i = 0; i < 16; i++) { xv0[i] = xb0[i]; xv1[i] = xb1[i]; xv2[i] = xb2[i]; xv3[i] = xb3[i]; }
			// w of node n: the tiles' buffers of the chunk, summed in consumer order
			const bool split = tpass <= kANC / 2;
			int wpar = 0;
			auto wof = [&](int n) -> double {
				const int nl = n & ((1 << CB) - 1);
				double s = 0.0;
#pragma unroll
				for (int c = 0; c < kANC; c++)
				{
					const bool wc = split ? ((c & 1) && (c >> 1) < tpass) : c < tpass;
					if (wc) s += (double)wbuf(wpar, c)[nl * kARow + lane];
				}
				return s;
			};
			if (!(ok = acquire())) break;
			publish(kAMsgItem, it.s, it.seg, np, it.start, blk * kWave);
#pragma unroll
			for (int d = 0; d < ((a.dim + 1) & ~1); d++) wb.DX[d * kWave + lane] = 0.0;
			wa.init(q0, q1, (double*)(lbase + L.pa), slot(npub), lane, gops, gskip);
			wb.init(q0, q1, gops, lane);
			int prev_chunk = -1, prev_msg = -1, prev_par = 0;
			const bool want_dx = a.dx != nullptr;
			// walk B through the previously published chunk, once its w is ready
			auto catch_up = [&]() -> bool {
				if (prev_chunk < 0 || !want_dx) return true;
				if (!wait_msg(prev_msg)) return false;
				wpar = prev_par;
				const uint64_t t0 = clk();
				wb.advance(prev_chunk, wof, lane, xv0, xv1, xv2, xv3);
				cyc_b += clk() - t0;
				return true;
			};
			for (;;)
			{
				int node = 0;
				double val = 0.0;
				const uint64_t ta = clk();
				const bool more = wa.run(nullptr, nullptr, q0, q1, lane, xb0, xb1, xb2, xb3, node, val);
				cyc_a += clk() - ta;
				if (wa.chunk >= 0)
				{
					st_chunks++;
					const int m = npub, par = nchunks & 1;
					publish(kAMsgChunk, wa.chunk, (int)wa.mask, 0, 0, 0);
					nchunks++;
					if (!(ok = catch_up())) break;
					prev_chunk = wa.chunk;
					prev_msg = m;
					prev_par = par;
					wa.chunk = -1;
				}
				if (!more) break;
				if (!(ok = acquire())) break;
				wa.B = slot(npub);
				wa.start_chunk(node, val, lane);
			}
			if (!ok) break;
			if (!want_dx) continue;
			if (prev_chunk >= 0 && !(ok = wait_msg(prev_msg))) break;
			// walk B ends with walk A's last chunk: a node left to emit beyond
			// it means the two walks disagreed
			wpar = prev_par;
			const uint64_t tb = clk();
			if (!wb.advance(prev_chunk, wof, lane, xv0, xv1, xv2, xv3))
				if (lane == 0) atomicOr(a.status, (uint32_t)HDDM_LB_STATUS_WALK_MISMATCH);
			wb.leave_from(0, lane, xv0, xv1, xv2, xv3);
			cyc_b += clk() - tb;
			st_items++;
			if (lane < np)
			{
				double* out = ldc(&sd->nseg) > 1
					? a.dxpart + ldc(&a.plans[it.s].dx_base) + ((int64_t)it.seg * it.count + blk * kWave + lane) * a.dim
					: a.dx + (int64_t)orig * a.lddx;
				for (int d = 0; d < a.dim; d++)
				{
					const double v = wb.DX[d * kWave + lane];
					out[d] = a.dt_base == 0 ? v : out[d] + v;
				}
			}
		}
		if (ok && acquire()) publish(kAMsgExit, 0, 0, 0, 0, 0);
		else a_release(&sy->abort, 1);
		if (st_on && lane == 0)
		{
			atomicAdd(&a.stats[0], (unsigned long long)cyc_a);
			atomicAdd(&a.stats[1], (unsigned long long)cyc_b);
			atomicAdd(&a.stats[2], (unsigned long long)cyc_w);
			atomicAdd(&a.stats[3], (unsigned long long)cyc_s);
			atomicAdd(&a.stats[7], (unsigned long long)st_items);
			atomicAdd(&a.stats[8], (unsigned long long)st_chunks);
		}
		return;
	}

	// ---- consumer ----
	const int c = wave - 1;
	AdjCons<PREC> cons;
	const StateDesc* sd = nullptr;
	bool active = false;
	int role = 0, tile = 0, dofs = 0, DP = 32;
	int nchunks = 0;
	double* dsb = nullptr;
	const bool st_on = a.stats != nullptr;
	uint64_t cyc_ds = 0, cyc_wc = 0, cyc_idle = 0;
	auto clk = [&]() -> uint64_t { return st_on ? __builtin_amdgcn_s_memtime() : 0; };
	for (int n = 0;; n++)
	{
		const uint64_t t0 = clk();
		const bool got = a_wait(&sy->produced, n + 1, sy, a.status);
		cyc_idle += clk() - t0;
		if (!got) break;
		const AMsg* m = msgs + (n & (kARing - 1));
		const int type = uniform(m->type);
		if (type == kAMsgChunk)
		{
			const int par = nchunks & 1;
			nchunks++;
			if (active)
			{
				const int chunk = uniform(m->a);
				const uint32_t mask = (uint32_t)uniform(m->b);
				const T* Bc = slot(n);
				const uint64_t t1 = clk();
				if ((role & kRoleDs) && a.ds) cons.ds(a, sd, dsb, Bc, chunk, mask, tile, dofs, lane);
				const uint64_t t2 = clk();
				if (role & kRoleW) cons.w(sd, wbuf(par, c), chunk, mask, tile, DP, lane);
				cyc_ds += t2 - t1;
				cyc_wc += clk() - t2;
			}
		}
		else if (type == kAMsgItem)
		{
			const int s = uniform(m->a);
			const int np = uniform(m->np);
			const int64_t start = (int64_t)uniform64((uint64_t)m->start);
			const int64_t lpos = (int64_t)uniform64((uint64_t)m->lpos);
			sd = a.states + s;
			dofs = ldc(&sd->dofs);
			DP = ldc(&sd->dofs_pad);
			const int tpass = min(kANC, (DP >> 5) - a.dt_base);
			const bool split = tpass <= kANC / 2;
			const int tl = split ? c >> 1 : c;
			active = tl < tpass;
			role = split ? ((c & 1) ? kRoleW : kRoleDs) : (kRoleDs | kRoleW);
			if (!a.dx) role &= ~kRoleW;
			tile = a.dt_base + tl;
			if (a.ds)
			{
				dsb = a.ds + ldc(&sd->ds_offset);
				if (a.ds_xcd > 0)
				{
					uint32_t xcc;
					asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(xcc));
					dsb += (int64_t)(xcc & 7u) * a.ds_xcd;
				}
			}
			if (active)
			{
				const uint32_t* perm = a.perm + start + lpos;
				auto orig = [&](int pt) -> uint32_t { return perm[pt < np ? pt : 0]; };
				cons.load(a, orig, np, tile, dofs, role, lane);
			}
		}
		else
		{
			a_release(&sy->done[c], n + 1);
			break;
		}
		a_release(&sy->done[c], n + 1);
	}
	if (st_on && lane == 0)
	{
		atomicAdd(&a.stats[4], (unsigned long long)cyc_wc);
		atomicAdd(&a.stats[5], (unsigned long long)cyc_ds);
		atomicAdd(&a.stats[6], (unsigned long long)cyc_idle);
	}
}

size_t adjoint_ws_lds_bytes(int prec, int depth, int dim)
{
	const int pm = depth < 8 ? 8 : 16;
	return prec == HDDM_LB_FP64 ? ALayout<HDDM_LB_FP64>(pm, dim).total : ALayout<HDDM_LB_F32>(pm, dim).total;
}

template <int PREC, int XB>
static const void* adj_ws_pm(int depth)
{
	return depth < 8 ? (const void*)&adjoint_ws_kernel<PREC, XB, 8> : (const void*)&adjoint_ws_kernel<PREC, XB, 16>;
}

template <int PREC>
static const void* adj_ws_xb(int xbanks, int depth)
{
	switch (xbanks)
	{
	case 1: return adj_ws_pm<PREC, 1>(depth);
	case 2: return adj_ws_pm<PREC, 2>(depth);
	case 3: case 4: return adj_ws_pm<PREC, 4>(depth);
	default: return nullptr;
	}
}

const void* adjoint_ws_kernel_ptr(int prec, int xbanks, int depth)
{
	if (prec == HDDM_LB_FP64) return adj_ws_xb<HDDM_LB_FP64>(xbanks, depth);
	if (prec == HDDM_LB_F32) return adj_ws_xb<HDDM_LB_F32>(xbanks, depth);
	return nullptr;
}

} // namespace hddm_lb
