// Wave-specialized adjoint (gfx950): one walker wave, NC = 4 consumer waves
// and two dx waves per workgroup. No counterpart in the reference.
//
// For y_p = sum_k B_k(x_p) S_k and an upstream gradient g_p:
//   dS_k  = sum_p B_k(x_p) g_p                    (dsurplus)
//   dx_pd = sum_k w_pk dB_k/dx_pd,  w_pk = g_p . S_k (dx)
// A work item is one 64-point block of one state against one node segment.
// Items come either one at a time from the work queue, or as owned units:
// a unit (state, segment, copy class k) is all blocks k, k + K, ... of the
// state against the segment, and its workgroup alone adds to those rows of
// copy k of dS, so it adds with plain loads and stores instead of atomics
// (the K copies are summed in a fixed order afterwards, which makes dS
// bitwise repeatable). The walker wave runs the forward's assembly walk
// (walk.h): surviving nodes' basis values go into chunks of an LDS ring (32
// nodes in FP32, 16 in FP64) that are published to the other waves.
// Consumer c takes one 32-dof tile of the surplus (or, when a state has at
// most two tiles in the pass, one of the two products of a tile) and, per
// published chunk, on MFMA:
//   dS   D[node][dof] = sum over the block's points of B[point][node] g[point][dof],
//        added to the unit's copy of dS (owned units) or with atomics to one
//        copy per XCD (copies summed after) or to dS itself;
//   w    W[point][node] = sum over the tile's dofs of g[point][dof] S[node][dof],
//        written to an LDS buffer of its own (two buffers, by the chunk's
//        parity within the item).
// A dx wave takes each chunk of its parity (counted from the item's start)
// once every consumer is done with it. For each surviving node k of the
// chunk it sums the tiles' w in consumer order, reads k's factors phi_f (dim
// d_f, i_f and the centre j_f / i_f) from the per-node factor table,
// recomputes them at the block's points in FP64 from u = (x - j/i) i (the
// walk's fma(x, i, -j) bit for bit: i is a power of two), and adds
//   w_pk (prod_{g != f} phi_g) phi_f'       (prefix and suffix products)
// to dx[d_f] in LDS rows of its own: no division, and a factor that is zero
// contributes exactly zero. Each wave adds in a fixed order and the two
// waves' rows are summed in a fixed order, so dx is bitwise repeatable; it is
// written out when the next item (or the exit) arrives.
// FP64: v_mfma_f64_16x16x4f64 throughout. F32: v_mfma_f32_32x32x16_bf16 on
// operands split into three bf16 parts (six terms per product, as accurate as
// F32 operands and products); dS partials of one chunk (64 points) are added
// in FP64; w is summed over tiles in FP64, or in F32 for F32 buffers when all
// four consumers hold w.
// Synchronization is the forward WS sweep's: an LDS message ring with
// release/acquire counters and bounded waits (sweep_ws.hip). The walker
// reuses a ring slot once all six reader waves are done with it; a consumer
// rewrites its w buffer of one parity only after the dx wave of that parity
// is done with the chunk that last used it.
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
	int type, a, b, np, copy, pad;
	int64_t start, lpos, count;
};

struct ASync
{
	int produced;
	int abort;
	int done[6];
};

template <int PREC> constexpr int adj_chunk_bits() { return PREC == HDDM_LB_FP64 ? 4 : 5; }
template <int PREC> using adj_t = typename std::conditional<PREC == HDDM_LB_FP64, double, float>::type;
// one chunk buffer: nodes x padded row
template <int PREC> constexpr size_t adj_buf_bytes() { return ((size_t)1 << adj_chunk_bits<PREC>()) * kARow * sizeof(adj_t<PREC>); }
// one consumer's w buffer: nodes x 64 points, point p of node n at n * 64 +
// (p ^ s(n)); the swizzle spreads the MFMA layout's column writes over banks
// without padding. FP64: s(n) = n. F32: s(n) = 4 n mod 64, which keeps each
// aligned group of four points together, so a lane writes four of them with
// one 16-byte store (the FP64 consumer writes single values and measures
// slightly faster with s(n) = n).
template <int PREC> constexpr size_t adj_wbuf_bytes() { return ((size_t)1 << adj_chunk_bits<PREC>()) * kWave * sizeof(adj_t<PREC>); }
typedef float f4v_t __attribute__((ext_vector_type(4)));
template <typename T> __device__ __forceinline__ int wpos(int node, int point)
{
	if constexpr (sizeof(T) == 4) return node * kWave + (point ^ ((node << 2) & (kWave - 1)));
	else return node * kWave + (point ^ node);
}

// LDS layout of the workgroup (bytes): kARing chunk slots, 2 x kANC w buffers,
// messages, counters, the walk's path stack [PM][64], ndx sets of dx rows
// [dim][64] (one per dx wave).
template <int PREC>
struct ALayout
{
	size_t slots, wbuf, msgs, sync, pa, dx, total;
	__host__ __device__ ALayout(int pm, int dim, int ndx)
	{
		slots = 0;
		wbuf = slots + kARing * adj_buf_bytes<PREC>();
		msgs = wbuf + 2 * kANC * adj_wbuf_bytes<PREC>();
		sync = msgs + kARing * sizeof(AMsg);
		pa = sync + 64;
		dx = pa + (size_t)pm * kWave * 8;
		total = dx + (size_t)ndx * dim * kWave * 8;
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
// twice the dim (the op's dim field). The indexing mode writes m0, declared
// clobbered (the compiler keeps nothing live in it across these reads).
template <int XB>
__device__ __forceinline__ double xpin(const d16_t& xv0, const d16_t& xv1, const d16_t& xv2, const d16_t& xv3,
	uint32_t idx2)
{
	double v;
#define HDDM_XPIN "s_set_gpr_idx_on %[i], gpr_idx(SRC0)\n" "v_mov_b64 %[v], v[64:65]\n" \
	"s_set_gpr_idx_off\n"
	if constexpr (XB == 1)
		asm volatile(HDDM_XPIN : [v] "=&v"(v) : [i] "s"(idx2), "{v[64:95]}"(xv0) : "m0");
	else if constexpr (XB == 2)
		asm volatile(HDDM_XPIN : [v] "=&v"(v) : [i] "s"(idx2), "{v[64:95]}"(xv0), "{v[96:127]}"(xv1) : "m0");
	else
		asm volatile(HDDM_XPIN : [v] "=&v"(v) : [i] "s"(idx2), "{v[64:95]}"(xv0), "{v[96:127]}"(xv1),
			"{v[128:159]}"(xv2), "{v[160:191]}"(xv3) : "m0");
#undef HDDM_XPIN
	(void)xv1; (void)xv2; (void)xv3;
	return v;
}

// Four coordinates at once (twice the dims in i0..i3), in one indexing run:
// s_set_gpr_idx_idx changes only the index between the moves.
template <int XB>
__device__ __forceinline__ void xpin4(const d16_t& xv0, const d16_t& xv1, const d16_t& xv2, const d16_t& xv3,
	uint32_t i0, uint32_t i1, uint32_t i2, uint32_t i3, double& x0, double& x1, double& x2, double& x3)
{
#define HDDM_XPI0 "s_set_gpr_idx_on %[i0], gpr_idx(SRC0)\n" "v_mov_b64 %[v0], v[64:65]\n"
#define HDDM_XPN(k) "s_set_gpr_idx_idx %[i" #k "]\n" "v_mov_b64 %[v" #k "], v[64:65]\n"
#define HDDM_XPIN4 HDDM_XPI0 HDDM_XPN(1) HDDM_XPN(2) HDDM_XPN(3) "s_set_gpr_idx_off\n"
#define HDDM_XPOUT [v0] "=&v"(x0), [v1] "=&v"(x1), [v2] "=&v"(x2), [v3] "=&v"(x3)
#define HDDM_XPIDX [i0] "s"(i0), [i1] "s"(i1), [i2] "s"(i2), [i3] "s"(i3)
	if constexpr (XB == 1)
		asm volatile(HDDM_XPIN4 : HDDM_XPOUT : HDDM_XPIDX, "{v[64:95]}"(xv0) : "m0");
	else if constexpr (XB == 2)
		asm volatile(HDDM_XPIN4 : HDDM_XPOUT : HDDM_XPIDX, "{v[64:95]}"(xv0), "{v[96:127]}"(xv1) : "m0");
	else
		asm volatile(HDDM_XPIN4 : HDDM_XPOUT : HDDM_XPIDX, "{v[64:95]}"(xv0), "{v[96:127]}"(xv1),
			"{v[128:159]}"(xv2), "{v[160:191]}"(xv3) : "m0");
#undef HDDM_XPIDX
#undef HDDM_XPOUT
#undef HDDM_XPIN4
#undef HDDM_XPI0
#undef HDDM_XPN
	(void)xv1; (void)xv2; (void)xv3;
}

// Two coordinates.
template <int XB>
__device__ __forceinline__ void xpin2(const d16_t& xv0, const d16_t& xv1, const d16_t& xv2, const d16_t& xv3,
	uint32_t i0, uint32_t i1, double& x0, double& x1)
{
#define HDDM_XPI0 "s_set_gpr_idx_on %[i0], gpr_idx(SRC0)\n" "v_mov_b64 %[v0], v[64:65]\n"
#define HDDM_XPN(k) "s_set_gpr_idx_idx %[i" #k "]\n" "v_mov_b64 %[v" #k "], v[64:65]\n"
#define HDDM_XPIN2 HDDM_XPI0 HDDM_XPN(1) "s_set_gpr_idx_off\n"
#define HDDM_XPOUT [v0] "=&v"(x0), [v1] "=&v"(x1)
#define HDDM_XPIDX [i0] "s"(i0), [i1] "s"(i1)
	if constexpr (XB == 1)
		asm volatile(HDDM_XPIN2 : HDDM_XPOUT : HDDM_XPIDX, "{v[64:95]}"(xv0) : "m0");
	else if constexpr (XB == 2)
		asm volatile(HDDM_XPIN2 : HDDM_XPOUT : HDDM_XPIDX, "{v[64:95]}"(xv0), "{v[96:127]}"(xv1) : "m0");
	else
		asm volatile(HDDM_XPIN2 : HDDM_XPOUT : HDDM_XPIDX, "{v[64:95]}"(xv0), "{v[96:127]}"(xv1),
			"{v[128:159]}"(xv2), "{v[160:191]}"(xv3) : "m0");
#undef HDDM_XPIDX
#undef HDDM_XPOUT
#undef HDDM_XPIN2
#undef HDDM_XPI0
#undef HDDM_XPN
	(void)xv1; (void)xv2; (void)xv3;
}

// Three coordinates.
template <int XB>
__device__ __forceinline__ void xpin3(const d16_t& xv0, const d16_t& xv1, const d16_t& xv2, const d16_t& xv3,
	uint32_t i0, uint32_t i1, uint32_t i2, double& x0, double& x1, double& x2)
{
#define HDDM_XPI0 "s_set_gpr_idx_on %[i0], gpr_idx(SRC0)\n" "v_mov_b64 %[v0], v[64:65]\n"
#define HDDM_XPN(k) "s_set_gpr_idx_idx %[i" #k "]\n" "v_mov_b64 %[v" #k "], v[64:65]\n"
#define HDDM_XPIN3 HDDM_XPI0 HDDM_XPN(1) HDDM_XPN(2) "s_set_gpr_idx_off\n"
#define HDDM_XPOUT [v0] "=&v"(x0), [v1] "=&v"(x1), [v2] "=&v"(x2)
#define HDDM_XPIDX [i0] "s"(i0), [i1] "s"(i1), [i2] "s"(i2)
	if constexpr (XB == 1)
		asm volatile(HDDM_XPIN3 : HDDM_XPOUT : HDDM_XPIDX, "{v[64:95]}"(xv0) : "m0");
	else if constexpr (XB == 2)
		asm volatile(HDDM_XPIN3 : HDDM_XPOUT : HDDM_XPIDX, "{v[64:95]}"(xv0), "{v[96:127]}"(xv1) : "m0");
	else
		asm volatile(HDDM_XPIN3 : HDDM_XPOUT : HDDM_XPIDX, "{v[64:95]}"(xv0), "{v[96:127]}"(xv1),
			"{v[128:159]}"(xv2), "{v[160:191]}"(xv3) : "m0");
#undef HDDM_XPIDX
#undef HDDM_XPOUT
#undef HDDM_XPIN3
#undef HDDM_XPI0
#undef HDDM_XPN
	(void)xv1; (void)xv2; (void)xv3;
}

// phi' of a factor at u = x i - j: -i sign(u) inside the support (phi =
// 1 - |u| > 0), 0 outside it and at the peak. Selects only (no branch).
// ih is the high word of i (i's low word is zero), so -i sign(u) is i with
// its sign bit flipped where u > 0: one xor on the high word, then the select.
__device__ __forceinline__ double dphi(double u, uint32_t ih, double phi)
{
	const uint32_t flip = ~(uint32_t)__double2hiint(u) & 0x80000000u;
	const bool on = (phi > 0.0) & (u != 0.0);
	return __hiloint2double(on ? (int)(ih ^ flip) : 0, 0);
}

// dx terms of the surviving nodes of one chunk at the block's points. The
// chunk's rows of the factor table (1 << CB nodes x FS slots, FS the state's
// stride) are read with one vector load per lane and broadcast with
// v_readlane, so the node loop issues no scalar-memory load. For a node with
// factors phi_f (dims d_f) and weight w,
//   w * (prod_{g < f} phi_g) * (prod_{g > f} phi_g) * phi_f'
// is added to row d_f. The code has no branch: padding slots (i = j = 0)
// evaluate to phi = 1 and phi' = 0 and add an exact zero to row 0, so two
// nodes at a time form one block of independent instructions.
template <int XB, int FS, int NS, int CB>
struct DxNodes
{
	static constexpr int kC = 1 << CB;
	static constexpr int kEPL = kC * FS / kWave;   // table entries per lane
	static_assert(kEPL >= 1, "chunk block smaller than a wave");
	static_assert(FS % 4 == 0 && NS <= FS, "factor slots come in fours; NS of them are read");
	u4_t fv[kEPL];

	__device__ __forceinline__ void load(const u4_t* ftab, int chunk, int lane)
	{
		const u4_t* src = ftab + ((size_t)chunk << CB) * FS + (size_t)lane * kEPL;
#pragma unroll
		for (int e = 0; e < kEPL; e++) fv[e] = src[e];
	}

	// Node nl: its first NS slots (the state's most factors per node).
	__device__ __forceinline__ void node(int nl, double w, double* DX, int lane,
		const d16_t& xv0, const d16_t& xv1, const d16_t& xv2, const d16_t& xv3)
	{
		uint32_t d2[NS];
		double phi[NS], dp[NS];
#pragma unroll
		for (int f0 = 0; f0 < NS; f0 += 4)
		{
			uint32_t ih[4], jh[4], di[4];
#pragma unroll
			for (int k = 0; k < 4; k++)
			{
				const int f = f0 + k;
				if (f < NS)
				{
					const int src = nl * (kWave / kC) + f / kEPL;
					const u4_t& v = fv[f % kEPL];
					d2[f] = di[k] = __builtin_amdgcn_readlane(v[0], src);
					ih[k] = __builtin_amdgcn_readlane(v[1], src);
					jh[k] = __builtin_amdgcn_readlane(v[2], src);
				}
				else
				{
					di[k] = di[0];
					ih[k] = jh[k] = 0u;
				}
			}
			// the group's coordinates: only as many as it has slots (the
			// count is a constant once the loop is unrolled)
			double xs[4];
			const int rem = NS - f0;
			if (rem >= 4) xpin4<XB>(xv0, xv1, xv2, xv3, di[0], di[1], di[2], di[3], xs[0], xs[1], xs[2], xs[3]);
			else if (rem == 3)
			{
				xpin3<XB>(xv0, xv1, xv2, xv3, di[0], di[1], di[2], xs[0], xs[1], xs[2]);
				xs[3] = xs[0];
			}
			else if (rem == 2)
			{
				xpin2<XB>(xv0, xv1, xv2, xv3, di[0], di[1], xs[0], xs[1]);
				xs[2] = xs[3] = xs[0];
			}
			else
			{
				xs[0] = xpin<XB>(xv0, xv1, xv2, xv3, di[0]);
				xs[1] = xs[2] = xs[3] = xs[0];
			}
#pragma unroll
			for (int k = 0; k < 4; k++)
			{
				if (f0 + k >= NS) continue;
				// u = (x - j / i) i: the table holds the centre, so the two
				// operations take their scalar operands directly (the walk's
				// fma(x, i, -j) gives the same value: i is a power of two)
				const double u = (xs[k] + f64_from_high(jh[k])) * f64_from_high(ih[k]);
				phi[f0 + k] = fmax(1.0 - fabs(u), 0.0);
				dp[f0 + k] = dphi(u, ih[k], phi[f0 + k]);
			}
		}
		double suf[NS];
		double acc = w;
#pragma unroll
		for (int f = NS - 1; f >= 0; f--)
		{
			suf[f] = acc;
			acc *= phi[f];
		}
		double pre = 1.0;
#pragma unroll
		for (int f = 0; f < NS; f++)
		{
			// d2 is twice the dim (even: the table's first field), so row
			// d2 / 2 starts at d2 * 32
			lds_add(DX + d2[f] * (kWave / 2) + lane, pre * suf[f] * dp[f]);
			pre *= phi[f];
		}
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

	// own: owned units, this wave is the only writer of the chunk's dS rows
	// in its copy: it loads them (before the MFMAs), adds and stores.
	// Otherwise FP64 atomics.
	__device__ __forceinline__ void ds(const AdjointArgs& a, const StateDesc* sd, double* dsb, const double* Bc,
		int chunk, uint32_t mask, int tile, int dofs, int lane, bool own)
	{
		const int col = lane & 15, kq = lane >> 4;
		const bool on = (mask >> col) & 1u;
		// the chunk's uploaded node indices, one coalesced load (lane i holds
		// node i % 16's), passed to the lanes that need them by shuffles
		const int32_t* order = ldc(&sd->order);
		const int ol = on ? order[chunk * 16 + col] : 0;
		(void)a;
		double* rowp[4];
		double old[4][2];
#pragma unroll
		for (int e = 0; e < 4; e++)
		{
			const int nl = 4 * e + kq;
			const bool live = (mask >> nl) & 1u;
			rowp[e] = dsb + (int64_t)__shfl(ol, nl, 64) * dofs;
#pragma unroll
			for (int db = 0; db < 2; db++)
			{
				const int dof = 32 * tile + 16 * db + col;
				old[e][db] = own && live && dof < dofs ? rowp[e][dof] : 0.0;
			}
		}
		d4_t acc[2] = { { 0.0, 0.0, 0.0, 0.0 }, { 0.0, 0.0, 0.0, 0.0 } };
#pragma unroll
		for (int s = 0; s < 16; s++)
		{
			const double bv = on ? Bc[col * kARow + 4 * s + kq] : 0.0;
			acc[0] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv, gd[0][s], acc[0], 0, 0, 0);
			acc[1] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv, gd[1][s], acc[1], 0, 0, 0);
		}
#pragma unroll
		for (int e = 0; e < 4; e++)
		{
			const int nl = 4 * e + kq;
			if (!((mask >> nl) & 1u)) continue;
#pragma unroll
			for (int db = 0; db < 2; db++)
			{
				const int dof = 32 * tile + 16 * db + col;
				if (dof >= dofs) continue;
				if (own) rowp[e][dof] = old[e][db] + acc[db][e];
				else unsafeAtomicAdd(rowp[e] + dof, acc[db][e]);
			}
		}
	}

	// The surplus rows of the w product, loaded by wrows() before the dS
	// product so that they arrive meanwhile.
	struct WRows { double v[8]; };

	__device__ __forceinline__ void wrows(const StateDesc* sd, int chunk, uint32_t mask, int tile, int DP, int lane,
		WRows& sv)
	{
		const int col = lane & 15, kq = lane >> 4;
		const double* s64 = ldc(&sd->s64);
		const bool on = (mask >> col) & 1u;
		const double* srow = s64 + (size_t)(chunk * 16 + col) * DP + 32 * tile + kq;
#pragma unroll
		for (int s = 0; s < 8; s++) sv.v[s] = on ? srow[4 * s] : 0.0;
	}

	__device__ __forceinline__ void w(const WRows& sr, double* Wb, int lane)
	{
		const int col = lane & 15, kq = lane >> 4;
		const double (&sv)[8] = sr.v;
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
			for (int e = 0; e < 4; e++) Wb[wpos<double>(col, 16 * pb + 4 * e + kq)] = acc[pb][e];
	}
};

// F32 through bf16 parts, v_mfma_f32_32x32x16_bf16. Every operand is split
// as x = x0 + x1 + x2, bf16 parts each rounded to nearest, so that
// |x - x0 - x1 - x2| <= 2^-24 |x|; a product accumulates (in F32) the six
// terms xa_p xb_q with p + q <= 2, smallest first, and the three left out are
// below 2^-23 of it. The bf16 MFMA does 16x the F32 one's flops per cycle
// (32,768 in 32 cycles against 4,096 in 64), so the six terms take 192
// cycles per 16 steps of k where F32 takes 512.
// A[i][k] from element k % 8 of lane i + 32 (k / 8), B[k][j] from element
// k % 8 of lane j + 32 (k / 8), D[i][j] in register 4 (i / 8) + i % 4 of lane
// j + 32 ((i / 4) % 2).
//   dS: D[node i][dof j] = sum over 4 steps of B[node i][point 16 s + k] g[point][dof j]
//   w:  D[point 32 h + i][node j] = sum over 2 steps of g[point][dof 16 s + k] S[node j][dof]
__device__ __forceinline__ void split3_bf16(const float (&v)[8], b8_t (&t)[3])
{
	uint32_t h[4], m[4], l[4];
#pragma unroll
	for (int p = 0; p < 4; p++)
	{
		h[p] = pack_b2(v[2 * p], v[2 * p + 1]);
		const float r0 = v[2 * p] - lo_half_bf16(h[p]), r1 = v[2 * p + 1] - hi_half_bf16(h[p]);
		m[p] = pack_b2(r0, r1);
		l[p] = pack_b2(r0 - lo_half_bf16(m[p]), r1 - hi_half_bf16(m[p]));
	}
	t[0] = as_vec8<b8_t>(h[0], h[1], h[2], h[3]);
	t[1] = as_vec8<b8_t>(m[0], m[1], m[2], m[3]);
	t[2] = as_vec8<b8_t>(l[0], l[1], l[2], l[3]);
}

// acc += a * b over the six terms of order <= 2 of the parts, smallest first
__device__ __forceinline__ f16_t mfma_bf16x3(const b8_t (&a)[3], const b8_t (&b)[3], f16_t acc)
{
	acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[2], b[0], acc, 0, 0, 0);
	acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[1], b[1], acc, 0, 0, 0);
	acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[0], b[2], acc, 0, 0, 0);
	acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[1], b[0], acc, 0, 0, 0);
	acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[0], b[1], acc, 0, 0, 0);
	return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[0], b[0], acc, 0, 0, 0);
}

template <>
struct AdjCons<HDDM_LB_F32>
{
	b8_t gd[4][3];      // g[point 16 s + 8 (lane / 32) + e][dof lane % 32], parts
	b8_t gw[2][2][3];   // g[point 32 h + lane % 32][dof 16 s + 8 (lane / 32) + e], parts

	template <typename ORIG>
	__device__ __forceinline__ void load(const AdjointArgs& a, ORIG orig, int np, int tile, int dofs, int role, int lane)
	{
		const int col = lane & 31, kh = lane >> 5;
#pragma unroll
		for (int s = 0; s < 4; s++)
		{
			float v[8];
#pragma unroll
			for (int e = 0; e < 8; e++)
			{
				const int pt = 16 * s + 8 * kh + e, dof = 32 * tile + col;
				v[e] = (role & kRoleDs) && pt < np && dof < dofs ? (float)a.g[(int64_t)orig(pt) * a.ldg + dof] : 0.0f;
			}
			split3_bf16(v, gd[s]);
		}
#pragma unroll
		for (int h = 0; h < 2; h++)
#pragma unroll
			for (int s = 0; s < 2; s++)
			{
				float v[8];
#pragma unroll
				for (int e = 0; e < 8; e++)
				{
					const int pt = 32 * h + col, dof = 32 * tile + 16 * s + 8 * kh + e;
					v[e] = (role & kRoleW) && pt < np && dof < dofs ? (float)a.g[(int64_t)orig(pt) * a.ldg + dof] : 0.0f;
				}
				split3_bf16(v, gw[h][s]);
			}
	}

	// own: as in the FP64 consumer, the owned rows are loaded (before the
	// MFMAs), added to and stored; otherwise FP64 atomics.
	__device__ __forceinline__ void ds(const AdjointArgs& a, const StateDesc* sd, double* dsb, const float* Bc,
		int chunk, uint32_t mask, int tile, int dofs, int lane, bool own)
	{
		const int col = lane & 31, kh = lane >> 5;
		const bool on = (mask >> col) & 1u;
		// the chunk's uploaded node indices, one coalesced load (lane i holds
		// node i % 32's), passed to the lanes that need them by shuffles
		const int32_t* order = ldc(&sd->order);
		const int ol = on ? order[chunk * 32 + col] : 0;
		(void)a;
		const int dof = 32 * tile + col;
		double old[16];
		if (own)
		{
#pragma unroll
			for (int r = 0; r < 16; r++)
			{
				const int nl = 8 * (r >> 2) + 4 * kh + (r & 3);
				const bool live = ((mask >> nl) & 1u) && dof < dofs;
				const int o = __shfl(ol, nl, 64);
				old[r] = live ? dsb[(int64_t)o * dofs + dof] : 0.0;
			}
		}
		f16_t acc;
#pragma unroll
		for (int r = 0; r < 16; r++) acc[r] = 0.0f;
#pragma unroll
		for (int s = 0; s < 4; s++)
		{
			float bv[8];
#pragma unroll
			for (int e = 0; e < 8; e++) bv[e] = on ? Bc[col * kARow + 16 * s + 8 * kh + e] : 0.0f;
			b8_t bp[3];
			split3_bf16(bv, bp);
			acc = mfma_bf16x3(bp, gd[s], acc);
		}
#pragma unroll
		for (int r = 0; r < 16; r++)
		{
			const int nl = 8 * (r >> 2) + 4 * kh + (r & 3);
			const int o = __shfl(ol, nl, 64);
			if (!((mask >> nl) & 1u) || dof >= dofs) continue;
			double* p = dsb + (int64_t)o * dofs + dof;
			if (own) *p = old[r] + (double)acc[r];
			else unsafeAtomicAdd(p, (double)acc[r]);
		}
	}

	// The surplus rows of the w product, loaded by wrows() before the dS
	// product so that they arrive meanwhile.
	// S[node][dof 32 tile + 16 s + 8 (lane / 32) + e], e = 0..7: two 16-byte
	// loads per step (rows of dofs_pad, a multiple of 32, floats)
	struct WRows { f4v_t v[2][2]; };

	__device__ __forceinline__ void wrows(const StateDesc* sd, int chunk, uint32_t mask, int tile, int DP, int lane,
		WRows& sv)
	{
		const int col = lane & 31, kh = lane >> 5;
		const float* s32r = ldc(&sd->s32r);
		const bool on = (mask >> col) & 1u;
		const f4v_t* srow = (const f4v_t*)(s32r + (size_t)(chunk * 32 + col) * DP + 32 * tile + 8 * kh);
#pragma unroll
		for (int s = 0; s < 2; s++)
#pragma unroll
			for (int q = 0; q < 2; q++) sv.v[s][q] = on ? srow[4 * s + q] : f4v_t{ 0.0f, 0.0f, 0.0f, 0.0f };
	}

	__device__ __forceinline__ void w(const WRows& sr, float* Wb, int lane)
	{
		const int col = lane & 31, kh = lane >> 5;
		b8_t sp[2][3];
#pragma unroll
		for (int s = 0; s < 2; s++)
		{
			float v[8];
#pragma unroll
			for (int e = 0; e < 8; e++) v[e] = sr.v[s][e >> 2][e & 3];
			split3_bf16(v, sp[s]);
		}
		f16_t acc[2];
#pragma unroll
		for (int h = 0; h < 2; h++)
		{
#pragma unroll
			for (int r = 0; r < 16; r++) acc[h][r] = 0.0f;
#pragma unroll
			for (int s = 0; s < 2; s++) acc[h] = mfma_bf16x3(gw[h][s], sp[s], acc[h]);
		}
		// W[node][point]: node = lane % 32, point = 32 h + 8 (r / 4) + 4 (lane / 32) + r % 4;
		// the four r of one r / 4 are four consecutive points, one store
#pragma unroll
		for (int h = 0; h < 2; h++)
#pragma unroll
			for (int q = 0; q < 4; q++)
				*(f4v_t*)(Wb + wpos<float>(col, 32 * h + 8 * q + 4 * kh)) =
					f4v_t{ acc[h][4 * q], acc[h][4 * q + 1], acc[h][4 * q + 2], acc[h][4 * q + 3] };
	}
};

} // namespace

// Owned units: (state, node segment, copy class k) for the states with points,
// in cost order, each state's units segment-major; unit (s, g, k) covers the
// blocks k, k + K, k + 2K, ... of state s against segment g and adds dS to
// copy k. No two units share a (copy, row), so one workgroup at a time adds
// to each row of a copy. False past the last unit.
__device__ __forceinline__ bool decode_unit(const AdjointArgs& a, int64_t unit, int K, int& s, int& seg, int& k)
{
	int64_t base = 0;
	for (int i = 0; i < a.nstates; i++)
	{
		const int si = ldc(&a.cost_order[i]);
		if (ldc(&a.plans[si].nblk) == 0) continue;
		const int64_t nu = (int64_t)ldc(&a.states[si].nseg) * K;
		if (unit < base + nu)
		{
			const int64_t local = unit - base;
			s = si;
			seg = (int)(local / K);
			k = (int)(local - (int64_t)seg * K);
			return true;
		}
		base += nu;
	}
	return false;
}

// Loads the block's coordinates into the four 16-dim banks (the pinned
// registers the walk and xpin read).
template <int XB>
__device__ __forceinline__ void load_xbanks(const double* xrow, int dim, double (&xb0)[16], double (&xb1)[16],
	double (&xb2)[16], double (&xb3)[16])
{
#pragma unroll
	for (int d = 0; d < 16; d++)
	{
		xb0[d] = xrow[d < dim ? d : 0];
		xb1[d] = XB > 1 ? xrow[16 + d < dim ? 16 + d : 0] : 0.0;
		xb2[d] = XB > 2 ? xrow[32 + d < dim ? 32 + d : 0] : 0.0;
		xb3[d] = XB > 2 ? xrow[48 + d < dim ? 48 + d : 0] : 0.0;
	}
}

template <int PREC, int XB, int PM>
__global__ void __launch_bounds__(64 * (kANC + 3)) adjoint_ws_kernel(AdjointArgs a)
{
	extern __shared__ __attribute__((aligned(16))) double lds[];
	typedef adj_t<PREC> T;
	constexpr int CB = adj_chunk_bits<PREC>();
	constexpr size_t kBuf = adj_buf_bytes<PREC>();
	constexpr size_t kWBuf = adj_wbuf_bytes<PREC>();
	constexpr int kDxWave = kANC + 1;    // first of the two dx waves
	constexpr int kReaders = kANC + 2;   // consumers and dx waves
	const int ndx = a.ndx;               // dx waves with work: 2, or 1 when two sets of rows exceed the LDS
	const ALayout<PREC> L(PM, a.dim, ndx);
	char* const lbase = (char*)lds;
	AMsg* const msgs = (AMsg*)(lbase + L.msgs);
	ASync* const sy = (ASync*)(lbase + L.sync);
	const int lane = threadIdx.x & 63;
	const int wave = uniform((int)(threadIdx.x >> 6));
	auto slot = [&](int n) { return (T*)(lbase + L.slots + (size_t)(n & (kARing - 1)) * kBuf); };
	auto wbuf = [&](int par, int c) { return (T*)(lbase + L.wbuf + ((size_t)(par & 1) * kANC + c) * kWBuf); };
	// the dx wave that takes the chunks of one parity, and its counter
	auto dxw = [&](int par) { return ndx == 2 ? (par & 1) : 0; };
	const bool st_on = a.stats != nullptr;
	auto clk = [&]() -> uint64_t { return st_on ? __builtin_amdgcn_s_memtime() : 0; };

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
		int npub = 0;
		bool ok = true;
		// statistics mode: shader-clock cycles of the walk and of waits for
		// ring slots; items and chunks
		uint64_t cyc_a = 0, cyc_s = 0, st_items = 0, st_chunks = 0;
		auto acquire = [&]() -> bool {
			if (npub < kARing) return true;
			const uint64_t t0 = clk();
			bool r = true;
#pragma unroll 1
			for (int c = 0; c < kReaders; c++)
				if (!a_wait(&sy->done[c], npub - kARing + 1, sy, a.status)) { r = false; break; }
			cyc_s += clk() - t0;
			return r;
		};
		auto publish = [&](int type, int a0, int b0, int np, int64_t start, int64_t lpos, int64_t count, int copy) {
			AMsg* m = msgs + (npub & (kARing - 1));
			if (lane == 0)
			{
				m->type = type; m->a = a0; m->b = b0; m->np = np; m->copy = copy;
				m->start = start; m->lpos = lpos; m->count = count;
			}
			a_release(&sy->produced, npub + 1);
			npub++;
		};
		// owned units (K > 0): the unit in hand and its next block
		const int K = a.ds_own;
		int u_s = 0, u_seg = 0, u_k = 0;
		int64_t u_blk = 0, u_nblk = 0, u_start = 0, u_count = 0;
		for (;;)
		{
			int s, seg, copy;
			int64_t blk, start, count;
			if (K > 0)
			{
				if (u_blk >= u_nblk)
				{
					bool found = false;
					for (;;)
					{
						uint64_t v = 0;
						if (lane == 0) v = atomicAdd(&a.hdr->counter, 1ull);
						if (!decode_unit(a, (int64_t)uniform64(v), K, u_s, u_seg, u_k)) break;
						u_nblk = ldc(&a.plans[u_s].nblk);
						if (u_k < u_nblk) { found = true; break; }
					}
					if (!found) break;
					u_blk = u_k;
					u_start = ldc(&a.plans[u_s].start);
					u_count = ldc(&a.plans[u_s].count);
				}
				s = u_s; seg = u_seg; copy = u_k; blk = u_blk; start = u_start; count = u_count;
				u_blk += K;
			}
			else
			{
				uint64_t v = 0;
				if (lane == 0) v = atomicAdd(&a.hdr->counter, 1ull);
				const int64_t item = (int64_t)uniform64(v);
				if (item >= nitems) break;
				ItemInfo it;
				decode_item(a, item, 1, it);
				s = it.s; seg = it.seg; copy = 0; blk = it.grp; start = it.start; count = it.count;
			}
			const StateDesc* sd = a.states + s;
			const int DP = ldc(&sd->dofs_pad);
			const int tpass = min(kANC, (DP >> 5) - a.dt_base);
			if (tpass <= 0) continue;
			const Op* const gops = ldc(&sd->ops);
			const uint32_t* const gskip = ldc(&sd->skip);
			const int32_t* const seg_op = ldc(&sd->seg_op);
			const int q0 = ldc(seg_op + seg), q1 = ldc(seg_op + seg + 1);
			const int64_t rem = count - blk * kWave;
			const int np = rem <= 0 ? 0 : (rem < kWave ? (int)rem : kWave);
			if (np == 0) continue;
			const int myp = lane < np ? lane : np - 1;
			const uint32_t orig = a.perm[start + blk * kWave + myp];
			double xb0[16], xb1[16], xb2[16], xb3[16];
			load_xbanks<XB>(a.x + (int64_t)orig * a.ldx, a.dim, xb0, xb1, xb2, xb3);
			if (!(ok = acquire())) break;
			publish(kAMsgItem, s, seg, np, start, blk * kWave, count, copy);
			wa.init(q0, q1, (double*)(lbase + L.pa), slot(npub), lane, gops, gskip);
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
					publish(kAMsgChunk, wa.chunk, (int)wa.mask, 0, 0, 0, 0, 0);
					wa.chunk = -1;
				}
				if (!more) break;
				if (!(ok = acquire())) break;
				wa.B = slot(npub);
				wa.start_chunk(node, val, lane);
			}
			if (!ok) break;
			st_items++;
		}
		if (ok && acquire()) publish(kAMsgExit, 0, 0, 0, 0, 0, 0, 0);
		else a_release(&sy->abort, 1);
		if (st_on && lane == 0)
		{
			atomicAdd(&a.stats[0], (unsigned long long)cyc_a);
			atomicAdd(&a.stats[3], (unsigned long long)cyc_s);
			atomicAdd(&a.stats[7], (unsigned long long)st_items);
			atomicAdd(&a.stats[8], (unsigned long long)st_chunks);
		}
		return;
	}

	if (wave >= kDxWave)
	{
		// ---- dx waves ----
		// Wave k takes the chunks of parity k (all chunks when ndx == 1)
		// into rows of its own; at the end of an item wave 0 writes out the
		// sum of both waves' rows (its own first), once wave 1 is done with
		// the item, and wave 1 clears its rows only after that.
		const int k = wave - kDxWave;
		double* const DX = (double*)(lbase + L.dx) + (size_t)(ndx == 2 ? k : 0) * a.dim * kWave;
		double* const DX1 = (double*)(lbase + L.dx) + (size_t)a.dim * kWave;
		int* const mydone = &sy->done[kANC + k];
		const bool want_dx = a.dx != nullptr && k < ndx;
		uint64_t cyc_dx = 0, cyc_wait = 0;
		int nitem_chunks = 0;
		// the item whose rows DX holds
		bool have = false;
		int np = 0, seg = 0, tpass = 0, fs = 4, nfac = 4;
		int64_t count = 0, lpos = 0;
		uint32_t orig = 0;
		const StateDesc* sd = nullptr;
		const u4_t* ftab = nullptr;
		bool split = false;
		d16_t xv0, xv1, xv2, xv3;
		// consumers whose buffers hold w for the current item (bit c)
		uint32_t wcons = 0;
		// w of node nl: the tiles' buffers summed in consumer order. rd()
		// issues the four reads; wsum() adds them (an idle consumer's value
		// selected away, which adds an exact zero). With every consumer
		// holding w (all4, the common case of states of four or more tiles)
		// the selects are left out; the order of the additions is the same.
		static_assert(kANC == 4, "the all-consumer sum is written for four consumers");
		auto rd = [&](int par, int nl, T (&v)[kANC]) {
			const T* b = wbuf(par, 0) + wpos<T>(nl, lane);
#pragma unroll
			for (int c = 0; c < kANC; c++) v[c] = b[c * (int)(kWBuf / sizeof(T))];
		};
		auto wsum = [&](const T (&v)[kANC], auto all4) -> double {
			if constexpr (decltype(all4)::value)
			{
				// F32 buffers are summed in F32 (three adds and one
				// conversion instead of four conversions and three FP64 adds)
				if constexpr (sizeof(T) == 4) return (double)(((v[0] + v[1]) + v[2]) + v[3]);
				else return ((v[0] + v[1]) + v[2]) + v[3];
			}
			else
			{
				double w = 0.0;
#pragma unroll
				for (int c = 0; c < kANC; c++) w += ((wcons >> c) & 1u) ? (double)v[c] : 0.0;
				return w;
			}
		};
		// One chunk (message n): its factor rows load while the consumers
		// finish it; then its nodes in order, two at a time (a lone last node
		// is paired with itself at w = 0, which adds zeros). The next pair's
		// w values are read before a pair's terms and summed only after them,
		// so the reads complete behind the FP64 work. False when a wait gave
		// up.
		auto run_chunk = [&](auto&& dn, int chunk, uint32_t mask, int par, int n, auto all4) -> bool {
			dn.load(ftab, chunk, lane);
			const uint64_t t1 = clk();
#pragma unroll 1
			for (int c = 0; c < kANC; c++)
				if (!a_wait(&sy->done[c], n + 1, sy, a.status)) return false;
			const uint64_t t2 = clk();
			cyc_wait += t2 - t1;
			// the next pair of surviving nodes (a lone one paired with itself)
			// and the reads of their w
			auto take = [&](int& n0, int& n1, bool& two, T (&r0)[kANC], T (&r1)[kANC]) {
				n0 = __builtin_ctz(mask);
				mask &= mask - 1u;
				two = mask != 0u;
				n1 = two ? __builtin_ctz(mask) : n0;
				mask &= mask - 1u;
				rd(par, n0, r0);
				rd(par, n1, r1);
			};
			if (!mask) return true;
			int n0, n1;
			bool two;
			T r0[kANC], r1[kANC];
			take(n0, n1, two, r0, r1);
			double w0 = wsum(r0, all4), w1 = two ? wsum(r1, all4) : 0.0;
			for (;;)
			{
				const bool more = mask != 0u;
				int m0 = n0, m1 = n1;
				bool two2 = false;
				if (more) take(m0, m1, two2, r0, r1);
				dn.node(n0, w0, DX, lane, xv0, xv1, xv2, xv3);
				dn.node(n1, w1, DX, lane, xv0, xv1, xv2, xv3);
				if (!more) break;
				// summed after the pair's terms (in the loop body, so the
				// wait counts only the pair's LDS adds issued since the reads)
				n0 = m0; n1 = m1; two = two2;
				w0 = wsum(r0, all4);
				w1 = two ? wsum(r1, all4) : 0.0;
			}
			cyc_dx += clk() - t2;
			return true;
		};
		// End of an item (message n): wave 0 writes out the rows. False when
		// a wait gave up.
		auto item_end = [&](int n) -> bool {
			if (!have || a.dx == nullptr) { have = false; return true; }
			have = false;
			if (k == 0)
			{
				if (ndx == 2)
				{
					const uint64_t t0 = clk();
					const bool got = a_wait(&sy->done[kANC + 1], n, sy, a.status);
					cyc_wait += clk() - t0;
					if (!got) return false;
				}
				if (lane < np)
				{
					double* out = ldc(&sd->nseg) > 1
						? a.dxpart + ldc(&a.plans[(int)(sd - a.states)].dx_base) + ((int64_t)seg * count + lpos + lane) * a.dim
						: a.dx + (int64_t)orig * a.lddx;
					for (int d = 0; d < a.dim; d++)
					{
						double v = DX[d * kWave + lane];
						if (ndx == 2) v += DX1[d * kWave + lane];
						out[d] = a.dt_base == 0 ? v : out[d] + v;
					}
				}
			}
			else if (ndx == 2)
			{
				// wave 0 reads these rows when it writes the item out
				const uint64_t t0 = clk();
				const bool got = a_wait(&sy->done[kANC], n + 1, sy, a.status);
				cyc_wait += clk() - t0;
				if (!got) return false;
			}
			return true;
		};
		for (int n = 0;; n++)
		{
			const uint64_t t0 = clk();
			const bool got = a_wait(&sy->produced, n + 1, sy, a.status);
			cyc_wait += clk() - t0;
			if (!got) break;
			const AMsg* m = msgs + (n & (kARing - 1));
			const int type = uniform(m->type);
			if (type == kAMsgChunk)
			{
				// the dx wave (and w buffer) of a chunk alternates with the
				// chunks of the item, counted from its start: the split of a
				// point's terms between the two waves, and the rounding of
				// their sum, depends only on the item's points (which the
				// stable point order fixes)
				const int par = nitem_chunks & 1;
				nitem_chunks++;
				if (have && want_dx && dxw(par) == k)
				{
					const int chunk = uniform(m->a);
					const uint32_t mask = (uint32_t)uniform(m->b);
					// the state's table stride (4, 8, 16 slots) and the slots
					// its nodes use (2, 3, 4 of 4; 5, 6, 8 of 8)
					bool ok = true;
					if (fs == 4)
					{
						const std::false_type some;
						const std::true_type all;
						if (wcons == 0xFu)
						{
							if (nfac <= 2) ok = run_chunk(DxNodes<XB, 4, 2, CB>(), chunk, mask, par, n, all);
							else if (nfac == 3) ok = run_chunk(DxNodes<XB, 4, 3, CB>(), chunk, mask, par, n, all);
							else ok = run_chunk(DxNodes<XB, 4, 4, CB>(), chunk, mask, par, n, all);
						}
						else if (nfac <= 2) ok = run_chunk(DxNodes<XB, 4, 2, CB>(), chunk, mask, par, n, some);
						else if (nfac == 3) ok = run_chunk(DxNodes<XB, 4, 3, CB>(), chunk, mask, par, n, some);
						else ok = run_chunk(DxNodes<XB, 4, 4, CB>(), chunk, mask, par, n, some);
					}
					else if (fs == 8)
					{
						const std::false_type some;
						if (nfac <= 5) ok = run_chunk(DxNodes<XB, 8, 5, CB>(), chunk, mask, par, n, some);
						else if (nfac == 6) ok = run_chunk(DxNodes<XB, 8, 6, CB>(), chunk, mask, par, n, some);
						else ok = run_chunk(DxNodes<XB, 8, 8, CB>(), chunk, mask, par, n, some);
					}
					// 16 factor slots next to four coordinate banks would spill:
					// the host runs such states on the per-tile kernels
					else if constexpr (XB < 4) ok = run_chunk(DxNodes<XB, 16, 16, CB>(), chunk, mask, par, n, std::false_type());
					if (!ok) break;
				}
			}
			else if (type == kAMsgItem)
			{
				if (!item_end(n)) break;
				nitem_chunks = 0;
				const int s = uniform(m->a);
				seg = uniform(m->b);
				np = uniform(m->np);
				const int64_t start = (int64_t)uniform64((uint64_t)m->start);
				lpos = (int64_t)uniform64((uint64_t)m->lpos);
				count = (int64_t)uniform64((uint64_t)m->count);
				sd = a.states + s;
				const int DP = ldc(&sd->dofs_pad);
				tpass = min(kANC, (DP >> 5) - a.dt_base);
				split = tpass <= kANC / 2;
				wcons = 0;
				for (int c = 0; c < kANC; c++)
					if (split ? ((c & 1) && (c >> 1) < tpass) : c < tpass) wcons |= 1u << c;
				ftab = (const u4_t*)ldc(&sd->ftab);
				fs = (int)ldc(&sd->fstride);
				nfac = ldc(&sd->max_depth);
				have = true;
				if (want_dx)
				{
					const int myp = lane < np ? lane : np - 1;
					orig = a.perm[start + lpos + myp];
					double xb0[16], xb1[16], xb2[16], xb3[16];
					load_xbanks<XB>(a.x + (int64_t)orig * a.ldx, a.dim, xb0, xb1, xb2, xb3);
#pragma unroll
					for (int i = 0; i < 16; i++) { xv0[i] = xb0[i]; xv1[i] = xb1[i]; xv2[i] = xb2[i]; xv3[i] = xb3[i]; }
					for (int d = 0; d < a.dim; d++) DX[d * kWave + lane] = 0.0;
				}
			}
			else
			{
				item_end(n);
				a_release(mydone, n + 1);
				break;
			}
			a_release(mydone, n + 1);
		}
		if (st_on && lane == 0)
		{
			atomicAdd(&a.stats[1], (unsigned long long)cyc_dx);
			atomicAdd(&a.stats[2], (unsigned long long)cyc_wait);
		}
		return;
	}

	// ---- consumer ----
	const int c = wave - 1;
	AdjCons<PREC> cons;
	const StateDesc* sd = nullptr;
	bool active = false;
	int role = 0, tile = 0, dofs = 0, DP = 32;
	int nitem_chunks = 0;
	int wmsg[2] = { -1, -1 };   // message of the last chunk whose w went to each buffer
	double* dsb = nullptr;
	uint64_t cyc_ds = 0, cyc_wc = 0, cyc_idle = 0;
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
			const int par = nitem_chunks & 1;   // as in the dx waves
			nitem_chunks++;
			if (active)
			{
				const int chunk = uniform(m->a);
				const uint32_t mask = (uint32_t)uniform(m->b);
				const T* Bc = slot(n);
				const uint64_t t1 = clk();
				// the w product's surplus rows load during the dS product
				typename AdjCons<PREC>::WRows sr;
				if (role & kRoleW) cons.wrows(sd, chunk, mask, tile, DP, lane, sr);
				if ((role & kRoleDs) && a.ds) cons.ds(a, sd, dsb, Bc, chunk, mask, tile, dofs, lane, a.ds_own > 0);
				const uint64_t t2 = clk();
				if (role & kRoleW)
				{
					// the dx wave of this parity must be done with the buffer's
					// previous chunk
					if (wmsg[par] >= 0 && !a_wait(&sy->done[kANC + dxw(par)], wmsg[par] + 1, sy, a.status)) break;
					wmsg[par] = n;
					cons.w(sr, wbuf(par, c), lane);
				}
				cyc_ds += t2 - t1;
				cyc_wc += clk() - t2;
			}
		}
		else if (type == kAMsgItem)
		{
			nitem_chunks = 0;
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
				if (a.ds_own > 0) dsb += (int64_t)uniform(m->copy) * a.ds_xcd;
				else if (a.ds_xcd > 0)
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

size_t adjoint_ws_lds_bytes(int prec, int depth, int dim, int ndx)
{
	const int pm = depth < 8 ? 8 : 16;
	return prec == HDDM_LB_FP64 ? ALayout<HDDM_LB_FP64>(pm, dim, ndx).total : ALayout<HDDM_LB_F32>(pm, dim, ndx).total;
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
