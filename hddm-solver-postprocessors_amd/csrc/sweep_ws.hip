// Wave-specialized forward sweep (gfx950): one walker wave and NC consumer
// waves per workgroup.
//
// Same math and work items as sweep_kernel (forward.hip): a work item is one
// 64-point block of one state against one node segment. The walker wave walks
// the segment's prefix trie for its block exactly as sweep_kernel does
// (walk.h) and writes the surviving nodes' basis values into 32-node chunks
// (16 on the FP64 path) held in an LDS ring of kRing slots. It never
// contracts: every finished chunk is published to the consumer waves, each of
// which owns one 32-dof tile (or one half of it) of the surplus and contracts
// the chunk on MFMA:
//   FP64    v_mfma_f64_16x16x4f64, FP64 accumulation over the whole item
//   F32     v_mfma_f32_32x32x2f32, each chunk's sum folded into FP64
//   F16X2   v_mfma_f32_32x32x16_f16, hi/lo splits, folded into FP64 per chunk
//   BF16X3  v_mfma_f32_32x32x16_bf16, 3-term splits, folded per chunk
// So a state of up to 4 dof tiles (128 dofs) is walked once instead of once
// per tile, and the walk of chunk n+1 overlaps the contraction of chunk n on
// other waves.
//
// Protocol (all in the workgroup's LDS): the walker fills slot n % kRing,
// writes message n (item start, chunk, or exit) and release-stores
// produced = n + 1; consumer c acquire-loads produced, handles message n and
// release-stores done[c] = n + 1; before reusing a slot for message n the
// walker waits for done[c] >= n - kRing + 1 for every consumer. Every wait is
// bounded: after kSpinLimit polls, or when another wave has given up, a wave
// sets the workgroup's abort flag, records HDDM_LB_STATUS_WAIT_LIMIT in the
// device status word and leaves, so a broken invariant ends the kernel with
// an error instead of a hang. A consumer leaves only on the exit message (or
// an abort), the walker after sending it: every wave reaches the end.
#include "sweep_common.h"
#include "kernels.h"
#include "walk.h"
#include "contract.h"

namespace hddm_lb {

namespace {

constexpr int kRing = 4;
constexpr int kSpinLimit = 1 << 22;

enum : int { kMsgItem = 1, kMsgChunk = 2, kMsgExit = 3 };

// One message of the ring, 32 bytes. Item: a = state, b = segment, np =
// points of the block, start = first sorted position of the state, lpos =
// the block's first point within the state. Chunk: a = chunk, b = mask of
// the positions written.
struct WsMsg
{
	int type, a, b, np;
	int64_t start, lpos;
};

struct WsSync
{
	int produced;
	int abort;
	int done[6];
};

template <int PREC> constexpr int ws_chunk_bits() { return PREC == HDDM_LB_FP64 ? 4 : 5; }
template <int PREC> using ws_bt = typename std::conditional<PREC == HDDM_LB_FP64, double, float>::type;
template <int PREC> constexpr size_t ws_slot_bytes() { return ((size_t)1 << ws_chunk_bits<PREC>()) * kWave * sizeof(ws_bt<PREC>); }
constexpr size_t kWsMsgOff(size_t slot_bytes) { return (size_t)kRing * slot_bytes; }
constexpr size_t kWsSyncOff(size_t slot_bytes) { return kWsMsgOff(slot_bytes) + kRing * sizeof(WsMsg); }
constexpr size_t kWsPathOff(size_t slot_bytes) { return kWsSyncOff(slot_bytes) + 64; }

__device__ __forceinline__ int lds_acquire(const int* p)
{
	return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
}

__device__ __forceinline__ void lds_release(int* p, int v)
{
	__hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Waits until *p >= target. False when the wait gave up (limit reached, or
// another wave of the workgroup gave up first).
__device__ __forceinline__ bool ws_wait(const int* p, int target, WsSync* sy, uint32_t* status)
{
	for (int spin = 0; spin < kSpinLimit; spin++)
	{
		if (uniform(lds_acquire(p)) >= target) return true;
		if (uniform(lds_acquire(&sy->abort))) return false;
		__builtin_amdgcn_s_sleep(2);
	}
	lds_release(&sy->abort, 1);
	if ((threadIdx.x & 63) == 0) atomicOr(status, (uint32_t)HDDM_LB_STATUS_WAIT_LIMIT);
	return false;
}

// One consumer's share of a work item: tile `tile` (dofs 32 tile ..) and the
// point halves (MFMA paths) or dof halves (FP64) selected by hsel (bit h set:
// half h is this consumer's).
template <int PREC>
struct WsAcc;

// FP64: D[dof i][point j] 16x16 tiles, 2 dof blocks x 4 point blocks; lane
// j + 16 (i % 4) holds element i / 4. A[dof][node] from lane dof + 16 node,
// B[node][point] from lane point + 16 node.
template <>
struct WsAcc<HDDM_LB_FP64>
{
	typedef double d4_t __attribute__((ext_vector_type(4)));
	d4_t acc[2][4];

	__device__ __forceinline__ void zero()
	{
#pragma unroll
		for (int db = 0; db < 2; db++)
#pragma unroll
			for (int pb = 0; pb < 4; pb++) acc[db][pb] = d4_t{ 0.0, 0.0, 0.0, 0.0 };
	}

	__device__ __forceinline__ void contract(const StateDesc* sd, const double* Bc, int chunk, uint32_t mask,
		int tile, int hsel, int DP, int lane, uint32_t& nmfma)
	{
		const double* s64 = ldc(&sd->s64);
		const int col = lane & 15, kq = lane >> 4;
		double av[4][2];
#pragma unroll
		for (int s = 0; s < 4; s++)
		{
			const int kn = 4 * s + kq;
			const bool on = (mask >> kn) & 1u;
			const double* row = s64 + (size_t)(chunk * 16 + kn) * DP + 32 * tile + col;
#pragma unroll
			for (int db = 0; db < 2; db++) av[s][db] = on && ((hsel >> db) & 1) ? row[16 * db] : 0.0;
		}
#pragma unroll
		for (int s = 0; s < 4; s++)
		{
			if (((mask >> (4 * s)) & 15u) == 0u) continue;
			const int kn = 4 * s + kq;
			const bool on = (mask >> kn) & 1u;
			double bv[4];
#pragma unroll
			for (int pb = 0; pb < 4; pb++) bv[pb] = on ? Bc[kn * kWave + 16 * pb + col] : 0.0;
#pragma unroll
			for (int db = 0; db < 2; db++)
				if ((hsel >> db) & 1)
				{
#pragma unroll
					for (int pb = 0; pb < 4; pb++)
						acc[db][pb] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[s][db], bv[pb], acc[db][pb], 0, 0, 0);
					nmfma += 4;
				}
		}
	}

	// out(point) -> row pointer of that point's outputs (dofs indexed absolutely)
	template <typename OUT>
	__device__ __forceinline__ void flush(OUT out, const StateDesc* sd, int tile, int hsel, int np, int dofs, int lane)
	{
		(void)sd;
#pragma unroll
		for (int pb = 0; pb < 4; pb++)
		{
			const int qp = 16 * pb + (lane & 15);
			if (qp >= np) continue;
			double* o = out(qp);
#pragma unroll
			for (int db = 0; db < 2; db++)
				if ((hsel >> db) & 1)
#pragma unroll
					for (int e = 0; e < 4; e++)
					{
						const int dof = 32 * tile + 16 * db + 4 * e + (lane >> 4);
						if (dof < dofs) o[dof] = acc[db][pb][e];
					}
		}
	}
};

// MFMA paths: C of v_mfma_f32_32x32x*: register r of lane l holds dof row
// 8 (r / 4) + 4 (l / 32) + r % 4 and point column l % 32 of a 32-point half.
// Each chunk's sum starts from zero and is folded into FP64.
template <int PREC>
struct WsAcc
{
	double acc64[2][16];

	__device__ __forceinline__ void zero()
	{
#pragma unroll
		for (int h = 0; h < 2; h++)
#pragma unroll
			for (int r = 0; r < 16; r++) acc64[h][r] = 0.0;
	}

	__device__ __forceinline__ void contract(const StateDesc* sd, const float* Bc, int chunk, uint32_t mask,
		int tile, int hsel, int DP, int lane, uint32_t& nmfma)
	{
		f16_t acc[2];
#pragma unroll
		for (int h = 0; h < 2; h++)
#pragma unroll
			for (int r = 0; r < 16; r++) acc[h][r] = 0.0f;
		const int pcol = lane & 31, khalf = lane >> 5;
		if constexpr (PREC == HDDM_LB_F32)
		{
			// 16 K-steps of 2 nodes over the 32-node chunk (two surplus chunks of 16)
			SOperands<HDDM_LB_F32, 1> so;
			so.load(sd, chunk, DP, tile, 1, lane);
#pragma unroll
			for (int q = 0; q < 16; q++)
			{
				if (((mask >> (2 * q)) & 3u) == 0u) continue;
				const int kk = 2 * q + khalf;
				const bool on = (mask >> kk) & 1u;
#pragma unroll
				for (int h = 0; h < 2; h++)
					if ((hsel >> h) & 1)
					{
						const float b = on ? Bc[kk * kWave + 32 * h + pcol] : 0.0f;
						acc[h] = __builtin_amdgcn_mfma_f32_32x32x2f32(so.A[0][q], b, acc[h], 0, 0, 0);
						nmfma++;
					}
			}
		}
		else
		{
			// two 16-node sub-chunks, K = 16 per MFMA: lane l holds nodes
			// 8 (l / 32) .. + 7 of the sub-chunk for point l % 32 of the half
			const int nch = ldc(&sd->nchunks);
#pragma unroll
			for (int sub = 0; sub < 2; sub++)
			{
				const uint32_t sm = (mask >> (16 * sub)) & 0xffffu;
				if (sm == 0u || 2 * chunk + sub >= nch) continue;
				SOperands<PREC, 1> so;
				so.load(sd, 2 * chunk + sub, DP, tile, 1, lane);
#pragma unroll
				for (int h = 0; h < 2; h++)
				{
					if (!((hsel >> h) & 1)) continue;
					float f[8];
#pragma unroll
					for (int m = 0; m < 8; m++)
					{
						const int kk = 8 * khalf + m;
						f[m] = (sm >> kk) & 1u ? Bc[(16 * sub + kk) * kWave + 32 * h + pcol] : 0.0f;
					}
					if constexpr (PREC == HDDM_LB_F16X2)
					{
						uint32_t hi[4], lo[4];
#pragma unroll
						for (int m = 0; m < 4; m++)
						{
							const float f0 = f[2 * m] * kBScale, f1 = f[2 * m + 1] * kBScale;
							hi[m] = pack_h2(f0, f1);
							lo[m] = pack_h2(f0 - lo_half_f16(hi[m]), f1 - hi_half_f16(hi[m]));
						}
						const h8_t Bh = as_vec8<h8_t>(hi[0], hi[1], hi[2], hi[3]);
						const h8_t Bl = as_vec8<h8_t>(lo[0], lo[1], lo[2], lo[3]);
						acc[h] = __builtin_amdgcn_mfma_f32_32x32x16_f16(so.A[0][1], Bh, acc[h], 0, 0, 0);
						acc[h] = __builtin_amdgcn_mfma_f32_32x32x16_f16(so.A[0][0], Bl, acc[h], 0, 0, 0);
						acc[h] = __builtin_amdgcn_mfma_f32_32x32x16_f16(so.A[0][0], Bh, acc[h], 0, 0, 0);
						nmfma += 3;
					}
					else
					{
						uint32_t t1[4], t2[4], t3[4];
#pragma unroll
						for (int m = 0; m < 4; m++)
						{
							const float f0 = f[2 * m], f1 = f[2 * m + 1];
							t1[m] = pack_b2(f0, f1);
							const float r0 = f0 - lo_half_bf16(t1[m]), r1 = f1 - hi_half_bf16(t1[m]);
							t2[m] = pack_b2(r0, r1);
							t3[m] = pack_b2(r0 - lo_half_bf16(t2[m]), r1 - hi_half_bf16(t2[m]));
						}
						const b8_t B1 = as_vec8<b8_t>(t1[0], t1[1], t1[2], t1[3]);
						const b8_t B2 = as_vec8<b8_t>(t2[0], t2[1], t2[2], t2[3]);
						const b8_t B3 = as_vec8<b8_t>(t3[0], t3[1], t3[2], t3[3]);
						// smallest products first
						acc[h] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(so.A[0][2], B1, acc[h], 0, 0, 0);
						acc[h] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(so.A[0][1], B2, acc[h], 0, 0, 0);
						acc[h] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(so.A[0][0], B3, acc[h], 0, 0, 0);
						acc[h] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(so.A[0][1], B1, acc[h], 0, 0, 0);
						acc[h] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(so.A[0][0], B2, acc[h], 0, 0, 0);
						acc[h] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(so.A[0][0], B1, acc[h], 0, 0, 0);
						nmfma += 6;
					}
				}
			}
		}
#pragma unroll
		for (int h = 0; h < 2; h++)
			if ((hsel >> h) & 1)
#pragma unroll
				for (int r = 0; r < 16; r++) acc64[h][r] += (double)acc[h][r];
	}

	template <typename OUT>
	__device__ __forceinline__ void flush(OUT out, const StateDesc* sd, int tile, int hsel, int np, int dofs, int lane)
	{
		const double* iscale = ldc(&sd->iscale);
		const int row = lane & 31, khalf = lane >> 5;
#pragma unroll
		for (int h = 0; h < 2; h++)
		{
			if (!((hsel >> h) & 1)) continue;
			const int qp = 32 * h + row;
			if (qp >= np) continue;
			double* o = out(qp);
#pragma unroll
			for (int r = 0; r < 16; r++)
			{
				const int dof = 32 * tile + 8 * (r >> 2) + 4 * khalf + (r & 3);
				if (dof < dofs)
				{
					double v = acc64[h][r];
					if constexpr (PREC == HDDM_LB_F16X2) v *= ldc(iscale + dof) * (1.0 / kBScale);
					o[dof] = v;
				}
			}
		}
	}
};

} // namespace

// NC consumers; HS = 1: consumer c takes half c % 2 of tile c / 2, else all
// of tile c. Tiles are counted from the pass's first tile a.dt_base.
template <int PREC, int XB, int PM, int NC, int HS>
__global__ void __launch_bounds__(64 * (NC + 1)) sweep_ws_kernel(SweepArgs a)
{
	extern __shared__ __attribute__((aligned(16))) double lds[];
	typedef ws_bt<PREC> BT;
	constexpr int CB = ws_chunk_bits<PREC>();
	constexpr size_t kSlot = ws_slot_bytes<PREC>();
	char* const lbase = (char*)lds;
	WsMsg* const msgs = (WsMsg*)(lbase + kWsMsgOff(kSlot));
	WsSync* const sy = (WsSync*)(lbase + kWsSyncOff(kSlot));
	double* const Pl = (double*)(lbase + kWsPathOff(kSlot));
	const int lane = threadIdx.x & 63;
	const int wave = uniform((int)(threadIdx.x >> 6));
	auto slot = [&](int n) { return (BT*)(lbase + (size_t)(n & (kRing - 1)) * kSlot); };

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
		Walker<XB, BT, kWalkSmem, PM, PREC == HDDM_LB_FP64, CB> w;
		int npub = 0;
		bool ok = true;
		uint64_t st_ops = 0, st_prunes = 0, st_items = 0;
		auto acquire = [&]() -> bool {
			if (npub < kRing) return true;
#pragma unroll 1
			for (int c = 0; c < NC; c++)
				if (!ws_wait(&sy->done[c], npub - kRing + 1, sy, a.status)) return false;
			return true;
		};
		auto publish = [&](int type, int a0, int b0, int np, int64_t start, int64_t lpos) {
			WsMsg* m = msgs + (npub & (kRing - 1));
			if (lane == 0)
			{
				m->type = type; m->a = a0; m->b = b0; m->np = np;
				m->start = start; m->lpos = lpos;
			}
			lds_release(&sy->produced, npub + 1);
			npub++;
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
			if (32 * a.dt_base >= DP) continue;
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
			if (!(ok = acquire())) break;
			publish(kMsgItem, it.s, it.seg, np, it.start, blk * kWave);
			w.init(q0, q1, Pl, slot(npub), lane, gops, gskip);
			for (;;)
			{
				int node = 0;
				double val = 0.0;
				if (!w.run(nullptr, nullptr, q0, q1, lane, xb0, xb1, xb2, xb3, node, val)) break;
				if (w.chunk >= 0) publish(kMsgChunk, w.chunk, (int)w.mask, 0, 0, 0);
				if (!(ok = acquire())) break;
				w.B = slot(npub);
				w.start_chunk(node, val, lane);
			}
			if (!ok) break;
			if (w.chunk >= 0)
			{
				if (!(ok = acquire())) break;   // keeps the invariant simple: never needed for a publish
				publish(kMsgChunk, w.chunk, (int)w.mask, 0, 0, 0);
			}
			st_ops += w.nops;
			st_prunes += w.nprunes;
			st_items++;
		}
		if (ok && acquire()) publish(kMsgExit, 0, 0, 0, 0, 0);
		else lds_release(&sy->abort, 1);
		if (a.stats && lane == 0)
		{
			atomicAdd(&a.stats[0], (unsigned long long)st_ops);
			atomicAdd(&a.stats[1], (unsigned long long)st_prunes);
			atomicAdd(&a.stats[3], (unsigned long long)st_items);
		}
		return;
	}

	// ---- consumer ----
	const int c = wave - 1;
	const int tile = a.dt_base + (HS ? c >> 1 : c);
	const int hsel = HS ? 1 << (c & 1) : 3;
	WsAcc<PREC> acc;
	const StateDesc* sd = nullptr;
	bool active = false;
	int np = 0, seg = 0, dofs = 0, DP = 32, s = 0;
	int64_t start = 0, lpos = 0;
	uint32_t st_chunks = 0, st_mfma = 0;
	uint64_t st_nnz = 0;
	auto flush = [&]() {
		if (!active) return;
		const int64_t count = ldc(&a.plans[s].count);
		double* const part = a.partial + ldc(&a.plans[s].partial_base);
		const bool split = ldc(&sd->nseg) > 1;
		auto out = [&](int qp) -> double* {
			if (split) return part + ((int64_t)seg * count + lpos + qp) * DP;
			return a.y + (int64_t)a.perm[start + lpos + qp] * a.ldy;
		};
		acc.flush(out, sd, tile, hsel, np, dofs, lane);
	};
	for (int n = 0;; n++)
	{
		if (!ws_wait(&sy->produced, n + 1, sy, a.status)) break;
		const WsMsg* m = msgs + (n & (kRing - 1));
		const int type = uniform(m->type);
		if (type == kMsgChunk)
		{
			if (active)
			{
				const int chunk = uniform(m->a);
				const uint32_t mask = (uint32_t)uniform(m->b);
				const BT* Bc = slot(n);
				acc.contract(sd, Bc, chunk, mask, tile, hsel, DP, lane, st_mfma);
				if (a.stats)
				{
					st_chunks++;
					// non-zero (point, node) pairs of the chunk, counted once (consumer 0)
					if (c == 0)
						for (uint32_t mm = mask; mm; mm &= mm - 1)
							st_nnz += __builtin_popcountll(
								__builtin_amdgcn_ballot_w64(Bc[__builtin_ctz(mm) * kWave + lane] != (BT)0 && lane < np));
				}
			}
		}
		else
		{
			flush();
			if (type == kMsgExit)
			{
				lds_release(&sy->done[c], n + 1);
				break;
			}
			s = uniform(m->a);
			seg = uniform(m->b);
			np = uniform(m->np);
			start = (int64_t)uniform64((uint64_t)m->start);
			lpos = (int64_t)uniform64((uint64_t)m->lpos);
			sd = a.states + s;
			dofs = ldc(&sd->dofs);
			DP = ldc(&sd->dofs_pad);
			active = 32 * tile < DP;
			acc.zero();
		}
		lds_release(&sy->done[c], n + 1);
	}
	if (a.stats && lane == 0)
	{
		if (c == 0) atomicAdd(&a.stats[2], (unsigned long long)st_chunks);
		atomicAdd(&a.stats[4], (unsigned long long)st_mfma);
		if (c == 0) atomicAdd(&a.stats[5], (unsigned long long)st_nnz);
	}
}

size_t sweep_ws_lds_bytes(int prec, int depth)
{
	const size_t slot = prec == HDDM_LB_FP64 ? ws_slot_bytes<HDDM_LB_FP64>() : ws_slot_bytes<HDDM_LB_F32>();
	return kWsPathOff(slot) + (size_t)(depth < 8 ? 8 : 16) * kWave * 8;
}

template <int PREC, int XB, int PM>
static const void* ws_nc(int nc)
{
	return nc == 2 ? (const void*)&sweep_ws_kernel<PREC, XB, PM, 2, 1> : (const void*)&sweep_ws_kernel<PREC, XB, PM, 4, 0>;
}

template <int PREC, int XB>
static const void* ws_pm(int depth, int nc)
{
	return depth < 8 ? ws_nc<PREC, XB, 8>(nc) : ws_nc<PREC, XB, 16>(nc);
}

template <int PREC>
static const void* ws_xb(int xbanks, int depth, int nc)
{
	switch (xbanks)
	{
	case 1: return ws_pm<PREC, 1>(depth, nc);
	case 2: return ws_pm<PREC, 2>(depth, nc);
	case 3: case 4: return ws_pm<PREC, 4>(depth, nc);
	default: return nullptr;
	}
}

const void* sweep_ws_kernel_ptr(int prec, int xbanks, int depth, int nc)
{
	if (nc != 2 && nc != 4) return nullptr;
	switch (prec)
	{
	case HDDM_LB_FP64: return ws_xb<HDDM_LB_FP64>(xbanks, depth, nc);
	case HDDM_LB_F32: return ws_xb<HDDM_LB_F32>(xbanks, depth, nc);
	case HDDM_LB_F16X2: return ws_xb<HDDM_LB_F16X2>(xbanks, depth, nc);
	case HDDM_LB_BF16X3: return ws_xb<HDDM_LB_BF16X3>(xbanks, depth, nc);
	}
	return nullptr;
}

} // namespace hddm_lb
