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
//                    64-point block x node segment, states in cost order)
//   4. scatter       counting-sort scatter: perm[sorted position] = point
//   5. sweep         persistent waves pop items; one point per lane; the wave
//                    walks the state's prefix trie of nodes, pruning every
//                    subtree whose prefix product is zero on all 64 lanes, and
//                    contracts the surviving nodes against the surplus
//                    (FP64 VALU, or MFMA on 16-node chunks with split operands)
//   6. reduce        sums node-segment partials of split states (fixed order)
#include "device_common.h"
#include "kernels.h"

namespace hddm_lb {

typedef _Float16 h2_t __attribute__((ext_vector_type(2)));
typedef _Float16 h8_t __attribute__((ext_vector_type(8)));
typedef __bf16 b2_t __attribute__((ext_vector_type(2)));
typedef __bf16 b8_t __attribute__((ext_vector_type(8)));
typedef float f16_t __attribute__((ext_vector_type(16)));
typedef uint32_t u4_t __attribute__((ext_vector_type(4)));

__global__ void fill_u32_kernel(uint32_t* p, int64_t n, uint32_t v)
{
	for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (int64_t)gridDim.x * blockDim.x)
		p[e] = v;
}

__global__ void bucket_key_kernel(int64_t n, const double* __restrict__ x, int64_t ldx,
	const int32_t* __restrict__ sid, const uint8_t* __restrict__ state_ok, int nstates,
	int dim, int bbits, uint32_t* __restrict__ bkey, uint32_t* __restrict__ hist)
{
	for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x)
	{
		int s = sid ? sid[p] : 0;
		if (s < 0 || s >= nstates || !state_ok[s]) s = nstates;  // never planned: row left untouched
		uint32_t key = ((uint32_t)s << bbits);
		if (bbits) key |= (uint32_t)morton_key(x + p * ldx, dim, bbits);
		bkey[p] = key;
		atomicAdd(&hist[key], 1u);
	}
}

// One workgroup of 1024 threads.
__global__ void __launch_bounds__(1024) scan_plan_kernel(const uint32_t* __restrict__ hist,
	uint32_t* __restrict__ cursor, int64_t nbuckets, int bbits, int nstates,
	const StateDesc* __restrict__ states, const int32_t* __restrict__ cost_order,
	StatePlan* __restrict__ plans, PlanHeader* __restrict__ hdr)
{
	__shared__ uint32_t part[1024];
	const int t = threadIdx.x;
	const int64_t per = (nbuckets + 1023) / 1024;
	const int64_t b0 = (int64_t)t * per, b1 = b0 + per < nbuckets ? b0 + per : nbuckets;
	uint32_t sum = 0;
	for (int64_t b = b0; b < b1; b++) sum += hist[b];
	part[t] = sum;
	__syncthreads();
	for (int off = 1; off < 1024; off <<= 1)
	{
		uint32_t v = t >= off ? part[t - off] : 0;
		__syncthreads();
		part[t] += v;
		__syncthreads();
	}
	uint32_t run = part[t] - sum;  // exclusive prefix of this thread's range
	for (int64_t b = b0; b < b1; b++) { uint32_t c = hist[b]; cursor[b] = run; run += c; }
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
		int64_t item = 0, pbase = 0;
		for (int i = 0; i < nstates; i++)
		{
			int s = cost_order[i];
			int nseg = states[s].nseg;
			plans[s].item_base = item;
			item += plans[s].nblk * nseg;
			plans[s].partial_base = pbase;
			if (nseg > 1) pbase += (int64_t)nseg * plans[s].count * states[s].dofs_pad;
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

// Trie walk state of one wave.
template <int FMAX>
struct Walker
{
	double stack[FMAX + 1];   // stack[t]: product of the first t factors of the current path
	int skip_end;             // nodes < skip_end are pruned
	bool fresh;               // the next evaluated node must rebuild its whole prefix
};

// Evaluates node n (not pruned by skip_end). Returns false and advances
// w.skip_end when a prefix product is zero on every lane; otherwise b is B_n.
template <int FMAX>
__device__ __forceinline__ bool eval_node(Walker<FMAX>& w, const uint32_t* rec, int rw,
	const Factor* fac, const double* xs, int lane, int n, double& b)
{
	const uint32_t* r = rec + (size_t)n * rw;
	const uint32_t w0 = ldc(r);
	const int F = (int)(w0 & 0xffu);
	int sh = (int)((w0 >> 8) & 0xffu);
	if (w.fresh) { sh = 0; w.fresh = false; }
	b = 1.0;
#pragma unroll
	for (int t = 0; t < FMAX; t++)
	{
		if (t >= sh && t < F)
		{
			const uint32_t fid = ldc(r + 2 + 2 * t);
			const Factor* f = fac + fid;
			const double fi = ldc(&f->i);
			const double fnj = ldc(&f->negj);
			const uint32_t xoff = ldc(&f->xoff);
			const double xv = xs[(xoff >> 3) + lane];
			// 1 - |x*i - j| as a multiply then an add, like the reference
			const double tt = __dadd_rn(__dmul_rn(xv, fi), fnj);
			const double phi = fmax(__dadd_rn(1.0, -fabs(tt)), 0.0);
			const double p = __dmul_rn(w.stack[t], phi);
			w.stack[t + 1] = p;
			b = p;
			if (!wave_any(p != 0.0))
			{
				w.skip_end = (int)ldc(r + 3 + 2 * t);
				return false;
			}
		}
	}
	return true;
}

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

struct ItemInfo
{
	int s, seg;
	int64_t blk, start, count;
	int np;
};

__device__ __forceinline__ bool next_item(const SweepArgs& a, int lane, ItemInfo& it)
{
	uint64_t item = 0;
	if (lane == 0) item = atomicAdd(&a.hdr->counter, 1ull);
	item = uniform64(item);
	const int64_t nitems = ldc(&a.hdr->nitems);
	if ((int64_t)item >= nitems) return false;
	int s = 0;
	int64_t base = 0, nblk = 1;
	for (int i = 0; i < a.nstates; i++)
	{
		const int si = ldc(&a.cost_order[i]);
		const int64_t ib = ldc(&a.plans[si].item_base);
		const int64_t nb = ldc(&a.plans[si].nblk);
		const int ns = ldc(&a.states[si].nseg);
		if ((int64_t)item < ib + nb * ns) { s = si; base = ib; nblk = nb; break; }
	}
	const int64_t local = (int64_t)item - base;
	it.s = s;
	it.seg = (int)(local / nblk);
	it.blk = local - (int64_t)it.seg * nblk;
	it.start = ldc(&a.plans[s].start);
	it.count = ldc(&a.plans[s].count);
	const int64_t rem = it.count - it.blk * kWave;
	it.np = rem < kWave ? (int)rem : kWave;
	return true;
}

} // namespace

template <int PREC, int DT, int FMAX>
__global__ void __launch_bounds__(256) sweep_kernel(SweepArgs a)
{
	extern __shared__ double lds_x[];
	const int lane = threadIdx.x & 63;
	const int wave = threadIdx.x >> 6;
	double* xs = lds_x + (size_t)wave * a.dim * kWave;

	ItemInfo it;
	while (next_item(a, lane, it))
	{
		const StateDesc* sd = a.states + it.s;
		const int myp = lane < it.np ? lane : it.np - 1;
		const uint32_t orig = a.perm[it.start + it.blk * kWave + myp];
		const double* xrow = a.x + (int64_t)orig * a.ldx;
		for (int d = 0; d < a.dim; d++) xs[d * kWave + lane] = xrow[d];

		const int nno = ldc(&sd->nno);
		const int dofs = ldc(&sd->dofs);
		const int DP = ldc(&sd->dofs_pad);
		const int rw = ldc(&sd->rec_words);
		const int seg_len = ldc(&sd->seg_len);
		const int nseg = ldc(&sd->nseg);
		const uint32_t* rec = ldc(&sd->rec);
		const Factor* fac = ldc(&sd->fac);
		const int dt0 = a.dt_base;
		const int ndt = min(DT, (DP >> 5) - dt0);
		if (ndt <= 0) continue;  // the state has no dofs in this pass
		const int n_begin = it.seg * seg_len;
		const int n_end = min(nno, n_begin + seg_len);

		Walker<FMAX> w;
		w.stack[0] = 1.0;
		w.skip_end = n_begin;
		w.fresh = true;

		// Output row of lane's point (or of point q for the MFMA layouts).
		const bool split = nseg > 1;
		double* const part = a.partial + ldc(&a.plans[it.s].partial_base);

		if constexpr (PREC == HDDM_LB_FP64)
		{
			const double* s64 = ldc(&sd->s64);
			double acc[DT * 32];
#pragma unroll
			for (int c = 0; c < DT * 32; c++) acc[c] = 0.0;
			for (int n = n_begin; n < n_end; )
			{
				if (n < w.skip_end) { n = w.skip_end; continue; }
				double b;
				if (!eval_node<FMAX>(w, rec, rw, fac, xs, lane, n, b)) { n = w.skip_end; continue; }
				const double* srow = s64 + (size_t)n * DP + 32 * dt0;
#pragma unroll
				for (int dt = 0; dt < DT; dt++)
					if (dt < ndt)
					{
#pragma unroll
						for (int c = 0; c < 32; c++)
							acc[dt * 32 + c] = fma(b, ldc(srow + dt * 32 + c), acc[dt * 32 + c]);
					}
				n++;
			}
			if (lane < it.np)
			{
				double* out;
				if (split)
					out = part + ((int64_t)it.seg * it.count + it.blk * kWave + lane) * DP;
				else
					out = a.y + (int64_t)orig * a.ldy;
				out += 32 * dt0;
				const int dlim = dofs - 32 * dt0;
#pragma unroll
				for (int dt = 0; dt < DT; dt++)
#pragma unroll
					for (int c = 0; c < 32; c++)
						if (dt * 32 + c < dlim) out[dt * 32 + c] = acc[dt * 32 + c];
			}
		}
		else
		{
			f16_t acc[2][DT];
			double acc64[2][DT][16];
#pragma unroll
			for (int t = 0; t < 2; t++)
#pragma unroll
				for (int dt = 0; dt < DT; dt++)
#pragma unroll
					for (int r = 0; r < 16; r++) { acc[t][dt][r] = 0.0f; acc64[t][dt][r] = 0.0; }
			int cur_win = -1;
			const int row = lane & 31;
			const int khalf = lane >> 5;

			for (int c0 = n_begin; c0 < n_end; c0 += kChunk)
			{
				if (w.skip_end >= c0 + kChunk) { c0 = (w.skip_end & ~(kChunk - 1)) - kChunk; continue; }
				double b[kChunk];
				bool any = false;
#pragma unroll
				for (int k = 0; k < kChunk; k++)
				{
					b[k] = 0.0;
					const int n = c0 + k;
					if (n < n_end && n >= w.skip_end)
					{
						double v;
						if (eval_node<FMAX>(w, rec, rw, fac, xs, lane, n, v)) { b[k] = v; any = true; }
					}
				}
				if (!any) continue;

				const int chunk = c0 / kChunk;
				const int win = chunk >> a.wlog;
				if (win != cur_win)
				{
#pragma unroll
					for (int t = 0; t < 2; t++)
#pragma unroll
						for (int dt = 0; dt < DT; dt++)
#pragma unroll
							for (int r = 0; r < 16; r++) { acc64[t][dt][r] += (double)acc[t][dt][r]; acc[t][dt][r] = 0.0f; }
					cur_win = win;
				}

				if constexpr (PREC == HDDM_LB_F16X2)
				{
					uint32_t hi[8], lo[8];
#pragma unroll
					for (int m = 0; m < 8; m++)
					{
						const float f0 = (float)b[2 * m], f1 = (float)b[2 * m + 1];
						hi[m] = pack_h2(f0, f1);
						lo[m] = pack_h2(f0 - lo_half_f16(hi[m]), f1 - hi_half_f16(hi[m]));
					}
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
					const uint16_t* sc = ldc(&sd->s16) + (size_t)chunk * 2 * DP * kChunk;
#pragma unroll
					for (int dt = 0; dt < DT; dt++)
						if (dt < ndt)
						{
							const int dr = (dt0 + dt) * 32 + row;
							const h8_t Ah = *(const h8_t*)(sc + ((size_t)dr) * kChunk + khalf * 8);
							const h8_t Al = *(const h8_t*)(sc + ((size_t)DP + dr) * kChunk + khalf * 8);
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
					const uint16_t* sc = ldc(&sd->sbf) + (size_t)chunk * 3 * DP * kChunk;
#pragma unroll
					for (int dt = 0; dt < DT; dt++)
						if (dt < ndt)
						{
							const int dr = (dt0 + dt) * 32 + row;
							b8_t A[3];
#pragma unroll
							for (int term = 0; term < 3; term++)
								A[term] = *(const b8_t*)(sc + ((size_t)term * DP + dr) * kChunk + khalf * 8);
#pragma unroll
							for (int t = 0; t < 2; t++)
							{
								// smallest products first
								acc[t][dt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A[2], B[t][0], acc[t][dt], 0, 0, 0);
								acc[t][dt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A[1], B[t][1], acc[t][dt], 0, 0, 0);
								acc[t][dt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A[0], B[t][2], acc[t][dt], 0, 0, 0);
								acc[t][dt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A[1], B[t][0], acc[t][dt], 0, 0, 0);
								acc[t][dt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A[0], B[t][1], acc[t][dt], 0, 0, 0);
								acc[t][dt] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A[0], B[t][0], acc[t][dt], 0, 0, 0);
							}
						}
				}
				else  // HDDM_LB_F32
				{
					float B0[8], B1[8];
#pragma unroll
					for (int q = 0; q < 8; q++)
					{
						auto r = __builtin_amdgcn_permlane32_swap(__builtin_bit_cast(uint32_t, (float)b[2 * q]),
							__builtin_bit_cast(uint32_t, (float)b[2 * q + 1]), false, false);
						B0[q] = __builtin_bit_cast(float, (uint32_t)r[0]);
						B1[q] = __builtin_bit_cast(float, (uint32_t)r[1]);
					}
					const float* sc = ldc(&sd->s32) + (size_t)chunk * DP * kChunk;
#pragma unroll
					for (int dt = 0; dt < DT; dt++)
						if (dt < ndt)
						{
							const float* ap = sc + (((size_t)((dt0 + dt) * 32 + row)) * 2 + khalf) * 8;
							float A[8];
#pragma unroll
							for (int q = 0; q < 8; q++) A[q] = ap[q];
#pragma unroll
							for (int q = 0; q < 8; q++)
							{
								acc[0][dt] = __builtin_amdgcn_mfma_f32_32x32x2f32(A[q], B0[q], acc[0][dt], 0, 0, 0);
								acc[1][dt] = __builtin_amdgcn_mfma_f32_32x32x2f32(A[q], B1[q], acc[1][dt], 0, 0, 0);
							}
						}
				}
			}
#pragma unroll
			for (int t = 0; t < 2; t++)
#pragma unroll
				for (int dt = 0; dt < DT; dt++)
#pragma unroll
					for (int r = 0; r < 16; r++) acc64[t][dt][r] += (double)acc[t][dt][r];

			// C layout of v_mfma_f32_32x32x*: register r of lane l holds
			// row 8*(r/4) + 4*(l/32) + r%4 (the dof) and column l%32 (the point).
			const double* iscale = ldc(&sd->iscale);
#pragma unroll
			for (int t = 0; t < 2; t++)
			{
				const int q = 32 * t + row;
				const uint32_t oq = (uint32_t)__shfl((int)orig, q, 64);
				if (q < it.np)
				{
					double* out = split ? part + ((int64_t)it.seg * it.count + it.blk * kWave + q) * DP
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
									if constexpr (PREC == HDDM_LB_F16X2) v *= ldc(iscale + dof);
									out[dof] = v;
								}
							}
						}
				}
			}
		}
	}
}

__global__ void reset_counter_kernel(PlanHeader* hdr)
{
	if (threadIdx.x == 0) hdr->counter = 0;
}

__global__ void reduce_partials_kernel(const StateDesc* __restrict__ states, const StatePlan* __restrict__ plans,
	const uint32_t* __restrict__ perm, const double* __restrict__ partial, double* __restrict__ y, int64_t ldy)
{
	const int s = blockIdx.y;
	const int nseg = states[s].nseg;
	if (nseg <= 1) return;
	const int dofs = states[s].dofs, DP = states[s].dofs_pad;
	const int64_t count = plans[s].count, start = plans[s].start;
	const double* base = partial + plans[s].partial_base;
	const int64_t total = count * dofs;
	for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x)
	{
		const int64_t lp = e / dofs;
		const int c = (int)(e - lp * dofs);
		double sum = 0.0;
		for (int g = 0; g < nseg; g++) sum += base[((int64_t)g * count + lp) * DP + c];
		y[(int64_t)perm[start + lp] * ldy + c] = sum;
	}
}

// ---- launchers ----

template <int PREC, int DT, int FMAX>
static const void* sweep_ptr() { return (const void*)&sweep_kernel<PREC, DT, FMAX>; }

const void* sweep_kernel_ptr(int prec, int dt, int fmax)
{
#define HDDM_FM(P, D) \
	switch (fmax) { \
	case 1: case 2: return sweep_ptr<P, D, 2>(); \
	case 3: return sweep_ptr<P, D, 3>(); \
	case 4: return sweep_ptr<P, D, 4>(); \
	case 5: case 6: return sweep_ptr<P, D, 6>(); \
	case 7: case 8: return sweep_ptr<P, D, 8>(); \
	default: return sweep_ptr<P, D, 15>(); }
#define HDDM_DT(P) \
	if (dt == 1) { HDDM_FM(P, 1) } else { HDDM_FM(P, 2) }
	switch (prec)
	{
	case HDDM_LB_FP64: HDDM_DT(HDDM_LB_FP64)
	case HDDM_LB_F16X2: HDDM_DT(HDDM_LB_F16X2)
	case HDDM_LB_BF16X3: HDDM_DT(HDDM_LB_BF16X3)
	case HDDM_LB_F32: HDDM_DT(HDDM_LB_F32)
	}
#undef HDDM_DT
#undef HDDM_FM
	return nullptr;
}

int fmax_bucket(int fmax)
{
	if (fmax <= 2) return 2;
	if (fmax <= 4) return fmax;
	if (fmax <= 6) return 6;
	if (fmax <= 8) return 8;
	return 15;
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
	int blocks = (int)((n + 255) / 256);
	if (blocks > 8192) blocks = 8192;
	if (blocks < 1) blocks = 1;
	hipLaunchKernelGGL(bucket_key_kernel, dim3(blocks), dim3(256), 0, st, n, x, ldx, sid, state_ok, nstates, dim, bbits, bkey, hist);
	return hipGetLastError();
}

hipError_t launch_scan_plan(const uint32_t* hist, uint32_t* cursor, int64_t nbuckets, int bbits, int nstates,
	const StateDesc* states, const int32_t* cost_order, StatePlan* plans, PlanHeader* hdr, hipStream_t st)
{
	hipLaunchKernelGGL(scan_plan_kernel, dim3(1), dim3(1024), 0, st, hist, cursor, nbuckets, bbits, nstates,
		states, cost_order, plans, hdr);
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
	hipLaunchKernelGGL(reduce_partials_kernel, dim3(512, nstates), dim3(256), 0, st, states, plans, perm, partial, y, ldy);
	return hipGetLastError();
}

} // namespace hddm_lb
