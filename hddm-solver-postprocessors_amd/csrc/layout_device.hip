// Device-side build of a state's node layout: the arrays build_layout
// (layout_build.cpp) makes on the host, bit for bit, made by kernels from the
// uploaded index and surplus matrices. Only the two small per-node arrays
// the host plans segments with (shared prefix lengths, first ops) come back.
//
// Steps, all data-parallel over nodes, ops or surplus elements:
//   factors   each node's active factors (dims with (i, j) != (0, 0)) in
//             ascending dim order, their count, the largest count;
//   sort      nodes in lexicographic order of their factor sequences: a
//             least-significant-first series of stable radix sorts, one per
//             factor position from the last to the first, on a 56-bit key
//             per factor (dim, i, j, each offset to be non-negative; an
//             absent factor is key 0 and sorts first, so a prefix sorts
//             before its extensions), starting from the identity, which is
//             std::stable_sort with lexicographical_compare;
//   shared    prefix length a node shares with its predecessor; ops per node
//             max(1, factors - shared) and their exclusive scan (first ops);
//   run ends  per depth t, the first node after n whose first t + 1 factors
//             differ: the start of the next run of nodes whose shared prefix
//             is at least t + 1 (a scan of run starts per depth);
//   ops       one op per unshared factor, its skip target the first op of
//             that run end; relation and children bits from the neighbours;
//   data      support masks, surplus rows in trie order (FP64, FP32), the
//             per-column f16 scale, f16 hi/lo and bf16 three-term splits,
//             the f32 chunk layout, and the per-node factor table.
#include "layout_build.h"
#include "context.h"
#include "errors.h"
#include "hddm_lb.h"

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cstring>
#include <vector>

namespace hddm_lb {

namespace {

constexpr int kMaxFactors = 15;
constexpr int kKeyBits = 56;

struct Factor
{
	int32_t d, i, j;
};

__device__ __forceinline__ uint64_t factor_key(const Factor& f)
{
	return ((uint64_t)(f.d + 1) << 48) | ((uint64_t)(f.i + (1 << 23)) << 24) | (uint64_t)(f.j + (1 << 23));
}

__device__ __forceinline__ uint32_t high_word(int64_t v)
{
	return (uint32_t)((uint64_t)__double_as_longlong((double)v) >> 32);
}

// Round-to-nearest-even FP64 -> f16 bits (with subnormals, overflow to
// infinity), from the binary representation: one rounding, as the host's
// conversion makes it.
__device__ uint16_t f16_bits(double v)
{
	const uint64_t u = (uint64_t)__double_as_longlong(v);
	const uint16_t sign = (uint16_t)((u >> 48) & 0x8000u);
	const int e = (int)((u >> 52) & 0x7ff);
	uint64_t m = u & 0xfffffffffffffull;
	if (e == 0x7ff) return sign | 0x7c00u | (m ? 0x200u : 0u);
	if (e == 0) return sign;   // FP64 subnormals are far below the f16 range
	const int ee = e - 1023 + 15;  // f16 biased exponent
	m |= 1ull << 52;               // 53-bit significand
	int shift;
	uint32_t eb;
	if (ee >= 1) { shift = 42; eb = (uint32_t)ee; }         // normal: keep 11 bits
	else
	{
		shift = 42 + (1 - ee);                                // subnormal
		eb = 0;
		if (shift > 54) return sign;                          // rounds to zero
	}
	const uint64_t keep = m >> shift;
	const uint64_t rem = m & ((1ull << shift) - 1);
	const uint64_t half = 1ull << (shift - 1);
	uint64_t r = keep + ((rem > half || (rem == half && (keep & 1))) ? 1 : 0);
	uint32_t bits;
	if (eb == 0) bits = (uint32_t)r;                          // may carry into the normal range
	else
	{
		if (r >> 11) { r >>= 1; eb++; }
		if (eb >= 31) return sign | 0x7c00u;
		bits = (eb << 10) | (uint32_t)(r & 0x3ff);
	}
	return sign | (uint16_t)bits;
}

__device__ double f16_value(uint16_t b)
{
	const uint32_t s = b >> 15, e = (b >> 10) & 31, m = b & 1023;
	double v;
	if (e == 0) v = ldexp((double)m, -24);
	else if (e == 31) v = m ? __longlong_as_double(0x7ff8000000000000ll) : __longlong_as_double(0x7ff0000000000000ll);
	else v = ldexp((double)(m | 1024), (int)e - 25);
	return s ? -v : v;
}

__device__ uint16_t bf16_bits(float f)
{
	uint32_t u = __float_as_uint(f);
	if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16);
	u += 0x7fffu + ((u >> 16) & 1u);
	return (uint16_t)(u >> 16);
}

__device__ double bf16_value(uint16_t b)
{
	return (double)__uint_as_float((uint32_t)b << 16);
}

__global__ void factors_kernel(int nno, int dim, const int32_t* ii, const int32_t* jj, int64_t ld, Factor* fact,
	int32_t* nf, int* fmax, int* bad)
{
	const int n = blockIdx.x * blockDim.x + threadIdx.x;
	if (n >= nno) return;
	int t = 0;
	for (int d = 0; d < dim; d++)
	{
		const int32_t i = ii[(int64_t)n * ld + d], j = jj[(int64_t)n * ld + d];
		if (!i && !j) continue;
		if (i <= -(1 << 21) || i >= (1 << 21) || j <= -(1 << 21) || j >= (1 << 21)) atomicOr(bad, 1);
		if (t < kMaxFactors) fact[(int64_t)n * kMaxFactors + t] = Factor{ d, i, j };
		t++;
	}
	nf[n] = t;
	atomicMax(fmax, t);
}

__global__ void iota_kernel(int n, uint32_t* p)
{
	const int k = blockIdx.x * blockDim.x + threadIdx.x;
	if (k < n) p[k] = (uint32_t)k;
}

__global__ void keys_kernel(int nno, int t, const Factor* fact, const int32_t* nf, const uint32_t* perm, uint64_t* key)
{
	const int n = blockIdx.x * blockDim.x + threadIdx.x;
	if (n >= nno) return;
	const uint32_t k = perm[n];
	key[n] = t < nf[k] ? factor_key(fact[(int64_t)k * kMaxFactors + t]) : 0ull;
}

// Factors shared with the trie predecessor, ops per node.
__global__ void shared_kernel(int nno, const Factor* fact, const int32_t* nf, const uint32_t* order, int32_t* F,
	int32_t* S, int32_t* nops_of)
{
	const int n = blockIdx.x * blockDim.x + threadIdx.x;
	if (n >= nno) return;
	const uint32_t k = order[n];
	const int f = nf[k];
	int s = 0;
	if (n > 0)
	{
		const uint32_t p = order[n - 1];
		const int fp = nf[p];
		const Factor* a = fact + (int64_t)k * kMaxFactors;
		const Factor* b = fact + (int64_t)p * kMaxFactors;
		while (s < f && s < fp && a[s].d == b[s].d && a[s].i == b[s].i && a[s].j == b[s].j) s++;
	}
	F[n] = f;
	S[n] = s;
	nops_of[n] = max(1, f - s);
}

// Run starts at depth t: nodes whose shared prefix is shorter than t + 1.
__global__ void run_flags_kernel(int nno, int t, const int32_t* S, int32_t* flag)
{
	const int n = blockIdx.x * blockDim.x + threadIdx.x;
	if (n < nno) flag[n] = (n == 0 || S[n] < t + 1) ? 1 : 0;
}

__global__ void run_starts_kernel(int nno, const int32_t* flag, const int32_t* runid, int32_t* start)
{
	const int n = blockIdx.x * blockDim.x + threadIdx.x;
	if (n < nno && flag[n]) start[runid[n] - 1] = n;
}

// end[t][n]: the first node of the run after n's at depth t.
__global__ void run_ends_kernel(int nno, int t, const int32_t* runid, const int32_t* start, int nruns, int32_t* end)
{
	const int n = blockIdx.x * blockDim.x + threadIdx.x;
	if (n >= nno) return;
	const int r = runid[n];   // 1-based
	end[(int64_t)t * nno + n] = r < nruns ? start[r] : nno;
}

__global__ void ops_kernel(int nno, const Factor* fact, const uint32_t* order, const int32_t* F, const int32_t* S,
	const int32_t* opstart, const int32_t* end, Op* ops, uint32_t* skip)
{
	const int n = blockIdx.x * blockDim.x + threadIdx.x;
	if (n >= nno) return;
	int q = opstart[n];
	const int f = F[n];
	if (f == 0)
	{
		ops[q] = Op{ 0u, kOpLast | ((uint32_t)n << kOpNodeShift), 0u, 0u };
		skip[q] = (uint32_t)opstart[n + 1];
		return;
	}
	const Factor* a = fact + (int64_t)order[n] * kMaxFactors;
	for (int t = S[n]; t < f; t++, q++)
	{
		Op op;
		op.dim = (uint32_t)a[t].d << 1;
		op.ctrl = (uint32_t)t | (t == f - 1 ? kOpLast : 0u) | ((uint32_t)n << kOpNodeShift);
		op.ihi = high_word(a[t].i);
		op.njhi = high_word(-(int64_t)a[t].j);
		ops[q] = op;
		skip[q] = (uint32_t)opstart[end[(int64_t)t * nno + n]];
	}
}

// Relation to the stream predecessor, children bit, relative skip; the
// sentinels after the stream.
__global__ void finish_ops_kernel(int nops, const Op* in, const uint32_t* skip, Op* out)
{
	const int q = blockIdx.x * blockDim.x + threadIdx.x;
	if (q >= nops + kOpPad) return;
	if (q >= nops)
	{
		out[q] = Op{ 0u, 0u, 0u, 0u };
		return;
	}
	Op o = in[q];
	const uint32_t t = o.ctrl & 15u;
	uint32_t rel = 2;
	if (q > 0)
	{
		const uint32_t pt = in[q - 1].ctrl & 15u;
		if (t == pt + 1) rel = 0;
		else if (t == pt) rel = 1;
	}
	const bool kids = q + 1 < nops && (in[q + 1].ctrl & 15u) == t + 1;
	o.ctrl |= rel << kOpRelShift;
	if (kids) o.ctrl |= kOpKids;
	o.dim |= (skip[q] - (uint32_t)q) << kOpSkipShift;
	out[q] = o;
}

__global__ void pad_skip_kernel(int nops, uint32_t* skip)
{
	const int q = threadIdx.x;
	if (q < kOpPad) skip[nops + q] = (uint32_t)nops;
}

__global__ void masks_kernel(int nno, int dim, const int32_t* ii, const int32_t* jj, int64_t ld, const uint32_t* order,
	uint64_t* lomask, uint64_t* himask)
{
	const int n = blockIdx.x * blockDim.x + threadIdx.x;
	if (n >= nno) return;
	const int64_t k = order[n];
	uint64_t lo = 0, hi = 0;
	for (int d = 0; d < dim && d < 64; d++)
	{
		const int64_t i = ii[k * ld + d], j = jj[k * ld + d];
		if (i <= 0) continue;
		if (2 * (j + 1) <= i) lo |= 1ull << d;
		if (2 * (j - 1) >= i) hi |= 1ull << d;
	}
	lomask[n] = lo;
	himask[n] = hi;
}

// Surplus rows in trie order and the columns' largest magnitudes (non-negative
// doubles order like their bit patterns).
__global__ void rows_kernel(int nno, int dofs, int DP, const double* S, int64_t lds, const uint32_t* order, double* s64,
	float* s32r, unsigned long long* colmax)
{
	const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (e >= (int64_t)nno * DP) return;
	const int64_t n = e / DP;
	const int c = (int)(e - n * DP);
	const double v = c < dofs ? S[(int64_t)order[n] * lds + c] : 0.0;
	s64[e] = v;
	s32r[e] = (float)v;
	if (c < dofs) atomicMax(colmax + c, (unsigned long long)__double_as_longlong(fabs(v)));
}

// Per-column power of two bringing max |S| into [2^14, 2^15).
__global__ void scales_kernel(int DP, int dofs, const unsigned long long* colmax, double* scale, double* iscale)
{
	const int c = blockIdx.x * blockDim.x + threadIdx.x;
	if (c >= DP) return;
	double sc = 1.0, is = 1.0;
	if (c < dofs)
	{
		const double m = __longlong_as_double((long long)colmax[c]);
		if (m > 0.0 && isfinite(m))
		{
			int ex;
			frexp(m, &ex);
			const int e = ex - 1;   // ilogb
			sc = ldexp(1.0, 14 - e);
			is = ldexp(1.0, e - 14);
		}
	}
	scale[c] = sc;
	iscale[c] = is;
}

__global__ void splits_kernel(int nno, int dofs, int DP, const double* s64, const double* scale, uint16_t* s16,
	uint16_t* sbf, float* s32)
{
	const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	if (e >= (int64_t)nno * dofs) return;
	const int n = (int)(e / dofs);
	const int dof = (int)(e - (int64_t)n * dofs);
	const int c = n / kChunk, pos = n % kChunk;
	const double v = s64[(int64_t)n * DP + dof];
	const double vs = v * scale[dof];
	const uint16_t hi = f16_bits(vs);
	const uint16_t lo = f16_bits(vs - f16_value(hi));
	s16[(((size_t)c * 2 + 0) * DP + dof) * kChunk + pos] = hi;
	s16[(((size_t)c * 2 + 1) * DP + dof) * kChunk + pos] = lo;
	const uint16_t t1 = bf16_bits((float)v);
	const double r1 = v - bf16_value(t1);
	const uint16_t t2 = bf16_bits((float)r1);
	const double r2 = r1 - bf16_value(t2);
	const uint16_t t3 = bf16_bits((float)r2);
	sbf[(((size_t)c * 3 + 0) * DP + dof) * kChunk + pos] = t1;
	sbf[(((size_t)c * 3 + 1) * DP + dof) * kChunk + pos] = t2;
	sbf[(((size_t)c * 3 + 2) * DP + dof) * kChunk + pos] = t3;
	const int h = pos & 1, qq = pos >> 1;
	s32[(((size_t)c * DP + dof) * 2 + h) * 8 + qq] = (float)v;
}

__global__ void ftab_kernel(int nno, int fstride, const Factor* fact, const int32_t* F, const uint32_t* order,
	uint32_t* ftab)
{
	const int n = blockIdx.x * blockDim.x + threadIdx.x;
	if (n >= nno) return;
	const Factor* a = fact + (int64_t)order[n] * kMaxFactors;
	uint32_t* e = ftab + (size_t)n * fstride * 4;
	for (int t = 0; t < F[n]; t++, e += 4)
	{
		e[0] = (uint32_t)a[t].d << 1;
		e[1] = high_word(a[t].i);
		// the node's centre negated, -j / i (exact: i is a power of two)
		e[2] = (uint32_t)((uint64_t)__double_as_longlong(-(double)a[t].j / (double)a[t].i) >> 32);
		e[3] = 1u;
	}
}

int blocks(int64_t n, int b = 256) { return (int)((n + b - 1) / b); }

// Device temporaries of one build, freed on every exit.
struct Temps
{
	std::vector<void*> p;
	template <typename T>
	T* get(size_t count, int& rc)
	{
		if (rc) return nullptr;
		void* q = nullptr;
		hipError_t e = hipMalloc(&q, count ? count * sizeof(T) : 1);
		if (e != hipSuccess)
		{
			(void)hipGetLastError();
			rc = set_error(HDDM_LB_ERROR_OUT_OF_MEMORY, "upload_state: device layout scratch of %zu bytes: %s",
				count * sizeof(T), hipGetErrorString(e));
			return nullptr;
		}
		p.push_back(q);
		return (T*)q;
	}
	~Temps() { for (void* q : p) (void)hipFree(q); }
};

} // namespace

#define HDDM_DL_TRY(call) do { hipError_t e_ = (call); if (e_ != hipSuccess) return hip_fail(e_, #call, __FILE__, __LINE__); } while (0)

int build_layout_device(int dim, int nno, int dofs,
	const int32_t* index_i, const int32_t* index_j, int64_t ld_index,
	const double* surplus, int64_t ld_surplus, const DeviceAlloc& alloc, DeviceLayout& out)
{
	if (dim < 1 || nno < 1 || dofs < 1 || ld_index < dim || ld_surplus < dofs || !index_i || !index_j || !surplus)
		return kLayoutHostCheck;
	if (nno >= (1 << (32 - kOpNodeShift))) return kLayoutHostCheck;
	int rc = HDDM_LB_SUCCESS;
	Temps tmp;
	const hipStream_t st = nullptr;
	const size_t nidx = (size_t)nno * ld_index;
	int32_t* d_ii = tmp.get<int32_t>(nidx, rc);
	int32_t* d_jj = tmp.get<int32_t>(nidx, rc);
	double* d_S = tmp.get<double>((size_t)nno * ld_surplus, rc);
	Factor* fact = tmp.get<Factor>((size_t)nno * kMaxFactors, rc);
	int32_t* nf = tmp.get<int32_t>(nno, rc);
	int* flags = tmp.get<int>(2, rc);   // [0] largest factor count, [1] an index out of range
	if (rc) return rc;
	HDDM_DL_TRY(hipMemcpy(d_ii, index_i, nidx * sizeof(int32_t), hipMemcpyHostToDevice));
	HDDM_DL_TRY(hipMemcpy(d_jj, index_j, nidx * sizeof(int32_t), hipMemcpyHostToDevice));
	HDDM_DL_TRY(hipMemcpy(d_S, surplus, (size_t)nno * ld_surplus * sizeof(double), hipMemcpyHostToDevice));
	HDDM_DL_TRY(hipMemset(flags, 0, 2 * sizeof(int)));
	hipLaunchKernelGGL(factors_kernel, dim3(blocks(nno)), dim3(256), 0, st, nno, dim, (const int32_t*)d_ii,
		(const int32_t*)d_jj, ld_index, fact, nf, flags, flags + 1);
	HDDM_DL_TRY(hipGetLastError());
	int hflags[2] = { 0, 0 };
	HDDM_DL_TRY(hipMemcpy(hflags, flags, sizeof(hflags), hipMemcpyDeviceToHost));
	// invalid indices or more than 15 factors: the host build reports them
	if (hflags[1] || hflags[0] > kMaxFactors) return kLayoutHostCheck;
	const int FM = hflags[0] > 1 ? hflags[0] : 1;

	// lexicographic order of the factor sequences
	uint32_t* perm = tmp.get<uint32_t>(nno, rc);
	uint32_t* perm2 = tmp.get<uint32_t>(nno, rc);
	uint64_t* key = tmp.get<uint64_t>(nno, rc);
	uint64_t* key2 = tmp.get<uint64_t>(nno, rc);
	if (rc) return rc;
	size_t sort_bytes = 0;
	HDDM_DL_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, key, key2, perm, perm2, nno, 0, kKeyBits, st));
	void* sort_tmp = tmp.get<char>(sort_bytes, rc);
	if (rc) return rc;
	hipLaunchKernelGGL(iota_kernel, dim3(blocks(nno)), dim3(256), 0, st, nno, perm);
	for (int t = FM - 1; t >= 0; t--)
	{
		hipLaunchKernelGGL(keys_kernel, dim3(blocks(nno)), dim3(256), 0, st, nno, t, (const Factor*)fact,
			(const int32_t*)nf, (const uint32_t*)perm, key);
		HDDM_DL_TRY(hipcub::DeviceRadixSort::SortPairs(sort_tmp, sort_bytes, key, key2, perm, perm2, nno, 0, kKeyBits, st));
		std::swap(perm, perm2);
	}
	HDDM_DL_TRY(hipGetLastError());

	// shared prefixes, first ops
	int32_t* F = tmp.get<int32_t>(nno, rc);
	int32_t* S = tmp.get<int32_t>(nno, rc);
	int32_t* nops_of = tmp.get<int32_t>((size_t)nno + 1, rc);
	int32_t* opstart = tmp.get<int32_t>((size_t)nno + 1, rc);
	if (rc) return rc;
	hipLaunchKernelGGL(shared_kernel, dim3(blocks(nno)), dim3(256), 0, st, nno, (const Factor*)fact,
		(const int32_t*)nf, (const uint32_t*)perm, F, S, nops_of);
	HDDM_DL_TRY(hipMemset(nops_of + nno, 0, sizeof(int32_t)));
	size_t scan_bytes = 0;
	HDDM_DL_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, nops_of, opstart, nno + 1, st));
	size_t scan2_bytes = 0;
	HDDM_DL_TRY(hipcub::DeviceScan::InclusiveSum(nullptr, scan2_bytes, nops_of, opstart, nno, st));
	void* scan_tmp = tmp.get<char>(std::max(scan_bytes, scan2_bytes), rc);
	if (rc) return rc;
	HDDM_DL_TRY(hipcub::DeviceScan::ExclusiveSum(scan_tmp, scan_bytes, nops_of, opstart, nno + 1, st));
	int32_t nops = 0;
	HDDM_DL_TRY(hipMemcpy(&nops, opstart + nno, sizeof(int32_t), hipMemcpyDeviceToHost));
	if ((int64_t)nops >= ((int64_t)1 << (32 - kOpSkipShift))) return kLayoutHostCheck;

	// run ends per depth
	int32_t* flag = tmp.get<int32_t>(nno, rc);
	int32_t* runid = tmp.get<int32_t>(nno, rc);
	int32_t* start = tmp.get<int32_t>(nno, rc);
	int32_t* end = tmp.get<int32_t>((size_t)FM * nno, rc);
	if (rc) return rc;
	for (int t = 0; t < FM; t++)
	{
		hipLaunchKernelGGL(run_flags_kernel, dim3(blocks(nno)), dim3(256), 0, st, nno, t, (const int32_t*)S, flag);
		HDDM_DL_TRY(hipcub::DeviceScan::InclusiveSum(scan_tmp, scan2_bytes, flag, runid, nno, st));
		hipLaunchKernelGGL(run_starts_kernel, dim3(blocks(nno)), dim3(256), 0, st, nno, (const int32_t*)flag,
			(const int32_t*)runid, start);
		int32_t nruns = 0;
		HDDM_DL_TRY(hipMemcpy(&nruns, runid + nno - 1, sizeof(int32_t), hipMemcpyDeviceToHost));
		hipLaunchKernelGGL(run_ends_kernel, dim3(blocks(nno)), dim3(256), 0, st, nno, t, (const int32_t*)runid,
			(const int32_t*)start, (int)nruns, end);
	}
	HDDM_DL_TRY(hipGetLastError());

	// outputs
	const int DP = (dofs + 31) / 32 * 32;
	const int NC = (nno + kChunk - 1) / kChunk;
	const int fstride = FM <= 4 ? 4 : (FM <= 8 ? 8 : 16);
	const size_t nrow = (size_t)((nno + 31) & ~31);
	out.nops = nops; out.fmax = FM; out.fstride = fstride; out.dofs_pad = DP; out.nchunks = NC;
	out.ops = (Op*)alloc(((size_t)nops + kOpPad) * sizeof(Op), rc);
	out.skip = (uint32_t*)alloc(((size_t)nops + kOpPad) * sizeof(uint32_t), rc);
	out.s64 = (double*)alloc((size_t)nno * DP * sizeof(double), rc);
	out.s16 = (uint16_t*)alloc((size_t)NC * 2 * DP * kChunk * sizeof(uint16_t), rc);
	out.sbf = (uint16_t*)alloc((size_t)NC * 3 * DP * kChunk * sizeof(uint16_t), rc);
	out.s32 = (float*)alloc((size_t)NC * DP * kChunk * sizeof(float), rc);
	out.s32r = (float*)alloc((size_t)nno * DP * sizeof(float), rc);
	out.iscale = (double*)alloc((size_t)DP * sizeof(double), rc);
	out.order = (int32_t*)alloc((size_t)nno * sizeof(int32_t), rc);
	out.lomask = (uint64_t*)alloc((size_t)nno * sizeof(uint64_t), rc);
	out.himask = (uint64_t*)alloc((size_t)nno * sizeof(uint64_t), rc);
	out.ftab = (uint32_t*)alloc(nrow * fstride * 4 * sizeof(uint32_t), rc);
	Op* ops_raw = tmp.get<Op>((size_t)nops + 1, rc);
	double* scale = tmp.get<double>(DP, rc);
	unsigned long long* colmax = tmp.get<unsigned long long>(DP, rc);
	if (rc) return rc;

	hipLaunchKernelGGL(ops_kernel, dim3(blocks(nno)), dim3(256), 0, st, nno, (const Factor*)fact,
		(const uint32_t*)perm, (const int32_t*)F, (const int32_t*)S, (const int32_t*)opstart, (const int32_t*)end,
		ops_raw, out.skip);
	hipLaunchKernelGGL(finish_ops_kernel, dim3(blocks((int64_t)nops + kOpPad)), dim3(256), 0, st, (int)nops,
		(const Op*)ops_raw, (const uint32_t*)out.skip, out.ops);
	hipLaunchKernelGGL(pad_skip_kernel, dim3(1), dim3(64), 0, st, (int)nops, out.skip);
	HDDM_DL_TRY(hipMemcpy(out.order, perm, (size_t)nno * sizeof(int32_t), hipMemcpyDeviceToDevice));
	hipLaunchKernelGGL(masks_kernel, dim3(blocks(nno)), dim3(256), 0, st, nno, dim, (const int32_t*)d_ii,
		(const int32_t*)d_jj, ld_index, (const uint32_t*)perm, out.lomask, out.himask);
	HDDM_DL_TRY(hipMemset(colmax, 0, (size_t)DP * sizeof(unsigned long long)));
	hipLaunchKernelGGL(rows_kernel, dim3(blocks((int64_t)nno * DP)), dim3(256), 0, st, nno, dofs, DP,
		(const double*)d_S, ld_surplus, (const uint32_t*)perm, out.s64, out.s32r, colmax);
	hipLaunchKernelGGL(scales_kernel, dim3(blocks(DP)), dim3(256), 0, st, DP, dofs,
		(const unsigned long long*)colmax, scale, out.iscale);
	HDDM_DL_TRY(hipMemset(out.s16, 0, (size_t)NC * 2 * DP * kChunk * sizeof(uint16_t)));
	HDDM_DL_TRY(hipMemset(out.sbf, 0, (size_t)NC * 3 * DP * kChunk * sizeof(uint16_t)));
	HDDM_DL_TRY(hipMemset(out.s32, 0, (size_t)NC * DP * kChunk * sizeof(float)));
	hipLaunchKernelGGL(splits_kernel, dim3(blocks((int64_t)nno * dofs)), dim3(256), 0, st, nno, dofs, DP,
		(const double*)out.s64, (const double*)scale, out.s16, out.sbf, out.s32);
	HDDM_DL_TRY(hipMemset(out.ftab, 0, nrow * fstride * 4 * sizeof(uint32_t)));
	hipLaunchKernelGGL(ftab_kernel, dim3(blocks(nno)), dim3(256), 0, st, nno, fstride, (const Factor*)fact,
		(const int32_t*)F, (const uint32_t*)perm, out.ftab);
	HDDM_DL_TRY(hipGetLastError());

	// the host plans node segments from the shared prefixes and first ops
	std::vector<int32_t> hS(nno);
	out.opstart.resize((size_t)nno + 1);
	HDDM_DL_TRY(hipMemcpy(hS.data(), S, (size_t)nno * sizeof(int32_t), hipMemcpyDeviceToHost));
	HDDM_DL_TRY(hipMemcpy(out.opstart.data(), opstart, ((size_t)nno + 1) * sizeof(int32_t), hipMemcpyDeviceToHost));
	out.shared.resize(nno);
	for (int n = 0; n < nno; n++) out.shared[n] = (uint8_t)hS[n];
	HDDM_DL_TRY(hipDeviceSynchronize());
	return HDDM_LB_SUCCESS;
}

} // namespace hddm_lb
