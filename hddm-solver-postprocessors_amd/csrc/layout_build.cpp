// Prefix-trie node layout of one state.
//
// Every node's active factors (the dims with (i, j) != (0, 0), the only ones
// whose phi is not identically 1) are listed in ascending dim order, the order
// in which the reference's gold kernel multiplies them
// (LinearBasis/gold/src/InterpolateArray.cpp:28-36). Nodes are then sorted
// lexicographically by that factor sequence, like the reference's chain sort
// (LinearBasis/x86/src/Data.cpp:866-902), which puts every node sharing a
// prefix into one contiguous run. A wave of 64 points walks the runs keeping
// the product of the current prefix, evaluates only the factors a node does
// not share with its predecessor, and skips a whole run when that product is
// zero for all 64 points.
#include "layout_build.h"
#include "hddm_lb.h"
#include "errors.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <tuple>

namespace hddm_lb {

namespace {

struct FactorKey
{
	int32_t d, i, j;
	bool operator<(const FactorKey& o) const
	{
		return std::tie(d, i, j) < std::tie(o.d, o.i, o.j);
	}
};

uint16_t f16_bits(double v)
{
	_Float16 h = (_Float16)v;
	uint16_t b;
	std::memcpy(&b, &h, 2);
	return b;
}

double f16_value(uint16_t b)
{
	_Float16 h;
	std::memcpy(&h, &b, 2);
	return (double)h;
}

// Round-to-nearest-even FP32 -> BF16.
uint16_t bf16_bits(float f)
{
	uint32_t u;
	std::memcpy(&u, &f, 4);
	if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16);
	u += 0x7fffu + ((u >> 16) & 1u);
	return (uint16_t)(u >> 16);
}

double bf16_value(uint16_t b)
{
	uint32_t u = (uint32_t)b << 16;
	float f;
	std::memcpy(&f, &u, 4);
	return (double)f;
}

} // namespace

namespace {

// High 32 bits of the FP64 value of v; requires the low 32 bits to be zero,
// which holds for every integer of magnitude below 2^21.
bool f64_high_word(int64_t v, uint32_t& hi)
{
	double d = (double)v;
	uint64_t u;
	std::memcpy(&u, &d, 8);
	hi = (uint32_t)(u >> 32);
	return (uint32_t)u == 0u;
}

// High word of -j / i, the node's centre negated (i a power of two: the
// quotient is exact and has j's significand, so its low word is zero too).
uint32_t f64_high_word_centre(int64_t i, int64_t j)
{
	const double c = -(double)j / (double)i;
	uint64_t u;
	std::memcpy(&u, &c, 8);
	return (uint32_t)(u >> 32);
}

} // namespace

std::vector<int32_t> segment_nodes(const HostLayout& L, int64_t seg_len)
{
	std::vector<int32_t> b;
	b.push_back(0);
	if (seg_len > 0 && seg_len < L.nno)
	{
		int64_t last = 0;
		for (int n = 1; n < L.nno; n++)
			if (n - last >= seg_len && L.shared[n] == 0)
			{
				b.push_back(n);
				last = n;
			}
	}
	b.push_back(L.nno);
	return b;
}

int build_layout(int dim, int nno, int dofs,
	const int32_t* index_i, const int32_t* index_j, int64_t ld_index,
	const double* surplus, int64_t ld_surplus, HostLayout& L)
{
	if (dim < 1 || nno < 1 || dofs < 1 || ld_index < dim || ld_surplus < dofs || !index_i || !index_j || !surplus)
		return set_error(HDDM_LB_ERROR_INVALID_VALUE,
			"upload_state: need dim, nno, dofs >= 1, ld_index >= dim, ld_surplus >= dofs and non-NULL arrays");
	if (nno >= (1 << (32 - kOpNodeShift)))
		return set_error(HDDM_LB_ERROR_UNSUPPORTED, "upload_state: nno %d exceeds %d", nno, 1 << (32 - kOpNodeShift));

	L.dim = dim; L.nno = nno; L.dofs = dofs;
	L.dofs_pad = (dofs + 31) / 32 * 32;
	L.nchunks = (nno + kChunk - 1) / kChunk;

	// Active factors of every node in ascending dim order (the order in which
	// the gold kernel multiplies them), as (dim, i, j) triples.
	std::vector<std::vector<FactorKey> > seq(nno);
	int fmax = 0;
	for (int k = 0; k < nno; k++)
	{
		for (int d = 0; d < dim; d++)
		{
			int32_t i = index_i[(int64_t)k * ld_index + d], j = index_j[(int64_t)k * ld_index + d];
			if (!i && !j) continue;
			uint32_t h;
			if (!f64_high_word(i, h) || !f64_high_word(-(int64_t)j, h))
				return set_error(HDDM_LB_ERROR_UNSUPPORTED,
					"upload_state: node %d dim %d has (i, j) = (%d, %d); supported are values whose FP64 low word is zero (every |v| < 2^21)", k, d, i, j);
			seq[k].push_back(FactorKey{ d, i, j });
		}
		fmax = std::max(fmax, (int)seq[k].size());
	}
	if (fmax > 15)
		return set_error(HDDM_LB_ERROR_UNSUPPORTED,
			"upload_state: a node has %d active dims; at most 15 are supported", fmax);
	L.fmax = std::max(fmax, 1);

	// Lexicographic order of the factor sequences (a prefix sorts first).
	L.order.resize(nno);
	for (int k = 0; k < nno; k++) L.order[k] = k;
	std::stable_sort(L.order.begin(), L.order.end(), [&](int a, int b) {
		return std::lexicographical_compare(seq[a].begin(), seq[a].end(), seq[b].begin(), seq[b].end());
	});

	// Shared-prefix lengths and subtree ends.
	std::vector<int> F(nno), S(nno);
	L.shared.assign(nno, 0);
	for (int n = 0; n < nno; n++)
	{
		const auto& cur = seq[L.order[n]];
		F[n] = (int)cur.size();
		int s = 0;
		if (n > 0)
		{
			const auto& prev = seq[L.order[n - 1]];
			while (s < (int)cur.size() && s < (int)prev.size() &&
				!(cur[s] < prev[s]) && !(prev[s] < cur[s])) s++;
		}
		S[n] = s;
		L.shared[n] = (uint8_t)s;
	}
	const int FM = L.fmax;
	// end[n][t]: first node after n whose first t+1 factors differ from n's.
	std::vector<int> end((size_t)nno * FM, 0);
	for (int n = nno - 1; n >= 0; n--)
		for (int t = 0; t < F[n]; t++)
			end[(size_t)n * FM + t] = (n + 1 < nno && S[n + 1] >= t + 1) ? end[(size_t)(n + 1) * FM + t] : n + 1;

	// Op stream: one op per factor a node does not share with its predecessor;
	// the root (no active factor) gets one op with i = j = 0 (phi == 1).
	L.opstart.assign(nno + 1, 0);
	int64_t nops = 0;
	for (int n = 0; n < nno; n++)
	{
		L.opstart[n] = (int32_t)nops;
		nops += std::max(1, F[n] - S[n]);
	}
	L.opstart[nno] = (int32_t)nops;
	if (nops >= ((int64_t)1 << (32 - kOpSkipShift)))
		return set_error(HDDM_LB_ERROR_UNSUPPORTED, "upload_state: too many trie edges (%lld)", (long long)nops);
	L.ops.assign(nops + 1, Op{ 0, 0, 0, 0 });
	L.skip.assign(nops + 1, 0);
	for (int n = 0; n < nno; n++)
	{
		const auto& cur = seq[L.order[n]];
		int64_t q = L.opstart[n];
		if (F[n] == 0)
		{
			L.ops[q] = Op{ 0u, kOpLast | ((uint32_t)n << kOpNodeShift), 0u, 0u };
			L.skip[q] = (uint32_t)L.opstart[n + 1];
			continue;
		}
		for (int t = S[n]; t < F[n]; t++, q++)
		{
			Op op;
			op.dim = (uint32_t)cur[t].d << 1;
			op.ctrl = (uint32_t)t | (t == F[n] - 1 ? kOpLast : 0u) | ((uint32_t)n << kOpNodeShift);
			f64_high_word(cur[t].i, op.ihi);
			f64_high_word(-(int64_t)cur[t].j, op.njhi);
			L.ops[q] = op;
			L.skip[q] = (uint32_t)L.opstart[end[(size_t)n * FM + t]];
		}
	}
	// Sentinels (never evaluated) keep a one-op read-ahead in bounds.
	L.ops[nops] = Op{ 0u, 0u, 0u, 0u };
	L.skip[nops] = (uint32_t)nops;

	// Relative skip, relation to the stream predecessor and children flag.
	for (int64_t q = 0; q < nops; q++)
	{
		Op& o = L.ops[q];
		const uint32_t t = o.ctrl & 15u;
		uint32_t rel = 2;
		if (q > 0)
		{
			const uint32_t pt = L.ops[q - 1].ctrl & 15u;
			if (t == pt + 1) rel = 0;
			else if (t == pt) rel = 1;
		}
		const bool kids = q + 1 < nops && (L.ops[q + 1].ctrl & 15u) == t + 1;
		o.ctrl |= rel << kOpRelShift;
		if (kids) o.ctrl |= kOpKids;
		o.dim |= (L.skip[q] - (uint32_t)q) << kOpSkipShift;
	}

	// Per-node factor table (the adjoint's dx terms), slots padded to 4, 8 or
	// 16, nodes padded to a multiple of 32 (the adjoint reads whole chunks).
	L.fstride = FM <= 4 ? 4 : (FM <= 8 ? 8 : 16);
	L.ftab.assign((size_t)((nno + 31) & ~31) * L.fstride * 4, 0u);
	for (int n = 0; n < nno; n++)
	{
		const auto& cur = seq[L.order[n]];
		uint32_t* e = &L.ftab[(size_t)n * L.fstride * 4];
		for (size_t t = 0; t < cur.size(); t++, e += 4)
		{
			e[0] = (uint32_t)cur[t].d << 1;
			f64_high_word(cur[t].i, e[1]);
			e[2] = f64_high_word_centre(cur[t].i, cur[t].j);
			e[3] = 1u;
		}
	}

	// Support masks: bit d of lomask (himask) when the node's support in dim d
	// lies in [0, 1/2] ([1/2, 1]); dims >= 64 are left out (conservative).
	L.lomask.assign(nno, 0);
	L.himask.assign(nno, 0);
	for (int n = 0; n < nno; n++)
	{
		int k = L.order[n];
		for (int d = 0; d < dim && d < 64; d++)
		{
			int64_t i = index_i[(int64_t)k * ld_index + d], j = index_j[(int64_t)k * ld_index + d];
			if (i <= 0) continue;
			if (2 * (j + 1) <= i) L.lomask[n] |= 1ull << d;
			if (2 * (j - 1) >= i) L.himask[n] |= 1ull << d;
		}
	}

	// Surplus layouts in trie order.
	const int DP = L.dofs_pad, NC = L.nchunks;
	L.s64.assign((size_t)nno * DP, 0.0);
	for (int n = 0; n < nno; n++)
		std::memcpy(&L.s64[(size_t)n * DP], surplus + (int64_t)L.order[n] * ld_surplus, sizeof(double) * dofs);

	// f16: per-column power-of-two scale bringing max |S| into [2^14, 2^15).
	L.iscale.assign(DP, 1.0);
	std::vector<double> scale(DP, 1.0);
	for (int c = 0; c < dofs; c++)
	{
		double m = 0.0;
		for (int n = 0; n < nno; n++) m = std::max(m, std::fabs(L.s64[(size_t)n * DP + c]));
		if (m > 0.0 && std::isfinite(m))
		{
			int e = std::ilogb(m);
			scale[c] = std::ldexp(1.0, 14 - e);
			L.iscale[c] = std::ldexp(1.0, e - 14);
		}
	}
	L.s16.assign((size_t)NC * 2 * DP * kChunk, 0);
	L.sbf.assign((size_t)NC * 3 * DP * kChunk, 0);
	L.s32.assign((size_t)NC * DP * kChunk, 0.0f);
	L.s32r.assign((size_t)nno * DP, 0.0f);
	for (size_t e = 0; e < L.s64.size(); e++) L.s32r[e] = (float)L.s64[e];
	for (int n = 0; n < nno; n++)
	{
		int c = n / kChunk, pos = n % kChunk;
		for (int dof = 0; dof < dofs; dof++)
		{
			double v = L.s64[(size_t)n * DP + dof];
			double vs = v * scale[dof];
			uint16_t hi = f16_bits(vs);
			uint16_t lo = f16_bits(vs - f16_value(hi));
			L.s16[(((size_t)c * 2 + 0) * DP + dof) * kChunk + pos] = hi;
			L.s16[(((size_t)c * 2 + 1) * DP + dof) * kChunk + pos] = lo;

			uint16_t t1 = bf16_bits((float)v);
			double r1 = v - bf16_value(t1);
			uint16_t t2 = bf16_bits((float)r1);
			double r2 = r1 - bf16_value(t2);
			uint16_t t3 = bf16_bits((float)r2);
			L.sbf[(((size_t)c * 3 + 0) * DP + dof) * kChunk + pos] = t1;
			L.sbf[(((size_t)c * 3 + 1) * DP + dof) * kChunk + pos] = t2;
			L.sbf[(((size_t)c * 3 + 2) * DP + dof) * kChunk + pos] = t3;

			// f32 layout [c][dof][h][q] holds position 2q + h.
			int h = pos & 1, q = pos >> 1;
			L.s32[(((size_t)c * DP + dof) * 2 + h) * 8 + q] = (float)v;
		}
	}
	return HDDM_LB_SUCCESS;
}

} // namespace hddm_lb

// ---- introspection of the host layout (used by CPU tests) ----

struct hddm_lb_layout
{
	hddm_lb::HostLayout L;
};

extern "C" int hddm_lb_layout_build(int dim, int nno, int dofs,
	const int32_t* index_i, const int32_t* index_j, int64_t ld_index,
	const double* surplus, int64_t ld_surplus, hddm_lb_layout** out)
{
	if (!out) return hddm_lb::set_error(HDDM_LB_ERROR_INVALID_VALUE, "layout_build: out is NULL");
	hddm_lb_layout* l = new hddm_lb_layout;
	int rc = hddm_lb::build_layout(dim, nno, dofs, index_i, index_j, ld_index, surplus, ld_surplus, l->L);
	if (rc) { delete l; *out = nullptr; return rc; }
	*out = l;
	return HDDM_LB_SUCCESS;
}

extern "C" int hddm_lb_layout_info(const hddm_lb_layout* l, int* max_factors, int* nops,
	int* nchunks, int* dofs_pad)
{
	if (!l) return hddm_lb::set_error(HDDM_LB_ERROR_INVALID_VALUE, "layout_info: NULL layout");
	if (max_factors) *max_factors = l->L.fmax;
	if (nops) *nops = (int)l->L.skip.size() - 1;
	if (nchunks) *nchunks = l->L.nchunks;
	if (dofs_pad) *dofs_pad = l->L.dofs_pad;
	return HDDM_LB_SUCCESS;
}

extern "C" int hddm_lb_layout_ops(const hddm_lb_layout* l, uint32_t* ops, uint32_t* skip, int32_t* opstart)
{
	if (!l || !ops || !skip || !opstart) return hddm_lb::set_error(HDDM_LB_ERROR_INVALID_VALUE, "layout_ops: NULL argument");
	const size_t nops = l->L.skip.size() - 1;
	std::memcpy(ops, l->L.ops.data(), nops * sizeof(hddm_lb::Op));
	std::memcpy(skip, l->L.skip.data(), nops * sizeof(uint32_t));
	std::memcpy(opstart, l->L.opstart.data(), l->L.opstart.size() * sizeof(int32_t));
	return HDDM_LB_SUCCESS;
}

extern "C" int hddm_lb_layout_order(const hddm_lb_layout* l, int32_t* order)
{
	if (!l || !order) return hddm_lb::set_error(HDDM_LB_ERROR_INVALID_VALUE, "layout_order: NULL argument");
	std::memcpy(order, l->L.order.data(), l->L.order.size() * sizeof(int32_t));
	return HDDM_LB_SUCCESS;
}

extern "C" int hddm_lb_layout_factors(const hddm_lb_layout* l, uint32_t* ftab, int* fstride)
{
	if (!l) return hddm_lb::set_error(HDDM_LB_ERROR_INVALID_VALUE, "layout_factors: NULL layout");
	if (fstride) *fstride = l->L.fstride;
	if (ftab) std::memcpy(ftab, l->L.ftab.data(), l->L.ftab.size() * sizeof(uint32_t));
	return HDDM_LB_SUCCESS;
}

extern "C" int hddm_lb_layout_free(hddm_lb_layout* l)
{
	delete l;
	return HDDM_LB_SUCCESS;
}
