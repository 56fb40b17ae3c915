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

int build_layout(int dim, int nno, int dofs,
	const int32_t* index_i, const int32_t* index_j, int64_t ld_index,
	const double* surplus, int64_t ld_surplus, HostLayout& L)
{
	if (dim < 1 || nno < 1 || dofs < 1 || ld_index < dim || ld_surplus < dofs || !index_i || !index_j || !surplus)
		return set_error(HDDM_LB_ERROR_INVALID_VALUE,
			"upload_state: need dim, nno, dofs >= 1, ld_index >= dim, ld_surplus >= dofs and non-NULL arrays");
	if ((int64_t)dim * kXRowBytes > (int64_t)UINT32_MAX)
		return set_error(HDDM_LB_ERROR_UNSUPPORTED, "upload_state: dim %d too large", dim);

	L.dim = dim; L.nno = nno; L.dofs = dofs;
	L.dofs_pad = (dofs + 31) / 32 * 32;
	L.nchunks = (nno + kChunk - 1) / kChunk;

	// Unique factors, ids ordered by (dim, i, j) so id order is factor order.
	std::map<FactorKey, int> ids;
	int fmax = 0;
	for (int k = 0; k < nno; k++)
	{
		int f = 0;
		for (int d = 0; d < dim; d++)
		{
			int32_t i = index_i[(int64_t)k * ld_index + d], j = index_j[(int64_t)k * ld_index + d];
			if (!i && !j) continue;
			ids.emplace(FactorKey{ d, i, j }, 0);
			f++;
		}
		fmax = std::max(fmax, f);
	}
	if (fmax > 15)
		return set_error(HDDM_LB_ERROR_UNSUPPORTED,
			"upload_state: a node has %d active dims; at most 15 are supported", fmax);
	L.fmax = std::max(fmax, 1);
	L.rec_words = 2 + 2 * L.fmax;
	L.fac.clear();
	for (auto& kv : ids)
	{
		kv.second = (int)L.fac.size();
		Factor f;
		std::memset(&f, 0, sizeof(f));
		f.i = (double)kv.first.i;
		f.negj = -(double)kv.first.j;
		f.dim = (uint32_t)kv.first.d;
		f.xoff = (uint32_t)kv.first.d * (uint32_t)kXRowBytes;
		L.fac.push_back(f);
	}

	// Factor-id sequence of every node.
	std::vector<int32_t> seq((size_t)nno * L.fmax, -1);
	std::vector<int> nf(nno, 0);
	for (int k = 0; k < nno; k++)
		for (int d = 0; d < dim; d++)
		{
			int32_t i = index_i[(int64_t)k * ld_index + d], j = index_j[(int64_t)k * ld_index + d];
			if (!i && !j) continue;
			seq[(size_t)k * L.fmax + nf[k]++] = ids[FactorKey{ d, i, j }];
		}

	// Lexicographic order of the sequences (a prefix sorts first).
	L.order.resize(nno);
	for (int k = 0; k < nno; k++) L.order[k] = k;
	const int FM = L.fmax;
	std::stable_sort(L.order.begin(), L.order.end(), [&](int a, int b) {
		const int32_t* sa = &seq[(size_t)a * FM];
		const int32_t* sb = &seq[(size_t)b * FM];
		for (int t = 0; t < FM; t++)
		{
			if (t >= nf[a] || t >= nf[b]) return nf[a] < nf[b];
			if (sa[t] != sb[t]) return sa[t] < sb[t];
		}
		return false;
	});

	// Shared-prefix lengths and subtree ends.
	std::vector<int> F(nno), S(nno);
	for (int n = 0; n < nno; n++)
	{
		int k = L.order[n];
		F[n] = nf[k];
		int s = 0;
		if (n > 0)
		{
			int kp = L.order[n - 1];
			while (s < nf[k] && s < nf[kp] && seq[(size_t)k * FM + s] == seq[(size_t)kp * FM + s]) s++;
		}
		S[n] = s;
	}
	// end[n][t]: first node after n whose first t+1 factors differ from n's.
	std::vector<int> end((size_t)nno * FM, 0);
	for (int n = nno - 1; n >= 0; n--)
		for (int t = 0; t < F[n]; t++)
			end[(size_t)n * FM + t] = (n + 1 < nno && S[n + 1] >= t + 1) ? end[(size_t)(n + 1) * FM + t] : n + 1;

	L.rec.assign((size_t)nno * L.rec_words, 0);
	for (int n = 0; n < nno; n++)
	{
		uint32_t* r = &L.rec[(size_t)n * L.rec_words];
		r[0] = (uint32_t)F[n] | ((uint32_t)S[n] << 8);
		int k = L.order[n];
		for (int t = 0; t < F[n]; t++)
		{
			r[2 + 2 * t] = (uint32_t)seq[(size_t)k * FM + t];
			r[3 + 2 * t] = (uint32_t)end[(size_t)n * FM + t];
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

extern "C" int hddm_lb_layout_info(const hddm_lb_layout* l, int* fmax, int* rec_words,
	int* nfactors, int* nchunks, int* dofs_pad)
{
	if (!l) return hddm_lb::set_error(HDDM_LB_ERROR_INVALID_VALUE, "layout_info: NULL layout");
	if (fmax) *fmax = l->L.fmax;
	if (rec_words) *rec_words = l->L.rec_words;
	if (nfactors) *nfactors = (int)l->L.fac.size();
	if (nchunks) *nchunks = l->L.nchunks;
	if (dofs_pad) *dofs_pad = l->L.dofs_pad;
	return HDDM_LB_SUCCESS;
}

extern "C" int hddm_lb_layout_records(const hddm_lb_layout* l, uint32_t* rec)
{
	if (!l || !rec) return hddm_lb::set_error(HDDM_LB_ERROR_INVALID_VALUE, "layout_records: NULL argument");
	std::memcpy(rec, l->L.rec.data(), l->L.rec.size() * sizeof(uint32_t));
	return HDDM_LB_SUCCESS;
}

extern "C" int hddm_lb_layout_factors(const hddm_lb_layout* l, double* i, double* negj, int32_t* dim)
{
	if (!l || !i || !negj || !dim) return hddm_lb::set_error(HDDM_LB_ERROR_INVALID_VALUE, "layout_factors: NULL argument");
	for (size_t f = 0; f < l->L.fac.size(); f++)
	{
		i[f] = l->L.fac[f].i; negj[f] = l->L.fac[f].negj; dim[f] = (int32_t)l->L.fac[f].dim;
	}
	return HDDM_LB_SUCCESS;
}

extern "C" int hddm_lb_layout_order(const hddm_lb_layout* l, int32_t* order)
{
	if (!l || !order) return hddm_lb::set_error(HDDM_LB_ERROR_INVALID_VALUE, "layout_order: NULL argument");
	std::memcpy(order, l->L.order.data(), l->L.order.size() * sizeof(int32_t));
	return HDDM_LB_SUCCESS;
}

extern "C" int hddm_lb_layout_free(hddm_lb_layout* l)
{
	delete l;
	return HDDM_LB_SUCCESS;
}
