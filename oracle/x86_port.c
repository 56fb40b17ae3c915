/*
 * CPU baseline: port of the reference's sparse "xps/chains" algorithm.
 *
 * TEST/BENCH INFRASTRUCTURE ONLY: timed by bench.py as the CPU baseline and
 * checked against the gold restatement in tests. Never linked by the product.
 *
 * Layout (LinearBasis/x86/src/Data.cpp:598-951):
 *   - xps: the unique (i, j, dim) factors over all nodes; entry 0 is the empty
 *     factor (0, 0, 0) whose value is 1.
 *   - chains: nno x nfreqs u16 xps ids, the active factors of each node in
 *     ascending dim order, 0-terminated; nfreqs = max active factors per node
 *     (maxRowPopulation, LinearBasis/gold/include/Data.h:182-207).
 *   - nodes are sorted lexicographically by chain and the surplus rows are
 *     permuted to match (LinearBasis/x86/src/Data.cpp:866-933).
 * xps ids here are assigned in first-appearance order instead of through the
 * reference's hash-ordered std::map; only the summation order of the DOF
 * update depends on that choice.
 *
 * Kernel (LinearBasis/x86/src/InterpolateArray.cpp:9-55): xpv[u] = max(0, phi)
 * for every unique factor, then per node the product of xpv over its chain
 * with an early exit on zero, then the DOF axpy.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

struct oracle_x86_state
{
	int dim, nno, dofs, nfreqs, nxps;
	int32_t* xi;      /* [nxps] */
	int32_t* xj;      /* [nxps] */
	int32_t* xdim;    /* [nxps] */
	uint16_t* chains; /* [nno * nfreqs] */
	double* surplus;  /* [nno * dofs], rows in chain order */
};

typedef struct { uint64_t key; int32_t id; } slot_t;

static uint64_t mix64(uint64_t z)
{
	z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
	z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
	return z ^ (z >> 31);
}

static int g_nfreqs;
static const uint16_t* g_chains;
static int cmp_chain(const void* a, const void* b)
{
	int ia = *(const int*)a, ib = *(const int*)b;
	const uint16_t* ca = g_chains + (int64_t)ia * g_nfreqs;
	const uint16_t* cb = g_chains + (int64_t)ib * g_nfreqs;
	for (int f = 0; f < g_nfreqs; f++)
		if (ca[f] != cb[f]) return ca[f] < cb[f] ? -1 : 1;
	return ia < ib ? -1 : (ia > ib);
}

oracle_x86_state* oracle_x86_build(int dim, int nno, int dofs,
	const int32_t* index_i, const int32_t* index_j, int64_t ld_index,
	const double* surplus, int64_t ld_surplus)
{
	oracle_x86_state* st = (oracle_x86_state*)calloc(1, sizeof(*st));
	st->dim = dim; st->nno = nno; st->dofs = dofs;

	int nfreqs = 0;
	for (int k = 0; k < nno; k++)
	{
		int f = 0;
		for (int d = 0; d < dim; d++)
			if (index_i[(int64_t)k * ld_index + d] || index_j[(int64_t)k * ld_index + d]) f++;
		if (f > nfreqs) nfreqs = f;
	}
	if (nfreqs == 0) nfreqs = 1;
	st->nfreqs = nfreqs;

	/* open-addressing table of unique factors */
	size_t cap = 1024;
	while (cap < (size_t)nno * 2) cap <<= 1;
	slot_t* table = (slot_t*)calloc(cap, sizeof(slot_t));
	int xcap = 1024, nx = 1;
	st->xi = (int32_t*)malloc(sizeof(int32_t) * xcap);
	st->xj = (int32_t*)malloc(sizeof(int32_t) * xcap);
	st->xdim = (int32_t*)malloc(sizeof(int32_t) * xcap);
	st->xi[0] = 0; st->xj[0] = 0; st->xdim[0] = 0;

	uint16_t* chains = (uint16_t*)calloc((size_t)nno * nfreqs, sizeof(uint16_t));
	for (int k = 0; k < nno; k++)
	{
		int f = 0;
		for (int d = 0; d < dim; d++)
		{
			int32_t i = index_i[(int64_t)k * ld_index + d], j = index_j[(int64_t)k * ld_index + d];
			if (!i && !j) continue;
			uint64_t key = ((uint64_t)(uint32_t)i << 40) ^ ((uint64_t)(uint32_t)j << 20) ^ (uint64_t)d;
			key = key * 2 + 1; /* never zero */
			size_t h = mix64(key) & (cap - 1);
			while (table[h].key && table[h].key != key) h = (h + 1) & (cap - 1);
			if (!table[h].key)
			{
				if (nx == xcap)
				{
					xcap *= 2;
					st->xi = (int32_t*)realloc(st->xi, sizeof(int32_t) * xcap);
					st->xj = (int32_t*)realloc(st->xj, sizeof(int32_t) * xcap);
					st->xdim = (int32_t*)realloc(st->xdim, sizeof(int32_t) * xcap);
				}
				table[h].key = key; table[h].id = nx;
				st->xi[nx] = i; st->xj[nx] = j; st->xdim[nx] = d;
				nx++;
			}
			chains[(int64_t)k * nfreqs + f++] = (uint16_t)table[h].id;
		}
	}
	free(table);
	st->nxps = nx;

	int* order = (int*)malloc(sizeof(int) * (size_t)nno);
	for (int k = 0; k < nno; k++) order[k] = k;
	g_nfreqs = nfreqs; g_chains = chains;
	qsort(order, (size_t)nno, sizeof(int), cmp_chain);

	st->chains = (uint16_t*)malloc(sizeof(uint16_t) * (size_t)nno * nfreqs);
	st->surplus = (double*)malloc(sizeof(double) * (size_t)nno * dofs);
	for (int k = 0; k < nno; k++)
	{
		memcpy(st->chains + (int64_t)k * nfreqs, chains + (int64_t)order[k] * nfreqs, sizeof(uint16_t) * nfreqs);
		memcpy(st->surplus + (int64_t)k * dofs, surplus + (int64_t)order[k] * ld_surplus, sizeof(double) * dofs);
	}
	free(order);
	free(chains);
	return st;
}

void oracle_x86_free(oracle_x86_state* st)
{
	if (!st) return;
	free(st->xi); free(st->xj); free(st->xdim); free(st->chains); free(st->surplus);
	free(st);
}

int oracle_x86_nfreqs(const oracle_x86_state* st) { return st->nfreqs; }
int oracle_x86_nxps(const oracle_x86_state* st) { return st->nxps; }

static void x86_point(const oracle_x86_state* st, const double* x, double* value, double* xpv)
{
	for (int u = 0; u < st->nxps; u++)
	{
		double xp = 1.0 - fabs(x[st->xdim[u]] * (double)st->xi[u] - (double)st->xj[u]);
		xpv[u] = fmax(0.0, xp);
	}
	const int dofs = st->dofs, nfreqs = st->nfreqs;
	memset(value, 0, sizeof(double) * (size_t)dofs);
	for (int k = 0; k < st->nno; k++)
	{
		const uint16_t* ch = st->chains + (int64_t)k * nfreqs;
		double temp = 1.0;
		for (int f = 0; f < nfreqs; f++)
		{
			uint16_t idx = ch[f];
			if (!idx) break;
			temp *= xpv[idx];
			if (temp == 0.0) goto next;
		}
		{
			const double* sk = st->surplus + (int64_t)k * dofs;
			for (int c = 0; c < dofs; c++) value[c] += temp * sk[c];
		}
	next:;
	}
}

void oracle_x86_forward_batch(const oracle_x86_state* st,
	int64_t npoints, const double* x, int64_t ldx, double* y, int64_t ldy, int nthreads)
{
#ifdef _OPENMP
	if (nthreads <= 0) nthreads = omp_get_max_threads();
	#pragma omp parallel num_threads(nthreads)
#endif
	{
		double* xpv = (double*)malloc(sizeof(double) * (size_t)st->nxps);
#ifdef _OPENMP
		#pragma omp for schedule(static)
#endif
		for (int64_t p = 0; p < npoints; p++)
			x86_point(st, x + p * ldx, y + p * ldy, xpv);
		free(xpv);
	}
	(void)nthreads;
}
