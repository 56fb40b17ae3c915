/*
 * CPU oracle: FP64 restatement of the reference's gold backend.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product). Built with
 * -ffp-contract=off so that x*i - j and the DOF update are a multiply followed
 * by an add, as in a non-FMA build of the reference.
 *
 * Pinning (how each function is checked independently of itself):
 *   oracle_phi               closed form of LinearBasis/include/LinearBasis.h:51 plus
 *                            special cases (level 1 -> 1, support edges -> 0) in
 *                            tests/test_oracle.py.
 *   oracle_forward_point     the reference's known-answer test
 *                            (LinearBasis/tests/InterpolateArray/test.cpp:24-215 with
 *                            its surplus.plt fixture, 118 values to 1e-3), the
 *                            closed form y = S_root + sum_d (1 - 2 x_d) S_{d,j=0}
 *                            that fixture reduces to, and exact-rational brute force
 *                            on tiny random grids (tests/test_oracle.py).
 *   oracle_forward_batch     equals oracle_forward_point row by row (test).
 *   oracle_forward_multistate  equals oracle_forward_point per state (test).
 *   oracle_adjoint_batch     central finite differences of oracle_forward_point
 *                            (h = 1e-4, points off the kinks) for dx, and the exact
 *                            rational B^T g for dS on tiny grids (test).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

double oracle_phi(double x, int i, int j)
{
	/* LinearBasis/include/LinearBasis.h:51: (1.0 - fabs((x) * (i) - (j))) */
	return 1.0 - fabs(x * (double)i - (double)j);
}

/* LinearBasis/gold/src/InterpolateArray.cpp:8-43: value zeroed, then for every
 * node the product over all dims in ascending order with an early exit on a
 * factor <= 0, then value[c] += temp * surplus(k, c). */
void oracle_forward_point(int dim, int nno, int dofs,
	const int32_t* index_i, const int32_t* index_j, int64_t ld_index,
	const double* surplus, int64_t ld_surplus,
	const double* x, double* value)
{
	memset(value, 0, sizeof(double) * (size_t)dofs);
	for (int k = 0; k < nno; k++)
	{
		const int32_t* ik = index_i + (int64_t)k * ld_index;
		const int32_t* jk = index_j + (int64_t)k * ld_index;
		double temp = 1.0;
		int zero = 0;
		for (int d = 0; d < dim; d++)
		{
			double xp = oracle_phi(x[d], ik[d], jk[d]);
			if (xp <= 0.0) { zero = 1; break; }
			temp *= xp;
		}
		if (zero) continue;
		const double* sk = surplus + (int64_t)k * ld_surplus;
		for (int c = 0; c < dofs; c++)
			value[c] += temp * sk[c];
	}
}

void oracle_forward_batch(int dim, int nno, int dofs,
	const int32_t* index_i, const int32_t* index_j, int64_t ld_index,
	const double* surplus, int64_t ld_surplus,
	int64_t npoints, const double* x, int64_t ldx, double* y, int64_t ldy,
	int nthreads)
{
#ifdef _OPENMP
	if (nthreads <= 0) nthreads = omp_get_max_threads();
	#pragma omp parallel for schedule(static) num_threads(nthreads)
#endif
	for (int64_t p = 0; p < npoints; p++)
		oracle_forward_point(dim, nno, dofs, index_i, index_j, ld_index,
			surplus, ld_surplus, x + p * ldx, y + p * ldy);
	(void)nthreads;
}

/* LinearBasis/gold/src/InterpolateArrayManyMultistate.cpp:8-49: the single-state
 * kernel applied to state many < count with that state's x, index, surplus and
 * value. The reference shares DOF_PER_NODE across states; here dofs is per state. */
void oracle_forward_multistate(int dim, int count, const int* nno, const int* dofs,
	const int32_t* const* index_i, const int32_t* const* index_j, const int64_t* ld_index,
	const double* const* surplus, const int64_t* ld_surplus,
	const double* const* x, double* const* value)
{
	for (int many = 0; many < count; many++)
		oracle_forward_point(dim, nno[many], dofs[many], index_i[many], index_j[many],
			ld_index[many], surplus[many], ld_surplus[many], x[many], value[many]);
}

/* Adjoint of oracle_forward_point for one point; dS accumulated into dsp. */
static void adjoint_point(int dim, int nno, int dofs,
	const int32_t* index_i, const int32_t* index_j, int64_t ld_index,
	const double* surplus, int64_t ld_surplus,
	const double* x, const double* g, double* dx,
	double* dsp, int64_t ld_dsurplus, double* phi, double* dphi)
{
	for (int d = 0; d < dim; d++) dx[d] = 0.0;
	for (int k = 0; k < nno; k++)
	{
		const int32_t* ik = index_i + (int64_t)k * ld_index;
		const int32_t* jk = index_j + (int64_t)k * ld_index;
		double temp = 1.0;
		int zero = 0;
		for (int d = 0; d < dim; d++)
		{
			double t = x[d] * (double)ik[d] - (double)jk[d];
			double xp = 1.0 - fabs(t);
			if (xp <= 0.0) { zero = 1; break; }
			temp *= xp;
			phi[d] = xp;
			/* piecewise-constant hat derivative inside the support */
			dphi[d] = (t > 0.0) ? -(double)ik[d] : (t < 0.0 ? (double)ik[d] : 0.0);
		}
		if (zero) continue;
		const double* sk = surplus + (int64_t)k * ld_surplus;
		double w = 0.0;
		for (int c = 0; c < dofs; c++) w += g[c] * sk[c];
		double* dsk = dsp + (int64_t)k * ld_dsurplus;
		for (int c = 0; c < dofs; c++) dsk[c] += temp * g[c];
		for (int d = 0; d < dim; d++)
		{
			if (ik[d] == 0) continue; /* inactive dim: dphi = 0 */
			double others = 1.0;
			for (int e = 0; e < dim; e++)
				if (e != d) others *= phi[e];
			dx[d] += w * others * dphi[d];
		}
	}
}

void oracle_adjoint_batch(int dim, int nno, int dofs,
	const int32_t* index_i, const int32_t* index_j, int64_t ld_index,
	const double* surplus, int64_t ld_surplus,
	int64_t npoints, const double* x, int64_t ldx,
	const double* gbar, int64_t ldg, double* dx, int64_t lddx,
	double* dsurplus, int64_t ld_dsurplus, int nthreads)
{
#ifdef _OPENMP
	if (nthreads <= 0) nthreads = omp_get_max_threads();
#else
	nthreads = 1;
#endif
	size_t dssz = (size_t)nno * (size_t)ld_dsurplus;
	double* priv = (double*)calloc(dssz * (size_t)nthreads, sizeof(double));
#ifdef _OPENMP
	#pragma omp parallel num_threads(nthreads)
#endif
	{
		int tid = 0;
#ifdef _OPENMP
		tid = omp_get_thread_num();
#endif
		double* phi = (double*)malloc(sizeof(double) * (size_t)dim * 2);
		double* dphi = phi + dim;
		double* dsp = priv + dssz * (size_t)tid;
#ifdef _OPENMP
		#pragma omp for schedule(static)
#endif
		for (int64_t p = 0; p < npoints; p++)
			adjoint_point(dim, nno, dofs, index_i, index_j, ld_index, surplus, ld_surplus,
				x + p * ldx, gbar + p * ldg, dx + p * lddx, dsp, ld_dsurplus, phi, dphi);
		free(phi);
	}
	/* deterministic reduction in thread order */
	for (int t = 0; t < nthreads; t++)
	{
		const double* src = priv + dssz * (size_t)t;
		for (size_t e = 0; e < dssz; e++) dsurplus[e] += src[e];
	}
	free(priv);
}
