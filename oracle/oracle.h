/*
 * CPU oracle for the HDDM LinearBasis interpolation path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker, never the product:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it. The product path (hddm-solver-postprocessors_amd) never links it
 * and has no CPU fallback.
 *
 * Layout conventions follow the reference's dense gold containers
 * (LinearBasis/gold/include/Data.h:127-208): the index matrix holds, per node
 * row, the "i" values (0 or 2^(l-1)) in columns [0, dim) and the "j" values
 * in columns [vdim, vdim + dim); the caller passes the two column blocks as
 * separate row-major pointers with a shared leading dimension. The surplus
 * matrix is row-major [nno x ld_surplus].
 *
 * Pinning status of every function is recorded in oracle.c next to it.
 */
#ifndef HDDM_LB_ORACLE_H
#define HDDM_LB_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* phi(x; i, j) = 1 - |x*i - j|  (LinearBasis/include/LinearBasis.h:51). */
double oracle_phi(double x, int i, int j);

/* One point, one state: restates LinearBasis/gold/src/InterpolateArray.cpp:8-43. */
void oracle_forward_point(int dim, int nno, int dofs,
	const int32_t* index_i, const int32_t* index_j, int64_t ld_index,
	const double* surplus, int64_t ld_surplus,
	const double* x, double* value);

/* Many points against one state, OpenMP over points (nthreads <= 0: all). */
void oracle_forward_batch(int dim, int nno, int dofs,
	const int32_t* index_i, const int32_t* index_j, int64_t ld_index,
	const double* surplus, int64_t ld_surplus,
	int64_t npoints, const double* x, int64_t ldx, double* y, int64_t ldy,
	int nthreads);

/* One point per state: restates LinearBasis/gold/src/InterpolateArrayManyMultistate.cpp:8-49,
 * generalised to per-state dofs. Arrays of per-state pointers. */
void oracle_forward_multistate(int dim, int count, const int* nno, const int* dofs,
	const int32_t* const* index_i, const int32_t* const* index_j, const int64_t* ld_index,
	const double* const* surplus, const int64_t* ld_surplus,
	const double* const* x, double* const* value);

/* Analytic FP64 adjoint of oracle_forward_batch (no reference counterpart):
 *   dx[p][d] = sum_k (sum_c gbar[p][c] S[k][c]) dB_k/dx_d(x_p),
 *   dS[k][c] = sum_p B_k(x_p) gbar[p][c],
 * with dphi/dx = -i*sign(x*i - j) strictly inside the support and 0 outside.
 * dS is accumulated (caller zeroes it); dx is overwritten. */
void oracle_adjoint_batch(int dim, int nno, int dofs,
	const int32_t* index_i, const int32_t* index_j, int64_t ld_index,
	const double* surplus, int64_t ld_surplus,
	int64_t npoints, const double* x, int64_t ldx,
	const double* gbar, int64_t ldg, double* dx, int64_t lddx,
	double* dsurplus, int64_t ld_dsurplus, int nthreads);

/* ---- CPU baseline: the reference's sparse "xps/chains" algorithm ----
 * Restates the x86 backend: layout LinearBasis/x86/src/Data.cpp:598-951 and
 * kernel LinearBasis/x86/src/InterpolateArray.cpp:9-55. The layout object is
 * opaque; oracle_x86_forward_batch partitions points statically over OpenMP
 * threads and runs the per-point kernel on each. */
typedef struct oracle_x86_state oracle_x86_state;

oracle_x86_state* oracle_x86_build(int dim, int nno, int dofs,
	const int32_t* index_i, const int32_t* index_j, int64_t ld_index,
	const double* surplus, int64_t ld_surplus);
void oracle_x86_free(oracle_x86_state* st);
int oracle_x86_nfreqs(const oracle_x86_state* st);
int oracle_x86_nxps(const oracle_x86_state* st);
void oracle_x86_forward_batch(const oracle_x86_state* st,
	int64_t npoints, const double* x, int64_t ldx, double* y, int64_t ldy, int nthreads);

#ifdef __cplusplus
}
#endif

#endif /* HDDM_LB_ORACLE_H */
