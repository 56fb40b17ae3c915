// Host-visible declarations of the HIP kernels and their launchers.
#pragma once

#include <hip/hip_runtime_api.h>
#include <stdint.h>

#include "hddm_lb.h"
#include "layout.h"

namespace hddm_lb {

struct SweepArgs
{
	const StateDesc* states;
	const StatePlan* plans;
	const int32_t* cost_order;
	PlanHeader* hdr;
	const uint32_t* perm;      // sorted position -> point
	const double* x;
	int64_t ldx;
	double* y;
	int64_t ldy;
	double* partial;           // node-segment partials of split states
	int nstates;
	int dim;
	int dt_base;               // first dof tile of 32 handled by this pass
	unsigned long long* stats; // optional: ops evaluated, prunes, chunks contracted, items, MFMAs, nnz
	uint32_t* status;          // device status bits (HDDM_LB_STATUS_*)
};

// Arguments of the adjoint kernels (adjoint.hip, adjoint_ws.hip).
struct AdjointArgs
{
	const StateDesc* states;
	const StatePlan* plans;
	const int32_t* cost_order;
	PlanHeader* hdr;
	const uint32_t* perm;
	const double* x;
	int64_t ldx;
	const double* g;
	int64_t ldg;
	double* dx;        // may be null
	int64_t lddx;
	double* dxpart;    // node-segment partials of dx for split states ([nseg][count][dim] per state)
	double* ds;        // may be null; state s starts at states[s].ds_offset
	int64_t ds_xcd;    // > 0: ds holds copies of this many elements (per XCD, or per owner class)
	int ds_own;        // wave-specialized adjoint, > 0: dS by owned units over this many copies
	int nstates;
	int dim;
	int dt_base;       // dof slab of 32 handled by this launch
	uint32_t* status;  // device status bits (HDDM_LB_STATUS_*)
	unsigned long long* stats;  // optional cycle and work counters of the wave-specialized adjoint
	int ndx;           // wave-specialized adjoint: dx waves with work (1 or 2)
};

// How a sweep reads the op stream: each wave through scalar loads, waves
// independent (kWalkSmem), or the workgroup through shared LDS tiles, its
// waves walking neighbouring point blocks in step (kWalkLds).
constexpr int kWalkLds = 0;
constexpr int kWalkSmem = 1;

// Kernel symbol of the sweep instantiation for (precision, dof tiles of 32
// per pass, register banks of 16 coordinates, max node depth, walk); NULL
// when unsupported. FP64 always runs one dof tile per pass.
const void* sweep_kernel_ptr(int prec, int dt, int xbanks, int depth, int walk);
// Wave-specialized sweep (sweep_ws.hip): a walker wave and nc consumer waves
// (2: one dof tile per pass, each consumer one half of it; 4: four tiles per
// pass); NULL when unsupported.
const void* sweep_ws_kernel_ptr(int prec, int xbanks, int depth, int nc);
size_t sweep_ws_lds_bytes(int prec, int depth);
// Wave-specialized adjoint (adjoint_ws.hip): a walker wave and 4 consumer
// waves; FP64 or F32; 4 dof tiles per launch. NULL when unsupported.
const void* adjoint_ws_kernel_ptr(int prec, int xbanks, int depth);
size_t adjoint_ws_lds_bytes(int prec, int depth, int dim, int ndx);
// Register banks needed for dim (0: dim above 64 is unsupported).
int xbanks_for(int dim);

hipError_t launch_fill_u32(uint32_t* p, int64_t n, uint32_t v, hipStream_t st);
hipError_t launch_bucket_key(int64_t n, const double* x, int64_t ldx, const int32_t* sid,
	const uint8_t* state_ok, int nstates, int dim, int bbits, uint32_t* bkey, uint32_t* hist, hipStream_t st);
hipError_t launch_scan_plan(const uint32_t* hist, uint32_t* cursor, int64_t nbuckets, int bbits, int nstates,
	const StateDesc* states, const int32_t* cost_order, StatePlan* plans, PlanHeader* hdr, int group, int dxw,
	hipStream_t st);
hipError_t launch_scatter(int64_t n, const uint32_t* bkey, uint32_t* cursor, uint32_t* perm, hipStream_t st);
hipError_t launch_reset_counter(PlanHeader* hdr, hipStream_t st);
hipError_t launch_sweep(const void* kern, int grid, int waves, size_t lds, const SweepArgs& args, hipStream_t st);
// Dynamic LDS of the sweep: the workgroup's op tiles (kWalkLds), then per
// wave the path stack and the chunk buffer.
size_t sweep_lds_bytes(int prec, int waves, int depth, int walk);
hipError_t launch_reduce(const StateDesc* states, const StatePlan* plans, const uint32_t* perm,
	const double* partial, double* y, int64_t ldy, int nstates, hipStream_t st);

} // namespace hddm_lb
