// Allocation-free forward: hddm_lb_forward with all of its per-call scratch
// (bucket keys, permutation, histogram and cursors, per-state plans, the
// work-queue header and the node-segment partials) carved out of one
// file-scope device array, for callers that cannot allocate device memory
// (code under hipGraph capture, tight serving loops).
//
// Each device the library runs on holds its own copy of `scratch` from module
// load until the process ends. In host code the name `scratch` is only a
// stand-in for the symbol; its device address comes from hipGetSymbolAddress
// once per context (with the context's device current) and is kept in
// hddm_lb_ctx::noalloc_scratch, which is what the launches use. Every call on
// a device shares that device's array: calls that overlap on one device
// (different streams, or threads) race. Everything the forward needs zeroed
// (the histogram) is cleared by a fill kernel inside the call.
#include "hddm_lb.h"
#include "errors.h"
#include "kernels.h"
#include "context.h"

#include <hip/hip_runtime.h>

#include <cstring>

namespace hddm_lb {

constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;

static __device__ __align__(256) char scratch[kNoAllocScratchBytes];

int noalloc_init(hddm_lb_ctx* ctx)
{
	void* p = nullptr;
	hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
	if (e != hipSuccess)
	{
		(void)hipGetLastError();
		return set_error(HDDM_LB_ERROR_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
	}
	ctx->noalloc_scratch = (char*)p;
	return HDDM_LB_SUCCESS;
}

namespace {

size_t align256(size_t v) { return (v + 255) & ~(size_t)255; }

} // namespace

} // namespace hddm_lb

using namespace hddm_lb;

extern "C" int hddm_lb_forward_noalloc(hddm_lb_ctx* ctx, int64_t npoints,
	const double* x, int64_t ldx, const int32_t* state_ids,
	double* y, int64_t ldy, int precision, void* stream)
{
	int rc = check_forward_args(ctx, npoints, x, ldx, y, ldy, precision);
	if (rc) return rc;
	if (npoints == 0) return HDDM_LB_SUCCESS;
	if (!ctx->noalloc_scratch)
		return set_error(HDDM_LB_ERROR_STATE, "forward_noalloc: the scratch array of device %d was not found", ctx->device);
	if (ctx->tables_dirty || ctx->tables_n != (int)ctx->states.size())
		return set_error(HDDM_LB_ERROR_STATE, "forward_noalloc: the state tables are not ready (upload a state first)");
	if (ctx->opt_stats)
		return set_error(HDDM_LB_ERROR_UNSUPPORTED, "forward_noalloc: statistics mode allocates; unset HDDM_LB_OPT_STATS");
	const int ns = ctx->tables_n;
	const int bb = bucket_bits(npoints, ns, ctx->opt_sort);
	const size_t nb = (size_t)(ns + 1) << bb;
	// layout of the scratch array, each piece 256-byte aligned
	size_t off = 0;
	auto take = [&](size_t bytes) { const size_t o = off; off = align256(off + bytes); return o; };
	const size_t o_bkey = take(sizeof(uint32_t) * (size_t)npoints);
	const size_t o_perm = take(sizeof(uint32_t) * (size_t)npoints);
	const size_t o_hist = take(sizeof(uint32_t) * nb);
	const size_t o_cursor = take(sizeof(uint32_t) * nb);
	const size_t o_plans = take(sizeof(StatePlan) * (size_t)ns);
	const size_t o_hdr = take(sizeof(PlanHeader));
	const size_t o_part = take(ctx->max_nseg_dp > 0 ? sizeof(double) * (size_t)npoints * (size_t)ctx->max_nseg_dp : 0);
	if (off > kNoAllocScratchBytes)
		return set_error(HDDM_LB_ERROR_INVALID_VALUE,
			"forward_noalloc: %lld points need %zu bytes of scratch; at most %zu are available",
			(long long)npoints, off, kNoAllocScratchBytes);
	char* base = ctx->noalloc_scratch;
	Scratch sc;
	sc.owned = false;
	sc.bkey = (uint32_t*)(base + o_bkey); sc.cap_bkey = (size_t)npoints;
	sc.perm = (uint32_t*)(base + o_perm); sc.cap_perm = (size_t)npoints;
	sc.hist = (uint32_t*)(base + o_hist); sc.cap_hist = nb;
	sc.cursor = (uint32_t*)(base + o_cursor); sc.cap_cursor = nb;
	sc.plans = (StatePlan*)(base + o_plans); sc.cap_plans = (size_t)ns;
	sc.hdr = (PlanHeader*)(base + o_hdr); sc.cap_hdr = 1;
	sc.partial = ctx->max_nseg_dp > 0 ? (double*)(base + o_part) : nullptr;
	sc.cap_partial = ctx->max_nseg_dp > 0 ? (size_t)npoints * (size_t)ctx->max_nseg_dp : 0;
	DeviceGuard g(ctx->device);
	// the default launch shape: tuning would create events and time candidates
	TuneConfig cfg = default_config(ctx, precision);
	rc = enqueue_forward(ctx, sc, npoints, x, ldx, state_ids, y, ldy, precision, (hipStream_t)stream, cfg, false);
	sc.bkey = sc.perm = sc.hist = sc.cursor = nullptr;
	sc.plans = nullptr;
	sc.hdr = nullptr;
	sc.partial = nullptr;
	return rc;
}
