// Autotuner of the sweep's launch shape, keyed by (precision, D, total nodes,
// max dofs, log2 of the point count). On the first forward of a key it times
// every candidate (op source x waves per workgroup x dof tiles per pass) on the caller's
// data with hipEvents and keeps the fastest. Tuning runs write y like the
// forward itself, so the caller's output is valid afterwards.
#include "context.h"
#include "kernels.h"
#include "errors.h"

#include <algorithm>
#include <vector>

namespace hddm_lb {

int autotune(hddm_lb_ctx* ctx, int64_t npoints, const double* x, int64_t ldx, const int32_t* sid,
	double* y, int64_t ldy, int prec, hipStream_t st, TuneConfig& cfg)
{
	int64_t nodes = 0;
	int maxdofs = 0;
	for (auto& s : ctx->states)
		if (s && s->loaded) { nodes += s->desc.nno; maxdofs = std::max(maxdofs, s->desc.dofs); }
	int lg = 0;
	while (((int64_t)1 << (lg + 1)) <= npoints) lg++;
	auto key = std::make_tuple(prec, ctx->dim, nodes, maxdofs, lg);
	auto hit = ctx->tuned.find(key);
	if (hit != ctx->tuned.end()) { cfg = hit->second; return HDDM_LB_SUCCESS; }

	std::vector<TuneConfig> cands;
	for (int walk : { kWalkSmem, kWalkLds })
	{
		if (ctx->opt_walk && walk != (ctx->opt_walk == 2 ? kWalkLds : kWalkSmem)) continue;
		for (int waves : { 1, 2, 4 })
		{
			if (ctx->opt_waves && waves != ctx->opt_waves) continue;
			for (int slab : { 1, 2 })
			{
				if (slab > ctx->dt_needed || (prec == HDDM_LB_FP64 && slab > 1)) continue;
				if (ctx->opt_slab && slab != ctx->opt_slab) continue;
				TuneConfig c = cfg;
				c.walk = walk;
				c.waves = waves;
				c.slab = slab;
				cands.push_back(c);
			}
		}
	}
	if (cands.empty()) { ctx->tuned[key] = cfg; return HDDM_LB_SUCCESS; }

	hipEvent_t e0, e1;
	if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
		return set_error(HDDM_LB_ERROR_HIP, "autotune: hipEventCreate failed");
	float best = 1e30f;
	TuneConfig bestc = cfg;
	int rc = HDDM_LB_SUCCESS;
	for (const TuneConfig& c : cands)
	{
		if ((rc = enqueue_forward(ctx, ctx->scratch, npoints, x, ldx, sid, y, ldy, prec, st, c, false))) break;
		(void)hipEventRecord(e0, st);
		if ((rc = enqueue_forward(ctx, ctx->scratch, npoints, x, ldx, sid, y, ldy, prec, st, c, false))) break;
		(void)hipEventRecord(e1, st);
		if (hipEventSynchronize(e1) != hipSuccess) { rc = set_error(HDDM_LB_ERROR_HIP, "autotune: event sync failed"); break; }
		float ms = 0.0f;
		(void)hipEventElapsedTime(&ms, e0, e1);
		if (ms < best) { best = ms; bestc = c; }
	}
	(void)hipEventDestroy(e0);
	(void)hipEventDestroy(e1);
	if (rc) return rc;
	ctx->tuned[key] = bestc;
	cfg = bestc;
	return HDDM_LB_SUCCESS;
}

} // namespace hddm_lb
