// Autotuner of the sweep's formulation and launch shape, keyed by (precision,
// D, total nodes, max dofs, log2 of the point count) and by the options that
// restrict the candidates (formulation, waves, slab tiles). On the first forward of
// a key it times every candidate on the caller's data with hipEvents and
// keeps the fastest:
//   * the sweep where every wave walks and contracts its own block, with 2
//     or 4 waves per workgroup and 1 or 2 dof tiles per pass (the split
//     paths; FP64 always takes one);
//   * the wave-specialized sweep (a walker wave feeding MFMA consumer waves).
// Candidates are timed on the first min(npoints, kTunePoints) points. The
// prefix has to be long: the per-block sweep hides its latency with many
// resident waves, and on a short prefix (2^16 points, four waves per CU) it
// loses to the wave-specialized sweep that it beats two to one on 2^20
// points (D=16, FP64). Those rows of y are written like any forward's and
// are overwritten by the forward that follows.
#include "context.h"
#include "kernels.h"
#include "errors.h"

#include <algorithm>
#include <vector>

namespace hddm_lb {

constexpr int64_t kTunePoints = (int64_t)1 << 20;

int autotune(hddm_lb_ctx* ctx, int64_t npoints, const double* x, int64_t ldx, const int32_t* sid,
	double* y, int64_t ldy, int prec, hipStream_t st, TuneConfig& cfg)
{
	int64_t nodes = 0;
	int maxdofs = 0;
	for (auto& s : ctx->states)
		if (s && s->loaded) { nodes += s->desc.nno; maxdofs = std::max(maxdofs, s->desc.dofs); }
	int lg = 0;
	while (((int64_t)1 << (lg + 1)) <= npoints) lg++;
	auto key = std::make_tuple(prec, ctx->dim, nodes, maxdofs, lg, ctx->opt_form, ctx->opt_waves, ctx->opt_slab);
	auto hit = ctx->tuned.find(key);
	if (hit != ctx->tuned.end()) { cfg = hit->second; ctx->tune_hits++; return HDDM_LB_SUCCESS; }

	std::vector<TuneConfig> cands;
	if (ctx->opt_form != 2)
		for (int waves : { 2, 4 })
		{
			if (ctx->opt_waves && waves != ctx->opt_waves) continue;
			for (int slab : { 1, 2 })
			{
				if (slab > ctx->dt_needed || (prec == HDDM_LB_FP64 && slab > 1)) continue;
				if (ctx->opt_slab && slab != ctx->opt_slab) continue;
				TuneConfig c = cfg;
				c.ws = 0;
				c.walk = kWalkSmem;
				c.waves = waves;
				c.slab = slab;
				cands.push_back(c);
			}
		}
	if (ctx->opt_form != 1)
	{
		TuneConfig c = cfg;
		c.ws = 1;
		c.nc = ctx->dt_needed >= 2 ? 4 : 2;
		cands.push_back(c);
	}
	if (cands.empty()) { ctx->tuned[key] = cfg; return HDDM_LB_SUCCESS; }

	const int64_t m = std::min(npoints, kTunePoints);
	hipEvent_t e0, e1;
	if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
		return set_error(HDDM_LB_ERROR_HIP, "autotune: hipEventCreate failed");
	float best = 1e30f;
	TuneConfig bestc = cfg;
	int rc = HDDM_LB_SUCCESS;
	for (const TuneConfig& c : cands)
	{
		if ((rc = enqueue_forward(ctx, ctx->scratch, m, x, ldx, sid, y, ldy, prec, st, c, false))) break;
		(void)hipEventRecord(e0, st);
		if ((rc = enqueue_forward(ctx, ctx->scratch, m, x, ldx, sid, y, ldy, prec, st, c, false))) break;
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
	ctx->tune_runs++;
	cfg = bestc;
	return HDDM_LB_SUCCESS;
}

} // namespace hddm_lb
