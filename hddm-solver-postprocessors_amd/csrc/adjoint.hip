// Adjoint kernels (placeholder until the adjoint sweep lands).
#include "context.h"
#include "errors.h"

extern "C" int hddm_lb_adjoint(hddm_lb_ctx* ctx, int64_t npoints,
	const double* x, int64_t ldx, const int32_t* state_ids,
	const float* gbar, int64_t ldg, float* dx, int64_t lddx,
	float* dsurplus, int precision, void* stream)
{
	(void)ctx; (void)npoints; (void)x; (void)ldx; (void)state_ids; (void)gbar; (void)ldg;
	(void)dx; (void)lddx; (void)dsurplus; (void)precision; (void)stream;
	return hddm_lb::set_error(HDDM_LB_ERROR_UNSUPPORTED, "adjoint: not available yet");
}
