// The Interpolator of the gfx950 postprocessor and its extern "C" exports
// (LinearBasis/cuda/src/Interpolator.cpp:28-133, priority.cpp, reset.cpp).
// Both interpolate() overloads run the engine's batched forward on the
// acquired device: one point of one state, or one point per state.
#include "plugin.h"

#include <memory>

#define PP_STR(x) #x
#define PP_STRINGIZE(x) PP_STR(x)

namespace NAMESPACE {

std::unique_ptr<Interpolator> interp;

const Parameters& Interpolator::getParameters() const { return params; }

Interpolator::Interpolator(const std::string& targetSuffix, const std::string& configFile) :
	jit(false), params(targetSuffix, configFile)
{ }

Interpolator::~Interpolator() { }

Interpolator* Interpolator::getInstance()
{
	static std::mutex mtx;
	std::lock_guard<std::mutex> lock(mtx);
	if (!interp) interp.reset(new Interpolator(PP_STRINGIZE(NAME)));
	return interp.get();
}

static void check_dofs(const Data* data, int DofPerNode, int total)
{
	(void)data;
	if (DofPerNode != total)
		fatal("Requested DofPerNode (%d) mismatches actual data TotalDof (%d)", DofPerNode, total);
}

void Interpolator::interpolate(Device* device, Data* data,
	const int istate, const real* x, const int DofPerNode, real* value)
{
	check_dofs(data, DofPerNode, data->TotalDof);
	if (istate < 0 || istate >= data->nstates || !data->states[istate].loaded)
		fatal("State %d is not loaded", istate);
	hddm_lb_ctx* ctx = data->context(device);
	const int32_t sid = istate;
	if (hddm_lb_forward_host(ctx, 1, x, data->dim, &sid, value, DofPerNode, plugin_precision()) != HDDM_LB_SUCCESS)
		fatal("interpolate: %s", hddm_lb_last_error());
}

void Interpolator::interpolate(Device* device, Data* data,
	const real** x, const int DofPerNode, real** value)
{
	check_dofs(data, DofPerNode, data->TotalDof);
	const int n = data->nstates, dim = data->dim;
	for (int s = 0; s < n; s++)
		if (!data->states[s].loaded) fatal("State %d is not loaded", s);
	hddm_lb_ctx* ctx = data->context(device);
	std::vector<double> xs((size_t)n * dim), ys((size_t)n * DofPerNode);
	std::vector<int32_t> sid(n);
	for (int s = 0; s < n; s++)
	{
		for (int d = 0; d < dim; d++) xs[(size_t)s * dim + d] = x[s][d];
		sid[s] = s;
	}
	if (hddm_lb_forward_host(ctx, n, xs.data(), dim, sid.data(), ys.data(), DofPerNode, plugin_precision()) != HDDM_LB_SUCCESS)
		fatal("interpolate: %s", hddm_lb_last_error());
	for (int s = 0; s < n; s++)
		for (int c = 0; c < DofPerNode; c++) value[s][c] = ys[(size_t)s * DofPerNode + c];
}

} // namespace NAMESPACE

using namespace NAMESPACE;

extern "C" Interpolator* getInterpolator()
{
	return Interpolator::getInstance();
}

extern "C" int getPriority()
{
	return Interpolator::getInstance()->getParameters().priority;
}

extern "C" void resetInterpolator()
{
	interp = nullptr;
}
