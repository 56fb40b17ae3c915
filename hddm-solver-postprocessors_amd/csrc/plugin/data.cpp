// Surplus states of the plugin: .plt loading with the reference's checks
// (LinearBasis/gold/src/Data.cpp:322-338, 597-600) and lazy upload to the
// engine context of each device that interpolates them.
#include "plugin.h"

namespace NAMESPACE {

Data::Data(int nstates_) : nstates(nstates_), dim(0), TotalDof(0)
{
	if (nstates < 1) fatal("Number of states must be positive, got %d", nstates);
	states.resize(nstates);
}

Data::~Data()
{
	for (auto& c : copies)
		if (c.second.ctx) hddm_lb_destroy(c.second.ctx);
}

void Data::load(const char* filename, int istate)
{
	const Parameters& params = Interpolator::getInstance()->getParameters();
	if (istate < 0 || istate >= nstates)
		fatal("State %d index out of range 0 .. %d", istate, nstates - 1);
	std::lock_guard<std::mutex> lock(mtx);
	if (states[istate].loaded)
		fatal("State %d data is already loaded", istate);

	int d = 0, nno = 0, dofs = 0, level = 0;
	int32_t* ii = nullptr;
	int32_t* jj = nullptr;
	double* s = nullptr;
	if (hddm_lb_plt_read(filename, params.nagents, &d, &nno, &dofs, &level, &ii, &jj, &s) != HDDM_LB_SUCCESS)
		fatal("Cannot load state %d from \"%s\": %s", istate, filename, hddm_lb_last_error());
	bool any = false;
	for (auto& st : states) any |= st.loaded;
	if (any && (d != dim || dofs != TotalDof))
	{
		hddm_lb_free(ii); hddm_lb_free(jj); hddm_lb_free(s);
		fatal("State %d in \"%s\" has dim %d and %d dofs; loaded states have dim %d and %d dofs",
			istate, filename, d, dofs, dim, TotalDof);
	}
	dim = d;
	TotalDof = dofs;
	State& st = states[istate];
	st.nno = nno;
	st.index_i.assign(ii, ii + (size_t)nno * d);
	st.index_j.assign(jj, jj + (size_t)nno * d);
	st.surplus.assign(s, s + (size_t)nno * dofs);
	hddm_lb_free(ii); hddm_lb_free(jj); hddm_lb_free(s);
	st.loaded = true;
	for (auto& c : copies)
		if ((int)c.second.uploaded.size() > istate) c.second.uploaded[istate] = false;
}

void Data::clear()
{
	std::lock_guard<std::mutex> lock(mtx);
	for (auto& st : states) st.loaded = false;
	for (auto& c : copies) c.second.uploaded.assign(nstates, false);
}

hddm_lb_ctx* Data::context(const Device* device)
{
	const int dev = device ? device->getIndex() : 0;
	std::lock_guard<std::mutex> lock(mtx);
	DeviceCopy& c = copies[dev];
	if (!c.ctx && hddm_lb_create(dev, &c.ctx) != HDDM_LB_SUCCESS)
		fatal("Cannot open HIP device %d: %s", dev, hddm_lb_last_error());
	c.uploaded.resize(nstates, false);
	for (int i = 0; i < nstates; i++)
	{
		const State& st = states[i];
		if (!st.loaded || c.uploaded[i]) continue;
		if (hddm_lb_upload_state(c.ctx, i, dim, st.nno, TotalDof, st.index_i.data(), st.index_j.data(), dim,
				st.surplus.data(), TotalDof) != HDDM_LB_SUCCESS)
			fatal("Cannot upload state %d to HIP device %d: %s", i, dev, hddm_lb_last_error());
		c.uploaded[i] = true;
	}
	return c.ctx;
}

} // namespace NAMESPACE

using namespace NAMESPACE;

extern "C" Data* getData(int nstates)
{
	return new Data(nstates);
}
