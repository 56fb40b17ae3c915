/*
 * interpolator.h -- the Interpolator API every postprocessor plugin implements (C++).
 *
 * Restates the host application's include/interpolator.h:1-43: the host
 * obtains the object from the plugin's getInterpolator() and calls the two
 * virtual interpolate() overloads through its vtable, so the field order
 * (jit, params), the virtual member order and the signatures must stay
 * exactly as declared there. This plugin is built with NAMESPACE=gfx950.
 */
#ifndef INTERPOLATOR_H
#define INTERPOLATOR_H

#include "parameters.h"

namespace NAMESPACE {

class Device;

enum BasisType
{
	LinearBasisType,
	PolynomialBasisType
};

class Data;

class Interpolator
{
	bool jit;

	const Parameters params;

public :
	const Parameters& getParameters() const;

	Interpolator(const std::string& targetSuffix, const std::string& configFile = "hddm-solver.cfg");

	static Interpolator* getInstance();

	// One point of one state: value[0:DofPerNode] = interpolant of state istate at x[0:dim].
	virtual void interpolate(Device* device, Data* data,
		const int istate, const real* x, const int DofPerNode, real* value);

	// One point per state: value[s][0:DofPerNode] at x[s][0:dim] for every loaded state s.
	virtual void interpolate(Device* device, Data* data,
		const real** x, const int DofPerNode, real** value);

	virtual ~Interpolator();
};

} // namespace NAMESPACE

#endif // INTERPOLATOR_H
