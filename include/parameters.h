/*
 * parameters.h -- configuration of a postprocessor plugin (C++).
 *
 * Restates the interface every LinearBasis backend declares
 * (LinearBasis/cuda/include/parameters.h:1-22 with the member list of
 * LinearBasis/cuda/include/parameters.c:1-17): a struct of references to
 * extern "C" globals, so that Fortran hosts see the same storage. The member
 * order and types must stay as they are: the host reads them through
 * Interpolator::getParameters().
 */
#ifndef HDDM_PARAMETERS_H
#define HDDM_PARAMETERS_H

#include <string>

typedef double real;

extern "C"
{
	extern int priority;              // interpolator's priority (cfg key priorityHIP)
	extern int nagents;               // number of agents in economy (cfg key n_agents or nagents)
	extern double surplusCutoff;      // small surplus values cutoff threshold
	extern bool surplusCutoffDefined; // set if surplusCutoff was given
}

namespace NAMESPACE {

struct Parameters
{
	int& priority;
	int& nagents;
	double& surplusCutoff;
	bool& surplusCutoffDefined;

	Parameters(const std::string& targetSuffix, const std::string& configFile = "hddm-solver.cfg");
};

} // namespace NAMESPACE

#endif // HDDM_PARAMETERS_H
