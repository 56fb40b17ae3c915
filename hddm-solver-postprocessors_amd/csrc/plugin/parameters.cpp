// Configuration of the gfx950 postprocessor: the same file format and keys
// as the reference backends (LinearBasis/cuda/src/parameters.cpp:37-156),
// with the priority under priorityHIP. One key is added: precisionHIP
// selects the contraction arithmetic (fp64, f16x2, bf16x3, f32).
#include "plugin.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>

extern "C"
{
	int priority = 0;
	int nagents = 0;
	double surplusCutoff = 0.0;
	bool surplusCutoffDefined = false;
}

namespace NAMESPACE {

static int g_precision = HDDM_LB_FP64;

int plugin_precision() { return g_precision; }

static int parse_precision(const char* v)
{
	if (!strcmp(v, "fp64")) return HDDM_LB_FP64;
	if (!strcmp(v, "f16x2")) return HDDM_LB_F16X2;
	if (!strcmp(v, "bf16x3")) return HDDM_LB_BF16X3;
	if (!strcmp(v, "f32")) return HDDM_LB_F32;
	return -1;
}

Parameters::Parameters(const std::string& targetSuffix, const std::string& configFile) :
	priority(::priority), nagents(::nagents), surplusCutoff(::surplusCutoff),
	surplusCutoffDefined(::surplusCutoffDefined)
{
	MPI_Process* process = nullptr;
	MPI_Process_get(&process);
	if (process->isMaster())
	{
		process->cout("------------------------------------\n");
		process->cout("LOADING \"%s\" POSTPROCESSOR CONFIG (%s)\n", targetSuffix.c_str(), configFile.c_str());
		process->cout("------------------------------------\n");
	}

	std::map<std::string, bool> undefined = { { "priority", true }, { "nagents", true } };

	FILE* cfg = fopen(configFile.c_str(), "r");
	if (!cfg)
	{
		process->cerr("Cannot open config file \"%s\"\n", configFile.c_str());
		process->abort();
		return;
	}
	bool cutoff = false;
	char* line = nullptr;
	size_t szline = 0;
	while (getline(&line, &szline, cfg) != -1)
	{
		char* save = nullptr;
		char* name = strtok_r(line, " \t\n", &save);
		if (!name || name[0] == '!') continue;
		char* value = strtok_r(nullptr, " \t\n", &save);
		if (!value || value[0] == '!')
		{
			if (process->isMaster()) process->cerr("Missing \"%s\" parameter value\n", name);
			process->abort();
			break;
		}
		const std::string key(name);
		if (key == "priorityHIP")
		{
			priority = atoi(value);
			undefined["priority"] = false;
			if (process->isMaster()) process->cout("priority : %d\n", priority);
		}
		else if (key == "n_agents" || key == "nagents")
		{
			nagents = atoi(value);
			undefined["nagents"] = false;
			if (process->isMaster()) process->cout("nagents : %d\n", nagents);
		}
		else if (key == "surplusCutoff")
		{
			surplusCutoff = atof(value);
			cutoff = true;
			if (process->isMaster()) process->cout("surplusCutoff : %e\n", surplusCutoff);
		}
		else if (key == "precisionHIP")
		{
			const int p = parse_precision(value);
			if (p < 0)
			{
				if (process->isMaster())
					process->cerr("precisionHIP must be fp64, f16x2, bf16x3 or f32, not \"%s\"\n", value);
				process->abort();
				break;
			}
			g_precision = p;
			if (process->isMaster()) process->cout("precisionHIP : %s\n", value);
		}
	}
	free(line);
	fclose(cfg);

	surplusCutoffDefined = cutoff;
	if (!cutoff && process->isMaster()) process->cout("surplusCutoff : unused\n");
	for (auto& u : undefined)
		if (u.second)
		{
			if (process->isMaster()) process->cerr("Required parameter \"%s\" is undefined\n", u.first.c_str());
			process->abort();
		}

	if (process->isMaster()) process->cout("%d GPU(s) available\n", registry().getCount());
}

} // namespace NAMESPACE
