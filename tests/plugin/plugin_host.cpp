// A minimal host for the drop-in plugin, standing in for the application
// that loads postprocessor plugins: it supplies MPI_Process (single process,
// master, abort() exits with status 3), dlopen()s lib/libpostprocessor.so,
// binds the extern "C" exports by name and calls the Interpolator and Data
// through their vtables, as a host compiled against include/interpolator.h
// does. Used by tests/test_plugin.py.
//
//   plugin_host <plugin.so> exports
//   plugin_host <plugin.so> kat <surplus.plt> [reps]            (cwd holds hddm-solver.cfg)
//   plugin_host <plugin.so> multistate <surplus.plt> <nstates>
//   plugin_host <plugin.so> reload <surplus.plt>                 (loads state 0 twice: abort path)
//   plugin_host <plugin.so> dofs <surplus.plt>                   (wrong DofPerNode: abort path)
//   plugin_host <plugin.so> latency <surplus.plt> <calls>        (per-point calls: prints us_per_call)
#include "interpolator.h"
#include "process.h"

#include <dlfcn.h>
#include <time.h>
#include <unistd.h>

#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

// ---- the host side of MPI_Process ----
static MPI_Process the_process;

bool MPI_Process::isMaster() const { return true; }
int MPI_Process::getRoot() const { return 0; }
int MPI_Process::getRank() const { return 0; }
int MPI_Process::getSize() const { return 1; }

void MPI_Process::abort()
{
	fprintf(stderr, "MPI_Process::abort\n");
	fflush(stdout);
	fflush(stderr);
	_exit(3);
}

void MPI_Process::cout(const char* format, ...)
{
	va_list ap;
	va_start(ap, format);
	vfprintf(stdout, format, ap);
	va_end(ap);
}

void MPI_Process::cerr(const char* format, ...)
{
	va_list ap;
	va_start(ap, format);
	vfprintf(stderr, format, ap);
	va_end(ap);
}

extern "C" int MPI_Process_get(MPI_Process** p)
{
	*p = &the_process;
	return 0;
}

// ---- what a host knows of the plugin's classes: the virtual members ----
namespace NAMESPACE {
class Data
{
public :
	virtual void load(const char* filename, int istate);
	virtual void clear();
	virtual ~Data();
};
class DeviceProperties;
}

using namespace NAMESPACE;

typedef Interpolator* (*GetInterpolator)();
typedef Data* (*GetData)(int);
typedef bool (*IsSupported)();
typedef int (*GetPriority)();
typedef void (*ResetInterpolator)();
typedef Device* (*TryAcquireDevice)();
typedef void (*ReleaseDevice)(Device*);
typedef void (*GetDeviceProperties)(DeviceProperties***, size_t*);

static void* bind(void* h, const char* name)
{
	void* f = dlsym(h, name);
	printf("export %s %s\n", name, f ? "found" : "MISSING");
	if (!f) exit(2);
	return f;
}

int main(int argc, char* argv[])
{
	if (argc < 3) { fprintf(stderr, "usage: plugin_host <plugin.so> <mode> ...\n"); return 2; }
	void* h = dlopen(argv[1], RTLD_NOW | RTLD_GLOBAL);
	if (!h) { fprintf(stderr, "dlopen: %s\n", dlerror()); return 2; }
	auto getInterpolator = (GetInterpolator)bind(h, "getInterpolator");
	auto getData = (GetData)bind(h, "getData");
	auto isSupported = (IsSupported)bind(h, "isSupported");
	auto getPriority = (GetPriority)bind(h, "getPriority");
	auto resetInterpolator = (ResetInterpolator)bind(h, "resetInterpolator");
	auto tryAcquireDevice = (TryAcquireDevice)bind(h, "tryAcquireDevice");
	auto releaseDevice = (ReleaseDevice)bind(h, "releaseDevice");
	auto getDeviceProperties = (GetDeviceProperties)bind(h, "getDeviceProperties");
	const char* mode = argv[2];

	bool supported = isSupported();
	printf("supported %d\n", (int)supported);
	DeviceProperties** props = nullptr;
	size_t nprops = 0;
	getDeviceProperties(&props, &nprops);
	printf("devices %zu\n", nprops);
	if (!strcmp(mode, "exports")) return 0;

	Interpolator* interp = getInterpolator();
	printf("priority %d\n", getPriority());
	const char* plt = argc > 3 ? argv[3] : "surplus.plt";
	Device* device = tryAcquireDevice();
	if (!device) { fprintf(stderr, "no device\n"); return 2; }
	Device* second = tryAcquireDevice();
	printf("second acquire %s\n", second ? "granted" : "refused");
	if (second) releaseDevice(second);

	const int dim = 59, dofs = 118;
	struct timespec t0 = {}, t1 = {};
	if (!strcmp(mode, "kat") || !strcmp(mode, "dofs") || !strcmp(mode, "reload"))
	{
		Data* data = getData(1);
		data->load(plt, 0);
		if (!strcmp(mode, "reload")) data->load(plt, 0);
		std::vector<double> x(dim), y(dofs, 0.0);
		for (int d = 0; d < dim; d++) x[d] = (d + 1) / 118.0;
		int reps = argc > 4 ? atoi(argv[4]) : 1;
		for (int r = 0; r < reps; r++)
			interp->interpolate(device, reinterpret_cast<NAMESPACE::Data*>(data), 0, x.data(),
				!strcmp(mode, "dofs") ? dofs - 1 : dofs, y.data());
		printf("value");
		for (int c = 0; c < dofs; c++) printf(" %.17g", y[c]);
		printf("\n");
		delete data;
	}
	else if (!strcmp(mode, "latency"))
	{
		// per-point calls, as the reference host makes them: reps calls after
		// 20 warm-up calls, each on a point moved slightly along the diagonal
		Data* data = getData(1);
		data->load(plt, 0);
		std::vector<double> x(dim), y(dofs, 0.0);
		const int reps = argc > 4 ? atoi(argv[4]) : 1000;
		for (int r = -20; r < reps; r++)
		{
			if (r == 0) clock_gettime(CLOCK_MONOTONIC, &t0);
			for (int d = 0; d < dim; d++) x[d] = (d + 1) / 118.0 + 1e-6 * ((r + 20) % 97);
			interp->interpolate(device, reinterpret_cast<NAMESPACE::Data*>(data), 0, x.data(), dofs, y.data());
		}
		clock_gettime(CLOCK_MONOTONIC, &t1);
		const double us = ((t1.tv_sec - t0.tv_sec) * 1e9 + (t1.tv_nsec - t0.tv_nsec)) * 1e-3 / (reps > 0 ? reps : 1);
		printf("us_per_call %.3f\n", us);
		printf("value");
		for (int c = 0; c < dofs; c++) printf(" %.17g", y[c]);
		printf("\n");
		delete data;
	}
	else if (!strcmp(mode, "multistate"))
	{
		const int n = argc > 4 ? atoi(argv[4]) : 4;
		Data* data = getData(n);
		for (int s = 0; s < n; s++) data->load(plt, s);
		std::vector<std::vector<double> > xs(n, std::vector<double>(dim)), ys(n, std::vector<double>(dofs, 0.0));
		std::vector<const double*> xp(n);
		std::vector<double*> yp(n);
		for (int s = 0; s < n; s++)
		{
			for (int d = 0; d < dim; d++) xs[s][d] = (double)((d + 1 + 7 * s) % 118) / 118.0;
			xp[s] = xs[s].data();
			yp[s] = ys[s].data();
		}
		interp->interpolate(device, reinterpret_cast<NAMESPACE::Data*>(data), xp.data(), dofs, yp.data());
		for (int s = 0; s < n; s++)
		{
			printf("x %d", s);
			for (int d = 0; d < dim; d++) printf(" %.17g", xs[s][d]);
			printf("\nvalue %d", s);
			for (int c = 0; c < dofs; c++) printf(" %.17g", ys[s][c]);
			printf("\n");
		}
		delete data;
	}
	releaseDevice(device);
	resetInterpolator();
	printf("done\n");
	return 0;
}
