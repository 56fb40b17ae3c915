// Internal classes of the gfx950 LinearBasis postprocessor plugin.
//
// The host sees Interpolator (include/interpolator.h), Data and Device only
// through pointers returned by the extern "C" exports and calls Data's and
// Interpolator's virtual members; everything else here is private to the
// plugin. All GPU work goes through the engine's C ABI (hddm_lb.h).
#pragma once

#include "interpolator.h"
#include "process.h"
#include "hddm_lb.h"

#include <cstddef>
#include <cstdint>
#include <map>
#include <mutex>
#include <string>
#include <vector>

namespace NAMESPACE {

// Prints through the host's process object and aborts it, like every fatal
// path of the reference backends (process->cerr + process->abort()).
[[noreturn]] void fatal(const char* format, ...);

// One visible HIP device; exclusive to one acquirer at a time
// (LinearBasis/cuda/include/Device.h:26-58, LinearBasis/cuda/src/Devices.cu:59-80).
class Device
{
	int available;
	int index;          // HIP device ordinal
	long long id;       // PCI bus id (low word) and PCI device id (high word)
	int warpSize;
	int cus;

public :
	long long getID() const;
	int getIndex() const;
	int getWarpSize() const;
	int getCUCount() const;

	Device();

	friend class Devices;
};

class Devices
{
	std::vector<Device> devices;
	std::mutex mtx;

public :
	Devices();

	int getCount();
	const Device* getDevice(int index) const;
	Device* tryAcquire();
	void release(Device* device);
};

Devices& registry();

// Vector length hint of a device for element sizes 4 and 8: one wavefront
// (LinearBasis/cuda/include/SIMDVector.h:10-38).
class SIMDVector
{
	const Device* device;

public :
	SIMDVector(const Device* device);

	virtual size_t getLength(size_t szelement) const;
};

class DeviceProperties
{
	SIMDVector simdVector;

public :
	DeviceProperties(const Device* device);

	virtual const SIMDVector* getSIMDVector() const;

	static std::vector<DeviceProperties*>& getDeviceProperties();
};

// Surplus states of one host process (LinearBasis/cuda/include/Data.h:482-552):
// load() reads a .plt file per state into host memory; the states are
// uploaded to a device's engine context the first time that device
// interpolates them.
class Data
{
	int nstates, dim, TotalDof;

	struct State
	{
		bool loaded = false;
		int nno = 0;
		std::vector<int32_t> index_i, index_j;
		std::vector<double> surplus;
	};
	std::vector<State> states;

	struct DeviceCopy
	{
		hddm_lb_ctx* ctx = nullptr;
		std::vector<bool> uploaded;
	};
	std::map<int, DeviceCopy> copies;
	std::mutex mtx;

	friend class Interpolator;

	// The engine context of a device with every loaded state uploaded.
	hddm_lb_ctx* context(const Device* device);

public :
	virtual void load(const char* filename, int istate);

	virtual void clear();

	Data(int nstates);

	virtual ~Data();
};

// Contraction arithmetic of the plugin's interpolate(): cfg key precisionHIP
// (fp64 default, f16x2, bf16x3 or f32).
int plugin_precision();

} // namespace NAMESPACE
