// Device registry, device properties and the device exports of the plugin
// (LinearBasis/cuda/src/Devices.cu:13-123, LinearBasis/cuda/src/DeviceProperties.cpp:1-51).
#include "plugin.h"

#include <hip/hip_runtime_api.h>

#include <cstdarg>
#include <cstdio>

namespace NAMESPACE {

void fatal(const char* format, ...)
{
	char buf[1024];
	va_list ap;
	va_start(ap, format);
	vsnprintf(buf, sizeof(buf), format, ap);
	va_end(ap);
	MPI_Process* process = nullptr;
	MPI_Process_get(&process);
	process->cerr("%s\n", buf);
	process->abort();
	std::abort();  // the host's abort() must not return
}

Device::Device() : available(1), index(-1), id(0), warpSize(64), cus(0) { }

long long Device::getID() const { return id; }
int Device::getIndex() const { return index; }
int Device::getWarpSize() const { return warpSize; }
int Device::getCUCount() const { return cus; }

Devices::Devices()
{
	int n = 0;
	if (hddm_lb_device_count(&n) != HDDM_LB_SUCCESS) n = 0;
	devices.resize(n);
	for (int i = 0; i < n; i++)
	{
		int bus = 0, dev = 0, ws = 64, cus = 0;
		if (hipDeviceGetAttribute(&bus, hipDeviceAttributePciBusId, i) != hipSuccess) bus = 0;
		if (hipDeviceGetAttribute(&dev, hipDeviceAttributePciDeviceId, i) != hipSuccess) dev = 0;
		if (hipDeviceGetAttribute(&ws, hipDeviceAttributeWarpSize, i) != hipSuccess) ws = 64;
		if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, i) != hipSuccess) cus = 0;
		(void)hipGetLastError();
		devices[i].index = i;
		devices[i].id = (long long)(uint32_t)bus | ((long long)(uint32_t)dev << 32);
		devices[i].warpSize = ws;
		devices[i].cus = cus;
	}
}

int Devices::getCount() { return (int)devices.size(); }

const Device* Devices::getDevice(int index) const { return &devices[index]; }

Device* Devices::tryAcquire()
{
	std::lock_guard<std::mutex> lock(mtx);
	for (auto& d : devices)
		if (d.available)
		{
			d.available = 0;
			return &d;
		}
	return nullptr;
}

void Devices::release(Device* device)
{
	if (!device) return;
	std::lock_guard<std::mutex> lock(mtx);
	device->available = 1;
}

Devices& registry()
{
	static Devices devs;
	return devs;
}

SIMDVector::SIMDVector(const Device* device_) : device(device_) { }

size_t SIMDVector::getLength(size_t szelement) const
{
	if (szelement == 4 || szelement == 8) return (size_t)device->getWarpSize();
	fatal("Unsupported vector element size = %zu", szelement);
}

DeviceProperties::DeviceProperties(const Device* device) : simdVector(device) { }

const SIMDVector* DeviceProperties::getSIMDVector() const { return &simdVector; }

std::vector<DeviceProperties*>& DeviceProperties::getDeviceProperties()
{
	// created once and kept for the life of the process, like the devices
	static std::mutex mtx;
	static std::vector<DeviceProperties*> props;
	static bool init = false;
	std::lock_guard<std::mutex> lock(mtx);
	if (!init)
	{
		Devices& devs = registry();
		for (int i = 0; i < devs.getCount(); i++) props.push_back(new DeviceProperties(devs.getDevice(i)));
		init = true;
	}
	return props;
}

} // namespace NAMESPACE

using namespace NAMESPACE;

extern "C" bool isSupported()
{
	return registry().getCount() > 0;
}

extern "C" Device* tryAcquireDevice()
{
	return registry().tryAcquire();
}

extern "C" void releaseDevice(Device* device)
{
	registry().release(device);
}

extern "C" void getDeviceProperties(DeviceProperties*** result, size_t* szresult)
{
	std::vector<DeviceProperties*>& v = DeviceProperties::getDeviceProperties();
	*result = v.empty() ? nullptr : &v[0];
	*szresult = v.size();
}
