/*
 * postprocessor.h -- the extern "C" exports of the drop-in plugin
 * lib/libpostprocessor.so (the gfx950 LinearBasis postprocessor).
 *
 * A host loads every backend's libpostprocessor.so and binds these names;
 * a replacement must export all of them with these signatures. Opaque types
 * here are the plugin's C++ classes (gfx950::Interpolator, gfx950::Data,
 * gfx950::Device, gfx950::DeviceProperties).
 *   getInterpolator      LinearBasis/cuda/src/Interpolator.cpp:130-133
 *   getData              LinearBasis/cuda/src/Data.cu:1210-1213
 *   isSupported          LinearBasis/cuda/src/supported.cpp:13-20
 *   getPriority          LinearBasis/cuda/src/priority.cpp:5-10 (cfg key priorityHIP
 *                        instead of priorityCUDA)
 *   resetInterpolator    LinearBasis/cuda/src/reset.cpp:14-17
 *   tryAcquireDevice     LinearBasis/cuda/src/Devices.cu:115-118
 *   releaseDevice        LinearBasis/cuda/src/Devices.cu:120-123
 *   getDeviceProperties  LinearBasis/cuda/src/DeviceProperties.cpp:46-51
 */
#ifndef HDDM_POSTPROCESSOR_H
#define HDDM_POSTPROCESSOR_H

#include <stddef.h>
#ifndef __cplusplus
#include <stdbool.h>
#endif

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pp_interpolator pp_interpolator;
typedef struct pp_data pp_data;
typedef struct pp_device pp_device;
typedef struct pp_device_properties pp_device_properties;

pp_interpolator* getInterpolator(void);
pp_data* getData(int nstates);
bool isSupported(void);
int getPriority(void);
void resetInterpolator(void);
pp_device* tryAcquireDevice(void);
void releaseDevice(pp_device* device);
void getDeviceProperties(pp_device_properties*** result, size_t* szresult);

#ifdef __cplusplus
}
#endif

#endif /* HDDM_POSTPROCESSOR_H */
