// The opened device (pxfdtd_device_open): its ordinal, properties and the device's address of
// the allocation-free flux's scratch array.
#pragma once
#include <hip/hip_runtime.h>

struct pxfdtd_device {
    int ordinal;
    hipDeviceProp_t prop;
    double *scratch = nullptr;
};

// flux.hip: the device's address of the flux's scratch array, looked up once per device
int flux_scratch_setup(int ordinal, double **out);
