// The opened device (pxfdtd_device_open): its ordinal and properties.
#pragma once
#include <hip/hip_runtime.h>

struct pxfdtd_device {
    int ordinal;
    hipDeviceProp_t prop;
};
