// hip_errors.h -- what the reference's harnesses take from src/cuda_errors.h,
// for the HIP drop-in: Is_GPU_present() (src/cuda_errors.h:17-39), which
// tests/cuda_diff.cu calls first.  Device queries go through the C ABI
// (include/galop_hip.h), so this header needs no HIP headers either.
#pragma once

#include <cstdio>
#include <cstdlib>

#include "galop_hip.h"

inline void Is_GPU_present() {
    const int count = galop_device_count();
    if (count <= 0) {
        std::printf("No HIP devices!\n");
        std::exit(1);
    }
    char name[64] = {0};
    int32_t n_cu = 0, clock_khz = 0;
    if (galop_device_info(0, name, (int32_t)sizeof(name), &n_cu, &clock_khz) != GALOP_OK) {
        std::printf("No HIP devices!\n");
        std::exit(1);
    }
    std::printf("Device count, devid: %d %d\n", count, 0);
    std::printf("Device: %s\n", name);
    std::printf("Compute units: %d, peak clock %d MHz\n\n", n_cu, clock_khz / 1000);
}
