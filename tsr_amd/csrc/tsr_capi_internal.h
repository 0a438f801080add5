// Internal declarations shared by the host C++ sources of the library.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/tsr_gpu.h"

struct tsr_error : std::runtime_error {
    tsr_status status;
    tsr_error(tsr_status s, const std::string &m) : std::runtime_error(m), status(s) {}
};

struct tsr_synthetic {
    std::vector<float> px, py, pz, nx, ny, nz, scale, conf;
    std::vector<uint32_t> cams;
    std::vector<float> cx, cy, cz;
    int width = 0, height = 0;
};

tsr_synthetic *tsr_synth_build(int config, uint64_t seed, double resScale);
