// synth_model.hpp — the seeded draws of one synthetic frame (scene, per-pulse phase error, range
// walk), shared by the CPU generator (synth.cpp) and the GPU one (synth_gpu.hip), so that both
// image the same scene for the same seed.
#pragma once

#include <complex>
#include <cstdint>
#include <vector>

#include "isar_synth.h"

struct isar_synth_model_t {
    std::vector<double> sx, sy, sz;            // scatterer positions (m)
    std::vector<std::complex<double>> amp;     // complex amplitudes
    std::vector<double> perr;                  // per-pulse phase error (rad)
    std::vector<double> walk;                  // per-pulse residual range walk (m)
};

// The draws isar_synth_frame makes from cfg->seed.
void isar_synth_model(const isar_scene_config* cfg, isar_synth_model_t* out);
// Seed of row `row`'s noise stream (SplitMix64 state before its first draw).
uint64_t isar_synth_noise_seed(uint64_t seed, int row);
