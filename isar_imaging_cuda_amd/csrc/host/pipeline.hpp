// pipeline.hpp — the imaging context: device workspaces sized once, the batch launch sequence.
#pragma once

#include <string>
#include <vector>

#include "../launchers.hpp"
#include "isar_imaging_ext.h"

namespace isar {

void set_error(const std::string& msg);
int fail(int code, const std::string& msg);

struct Context {
    isar_radar_params p{};
    int max_frames = 0;
    int F = 0;
    Job tmpl{};                       // workspace pointers; per-run fields filled in run()
    std::vector<void*> allocations;   // everything allocated at creation
    float2* czt_table = nullptr;
    int last_frames = 0;
    // device staging for isar_run_host: echo, velocity, phi0, image (allocated on first use)
    void* stage_dev[4] = {nullptr, nullptr, nullptr, nullptr};

    ~Context();
    int init(const isar_radar_params& params, int frames);
    int run(int frames, const void* echo, const double* velocity, const void* phi0, float* img,
            const isar_options& opt, hipStream_t stream);
    int trace(int frame, isar_frame_trace* out);
};

int next_pow2_ref(int n);   // common.cu:521-529
size_t workspace_bytes(const isar_radar_params& p, int frames);
int validate_params(const isar_radar_params& p);

}  // namespace isar
