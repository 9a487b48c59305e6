// pipeline.hpp — the imaging context: device workspaces sized once, the batch launch sequence.
#pragma once

#include <string>
#include <vector>

#include "../launchers.hpp"
#include "isar_imaging_ext.h"

namespace isar {

void set_error(const std::string& msg);
int fail(int code, const std::string& msg);

// tables shared by every frame of a context: chirp-z filter spectra and chirps, range window
struct SharedTables {
    float2* czt_table;
    float2* czt_pre;
    float2* czt_post;
    float* win_range;
};

struct Context {
    isar_radar_params p{};
    int max_frames = 0;
    int F = 0;
    Job tmpl{};                       // workspace pointers; per-run fields filled in run()
    std::vector<void*> allocations;   // everything allocated at creation
    SharedTables tables{};            // inside the workspace
    int last_frames = 0;
    int front_frames = 0;             // batch size of the last run that formed the cut
    // kernel probe: event pairs recorded around each launch of one kernel family
    int probe_stage = -1;
    std::vector<hipEvent_t> probe_events;   // 2 per slot
    int probe_used = 0;
    double probe_ms = 0.0;
    int probe_launches = 0;
    int probe_enable(int stage);
    int probe_read(double* ms, int* launches);
    int probe_collect();
    template <typename F>
    void probed(int stage, hipStream_t s, F&& launch) {
        if (stage == probe_stage && probe_used * 2 + 1 < static_cast<int>(probe_events.size())) {
            (void)hipEventRecord(probe_events[2 * probe_used], s);
            launch();
            (void)hipEventRecord(probe_events[2 * probe_used + 1], s);
            ++probe_used;
        } else {
            launch();
        }
    }
    // device staging for isar_run_host: echo, velocity, phi0, image (allocated on first use)
    void* stage_dev[4] = {nullptr, nullptr, nullptr, nullptr};

    ~Context();
    int init(const isar_radar_params& params, int frames);
    // the same over a caller's workspace of workspace_bytes(params, frames) bytes at base (256-byte
    // aligned): no allocation; the tables are formed by kernels on stream s, asynchronously
    int init_in(const isar_radar_params& params, int frames, void* base, hipStream_t s);
    int run(int frames, const void* echo, const double* velocity, const void* phi0, float* img,
            const isar_options& opt, hipStream_t stream);
    // frame f reads echo rows frame_row[f] .. + M - 1 and writes its image at img + frame_img[f]
    // (either table may be null: contiguous frames); both tables are device arrays of K entries
    int run_frames(int frames, const void* echo, const double* velocity, const long long* frame_row,
                   const void* phi0, float* img, const long long* frame_img, const isar_options& opt,
                   hipStream_t stream);
    // the launches of one run, without graph replay
    int enqueue(int frames, const void* echo, const double* velocity, const long long* frame_row,
                const void* phi0, float* img, const long long* frame_img, const isar_options& opt,
                hipStream_t stream);
    // graph replay of full runs whose arguments repeat (isar_set_graph_mode)
    struct GraphKey {
        int frames = 0;
        const void* echo = nullptr;
        const double* velocity = nullptr;
        const long long* frame_row = nullptr;
        const void* phi0 = nullptr;
        float* img = nullptr;
        const long long* frame_img = nullptr;
        isar_options opt{};
        int ra_chunk = 0, af_mode = 0, ra_split = 0, czt_len = 0, czt_threads = 0, af_wait_polls = 0;
        const float2* hrrp = nullptr;
    };
    bool graph_mode = false;
    GraphKey graph_key{};
    Job graph_job{};
    hipStream_t graph_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    void drop_graph();
    // the last run, for isar_stage_read: its job (with the caller's buffer pointers) and stage
    // mode, and scratch for rebuilding intermediate arrays (allocated on the first stage read)
    Job last_job{};
    int last_mode = -1;
    float2* stage_scratch = nullptr;
    float2* hrrp_out = nullptr;           // max_frames x M x N profiles written by if_hrrp runs, or null
    int af_mode = ISAR_AF_AUTO;           // autofocus form: auto, one workgroup per frame, split
    float tuned_ms = 0.0f;                // time of one run in the configuration the autotuner kept
    float mode_ms[5] = {0, 0, 0, 0, 0};   // time of one run per autofocus form (index: isar_autofocus_mode)
    int ra_chunk = kRaChunkRowsDefault;   // tunable: rows per workgroup in the alignment update
    int ra_split = 0;                     // alignment update grids per level (isar_set_ra_split)
    int czt_threads = 256;                // chirp-z workgroup width (isar_set_czt_threads)
    int af_wait_polls = 1 << 20;          // bound of each resident-autofocus wait (isar_set_af_wait_polls)
    // chirp-z transform length (one of czt_length_candidates; 0 = the default): waits for the
    // device, then re-forms the chirp-z tables
    int set_czt_length(int length);
    // times every range-alignment chunk size on these frames and keeps the fastest (synchronous)
    // (with_resident: the resident autofocus form is among the candidates)
    int autotune(int frames, const void* echo, const double* velocity, const long long* frame_row, float* img,
                 const long long* frame_img, const isar_options& opt, hipStream_t s, int repeats, int* chosen,
                 bool with_resident = true);
    int trace(int frame, isar_frame_trace* out);
    int phases(int frames, void* d_start, void* d_final, hipStream_t s);
};

int next_pow2_ref(int n);   // common.cu:521-529
// chirp-z transform lengths with a plan for M pulses: nextPow2(2M - 1) first (the reference's
// length, mtrc.cu:125-172), then any shorter 3 * 2^k >= 2M - 1; returns how many (at most 2)
int czt_length_candidates(int M, int* out);
size_t workspace_bytes(const isar_radar_params& p, int frames);
size_t carve_workspace(const isar_radar_params& p, int frames, char* base, Job* J, SharedTables* tables);
int validate_params(const isar_radar_params& p);

}  // namespace isar

// the C API's opaque context handle
struct isar_context : isar::Context {};
