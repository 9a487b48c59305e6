// pipeline.cpp — imaging context and the batch launch sequence (isar_main.cu:3-227).
#include "pipeline.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>   // std::call_once
#include <string>
#include <vector>

namespace isar {

hipError_t configure_ra_kernels();
hipError_t configure_center_kernels();
hipError_t configure_pa_kernels();
hipError_t configure_mtrc_kernels();
hipError_t configure_ifds_kernels();
hipError_t configure_debug_kernels();

namespace {
thread_local std::string t_last_error;   // per calling thread, like errno
constexpr double kLightSpeed = 3e8;   // LIGHT_SPEED (common.cuh:69)
constexpr double kPi = 3.14159265358979323846;
}  // namespace

void set_error(const std::string& msg) { t_last_error = msg; }
int fail(int code, const std::string& msg) {
    set_error(msg);
    return code;
}
const char* last_error_cstr() { return t_last_error.c_str(); }

std::string g_config_error;
void report_config_error(const char* file, size_t lds_bytes) {
    if (g_config_error.empty())
        g_config_error = std::string(file) + ": dynamic LDS of " + std::to_string(lds_bytes) + " bytes rejected";
}

hipError_t configure_kernels() {
    static std::once_flag once;
    static hipError_t result = hipSuccess;
    std::call_once(once, [] {
        hipError_t errs[] = {configure_ra_kernels(), configure_center_kernels(), configure_pa_kernels(),
                             configure_mtrc_kernels(), configure_debug_kernels(), configure_ifds_kernels()};
        for (hipError_t e : errs)
            if (e != hipSuccess) result = e;
    });
    return result;
}

int next_pow2_ref(int N) {
    int n = 1;
    while (N >> 1) {
        n <<= 1;
        N >>= 1;
    }
    return n << 1;
}

int validate_params(const isar_radar_params& p) {
    // imagingMemInit (isar_imaging_export.cpp:144-155) plus the sizes the kernels exist for
    if (p.range_num_cut != kRangeCut) return fail(ISAR_ERR_INVALID, "range_num_cut must be 512");
    if (p.echo_num > 1024) return fail(ISAR_ERR_INVALID, "echo_num > 1024");
    if (p.range_num < p.range_num_cut) return fail(ISAR_ERR_INVALID, "range_num < range_num_cut");
    if (p.range_num < p.echo_num) return fail(ISAR_ERR_INVALID, "range_num < echo_num");
    if (!pulse_count_supported(p.echo_num))
        return fail(ISAR_ERR_INVALID, "echo_num " + std::to_string(p.echo_num) + " not supported");
    if (!range_size_supported(p.range_num))
        return fail(ISAR_ERR_INVALID, "range_num " + std::to_string(p.range_num) + " not supported");
    if (!(p.band_width > 0) || !(p.fc > 0) || !(p.Tp > 0) || !(p.Fs > 0))
        return fail(ISAR_ERR_INVALID, "radar parameters must be positive");
    return ISAR_OK;
}

// Places the per-batch buffers of K frames and the shared tables in one region starting at base
// (256-byte aligned), each buffer at a 256-byte boundary and sized K times its per-frame bytes
// rounded up to 256, so the total grows linearly with K. With base == nullptr it only counts.
// Returns the bytes the region needs; fills J's workspace pointers and *tables when base is set.
size_t carve_workspace(const isar_radar_params& p, int K_, char* base, Job* J, SharedTables* tables) {
    const size_t K = K_, M = p.echo_num, N = p.range_num, W = p.range_num_cut;
    const size_t F = next_pow2_ref(2 * p.echo_num - 1);
    const size_t nchunks = M / 2;   // final-row chunks are indexed by head row / 2
    const size_t nbands = (M + kCutBandRows - 1) / kCutBandRows;
    auto r256 = [](size_t b) { return (b + 255) & ~static_cast<size_t>(255); };
    size_t off = 0;
    auto take = [&](size_t per_frame, size_t count) -> void* {
        void* ptr = base ? base + off : nullptr;
        off += count * r256(per_frame ? per_frame : 16);
        return ptr;
    };
    auto frame = [&](size_t bytes) { return take(bytes, K); };
    auto shared = [&](size_t bytes) { return take(bytes, 1); };
    Job tmp{};
    Job& j = J ? *J : tmp;
    j.pre = static_cast<float2*>(frame(M * N * sizeof(float2)));
    j.st = j.pre;   // each row is read completely before its aligned profile overwrites it
    j.egrp = static_cast<float*>(frame(M * N * sizeof(float)));
    j.epart = static_cast<float*>(frame(M * N * sizeof(float)));
    j.arp_part = static_cast<float*>(frame(nchunks * N * sizeof(float)));
    j.max_part = static_cast<float*>(frame(nchunks * sizeof(float)));
    j.lagsum = static_cast<double*>(frame(M * sizeof(double)));
    j.lags = static_cast<double*>(frame(M * sizeof(double)));
    j.imax = static_cast<int32_t*>(frame(M * sizeof(int32_t)));
    j.dec = static_cast<FrameDecisions*>(frame(sizeof(FrameDecisions)));
    j.x = static_cast<float2*>(frame(M * W * sizeof(float2)));
    j.rowdop = static_cast<float2*>(frame(M * sizeof(float2)));
    j.dopphi = static_cast<float2*>(frame(M * sizeof(float2)));
    j.colmax_part = static_cast<float*>(frame(nbands * W * sizeof(float)));
    j.colmax = static_cast<float*>(frame(W * sizeof(float)));
    j.colE = static_cast<float*>(frame(W * sizeof(float)));
    j.colH = static_cast<float*>(frame(W * sizeof(float)));
    j.sel = static_cast<int32_t*>(frame(W * sizeof(int32_t)));
    j.su = static_cast<float2*>(frame(W * M * sizeof(float2)));
    j.phi0 = static_cast<float2*>(frame(M * sizeof(float2)));
    j.phi = static_cast<float2*>(frame(M * sizeof(float2)));
    j.stm = static_cast<float2*>(frame(M * W * sizeof(float2)));
    j.czt = static_cast<float2*>(frame(M * W * sizeof(float2)));
    j.colsum = static_cast<float2*>(frame(W * sizeof(float2)));
    j.delta = static_cast<float2*>(frame(W * sizeof(float2)));
    j.af_part = static_cast<float2*>(frame(4 * (W / 8) * M * sizeof(float2)));   // split: half; resident: all
    j.af_sync = static_cast<unsigned*>(frame(2 * sizeof(unsigned)));
    SharedTables t{};
    t.czt_table = static_cast<float2*>(shared(W * F * sizeof(float2)));
    t.czt_pre = static_cast<float2*>(shared(W * M * sizeof(float2)));
    t.czt_post = static_cast<float2*>(shared(W * M * sizeof(float2)));
    t.win_range = static_cast<float*>(shared(N * sizeof(float)));
    if (tables) *tables = t;
    return off;
}

size_t workspace_bytes(const isar_radar_params& p, int K) { return carve_workspace(p, K, nullptr, nullptr, nullptr); }

Context::~Context() {
    drop_graph();
    if (graph_stream) (void)hipStreamDestroy(graph_stream);
    for (hipEvent_t e : probe_events) (void)hipEventDestroy(e);
    for (void* a : allocations) (void)hipFree(a);
}

int Context::probe_enable(int stage) {
    probe_collect();
    probe_stage = stage;
    probe_ms = 0.0;
    probe_launches = 0;
    if (stage >= 0 && probe_events.empty()) {
        // 128 launches between collections (a run probes at most one launch per merge level of
        // a stage, and run_frames collects before the pool can run out)
        probe_events.resize(2 * 128);
        for (auto& e : probe_events)
            if (hipEventCreate(&e) != hipSuccess) return fail(ISAR_ERR_DEVICE, "hipEventCreate failed");
    }
    return ISAR_OK;
}

int Context::probe_collect() {
    if (probe_used == 0) return ISAR_OK;
    for (int i = 0; i < probe_used; ++i) {
        if (hipEventSynchronize(probe_events[2 * i + 1]) != hipSuccess) return fail(ISAR_ERR_DEVICE, "probe sync");
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, probe_events[2 * i], probe_events[2 * i + 1]) != hipSuccess)
            return fail(ISAR_ERR_DEVICE, "probe elapsed time");
        probe_ms += ms;
    }
    probe_launches += probe_used;
    probe_used = 0;
    return ISAR_OK;
}

int Context::probe_read(double* ms, int* launches) {
    int rc = probe_collect();
    if (ms) *ms = probe_ms;
    if (launches) *launches = probe_launches;
    probe_ms = 0.0;
    probe_launches = 0;
    return rc;
}

int Context::init(const isar_radar_params& params, int frames) {
    int rc = validate_params(params);
    if (rc) return rc;
    if (frames < 1) return fail(ISAR_ERR_INVALID, "max_frames must be >= 1");
    const size_t bytes = workspace_bytes(params, frames);
    void* base = nullptr;
    if (hipMalloc(&base, bytes) != hipSuccess)
        return fail(ISAR_ERR_MEMORY, "device allocation failed for " + std::to_string(frames) + " frames");
    allocations.push_back(base);
    rc = init_in(params, frames, base, nullptr);
    if (rc) return rc;
    const hipError_t e = hipStreamSynchronize(nullptr);
    if (e != hipSuccess) return fail(ISAR_ERR_DEVICE, std::string("workspace tables: ") + hipGetErrorString(e));
    return ISAR_OK;
}

int Context::init_in(const isar_radar_params& params, int frames, void* base, hipStream_t s) {
    int rc = validate_params(params);
    if (rc) return rc;
    if (frames < 1) return fail(ISAR_ERR_INVALID, "max_frames must be >= 1");
    hipError_t ce = configure_kernels();
    if (ce != hipSuccess)
        return fail(ISAR_ERR_DEVICE, std::string("kernel configuration: ") + hipGetErrorString(ce) + " " + g_config_error);
    p = params;
    max_frames = frames;
    Job& J = tmpl;
    J.M = p.echo_num;
    J.N = p.range_num;
    J.W = p.range_num_cut;
    F = next_pow2_ref(2 * p.echo_num - 1);
    J.F = F;
    SharedTables t{};
    carve_workspace(p, frames, static_cast<char*>(base), &J, &t);
    // constants (high_speed_compensation.cu:16-18, mtrc.cu:10-12, 47)
    const double B = static_cast<double>(p.band_width), fc = static_cast<double>(p.fc), Fs = static_cast<double>(p.Fs);
    J.hpc_k = -4.0 * (B / p.Tp) / kLightSpeed / (Fs * Fs);
    J.czt_c0 = B / (2.0 * fc);
    J.czt_posa = B / (fc * (static_cast<double>(J.W) - 1.0));
    tables = t;
    J.czt_table = t.czt_table;
    J.czt_pre = t.czt_pre;
    J.czt_post = t.czt_post;
    J.win_range = t.win_range;
    // the range Hamming window and the chirp-z tables, formed on the device on stream s
    launch_range_window(J.N, t.win_range, s);
    launch_czt_table(J.M, J.W, F, J.czt_c0, J.czt_posa, t.czt_table, s);
    launch_czt_chirps(J.M, J.W, F, J.czt_c0, J.czt_posa, t.czt_pre, t.czt_post, s);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(ISAR_ERR_DEVICE, std::string("workspace tables: ") + hipGetErrorString(e));
    return ISAR_OK;
}

int czt_length_candidates(int M, int* out) {
    const int n = 2 * M - 1;
    const int p2 = next_pow2_ref(n);
    int c = 0;
    out[c++] = p2;
    for (int f : {384, 768})   // the 3 * 2^k lengths with a transform plan
        if (f >= n && f < p2) out[c++] = f;
    return c;
}

int Context::set_czt_length(int length) {
    int cand[2];
    const int nc = czt_length_candidates(p.echo_num, cand);
    if (length == 0) length = cand[0];
    if (std::find(cand, cand + nc, length) == cand + nc)
        return fail(ISAR_ERR_INVALID, "chirp-z length " + std::to_string(length) + " is not available for " +
                                          std::to_string(p.echo_num) + " pulses");
    if (length == F) return ISAR_OK;
    // the tables may be in use by runs still in flight on any stream
    hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) return fail(ISAR_ERR_DEVICE, std::string("synchronize: ") + hipGetErrorString(e));
    F = length;
    tmpl.F = F;
    drop_graph();
    (void)hipGetLastError();
    launch_czt_table(tmpl.M, tmpl.W, F, tmpl.czt_c0, tmpl.czt_posa, tables.czt_table, nullptr);
    launch_czt_chirps(tmpl.M, tmpl.W, F, tmpl.czt_c0, tmpl.czt_posa, tables.czt_pre, tables.czt_post, nullptr);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
    if (e != hipSuccess) return fail(ISAR_ERR_DEVICE, std::string("chirp-z tables: ") + hipGetErrorString(e));
    return ISAR_OK;
}

int Context::run(int frames, const void* echo, const double* velocity, const void* phi0, float* img,
                 const isar_options& opt, hipStream_t s) {
    return run_frames(frames, echo, velocity, nullptr, phi0, img, nullptr, opt, s);
}

void Context::drop_graph() {
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
    graph_exec = nullptr;
    graph = nullptr;
}

int Context::run_frames(int frames, const void* echo, const double* velocity, const long long* frame_row,
                        const void* phi0, float* img, const long long* frame_img, const isar_options& opt,
                        hipStream_t s) {
    if (!graph_mode || opt.stage_mode != ISAR_RUN_FULL || probe_stage >= 0)
        return enqueue(frames, echo, velocity, frame_row, phi0, img, frame_img, opt, s);
    GraphKey k;
    k.frames = frames;
    k.echo = echo;
    k.velocity = velocity;
    k.frame_row = frame_row;
    k.phi0 = phi0;
    k.img = img;
    k.frame_img = frame_img;
    k.opt = opt;
    k.ra_chunk = ra_chunk;
    k.af_mode = af_mode;
    k.ra_split = ra_split;
    k.czt_len = F;
    k.czt_threads = czt_threads;
    k.af_wait_polls = af_wait_polls;
    k.hrrp = hrrp_out;
    const GraphKey& g0 = graph_key;
    const bool same = graph_exec && k.frames == g0.frames && k.echo == g0.echo && k.velocity == g0.velocity &&
                      k.frame_row == g0.frame_row && k.phi0 == g0.phi0 && k.img == g0.img &&
                      k.frame_img == g0.frame_img && k.ra_chunk == g0.ra_chunk && k.af_mode == g0.af_mode && k.ra_split == g0.ra_split &&
                      k.czt_len == g0.czt_len && k.czt_threads == g0.czt_threads && k.af_wait_polls == g0.af_wait_polls &&
                      k.hrrp == g0.hrrp && std::memcmp(&k.opt, &g0.opt, sizeof(isar_options)) == 0;
    if (!same) {
        // capture the launches on an internal stream; the graph is then launched on the caller's
        drop_graph();
        if (!graph_stream && hipStreamCreateWithFlags(&graph_stream, hipStreamNonBlocking) != hipSuccess)
            return fail(ISAR_ERR_DEVICE, "graph stream");
        if (hipStreamBeginCapture(graph_stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
            return fail(ISAR_ERR_DEVICE, "begin capture");
        const int rc = enqueue(frames, echo, velocity, frame_row, phi0, img, frame_img, opt, graph_stream);
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(graph_stream, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (ec != hipSuccess || hipGraphInstantiate(&graph_exec, g, nullptr, nullptr, 0) != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            graph_exec = nullptr;
            return fail(ISAR_ERR_DEVICE, "graph capture");
        }
        graph = g;
        graph_key = k;
        graph_job = last_job;
    }
    if (hipGraphLaunch(graph_exec, s) != hipSuccess) return fail(ISAR_ERR_DEVICE, "graph launch");
    // the bookkeeping of the captured run
    last_frames = frames;
    front_frames = frames;
    last_job = graph_job;
    last_mode = ISAR_RUN_FULL;
    return ISAR_OK;
}

int Context::enqueue(int frames, const void* echo, const double* velocity, const long long* frame_row,
                     const void* phi0, float* img, const long long* frame_img, const isar_options& opt,
                     hipStream_t s) {
    if (frames < 1 || frames > max_frames)
        return fail(ISAR_ERR_INVALID, "frames must be in [1, " + std::to_string(max_frames) + "]");
    const int mode = opt.stage_mode;
    if (mode != ISAR_RUN_FULL && mode != ISAR_RUN_FRONT && mode != ISAR_RUN_BACK)
        return fail(ISAR_ERR_INVALID, "stage_mode must be 0 (full), 1 (front) or 2 (back)");
    if (mode != ISAR_RUN_BACK && !echo) return fail(ISAR_ERR_INVALID, "null echo pointer");
    if (mode == ISAR_RUN_FULL && !img) return fail(ISAR_ERR_INVALID, "null image pointer");
    if (mode == ISAR_RUN_BACK && frames != front_frames)
        return fail(ISAR_ERR_STATE, "a back run needs a front or full run of the same frame count before it");
    if (opt.data_type != 1 && opt.data_type != 2) return fail(ISAR_ERR_INVALID, "data_type must be 1 (IFDS) or 2 (STRETCH)");
    const bool hpc = opt.if_hpc && opt.data_type == 2;
    if (hpc && !velocity && mode != ISAR_RUN_BACK)
        return fail(ISAR_ERR_INVALID, "high-speed compensation needs the velocity vector");
    if (opt.warm_start && !phi0) return fail(ISAR_ERR_INVALID, "warm start needs phi0");
    if (opt.iterations < 0) return fail(ISAR_ERR_INVALID, "iterations must be >= 0");
    Job J = tmpl;
    J.K = frames;
    J.data_type = opt.data_type;
    J.if_hpc = opt.if_hpc ? 1 : 0;
    J.if_hrrp = opt.if_hrrp ? 1 : 0;
    J.if_mtrc = opt.if_mtrc ? 1 : 0;
    J.quirk_q1 = opt.quirk_q1 ? 1 : 0;
    J.iterations = opt.iterations;
    J.phi0_given = opt.warm_start ? 1 : 0;
    J.echo = static_cast<const float2*>(echo);
    J.velocity = velocity;
    J.frame_row = frame_row;
    J.frame_img = frame_img;
    J.ra_chunk = ra_chunk;
    J.ra_split = ra_split;
    J.czt_threads = czt_threads;
    J.af_wait_polls = af_wait_polls;
    J.phi0_in = static_cast<const float2*>(phi0);
    J.img = img;
    J.hrrp = opt.if_hrrp && mode != ISAR_RUN_BACK ? hrrp_out : nullptr;
    const int M = J.M;

    (void)hipGetLastError();
    // a family launches at most log2(M) + 1 times per run; drain the ring only when it could overflow
    if (probe_stage >= 0 && probe_used + 16 > static_cast<int>(probe_events.size()) / 2) probe_collect();
    if (mode != ISAR_RUN_BACK) {
        if (J.data_type == 1) probed(ISAR_STAGE_IFDS_PRE, s, [&] { launch_ifds_pre(J, s); });
        else if (J.hrrp) probed(ISAR_STAGE_HRRP, s, [&] { launch_hrrp(J, s); });
        probed(ISAR_STAGE_RA_PAIR0, s, [&] { launch_ra_pair0(J, s); });
        int pair_base = M / 2;
        for (int st = 2; st < M; st *= 2) {
            const int ngroups = (M + st - 1) / st, npairs = ngroups / 2, unpaired = ngroups & 1;
            probed(ISAR_STAGE_RA_CORR, s, [&] { launch_ra_corr(J, st, pair_base, npairs + unpaired, s); });
            const int R = std::min(st, J.ra_chunk), cpp = (st + R - 1) / R;
            if (npairs > 0) probed(ISAR_STAGE_RA_UPD, s, [&] { launch_ra_upd(J, st, pair_base, npairs * cpp, s); });
            pair_base += npairs;
        }
        probed(ISAR_STAGE_CENTER, s, [&] { launch_center(J, s); });
        probed(ISAR_STAGE_CUT, s, [&] { launch_cut(J, s); });
        probed(ISAR_STAGE_PA_COLFFT, s, [&] { launch_pa_colfft(J, s); });
        probed(ISAR_STAGE_PA_SELECT, s, [&] { launch_pa_select(J, s); });
    }
    // the resident form's exchange tags hold the iteration in 7 bits; above that it runs split
    const bool af_resident = af_mode == ISAR_AF_RESIDENT && J.iterations <= 128;
    const bool af_split = af_mode == ISAR_AF_SPLIT || af_mode == ISAR_AF_SPLIT_PHASE || af_mode == ISAR_AF_RESIDENT ||
                          (af_mode == ISAR_AF_AUTO && frames <= kAfSplitFrames);
    probed(ISAR_STAGE_AUTOFOCUS, s, [&] {
        if (af_resident) launch_autofocus_resident(J, s);
        else if (af_split) launch_autofocus_split(J, af_mode != ISAR_AF_SPLIT_PHASE, s);
        else launch_autofocus(J, s);
    });
    if (mode == ISAR_RUN_FRONT || (mode == ISAR_RUN_BACK && !img)) {
        // nothing after the autofocus
    } else if (J.if_mtrc) {
        probed(ISAR_STAGE_MTRC_ROWS, s, [&] { launch_mtrc_rows(J, s); });
        probed(ISAR_STAGE_MTRC_CZT, s, [&] { launch_mtrc_czt(J, s); });
        if (!J.quirk_q1) probed(ISAR_STAGE_COL_WINDOW, s, [&] { launch_col_window(J, s); });
        else probed(ISAR_STAGE_MTRC_IMAGE, s, [&] { launch_mtrc_delta(J, s); });
        probed(ISAR_STAGE_MTRC_IMAGE, s, [&] { launch_mtrc_image(J, s); });
    } else {
        probed(ISAR_STAGE_POST_NOMTRC, s, [&] { launch_post_nomtrc(J, s); });
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(ISAR_ERR_DEVICE, std::string("kernel launch: ") + hipGetErrorString(e));
    last_frames = frames;
    if (mode != ISAR_RUN_BACK) front_frames = frames;
    last_job = J;
    last_mode = mode;
    return ISAR_OK;
}

int Context::autotune(int frames, const void* echo, const double* velocity, const long long* frame_row, float* img,
                      const long long* frame_img, const isar_options& opt, hipStream_t s, int repeats, int* chosen,
                      bool with_resident) {
    if (repeats < 1) return fail(ISAR_ERR_INVALID, "autotune needs at least one timed run");
    hipEvent_t a = nullptr, b = nullptr;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess)
        return fail(ISAR_ERR_DEVICE, "autotune: events");
    int rc = ISAR_OK;
    // one warm-up and `repeats` timed runs of the current configuration; negative on failure
    auto time_runs = [&]() -> float {
        rc = run_frames(frames, echo, velocity, frame_row, nullptr, img, frame_img, opt, s);
        if (rc) return -1.0f;
        (void)hipEventRecord(a, s);
        for (int r = 0; r < repeats && rc == ISAR_OK; ++r)
            rc = run_frames(frames, echo, velocity, frame_row, nullptr, img, frame_img, opt, s);
        (void)hipEventRecord(b, s);
        if (rc) return -1.0f;
        if (hipEventSynchronize(b) != hipSuccess) {
            rc = fail(ISAR_ERR_DEVICE, "autotune: timing");
            return -1.0f;
        }
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, a, b);
        return ms;
    };
    // range-alignment chunk rows
    float best_ms = 0.0f;
    int best = ra_chunk;
    for (int c : {2, 4, 8, 16}) {
        if (c > p.echo_num / 2) continue;
        ra_chunk = c;
        const float ms = time_runs();
        if (ms < 0.0f) break;
        if (best_ms == 0.0f || ms < best_ms) {
            best_ms = ms;
            best = c;
        }
    }
    ra_chunk = best;
    // autofocus form, with that chunk size
    if (rc == ISAR_OK) {
        float best_af = 0.0f;
        int best_mode = af_mode;
        for (int mode : {ISAR_AF_FUSED, ISAR_AF_SPLIT, ISAR_AF_SPLIT_PHASE, ISAR_AF_RESIDENT}) {
            if (mode == ISAR_AF_RESIDENT && !with_resident) continue;
            af_mode = mode;
            const float ms = time_runs();
            if (ms < 0.0f) break;
            mode_ms[mode] = ms / static_cast<float>(repeats);
            if (best_af == 0.0f || ms < best_af) {
                best_af = ms;
                best_mode = mode;
            }
        }
        af_mode = best_mode;
        tuned_ms = (best_af > 0.0f ? best_af : best_ms) / static_cast<float>(repeats);
    }
    // chirp-z length and workgroup width, timed on the chirp-z launches alone (a whole run is too
    // noisy for a stage of a few percent); then the kept configuration is timed again
    if (rc == ISAR_OK && opt.if_mtrc && opt.stage_mode == ISAR_RUN_FULL) {
        const int saved_probe = probe_stage;
        int cand[2];
        const int nc = czt_length_candidates(p.echo_num, cand);
        double best_czt = -1.0;
        int best_len = F, best_threads = czt_threads;
        for (int i = 0; i < nc && rc == ISAR_OK; ++i) {
            for (int t : {256, 512}) {
                rc = set_czt_length(cand[i]);
                if (rc) break;
                czt_threads = t;
                rc = probe_enable(ISAR_STAGE_MTRC_CZT);
                if (rc == ISAR_OK) rc = run_frames(frames, echo, velocity, frame_row, nullptr, img, frame_img, opt, s);
                if (rc == ISAR_OK) rc = probe_read(nullptr, nullptr);   // the warm-up run
                for (int r = 0; r < repeats && rc == ISAR_OK; ++r)
                    rc = run_frames(frames, echo, velocity, frame_row, nullptr, img, frame_img, opt, s);
                double ms = 0.0;
                int launches = 0;
                if (rc == ISAR_OK) rc = probe_read(&ms, &launches);
                if (rc) break;
                if (launches > 0 && (best_czt < 0.0 || ms / launches < best_czt)) {
                    best_czt = ms / launches;
                    best_len = cand[i];
                    best_threads = t;
                }
            }
        }
        const int prc = probe_enable(saved_probe);
        if (rc == ISAR_OK) rc = prc;
        if (rc == ISAR_OK) rc = set_czt_length(best_len);
        czt_threads = best_threads;
        if (rc == ISAR_OK) {
            const float ms = time_runs();
            if (ms >= 0.0f) tuned_ms = ms / static_cast<float>(repeats);
        }
    }
    if (chosen) *chosen = best;
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return rc;
}

int Context::phases(int frames, void* d_start, void* d_final, hipStream_t s) {
    if (frames < 1 || frames > last_frames) return fail(ISAR_ERR_INVALID, "frames exceeds the last run's batch");
    const size_t bytes = static_cast<size_t>(frames) * p.echo_num * sizeof(float2);
    if (d_start && hipMemcpyAsync(d_start, tmpl.phi0, bytes, hipMemcpyDeviceToDevice, s) != hipSuccess)
        return fail(ISAR_ERR_DEVICE, "phase copy failed");
    if (d_final && hipMemcpyAsync(d_final, tmpl.phi, bytes, hipMemcpyDeviceToDevice, s) != hipSuccess)
        return fail(ISAR_ERR_DEVICE, "phase copy failed");
    return ISAR_OK;
}

int Context::trace(int frame, isar_frame_trace* out) {
    if (!out || frame < 0 || frame >= last_frames) return fail(ISAR_ERR_INVALID, "no such frame in the last batch");
    const Job& J = tmpl;
    const size_t M = J.M, W = J.W;
    hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) return fail(ISAR_ERR_DEVICE, std::string("synchronize: ") + hipGetErrorString(e));
    FrameDecisions d{};
    e = hipMemcpy(&d, J.dec + frame, sizeof(d), hipMemcpyDeviceToHost);
    auto cp = [&](void* dst, const void* src, size_t bytes) {
        if (dst && e == hipSuccess) e = hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost);
    };
    cp(out->ra_lag, J.lags + frame * (M - 1), (M - 1) * sizeof(double));
    cp(out->ra_imax, J.imax + frame * (M - 1), (M - 1) * sizeof(int32_t));
    cp(out->pa_colmax, J.colmax + frame * W, W * sizeof(float));
    cp(out->pa_colE, J.colE + frame * W, W * sizeof(float));
    cp(out->pa_colH, J.colH + frame * W, W * sizeof(float));
    cp(out->select_bin, J.sel + frame * W, std::min<size_t>(W, std::max(d.unit2_num, 0)) * sizeof(int32_t));
    cp(out->phi0, J.phi0 + frame * M, M * sizeof(float2));
    cp(out->phi, J.phi + frame * M, M * sizeof(float2));
    if (e != hipSuccess) return fail(ISAR_ERR_DEVICE, std::string("trace copy: ") + hipGetErrorString(e));
    out->n_pairs = static_cast<int32_t>(M - 1);
    out->center_applied = d.center_applied;
    out->center_shift = d.center_shift;
    out->idx1_len = d.idx1_len;
    out->idx2_len = d.idx2_len;
    out->diff_imin = d.diff_imin;
    out->cut_offset = d.cut_offset;
    out->tgt_num = d.tgt_num;
    out->unit2_num = d.unit2_num;
    out->status = d.status;
    out->center_thr = d.center_thr;
    out->pa_threshold = d.pa_threshold;
    return ISAR_OK;
}

}  // namespace isar

// ------------------------------------------------------------------------------------------
// C ABI (isar_imaging_ext.h)
// ------------------------------------------------------------------------------------------
using isar::fail;

namespace {
// Host-fed device transforms for tests: one per workgroup (debug_fft) or several per workgroup,
// each within one wave (debug_fft_grouped).
int fft_host(int n, int dir, const float* h_in, float* h_out, int batch, bool grouped) {
    if (!h_in || !h_out || batch < 1 || (dir != -1 && dir != 1)) return fail(ISAR_ERR_INVALID, "bad fft arguments");
    hipError_t ce = isar::configure_kernels();
    if (ce != hipSuccess) return fail(ISAR_ERR_DEVICE, std::string("kernel configuration failed: ") + isar::g_config_error);
    const size_t bytes = static_cast<size_t>(n) * batch * sizeof(float2);
    void *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) {
        (void)hipFree(a);
        return fail(ISAR_ERR_MEMORY, "fft buffers");
    }
    hipError_t e = hipMemcpy(a, h_in, bytes, hipMemcpyHostToDevice);
    const auto* in = static_cast<const float2*>(a);
    auto* out = static_cast<float2*>(b);
    bool ok = e == hipSuccess && (grouped ? isar::debug_fft_grouped(n, dir, in, out, batch, nullptr)
                                          : isar::debug_fft(n, dir, in, out, batch, nullptr));
    if (ok) {
        e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpy(h_out, b, bytes, hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = hipDeviceSynchronize();
    }
    (void)hipFree(a);
    (void)hipFree(b);
    if (!ok) return fail(ISAR_ERR_INVALID, "unsupported fft length " + std::to_string(n));
    if (e != hipSuccess) return fail(ISAR_ERR_DEVICE, std::string("fft: ") + hipGetErrorString(e));
    return ISAR_OK;
}
}  // namespace

extern "C" {

ISAR_API void isar_default_options(isar_options* opt) {
    if (!opt) return;
    std::memset(opt, 0, sizeof(*opt));
    opt->data_type = 2;
    opt->option_alignment = 0;
    opt->option_phase = 1;
    opt->if_hrrp = 0;
    opt->if_hpc = 1;
    opt->if_mtrc = 1;
    opt->quirk_q1 = 1;
    opt->iterations = 120;   // FAST_ENTROPY_ITERATION_NUM (common.cuh:73)
    opt->warm_start = 0;
    opt->stage_mode = ISAR_RUN_FULL;
}

ISAR_API size_t isar_workspace_bytes(const isar_radar_params* p, int max_frames) {
    return p ? isar::workspace_bytes(*p, max_frames) : 0;
}

ISAR_API int isar_create(const isar_radar_params* p, int max_frames, isar_context** ctx) {
    if (!p || !ctx) return fail(ISAR_ERR_INVALID, "null argument");
    *ctx = nullptr;
    auto* c = new isar_context();
    int rc = c->init(*p, max_frames);
    if (rc) {
        delete c;
        return rc;
    }
    *ctx = c;
    return ISAR_OK;
}

ISAR_API int isar_destroy(isar_context* ctx) {
    if (!ctx) return fail(ISAR_ERR_INVALID, "null context");
    (void)hipDeviceSynchronize();
    delete ctx;
    return ISAR_OK;
}

ISAR_API int isar_run_device(isar_context* ctx, int frames, const void* d_echo, const double* d_velocity,
                             const void* d_phi0, float* d_img, const isar_options* opt, void* stream) {
    if (!ctx || !opt) return fail(ISAR_ERR_INVALID, "null context or options");
    return ctx->run(frames, d_echo, d_velocity, d_phi0, d_img, *opt, static_cast<hipStream_t>(stream));
}

ISAR_API int isar_run_host(isar_context* ctx, int frames, const float* h_echo, const double* h_velocity,
                           const float* h_phi0, float* h_img, const isar_options* opt) {
    if (!ctx || !opt || !h_echo || !h_img) return fail(ISAR_ERR_INVALID, "null argument");
    if (frames < 1 || frames > ctx->max_frames) return fail(ISAR_ERR_INVALID, "frames out of range");
    const size_t M = ctx->p.echo_num, N = ctx->p.range_num, W = ctx->p.range_num_cut;
    // device staging buffers are allocated on first use and kept for the context's lifetime
    if (!ctx->stage_dev[0]) {
        const size_t K = ctx->max_frames;
        const size_t bytes[4] = {K * M * N * sizeof(float2), K * M * sizeof(double), K * M * sizeof(float2),
                                 K * M * W * sizeof(float)};
        for (int i = 0; i < 4; ++i) {
            void* ptr = nullptr;
            if (hipMalloc(&ptr, bytes[i]) != hipSuccess) return fail(ISAR_ERR_MEMORY, "staging allocation failed");
            ctx->allocations.push_back(ptr);
            ctx->stage_dev[i] = ptr;
        }
    }
    void* d_echo = ctx->stage_dev[0];
    double* d_vel = static_cast<double*>(ctx->stage_dev[1]);
    void* d_phi0 = ctx->stage_dev[2];
    float* d_img = static_cast<float*>(ctx->stage_dev[3]);
    hipError_t e = hipMemcpy(d_echo, h_echo, frames * M * N * sizeof(float2), hipMemcpyHostToDevice);
    if (e == hipSuccess && h_velocity) e = hipMemcpy(d_vel, h_velocity, frames * M * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess && h_phi0) e = hipMemcpy(d_phi0, h_phi0, frames * M * sizeof(float2), hipMemcpyHostToDevice);
    if (e != hipSuccess) return fail(ISAR_ERR_DEVICE, std::string("upload: ") + hipGetErrorString(e));
    int rc = ctx->run(frames, d_echo, h_velocity ? d_vel : nullptr, h_phi0 ? d_phi0 : nullptr, d_img, *opt, nullptr);
    if (rc) return rc;
    e = hipMemcpy(h_img, d_img, frames * M * W * sizeof(float), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) return fail(ISAR_ERR_DEVICE, std::string("run/download: ") + hipGetErrorString(e));
    return ISAR_OK;
}

ISAR_API int isar_run_frames(isar_context* ctx, int frames, const void* d_echo, const double* d_velocity,
                             const int64_t* d_frame_rows, const void* d_phi0, float* d_img,
                             const int64_t* d_frame_img, const isar_options* opt, void* stream) {
    if (!ctx || !opt) return isar::fail(ISAR_ERR_INVALID, "null context or options");
    static_assert(sizeof(long long) == sizeof(int64_t), "frame tables are int64");
    return ctx->run_frames(frames, d_echo, d_velocity, reinterpret_cast<const long long*>(d_frame_rows), d_phi0,
                           d_img, reinterpret_cast<const long long*>(d_frame_img), *opt,
                           static_cast<hipStream_t>(stream));
}

ISAR_API int isar_pulse_compress(const isar_radar_params* p, int echoes, int len, const int16_t* d_iq,
                                 const double* d_velocity, void* d_out, int32_t* d_offsets, void* stream) {
    if (!p || !d_iq || !d_velocity || !d_out || echoes < 1 || len < 1)
        return isar::fail(ISAR_ERR_INVALID, "isar_pulse_compress: bad arguments");
    const int nfft = isar::next_pow2_ref(len);
    const int S = static_cast<int>(static_cast<double>(p->Fs) * p->Tp);
    if (!isar::pulse_compress_supported(nfft))
        return isar::fail(ISAR_ERR_INVALID, "isar_pulse_compress: echo length " + std::to_string(len) +
                                          " gives an unsupported transform size " + std::to_string(nfft));
    if (S < 2 || S > nfft) return isar::fail(ISAR_ERR_INVALID, "isar_pulse_compress: reference chirp (Fs Tp samples) out of range");
    hipError_t ce = isar::configure_kernels();
    if (ce != hipSuccess) return isar::fail(ISAR_ERR_DEVICE, std::string("kernel configuration: ") + hipGetErrorString(ce));
    isar::PcJob P{};
    P.iq = d_iq;
    P.velocity = d_velocity;
    P.out = static_cast<float2*>(d_out);
    P.offsets = d_offsets;
    P.echoes = echoes;
    P.len = len;
    P.S = S;
    P.cut = isar::kRangeNumIfdsPc;
    P.fc = static_cast<double>(p->fc);
    P.band_width = static_cast<double>(p->band_width);
    P.tp = p->Tp;
    (void)hipGetLastError();
    isar::launch_pulse_compress(P, nfft, static_cast<hipStream_t>(stream));
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return isar::fail(ISAR_ERR_DEVICE, std::string("isar_pulse_compress: ") + hipGetErrorString(e));
    return ISAR_OK;
}

ISAR_API int isar_set_ra_chunk(isar_context* ctx, int rows) {
    if (!ctx) return isar::fail(ISAR_ERR_INVALID, "null context");
    if (rows != 2 && rows != 4 && rows != 8 && rows != 16)
        return isar::fail(ISAR_ERR_INVALID, "ra chunk must be 2, 4, 8 or 16 rows");
    ctx->ra_chunk = rows;
    return ISAR_OK;
}

ISAR_API int isar_get_ra_chunk(isar_context* ctx) { return ctx ? ctx->ra_chunk : 0; }

ISAR_API int isar_set_ra_split(isar_context* ctx, int mode) {
    if (!ctx) return isar::fail(ISAR_ERR_INVALID, "null context");
    if (mode < 0 || mode > 2)
        return isar::fail(ISAR_ERR_INVALID, "ra split must be 0 (by batch size), 1 (one grid) or 2 (two grids)");
    ctx->ra_split = mode;
    return ISAR_OK;
}

ISAR_API int isar_get_ra_split(isar_context* ctx) { return ctx ? ctx->ra_split : -1; }

ISAR_API int isar_set_czt_length(isar_context* ctx, int length) {
    if (!ctx) return isar::fail(ISAR_ERR_INVALID, "null context");
    return ctx->set_czt_length(length);
}

ISAR_API int isar_get_czt_length(isar_context* ctx) { return ctx ? ctx->F : 0; }

ISAR_API int isar_czt_lengths(isar_context* ctx, int32_t* out, int max) {
    if (!ctx || !out || max < 1) return isar::fail(ISAR_ERR_INVALID, "isar_czt_lengths: bad arguments");
    int cand[2];
    const int n = isar::czt_length_candidates(ctx->p.echo_num, cand);
    for (int i = 0; i < std::min(n, max); ++i) out[i] = cand[i];
    return std::min(n, max);
}

ISAR_API int isar_set_czt_threads(isar_context* ctx, int threads) {
    if (!ctx) return isar::fail(ISAR_ERR_INVALID, "null context");
    if (threads != 256 && threads != 512) return isar::fail(ISAR_ERR_INVALID, "chirp-z workgroups have 256 or 512 threads");
    ctx->czt_threads = threads;
    return ISAR_OK;
}

ISAR_API int isar_get_czt_threads(isar_context* ctx) { return ctx ? ctx->czt_threads : 0; }

ISAR_API int isar_set_autofocus_mode(isar_context* ctx, int mode) {
    if (!ctx) return isar::fail(ISAR_ERR_INVALID, "null context");
    if (mode != ISAR_AF_AUTO && mode != ISAR_AF_FUSED && mode != ISAR_AF_SPLIT && mode != ISAR_AF_SPLIT_PHASE &&
        mode != ISAR_AF_RESIDENT)
        return isar::fail(ISAR_ERR_INVALID, "autofocus mode must be 0 (auto), 1 (one workgroup per frame), 2 (split), "
                                            "3 (split, phase launches) or 4 (resident split)");
    ctx->af_mode = mode;
    return ISAR_OK;
}

ISAR_API int isar_get_autofocus_mode(isar_context* ctx) { return ctx ? ctx->af_mode : -1; }

ISAR_API int isar_set_af_wait_polls(isar_context* ctx, int polls) {
    if (!ctx) return isar::fail(ISAR_ERR_INVALID, "null context");
    ctx->af_wait_polls = polls;
    return ISAR_OK;
}

ISAR_API int isar_set_graph_mode(isar_context* ctx, int on) {
    if (!ctx) return isar::fail(ISAR_ERR_INVALID, "null context");
    ctx->graph_mode = on != 0;
    if (!ctx->graph_mode) ctx->drop_graph();
    return ISAR_OK;
}

ISAR_API int isar_set_hrrp_output(isar_context* ctx, void* d_hrrp) {
    if (!ctx) return isar::fail(ISAR_ERR_INVALID, "null context");
    ctx->hrrp_out = static_cast<float2*>(d_hrrp);
    return ISAR_OK;
}

ISAR_API int isar_autotune(isar_context* ctx, int frames, const void* d_echo, const double* d_velocity, float* d_img,
                           const isar_options* opt, int repeats, int32_t* chosen, void* stream) {
    if (!ctx || !opt) return isar::fail(ISAR_ERR_INVALID, "null context or options");
    int c = 0;
    const int rc = ctx->autotune(frames, d_echo, d_velocity, nullptr, d_img, nullptr, *opt,
                                 static_cast<hipStream_t>(stream), repeats, &c);
    if (chosen) *chosen = c;
    return rc;
}

ISAR_API int isar_get_phases(isar_context* ctx, int frames, void* d_phi_start, void* d_phi_final, void* stream) {
    if (!ctx) return isar::fail(ISAR_ERR_INVALID, "null context");
    return ctx->phases(frames, d_phi_start, d_phi_final, static_cast<hipStream_t>(stream));
}

ISAR_API int isar_get_trace(isar_context* ctx, int frame, isar_frame_trace* out) {
    if (!ctx) return fail(ISAR_ERR_INVALID, "null context");
    return ctx->trace(frame, out);
}


ISAR_API int isar_fft_host(int n, int dir, const float* h_in, float* h_out, int batch) {
    return fft_host(n, dir, h_in, h_out, batch, false);
}

ISAR_API int isar_fft_host_grouped(int n, int dir, const float* h_in, float* h_out, int batch) {
    return fft_host(n, dir, h_in, h_out, batch, true);
}

ISAR_API int isar_probe_enable(isar_context* ctx, int stage) {
    if (!ctx) return fail(ISAR_ERR_INVALID, "null context");
    if (stage >= ISAR_STAGE_COUNT) return fail(ISAR_ERR_INVALID, "unknown stage");
    return ctx->probe_enable(stage);
}

ISAR_API int isar_probe_read(isar_context* ctx, double* total_ms, int* launches) {
    if (!ctx) return fail(ISAR_ERR_INVALID, "null context");
    return ctx->probe_read(total_ms, launches);
}

ISAR_API const char* isar_last_error(void) { return isar::last_error_cstr(); }

ISAR_API int isar_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

}  // extern "C"
