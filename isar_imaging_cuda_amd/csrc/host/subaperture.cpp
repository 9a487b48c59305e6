// subaperture.cpp — sub-aperture plans behind isar_subaperture.h: windows grouped by pulse
// count, one imaging context per group reading straight out of the dwell, the launches of all
// groups captured into one HIP graph, and a per-group autotuner for the range-alignment chunk
// and the autofocus form.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "isar_subaperture.h"
#include "pipeline.hpp"

namespace isar {

struct SaGroup {
    int M = 0;
    std::vector<int> windows;           // window indices in this group
    Context ctx;
    long long* d_rows = nullptr;        // first dwell row of each window
    long long* d_img = nullptr;         // image offset of each window
    double cost = 0.0;                  // expected time: a model at creation, measured by calibration or the autotuner
    int lane = 0;                       // branch stream the group is captured on
};

// Hardware queues a process gets (HIP's default, or GPU_MAX_HW_QUEUES when set): streams beyond
// that share queues, and the groups on a shared queue run one after the other.
inline int hw_queues() {
    const char* v = std::getenv("GPU_MAX_HW_QUEUES");
    const int n = v ? std::atoi(v) : 0;
    return n > 0 ? n : 4;
}

struct SaPlan {
    isar_radar_params p{};
    int dwell = 0;
    std::vector<int32_t> first, pulses;
    std::vector<long long> img_off;
    size_t img_floats = 0;
    std::vector<SaGroup*> groups;
    hipStream_t stream = nullptr;       // capture and replay stream (default priority)
    std::vector<hipStream_t> branch;    // independent graph branches, at most one per hardware queue
    std::vector<hipEvent_t> joins;
    hipEvent_t ev_in = nullptr, ev_out = nullptr, ev_fork = nullptr;
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    // what the captured graph was recorded for
    const void* g_dwell = nullptr;
    const double* g_vel = nullptr;
    float* g_img = nullptr;
    isar_options g_opt{};

    void drop_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
    ~SaPlan() {
        drop_graph();
        for (SaGroup* g : groups) {
            if (g->d_rows) (void)hipFree(g->d_rows);
            if (g->d_img) (void)hipFree(g->d_img);
            delete g;
        }
        if (ev_in) (void)hipEventDestroy(ev_in);
        if (ev_out) (void)hipEventDestroy(ev_out);
        if (ev_fork) (void)hipEventDestroy(ev_fork);
        for (hipEvent_t e : joins) (void)hipEventDestroy(e);
        for (hipStream_t b : branch) (void)hipStreamDestroy(b);
        if (stream) (void)hipStreamDestroy(stream);
    }

    // Groups onto branches, longest expected time first onto the least loaded branch, so that
    // the branch that finishes last carries as little as possible.
    void assign_lanes() {
        std::vector<SaGroup*> order(groups);
        std::stable_sort(order.begin(), order.end(), [](const SaGroup* a, const SaGroup* b) { return a->cost > b->cost; });
        std::vector<double> load(branch.size(), 0.0);
        for (SaGroup* g : order) {
            const size_t b = static_cast<size_t>(std::min_element(load.begin(), load.end()) - load.begin());
            g->lane = static_cast<int>(b);
            load[b] += g->cost;
        }
    }

    // During capture on `stream`: each branch stream is forked from the capture stream, runs its
    // groups in turn, and is joined back, so the graph holds one independent branch per stream.
    int capture_branches(const void* dwell_ptr, const double* vel, float* img, const isar_options& opt) {
        if (hipEventRecord(ev_fork, stream) != hipSuccess) return fail(ISAR_ERR_DEVICE, "fork");
        for (size_t b = 0; b < branch.size(); ++b) {
            if (hipStreamWaitEvent(branch[b], ev_fork, 0) != hipSuccess) return fail(ISAR_ERR_DEVICE, "fork wait");
            for (SaGroup* g : groups) {
                if (g->lane != static_cast<int>(b)) continue;
                const int rc = g->ctx.run_frames(static_cast<int>(g->windows.size()), dwell_ptr, vel, g->d_rows, nullptr,
                                                 img, g->d_img, opt, branch[b]);
                if (rc) return rc;
            }
            if (hipEventRecord(joins[b], branch[b]) != hipSuccess || hipStreamWaitEvent(stream, joins[b], 0) != hipSuccess)
                return fail(ISAR_ERR_DEVICE, "join");
        }
        return ISAR_OK;
    }

    int launch_all(const void* dwell_ptr, const double* vel, float* img, const isar_options& opt, hipStream_t s) {
        for (SaGroup* g : groups) {
            const int rc = g->ctx.run_frames(static_cast<int>(g->windows.size()), dwell_ptr, vel, g->d_rows, nullptr,
                                             img, g->d_img, opt, s);
            if (rc) return rc;
        }
        return ISAR_OK;
    }
};

}  // namespace isar

struct isar_sa_plan : isar::SaPlan {};

using isar::fail;

extern "C" {

ISAR_API int isar_sa_plan_create(const isar_radar_params* p, int dwell_pulses, int n_windows,
                                 const int32_t* first_pulse, const int32_t* pulses, isar_sa_plan** out) {
    if (!p || !out || !first_pulse || !pulses || n_windows < 1 || dwell_pulses < 1)
        return fail(ISAR_ERR_INVALID, "isar_sa_plan_create: bad arguments");
    for (int i = 0; i < n_windows; ++i) {
        if (!isar::pulse_count_supported(pulses[i]))
            return fail(ISAR_ERR_INVALID, "isar_sa_plan_create: unsupported pulse count " + std::to_string(pulses[i]));
        if (first_pulse[i] < 0 || first_pulse[i] + pulses[i] > dwell_pulses)
            return fail(ISAR_ERR_INVALID, "isar_sa_plan_create: window " + std::to_string(i) + " leaves the dwell");
    }
    auto* plan = new isar_sa_plan();
    plan->p = *p;
    plan->dwell = dwell_pulses;
    plan->first.assign(first_pulse, first_pulse + n_windows);
    plan->pulses.assign(pulses, pulses + n_windows);
    const long long W = p->range_num_cut;
    for (int i = 0; i < n_windows; ++i) {
        plan->img_off.push_back(static_cast<long long>(plan->img_floats));
        plan->img_floats += static_cast<size_t>(pulses[i]) * W;
    }
    std::map<int, std::vector<int>> by_m;
    for (int i = 0; i < n_windows; ++i) by_m[pulses[i]].push_back(i);
    for (auto& [M, wins] : by_m) {
        auto* g = new isar::SaGroup();
        plan->groups.push_back(g);
        g->M = M;
        g->windows = wins;
        // cost model before any measurement: a per-iteration autofocus latency that grows with
        // the pulse count plus range-alignment work per window
        g->cost = static_cast<double>(M) * (static_cast<double>(wins.size()) + 16.0);
        isar_radar_params gp = *p;
        gp.echo_num = M;
        int rc = g->ctx.init(gp, static_cast<int>(wins.size()));
        if (rc) {
            delete plan;
            return rc;
        }
        std::vector<long long> rows, offs;
        for (int w : wins) {
            rows.push_back(first_pulse[w]);
            offs.push_back(plan->img_off[w]);
        }
        const size_t bytes = wins.size() * sizeof(long long);
        if (hipMalloc(&g->d_rows, bytes) != hipSuccess || hipMalloc(&g->d_img, bytes) != hipSuccess ||
            hipMemcpy(g->d_rows, rows.data(), bytes, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(g->d_img, offs.data(), bytes, hipMemcpyHostToDevice) != hipSuccess) {
            delete plan;
            return fail(ISAR_ERR_MEMORY, "isar_sa_plan_create: window tables");
        }
    }
    bool ok = hipStreamCreateWithFlags(&plan->stream, hipStreamNonBlocking) == hipSuccess &&
              hipEventCreateWithFlags(&plan->ev_in, hipEventDisableTiming) == hipSuccess &&
              hipEventCreateWithFlags(&plan->ev_out, hipEventDisableTiming) == hipSuccess &&
              hipEventCreateWithFlags(&plan->ev_fork, hipEventDisableTiming) == hipSuccess;
    const size_t nbranch = std::min(plan->groups.size(), static_cast<size_t>(isar::hw_queues()));
    for (size_t bi = 0; ok && bi < nbranch; ++bi) {
        hipStream_t b = nullptr;
        hipEvent_t e = nullptr;
        ok = hipStreamCreateWithFlags(&b, hipStreamNonBlocking) == hipSuccess;
        if (ok) plan->branch.push_back(b);
        ok = ok && hipEventCreateWithFlags(&e, hipEventDisableTiming) == hipSuccess;
        if (ok) plan->joins.push_back(e);
    }
    if (!ok) {
        delete plan;
        return fail(ISAR_ERR_DEVICE, "isar_sa_plan_create: streams");
    }
    plan->assign_lanes();
    *out = plan;
    return ISAR_OK;
}

ISAR_API int isar_sa_plan_destroy(isar_sa_plan* plan) {
    delete plan;
    return ISAR_OK;
}

ISAR_API size_t isar_sa_plan_image_floats(const isar_sa_plan* plan) { return plan ? plan->img_floats : 0; }

ISAR_API int isar_sa_plan_image_offsets(const isar_sa_plan* plan, int64_t* offsets) {
    if (!plan || !offsets) return fail(ISAR_ERR_INVALID, "isar_sa_plan_image_offsets: null argument");
    for (size_t i = 0; i < plan->img_off.size(); ++i) offsets[i] = plan->img_off[i];
    return ISAR_OK;
}

ISAR_API int isar_sa_plan_groups(const isar_sa_plan* plan, int32_t* pulses, int32_t* windows) {
    if (!plan) return fail(ISAR_ERR_INVALID, "isar_sa_plan_groups: null plan");
    for (size_t g = 0; g < plan->groups.size(); ++g) {
        if (pulses) pulses[g] = plan->groups[g]->M;
        if (windows) windows[g] = static_cast<int32_t>(plan->groups[g]->windows.size());
    }
    return static_cast<int>(plan->groups.size());
}

ISAR_API int isar_sa_plan_group_tuning(const isar_sa_plan* plan, int group, int32_t* tuning) {
    if (!plan || !tuning || group < 0 || group >= static_cast<int>(plan->groups.size()))
        return fail(ISAR_ERR_INVALID, "isar_sa_plan_group_tuning: bad arguments");
    const isar::Context& c = plan->groups[group]->ctx;
    tuning[0] = c.ra_chunk;
    tuning[1] = c.af_mode;
    tuning[2] = c.F;
    tuning[3] = c.czt_threads;
    return ISAR_OK;
}

ISAR_API int isar_sa_plan_set_group_tuning(isar_sa_plan* plan, int group, const int32_t* tuning) {
    if (!plan || !tuning || group < 0 || group >= static_cast<int>(plan->groups.size()))
        return fail(ISAR_ERR_INVALID, "isar_sa_plan_set_group_tuning: bad arguments");
    const int rows = tuning[0], form = tuning[1], threads = tuning[3];
    if (rows != 2 && rows != 4 && rows != 8 && rows != 16)
        return fail(ISAR_ERR_INVALID, "isar_sa_plan_set_group_tuning: chunk rows must be 2, 4, 8 or 16");
    if (form != ISAR_AF_FUSED && form != ISAR_AF_SPLIT && form != ISAR_AF_SPLIT_PHASE && form != ISAR_AF_RESIDENT)
        return fail(ISAR_ERR_INVALID, "isar_sa_plan_set_group_tuning: autofocus form must be 1, 2, 3 or 4");
    if (threads != 256 && threads != 512)
        return fail(ISAR_ERR_INVALID, "isar_sa_plan_set_group_tuning: chirp-z threads must be 256 or 512");
    isar::Context& c = plan->groups[group]->ctx;
    const int rc = c.set_czt_length(tuning[2]);
    if (rc) return rc;
    plan->drop_graph();
    c.ra_chunk = rows;
    c.af_mode = form;
    c.czt_threads = threads;
    return ISAR_OK;
}

ISAR_API int isar_sa_plan_group_lane(const isar_sa_plan* plan, int group) {
    if (!plan || group < 0 || group >= static_cast<int>(plan->groups.size())) return -1;
    return plan->groups[group]->lane;
}

ISAR_API int isar_sa_plan_calibrate(isar_sa_plan* plan, const void* d_dwell, const double* d_velocity, float* d_img,
                                    const isar_options* opt, int repeats) {
    if (!plan || !d_dwell || !d_img || !opt || repeats < 1)
        return fail(ISAR_ERR_INVALID, "isar_sa_plan_calibrate: bad arguments");
    hipEvent_t a = nullptr, b = nullptr;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
        if (a) (void)hipEventDestroy(a);
        return fail(ISAR_ERR_DEVICE, "isar_sa_plan_calibrate: events");
    }
    plan->drop_graph();
    int rc = ISAR_OK;
    for (isar::SaGroup* g : plan->groups) {
        const int frames = static_cast<int>(g->windows.size());
        auto one = [&]() {
            return g->ctx.run_frames(frames, d_dwell, d_velocity, g->d_rows, nullptr, d_img, g->d_img, *opt, plan->stream);
        };
        rc = one();   // warm-up
        if (rc == ISAR_OK && hipEventRecord(a, plan->stream) != hipSuccess) rc = fail(ISAR_ERR_DEVICE, "calibrate: record");
        for (int r = 0; r < repeats && rc == ISAR_OK; ++r) rc = one();
        if (rc == ISAR_OK && (hipEventRecord(b, plan->stream) != hipSuccess || hipEventSynchronize(b) != hipSuccess))
            rc = fail(ISAR_ERR_DEVICE, "calibrate: timing");
        if (rc) break;
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, a, b);
        g->cost = ms / static_cast<float>(repeats);
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (rc == ISAR_OK) plan->assign_lanes();
    return rc;
}

ISAR_API int isar_sa_plan_run(isar_sa_plan* plan, const void* d_dwell, const double* d_velocity, float* d_img,
                              const isar_options* opt, int use_graph, void* stream) {
    if (!plan || !d_dwell || !d_img || !opt) return fail(ISAR_ERR_INVALID, "isar_sa_plan_run: null argument");
    if (opt->stage_mode != ISAR_RUN_FULL || opt->warm_start)
        return fail(ISAR_ERR_INVALID, "isar_sa_plan_run: full runs without warm start only");
    const hipStream_t user = static_cast<hipStream_t>(stream);
    if (!use_graph) return plan->launch_all(d_dwell, d_velocity, d_img, *opt, user);
    const bool same = plan->exec && plan->g_dwell == d_dwell && plan->g_vel == d_velocity && plan->g_img == d_img &&
                      std::memcmp(&plan->g_opt, opt, sizeof(isar_options)) == 0;
    if (!same) {
        plan->drop_graph();
        if (hipStreamBeginCapture(plan->stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
            return fail(ISAR_ERR_DEVICE, "isar_sa_plan_run: begin capture");
        const int rc = plan->capture_branches(d_dwell, d_velocity, d_img, *opt);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(plan->stream, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess || !g) return fail(ISAR_ERR_DEVICE, "isar_sa_plan_run: end capture");
        plan->graph = g;
        if (hipGraphInstantiate(&plan->exec, g, nullptr, nullptr, 0) != hipSuccess) {
            plan->drop_graph();
            return fail(ISAR_ERR_DEVICE, "isar_sa_plan_run: graph instantiation");
        }
        plan->g_dwell = d_dwell;
        plan->g_vel = d_velocity;
        plan->g_img = d_img;
        plan->g_opt = *opt;
    }
    // order the replay after the caller's stream and the caller's stream after the replay
    if (hipEventRecord(plan->ev_in, user) != hipSuccess || hipStreamWaitEvent(plan->stream, plan->ev_in, 0) != hipSuccess ||
        hipGraphLaunch(plan->exec, plan->stream) != hipSuccess || hipEventRecord(plan->ev_out, plan->stream) != hipSuccess ||
        hipStreamWaitEvent(user, plan->ev_out, 0) != hipSuccess)
        return fail(ISAR_ERR_DEVICE, "isar_sa_plan_run: graph launch");
    return ISAR_OK;
}

ISAR_API int isar_sa_plan_autotune(isar_sa_plan* plan, const void* d_dwell, const double* d_velocity,
                                   float* d_img, const isar_options* opt, int repeats, int32_t* chosen) {
    if (!plan || !d_dwell || !d_img || !opt || repeats < 1)
        return fail(ISAR_ERR_INVALID, "isar_sa_plan_autotune: bad arguments");
    plan->drop_graph();
    for (size_t gi = 0; gi < plan->groups.size(); ++gi) {
        isar::SaGroup* g = plan->groups[gi];
        int c = 0;
        // the resident autofocus form is not offered to sub-aperture plans (see below)
        const int rc = g->ctx.autotune(static_cast<int>(g->windows.size()), d_dwell, d_velocity, g->d_rows, d_img,
                                       g->d_img, *opt, plan->stream, repeats, &c, false);
        if (rc) return rc;
        if (chosen) chosen[gi] = c;
        g->cost = g->ctx.tuned_ms;
    }
    plan->assign_lanes();
    // The groups share the GPU when the graph runs, so the autofocus form each group won alone
    // is checked against one form for every group, timing whole graph replays; each candidate
    // places the groups on branches by their measured times in that form.
    hipStream_t ts = nullptr;
    hipEvent_t a = nullptr, b = nullptr;
    if (hipStreamCreateWithFlags(&ts, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&a) != hipSuccess ||
        hipEventCreate(&b) != hipSuccess)
        return fail(ISAR_ERR_DEVICE, "isar_sa_plan_autotune: timing objects");
    // candidates: each group's fastest form alone among a subset of the forms, and one form for all
    // The resident form is left out: with it among the candidates, this tuner's process once ended
    // with a host segmentation fault (no GPU fault) whose cause was not found.
    const std::vector<std::vector<int>> subsets = {
        {ISAR_AF_FUSED, ISAR_AF_SPLIT, ISAR_AF_SPLIT_PHASE}, {ISAR_AF_FUSED, ISAR_AF_SPLIT_PHASE},
        {ISAR_AF_FUSED, ISAR_AF_SPLIT}, {ISAR_AF_FUSED}, {ISAR_AF_SPLIT}, {ISAR_AF_SPLIT_PHASE}};
    auto pick = [&](const std::vector<int>& forms) {
        for (isar::SaGroup* g : plan->groups) {
            int m = forms[0];
            for (int f : forms)
                if (g->ctx.mode_ms[f] > 0.0f && g->ctx.mode_ms[f] < g->ctx.mode_ms[m]) m = f;
            g->ctx.af_mode = m;
            g->cost = g->ctx.mode_ms[m];
        }
        plan->assign_lanes();
    };
    int rc = ISAR_OK;
    float best_ms = 0.0f;
    size_t best = 0;
    for (size_t c = 0; c < subsets.size(); ++c) {
        pick(subsets[c]);
        plan->drop_graph();
        rc = isar_sa_plan_run(plan, d_dwell, d_velocity, d_img, opt, 1, ts);   // capture
        if (rc) break;
        (void)hipEventRecord(a, ts);
        for (int r = 0; r < repeats && rc == ISAR_OK; ++r) rc = isar_sa_plan_run(plan, d_dwell, d_velocity, d_img, opt, 1, ts);
        (void)hipEventRecord(b, ts);
        if (rc) break;
        if (hipEventSynchronize(b) != hipSuccess) {
            rc = fail(ISAR_ERR_DEVICE, "isar_sa_plan_autotune: timing");
            break;
        }
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, a, b);
        if (best_ms == 0.0f || ms < best_ms) {
            best_ms = ms;
            best = c;
        }
    }
    pick(subsets[best]);
    plan->drop_graph();
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    (void)hipStreamDestroy(ts);
    return rc;
}

}  // extern "C"
