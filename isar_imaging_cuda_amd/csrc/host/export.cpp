// export.cpp — the reference's seven exported entry points (isar_imaging_export.h:37-130) over
// the imaging context, plus the STRETCH .wbd ingest they need (common.cu:750-1191).
//
// One session per process, as in the reference (all state is file-scope, isar_imaging_export.cpp:
// 6-35). Errors never terminate the process: they set isar_last_error() and make int entry points
// return EXIT_FAILURE.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <regex>
#include <string>
#include <vector>

#include "isar_imaging_export.h"
#include "pipeline.hpp"

namespace fs = std::filesystem;

namespace {

constexpr int kIfds = 1, kStretch = 2;
constexpr int kRangeNumIfds = 4096;   // RANGE_NUM_IFDS_PC (common.cuh:72)
constexpr float kPiF = 3.14159265358979f;

struct Session {
    // ingest (ioOperation, common.cuh:700-880)
    std::vector<std::string> files;
    int data_type = 0;
    int polar_type = 0;
    int frame_len = 0;
    int frame_num = 0;
    isar_radar_params paras{};
    bool have_paras = false;
    // imaging
    isar_context* ctx = nullptr;
    float2* d_pp[2] = {nullptr, nullptr};
    int active = -1;            // ping-pong buffer holding the current window, -1 before the first
    float2* d_proc = nullptr;   // window handed to the pipeline (reference d_data_proc)
    double* d_velocity = nullptr;
    float* d_img = nullptr;
    float2* d_hrrp = nullptr;       // M x N range profiles when imagingMemInit had if_hrrp
    std::vector<double> h_velocity;
    // IFDS: raw echoes staged for pulse compression (common.cu:583-604)
    int16_t* d_iq = nullptr;
    double* d_pc_velocity = nullptr;
    int iq_len = 0;
    // what imagingMemInit sized the buffers for; dataExtracting must match it
    int init_data_type = 0;
    int init_frame_len = 0;
};
Session g;

void err(const std::string& m) { isar::set_error(m); }

// ---- turn-angle curve (common.cu:902-1020) ----
float turn_angle(float az1, float pt1, float az2, float pt2) {
    const float v1[3] = {std::sin(pt1 / 180 * kPiF), std::cos(pt1 / 180 * kPiF) * std::cos(az1 / 180 * kPiF),
                         std::cos(pt1 / 180 * kPiF) * std::sin(az1 / 180 * kPiF)};
    const float v2[3] = {std::sin(pt2 / 180 * kPiF), std::cos(pt2 / 180 * kPiF) * std::cos(az2 / 180 * kPiF),
                         std::cos(pt2 / 180 * kPiF) * std::sin(az2 / 180 * kPiF)};
    return std::acos(v1[0] * v2[0] + v1[1] * v2[1] + v1[2] * v2[2]) / kPiF * 180;
}

float interpolate(const std::vector<int>& xs, const std::vector<float>& ys, int x) {
    const int n = static_cast<int>(xs.size());
    int i = 0;
    if (x >= xs[n - 2]) i = n - 2;
    else
        while (x > xs[i + 1]) ++i;
    const float xL = static_cast<float>(xs[i]), xR = static_cast<float>(xs[i + 1]);
    float yL = ys[i], yR = ys[i + 1];
    if (x < xL) yR = yL;   // no extrapolation beyond the ends
    if (x > xR) yL = yR;
    return yL + (yR - yL) / (xR - xL) * (static_cast<float>(x) - xL);
}

void turn_angle_line(std::vector<float>* out, const std::vector<float>& az, const std::vector<float>& pt) {
    const int P = static_cast<int>(pt.size());
    std::vector<int> jumps;
    for (int i = 0; i + 1 < P; ++i)
        if (std::abs(pt[i + 1] - pt[i]) > 0.2f) jumps.push_back(i);
    std::vector<int> begin{0}, end;
    for (int j : jumps) begin.push_back(j + 1);
    end = jumps;
    end.push_back(P - 1);
    std::vector<float>& ta = *out;
    ta.assign(P, 0.0f);
    for (size_t b = 0; b < begin.size(); ++b) {
        const int n = end[b] - begin[b] + 1;
        const int stride = n < 21 ? 1 : 20;
        for (int i = stride; i < n; i += stride) {
            const int cur = begin[b] + i;
            ta[cur] = ta[cur - stride] + turn_angle(az[cur - stride], pt[cur - stride], az[cur], pt[cur]);
        }
        for (float& v : ta) v = std::abs(v);
        if (n >= 21) {
            std::vector<int> xs;
            for (int i = 0; i + stride <= n; i += stride) xs.push_back(i);
            if (xs.size() >= 2) {   // the reference indexes xs[size-2] unguarded
                std::vector<float> ys;
                // the reference samples the curve at the block-relative positions xs[i] taken as
                // absolute indices (common.cu:976); kept as is
                for (int x : xs) ys.push_back(ta[x]);
                std::vector<float> interp(n);
                for (int i = 0; i < n; ++i) interp[i] = interpolate(xs, ys, i);
                std::copy(interp.begin(), interp.end(), ta.begin() + begin[b]);
            }
        }
        if (b > 0)
            for (int i = begin[b]; i <= end[b]; ++i) ta[i] += ta[end[b - 1]];
    }
}

int read_headers(std::vector<double>* dataN, std::vector<float>* turnAngle) {
    // getSystemParas (common.cu:796-833)
    std::ifstream f0(g.files[0], std::ios::binary);
    if (!f0) return err("cannot open " + g.files[0]), EXIT_FAILURE;
    uint32_t t[36] = {};
    f0.read(reinterpret_cast<char*>(t), sizeof(t));
    if (!f0) return err("short system header in " + g.files[0]), EXIT_FAILURE;
    g.frame_len = static_cast<int>(t[4]) * 4;
    if (g.frame_len <= 256) return err("invalid frame length in " + g.files[0]), EXIT_FAILURE;
    g.paras.fc = static_cast<int64_t>(t[12]) * 1000000LL;
    g.paras.band_width = static_cast<int64_t>(t[13]) * 1000000LL;
    g.paras.Tp = static_cast<double>(t[15]) / 1000000.0;
    const int64_t fs_mhz = static_cast<int64_t>(t[17] % (1u << 16));
    g.paras.Fs = g.data_type == kIfds ? fs_mhz * 1000000LL / 2 : fs_mhz * 1000000LL;
    g.frame_num = static_cast<int>(fs::file_size(g.files[0]) / static_cast<uintmax_t>(g.frame_len));
    g.have_paras = true;
    // readKuIFDSAllNB (common.cu:836-899)
    const int total = g.frame_num * static_cast<int>(g.files.size());
    dataN->assign(static_cast<size_t>(total) * 4, 0.0);
    std::vector<float> az(total), pt(total);
    auto s32 = [](uint32_t w) { return static_cast<double>(w) - (static_cast<double>(w) > 2147483648.0 ? 4294967296.0 : 0.0); };
    for (size_t fi = 0; fi < g.files.size(); ++fi) {
        std::ifstream f(g.files[fi], std::ios::binary);
        if (!f) return err("cannot open " + g.files[fi]), EXIT_FAILURE;
        for (int k = 0; k < g.frame_num; ++k) {
            const int i = static_cast<int>(fi) * g.frame_num + k;
            uint32_t h[11] = {};
            f.seekg(static_cast<std::streamoff>(k) * g.frame_len + 48);
            f.read(reinterpret_cast<char*>(h), sizeof(h));
            if (!f) return err("short frame header in " + g.files[fi]), EXIT_FAILURE;
            const double range = static_cast<double>(h[7]) * 0.1;
            const double vel = s32(h[8]) * 0.1;
            double a = s32(h[9]) * (360.0 / 16777216.0);
            if (a < 0) a += 360.0;
            double p = s32(h[10]) * (360.0 / 16777216.0);
            if (p < 0) p += 360.0;
            (*dataN)[i] = range;
            (*dataN)[i + total] = vel;
            (*dataN)[i + 2 * static_cast<size_t>(total)] = a;
            (*dataN)[i + 3 * static_cast<size_t>(total)] = p;
            az[i] = static_cast<float>(a);
            pt[i] = static_cast<float>(p);
        }
    }
    turn_angle_line(turnAngle, az, pt);
    return EXIT_SUCCESS;
}

void free_buffers() {
    if (g.ctx) isar_destroy(g.ctx);
    g.ctx = nullptr;
    for (void* p : {static_cast<void*>(g.d_pp[0]), static_cast<void*>(g.d_pp[1]), static_cast<void*>(g.d_proc),
                    static_cast<void*>(g.d_velocity), static_cast<void*>(g.d_img), static_cast<void*>(g.d_hrrp),
                    static_cast<void*>(g.d_iq),
                    static_cast<void*>(g.d_pc_velocity)})
        if (p) (void)hipFree(p);
    g.d_pp[0] = g.d_pp[1] = g.d_proc = nullptr;
    g.d_velocity = nullptr;
    g.d_img = nullptr;
    g.d_hrrp = nullptr;
    g.d_iq = nullptr;
    g.d_pc_velocity = nullptr;
    g.iq_len = 0;
    g.active = -1;
    g.init_data_type = 0;
    g.init_frame_len = 0;
}

}  // namespace

extern "C" {

DLL_EXPORT_API int gpuDevInit() {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1) return err("no GPU visible"), EXIT_FAILURE;
    // the reference picks the device with the highest peak throughput; with identical devices
    // that is device 0
    if (hipSetDevice(0) != hipSuccess) return err("hipSetDevice(0) failed"), EXIT_FAILURE;
    return EXIT_SUCCESS;
}

DLL_EXPORT_API int dataParsing(vec1D_DBL* dataN, vec1D_FLT* turnAngle, int* frame_len, int* frame_num,
                               const std::string& dir_path, const int& polar_type, const int& data_type) {
    // ioInit (common.cu:750-793): files matching the polarisation pattern, in sorted order
    if (!dataN || !turnAngle || !frame_len || !frame_num) return err("null output pointer"), EXIT_FAILURE;
    g.files.clear();
    g.data_type = data_type;
    g.polar_type = polar_type;
    std::error_code ec;
    if (!fs::is_directory(dir_path, ec)) return err("invalid directory path: " + dir_path), EXIT_FAILURE;
    std::string pattern;
    if (data_type == kIfds) pattern = polar_type == 0 ? R"(.*01_1100.*\.wbd)" : R"(.*01_1101.*\.wbd)";
    else if (data_type == kStretch) pattern = polar_type == 0 ? R"(.*00_1100\.wbd)" : R"(.*00_1101\.wbd)";
    else return err("invalid data type"), EXIT_FAILURE;
    const std::regex re(pattern);
    for (const auto& e : fs::directory_iterator(dir_path, ec))
        if (e.is_regular_file() && std::regex_match(e.path().string(), re)) g.files.push_back(e.path().string());
    std::sort(g.files.begin(), g.files.end());
    if (g.files.empty()) return err("no valid file in " + dir_path), EXIT_FAILURE;
    if (read_headers(dataN, turnAngle) != EXIT_SUCCESS) return EXIT_FAILURE;
    *frame_len = g.frame_len;
    *frame_num = g.frame_num;
    return EXIT_SUCCESS;
}

DLL_EXPORT_API void imagingMemInit(vec1D_FLT* img, vec1D_INT* dataWFileSn, vec1D_DBL* dataNOut, vec1D_FLT* turnAngleOut,
                                   vec1D_COM_FLT* dataW, const int& window_len, const int& frame_len, const int& data_type,
                                   const bool& if_hpc, const bool& if_hrrp) {
    (void)if_hpc;
    free_buffers();
    // size rules and validation (isar_imaging_export.cpp:129-155)
    isar_radar_params p = g.paras;
    p.echo_num = window_len;
    p.range_num = data_type == kIfds ? kRangeNumIfds : (frame_len - 256) / 4;
    p.range_num_cut = isar::kRangeCut;
    if (!g.have_paras) { err("imagingMemInit before dataParsing"); return; }
    if (isar::validate_params(p) != ISAR_OK) return;   // message already recorded
    const size_t M = p.echo_num, N = p.range_num, W = p.range_num_cut;
    img->resize(M * W);
    dataWFileSn->resize(M);
    dataNOut->resize(M * 4);
    turnAngleOut->resize(M);
    if (data_type == kStretch) dataW->resize(M * N);
    if (isar_create(&p, 1, &g.ctx) != ISAR_OK) return;
    (void)isar_set_graph_mode(g.ctx, 1);   // every window runs with the same buffers and options
    // one frame per call: two rows per alignment-update workgroup give the most workgroups per
    // merge level (1.35 against 1.46 ms per 256 x 4096 frame with eight)
    (void)isar_set_ra_chunk(g.ctx, 2);
    // all autofocus iterations in one launch (1.16 against 1.35 ms per 256 x 4096 frame with a
    // launch per iteration); isarMainSingle checks the frame's status for a wait that hit its bound
    (void)isar_set_autofocus_mode(g.ctx, ISAR_AF_RESIDENT);
    bool ok = hipMalloc(&g.d_pp[0], M * N * sizeof(float2)) == hipSuccess &&
              hipMalloc(&g.d_pp[1], M * N * sizeof(float2)) == hipSuccess &&
              hipMalloc(&g.d_proc, M * N * sizeof(float2)) == hipSuccess &&
              hipMalloc(&g.d_velocity, M * sizeof(double)) == hipSuccess &&
              hipMalloc(&g.d_img, M * W * sizeof(float)) == hipSuccess;
    // isar_imaging_export.cpp:176-178: the profile buffer exists only when if_hrrp
    if (ok && if_hrrp)
        ok = hipMalloc(&g.d_hrrp, M * N * sizeof(float2)) == hipSuccess && isar_set_hrrp_output(g.ctx, g.d_hrrp) == ISAR_OK;
    if (ok && data_type == kIfds) {
        g.iq_len = (frame_len - 256) / 4;
        ok = hipMalloc(&g.d_iq, M * static_cast<size_t>(g.iq_len) * 2 * sizeof(int16_t)) == hipSuccess &&
             hipMalloc(&g.d_pc_velocity, M * sizeof(double)) == hipSuccess;
    }
    if (!ok) {
        free_buffers();
        err("device allocation failed in imagingMemInit");
        return;
    }
    g.paras = p;
    g.h_velocity.assign(M, 0.0);
    g.init_data_type = data_type;
    g.init_frame_len = frame_len;
}

DLL_EXPORT_API int dataExtracting(vec1D_INT* dataWFileSn, vec1D_DBL* dataNOut, vec1D_FLT* turnAngleOut, vec1D_COM_FLT* dataW,
                                  const vec1D_DBL& dataN, const vec1D_FLT& turnAngle, const int& frame_len,
                                  const int& frame_num, const int& sampling_stride, const int& window_head,
                                  const int& window_len, int& imaging_stride, const int& data_type) {
    if (!g.ctx) return err("dataExtracting before imagingMemInit"), EXIT_FAILURE;
    if (!dataWFileSn || !dataNOut || !turnAngleOut || !dataW) return err("null output vector"), EXIT_FAILURE;
    if (data_type != kStretch && data_type != kIfds) return err("unknown data type"), EXIT_FAILURE;
    // the buffers were sized by imagingMemInit for one data type and frame length, and the files
    // hold frames of the length dataParsing read: anything else would overrun them
    if (data_type != g.init_data_type) return err("data_type differs from imagingMemInit"), EXIT_FAILURE;
    if (frame_len != g.init_frame_len || frame_len != g.frame_len)
        return err("frame_len differs from imagingMemInit or from the files"), EXIT_FAILURE;
    const int M = g.paras.echo_num, N = g.paras.range_num;
    if (window_len != M) return err("window_len differs from imagingMemInit"), EXIT_FAILURE;
    const size_t total_frames = static_cast<size_t>(frame_num) * g.files.size();
    if (frame_num != g.frame_num || dataN.size() < 4 * total_frames || turnAngle.size() < total_frames)
        return err("dataN / turnAngle / frame_num do not match dataParsing"), EXIT_FAILURE;
    if (dataWFileSn->size() < static_cast<size_t>(M) || dataNOut->size() < 4 * static_cast<size_t>(M) ||
        turnAngleOut->size() < static_cast<size_t>(M) ||
        (data_type == kStretch && dataW->size() < static_cast<size_t>(M) * N))
        return err("output vectors smaller than imagingMemInit sized them"), EXIT_FAILURE;
    // uniformSampling (common.cu:1165-1185)
    const int total = frame_num * static_cast<int>(g.files.size());
    for (int i = 0; i < M; ++i) {
        const int sn = window_head + i * sampling_stride;
        if (sn < 0 || sn >= total) return err("window reaches past the last frame"), EXIT_FAILURE;
        (*dataWFileSn)[i] = sn;
        for (int r = 0; r < 4; ++r) (*dataNOut)[static_cast<size_t>(r) * M + i] = dataN[sn + static_cast<size_t>(r) * total];
        (*turnAngleOut)[i] = std::abs(turnAngle[sn]);
    }
    // ping-pong with reuse of the overlapping pulses (isar_imaging_export.cpp:89-106)
    int overlap = 0, old = -1;
    if (g.active < 0) {
        g.active = 0;
    } else {
        old = g.active;
        g.active ^= 1;
        overlap = M - imaging_stride;
        if (overlap < 0 || overlap > M) return err("imaging_stride outside [0, window_len]"), EXIT_FAILURE;
    }
    const int fresh = M - overlap;
    float2* cur = g.d_pp[g.active];
    hipError_t e = hipSuccess;
    if (overlap > 0)
        e = hipMemcpy(cur, g.d_pp[old] + static_cast<size_t>(fresh) * N, sizeof(float2) * overlap * N, hipMemcpyDeviceToDevice);
    for (int i = 0; i < M; ++i) g.h_velocity[i] = (*dataNOut)[static_cast<size_t>(M) + i];
    // int16 I/Q pairs after the 256-byte frame header (common.cu:1044-1100)
    std::vector<int16_t> ad(static_cast<size_t>(frame_len - 256) / 2);
    std::vector<int16_t> iq;   // IFDS: the fresh echoes, raw
    if (data_type == kIfds) iq.resize(static_cast<size_t>(fresh) * ad.size());
    std::vector<std::ifstream> fin(g.files.size());
    for (int i = overlap; i < M && e == hipSuccess; ++i) {
        const int sn = (*dataWFileSn)[i];
        const int fi = sn / frame_num;
        if (!fin[fi].is_open()) fin[fi].open(g.files[fi], std::ios::binary);
        fin[fi].seekg(static_cast<std::streamoff>(sn - fi * frame_num) * frame_len + 256);
        fin[fi].read(reinterpret_cast<char*>(ad.data()), static_cast<std::streamsize>(ad.size() * sizeof(int16_t)));
        if (!fin[fi]) return err("short read in " + g.files[fi]), EXIT_FAILURE;
        if (data_type == kIfds) {
            std::copy(ad.begin(), ad.end(), iq.begin() + static_cast<size_t>(i - overlap) * ad.size());
        } else {
            std::complex<float>* row = dataW->data() + static_cast<size_t>(i) * N;
            for (size_t j = 0; j + 1 < ad.size(); j += 2) row[j / 2] = {static_cast<float>(ad[j]), static_cast<float>(ad[j + 1])};
        }
    }
    if (data_type == kIfds && e == hipSuccess && fresh > 0) {
        // pulse compression on the device (multiThreadIFDS, common.cu:583-604). The reference
        // takes echo i's velocity from dataNOut[M + dataWFileSn[i]], an index into the window's
        // 4 x M table by absolute frame number (out of range once the window moves); here it is
        // echo i's own velocity, dataNOut[M + i].
        e = hipMemcpy(g.d_iq, iq.data(), iq.size() * sizeof(int16_t), hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hipMemcpy(g.d_pc_velocity, g.h_velocity.data() + overlap, sizeof(double) * fresh, hipMemcpyHostToDevice);
        if (e == hipSuccess &&
            isar_pulse_compress(&g.paras, fresh, g.iq_len, g.d_iq, g.d_pc_velocity, cur + static_cast<size_t>(overlap) * N,
                                nullptr, nullptr) != ISAR_OK)
            return EXIT_FAILURE;   // reason recorded by isar_pulse_compress
    } else if (e == hipSuccess && fresh > 0) {
        e = hipMemcpy(cur + static_cast<size_t>(overlap) * N, dataW->data() + static_cast<size_t>(overlap) * N,
                      sizeof(float2) * fresh * N, hipMemcpyHostToDevice);
    }
    if (e == hipSuccess) e = hipMemcpy(g.d_velocity, g.h_velocity.data(), sizeof(double) * M, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(g.d_proc, cur, sizeof(float2) * static_cast<size_t>(M) * N, hipMemcpyDeviceToDevice);
    if (e != hipSuccess) return err(std::string("dataExtracting copy: ") + hipGetErrorString(e)), EXIT_FAILURE;
    return EXIT_SUCCESS;
}

ISAR_API int isar_session_hrrp(float* h_hrrp) {
    if (!g.ctx || !g.d_hrrp) return err("no range-profile buffer: imagingMemInit without if_hrrp"), ISAR_ERR_STATE;
    if (!h_hrrp) return err("null output pointer"), ISAR_ERR_INVALID;
    const size_t bytes = sizeof(float2) * static_cast<size_t>(g.paras.echo_num) * g.paras.range_num;
    const hipError_t e = hipMemcpy(h_hrrp, g.d_hrrp, bytes, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return err(std::string("isar_session_hrrp: ") + hipGetErrorString(e)), ISAR_ERR_DEVICE;
    return ISAR_OK;
}

ISAR_API int isar_session_dump(unsigned mask, const char* dir) {
    if (!g.ctx) return err("isar_session_dump before imagingMemInit"), ISAR_ERR_STATE;
    return isar_stage_dump(g.ctx, 0, mask, dir);
}

DLL_EXPORT_API void isarMainSingle(float* h_img, const int& data_type, const int& option_alignment, const int& option_phase,
                                   const bool& if_hrrp, const bool& if_hpc, const bool& if_mtrc) {
    if (!g.ctx || !h_img) { err("isarMainSingle before imagingMemInit"); return; }
    isar_options o;
    isar_default_options(&o);
    o.data_type = data_type;
    o.option_alignment = option_alignment;
    o.option_phase = option_phase;
    o.if_hrrp = if_hrrp;
    o.if_hpc = if_hpc;
    o.if_mtrc = if_mtrc;
    const size_t bytes = sizeof(float) * g.paras.echo_num * g.paras.range_num_cut;
    auto image = [&]() -> bool {
        if (isar_run_device(g.ctx, 1, g.d_proc, g.d_velocity, nullptr, g.d_img, &o, nullptr) != ISAR_OK) return false;
        hipError_t e = hipMemcpy(h_img, g.d_img, bytes, hipMemcpyDeviceToHost);
        if (e == hipSuccess) e = hipDeviceSynchronize();
        if (e != hipSuccess) err(std::string("isarMainSingle: ") + hipGetErrorString(e));
        return e == hipSuccess;
    };
    if (!image()) return;
    isar_frame_trace tr{};   // decisions only (no arrays requested)
    const int mode = isar_get_autofocus_mode(g.ctx);
    if (mode == ISAR_AF_RESIDENT && isar_get_trace(g.ctx, 0, &tr) == ISAR_OK && (tr.status & 2)) {
        std::fprintf(stderr, "isarMainSingle: a resident autofocus wait hit its bound; imaging the frame again with the "
                             "split form\n");
        (void)isar_set_autofocus_mode(g.ctx, ISAR_AF_SPLIT);
        (void)image();
        (void)isar_set_autofocus_mode(g.ctx, mode);
    }
}

ISAR_API isar_context* isar_session_context(void) { return g.ctx; }

DLL_EXPORT_API void writeFileFLT(const std::string& outFilePath, const float* data, const size_t& data_size) {
    std::ofstream ofs(outFilePath);
    if (!ofs.is_open()) { err("cannot open " + outFilePath); return; }
    for (size_t i = 0; i < data_size; ++i) ofs << std::fixed << std::setprecision(5) << data[i] << "\n";
}

DLL_EXPORT_API void imagingMemDest(const int& data_type, const bool& if_hpc, const bool& if_hrrp) {
    (void)data_type;
    (void)if_hpc;
    free_buffers();
}

}  // extern "C"
