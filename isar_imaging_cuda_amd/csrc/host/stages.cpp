// stages.cpp — intermediate arrays of the last run (isar_stage_read / isar_stage_dump), the
// runtime counterpart of the reference's DATA_WRITE_BACK_* dumps (isar_main.cu:10-12, 51-53,
// 76-78, 101-103, 152-154, 178-180, 221-223; ioOperation::dataWriteBack, common.cuh:866-879).
#include <fstream>
#include <iomanip>
#include <string>
#include <vector>

#include "../launchers.hpp"
#include "pipeline.hpp"

namespace isar {

namespace {

const char* const kStageFile[ISAR_DUMP_COUNT] = {"dataW.dat", "hpc.dat", "hrrp.dat", "ra.dat",
                                                 "pc.dat",    "mtrc.dat", "final.dat"};

size_t stage_bytes(const Job& J, int stage) {
    const size_t MN = static_cast<size_t>(J.M) * J.N, MW = static_cast<size_t>(J.M) * J.W;
    switch (stage) {
    case ISAR_DUMP_DATAW: case ISAR_DUMP_HPC: case ISAR_DUMP_HRRP: case ISAR_DUMP_RA: return MN * sizeof(float2);
    case ISAR_DUMP_PC: case ISAR_DUMP_MTRC: return MW * sizeof(float2);
    case ISAR_DUMP_FINAL: return MW * sizeof(float);
    default: return 0;
    }
}

template <typename T>
int read_entry(const T* dev, int f, T* out) {
    return hipMemcpy(out, dev + f, sizeof(T), hipMemcpyDeviceToHost) == hipSuccess ? ISAR_OK : ISAR_ERR_DEVICE;
}

int rebuild(Context& c, int stage, int f, void* h_out) {
    const Job& L = c.last_job;
    if (c.last_mode != ISAR_RUN_FULL) return fail(ISAR_ERR_STATE, "stage reads need a full run before them");
    if (stage < 0 || stage >= ISAR_DUMP_COUNT) return fail(ISAR_ERR_INVALID, "unknown stage");
    if (f < 0 || f >= L.K) return fail(ISAR_ERR_INVALID, "frame outside the last run");
    if (!h_out) return fail(ISAR_ERR_INVALID, "null output pointer");
    const int M = L.M, N = L.N, W = L.W;
    hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) return fail(ISAR_ERR_DEVICE, std::string("stage read: ") + hipGetErrorString(e));
    const size_t cap = static_cast<size_t>(M) * (N > W ? N : W);
    if (!c.stage_scratch) {
        if (hipMalloc(&c.stage_scratch, 3 * cap * sizeof(float2)) != hipSuccess) {
            c.stage_scratch = nullptr;
            return fail(ISAR_ERR_MEMORY, "stage scratch allocation failed");
        }
        c.allocations.push_back(c.stage_scratch);
    }
    float2* b0 = c.stage_scratch;
    float2* b1 = b0 + cap;
    long long first = static_cast<long long>(f) * M;
    if (L.frame_row && read_entry(L.frame_row, f, &first) != ISAR_OK) return fail(ISAR_ERR_DEVICE, "frame table read");
    const float2* echo = L.echo + first * N;
    const double* vel = L.velocity ? L.velocity + first : nullptr;
    const size_t fMN = static_cast<size_t>(f) * M * N, fMW = static_cast<size_t>(f) * M * W;
    const hipStream_t s = nullptr;
    const void* src = b0;
    Job one = L;   // the frame as a batch of one, for the imaging kernels reused here
    one.K = 1;
    one.echo = echo;
    one.velocity = vel;
    one.frame_row = nullptr;
    one.hrrp = nullptr;
    switch (stage) {
    case ISAR_DUMP_DATAW:
        src = echo;
        break;
    case ISAR_DUMP_HPC:
        if (L.data_type == 1) {
            one.pre = b0;
            launch_ifds_pre(one, s);
        } else if (L.if_hpc) {
            if (!vel) return fail(ISAR_ERR_STATE, "the last run had no velocity vector");
            launch_stage_hpc(echo, vel, L.hpc_k, M, N, b0, s);
        } else {
            src = echo;
        }
        break;
    case ISAR_DUMP_HRRP:
        if (L.data_type == 1) {
            src = echo;   // getHRRP undoes the IFDS pre-stage
        } else {
            if (L.if_hpc && !vel) return fail(ISAR_ERR_STATE, "the last run had no velocity vector");
            one.hrrp = b0;
            launch_hrrp(one, s);
        }
        break;
    case ISAR_DUMP_RA: {
        FrameDecisions d{};
        if (read_entry(L.dec, f, &d) != ISAR_OK) return fail(ISAR_ERR_DEVICE, "decision read");
        launch_stage_roll(L.st + fMN, M, N, d.center_shift, b0, s);
        break;
    }
    case ISAR_DUMP_PC:
        launch_stage_rowphase(L.x + fMW, L.phi + static_cast<size_t>(f) * M, M, W, b0, s);
        break;
    case ISAR_DUMP_MTRC:
        if (!L.if_mtrc) {
            launch_stage_rowphase(L.x + fMW, L.phi + static_cast<size_t>(f) * M, M, W, b0, s);
            break;
        }
        // mtrc.cu:53-75 after the chirp-z transform: ifftshift, IFFT / M and ifftshift along slow
        // time (through a transpose), FFT along range, ifftshift of rows and columns, fliplr for IFDS
        launch_stage_transpose(L.czt + fMW, M, W, M / 2, 0, 1.0f, b0, s);
        if (!debug_fft(M, +1, b0, b1, W, s)) return fail(ISAR_ERR_INVALID, "unsupported pulse count");
        launch_stage_transpose(b1, W, M, 0, M / 2, 1.0f / static_cast<float>(M), b0, s);
        if (!debug_fft(W, -1, b0, b1, M, s)) return fail(ISAR_ERR_INVALID, "unsupported cut width");
        launch_stage_mtrc_out(b1, M, W, L.data_type == 1, b0, s);
        break;
    case ISAR_DUMP_FINAL: {
        long long off = static_cast<long long>(f) * M * W;
        if (L.frame_img && read_entry(L.frame_img, f, &off) != ISAR_OK) return fail(ISAR_ERR_DEVICE, "image table read");
        src = L.img + off;
        break;
    }
    default:
        break;
    }
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpy(h_out, src, stage_bytes(L, stage), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return fail(ISAR_ERR_DEVICE, std::string("stage read: ") + hipGetErrorString(e));
    return ISAR_OK;
}

}  // namespace

}  // namespace isar

extern "C" {

ISAR_API size_t isar_stage_bytes(const isar_context* ctx, int stage) {
    return ctx ? isar::stage_bytes(ctx->tmpl, stage) : 0;
}

ISAR_API int isar_stage_read(isar_context* ctx, int stage, int frame, void* h_out) {
    if (!ctx) return isar::fail(ISAR_ERR_INVALID, "null context");
    return isar::rebuild(*ctx, stage, frame, h_out);
}

ISAR_API int isar_stage_dump(isar_context* ctx, int frame, unsigned mask, const char* dir) {
    if (!ctx || !dir) return isar::fail(ISAR_ERR_INVALID, "null context or directory");
    for (int st = 0; st < ISAR_DUMP_COUNT; ++st) {
        if (!(mask & (1u << st))) continue;
        const size_t bytes = isar::stage_bytes(ctx->tmpl, st);
        std::vector<float> h(bytes / sizeof(float));
        const int rc = isar::rebuild(*ctx, st, frame, h.data());
        if (rc != ISAR_OK) return rc;
        const std::string path = std::string(dir) + "/" + isar::kStageFile[st];
        std::ofstream out(path);
        if (!out) return isar::fail(ISAR_ERR_IO, "cannot open " + path);
        out << std::fixed << std::setprecision(5);
        for (float v : h) out << v << "\n";   // complex arrays: real part, then imaginary part
        if (!out) return isar::fail(ISAR_ERR_IO, "cannot write " + path);
    }
    return ISAR_OK;
}

}  // extern "C"
