// stage.hip — intermediate results of a finished run in the form the reference dumps them with
// its DATA_WRITE_BACK_* switches (isar_main.cu:10-12, 51-53, 101-103, 152-154, 178-180). The
// fused kernels never hold these arrays whole; these kernels rebuild them on request from what
// the run kept (echo rows, aligned profiles, cut, phases, chirp-z output). Not on the imaging path.
#include "../dev_common.hpp"
#include "../launchers.hpp"

namespace isar {

namespace {
constexpr int kStageThreads = 256;
}

// Rows after high-speed compensation (high_speed_compensation.cu:4-54): x[m, n] exp(j pi k v_m n^2),
// the same FP64-reduced phase the range-alignment loads apply.
__global__ __launch_bounds__(kStageThreads) void k_stage_hpc(const float2* echo, const double* vel, double hpc_k,
                                                             int N, float2* out) {
    const int m = blockIdx.y, n = blockIdx.x * kStageThreads + threadIdx.x;
    if (n >= N) return;
    const size_t i = static_cast<size_t>(m) * N + n;
    const double dn = static_cast<double>(n);
    out[i] = cmul(echo[i], expjpi_reduced(hpc_k * vel[m] * dn * dn));
}

// Centred range profiles (HRRPCenter, range_alignment.cu:371-465): row m of the aligned profiles
// circularly shifted by `shift` cells, out[m, n] = st[m, (n - shift) mod N].
__global__ __launch_bounds__(kStageThreads) void k_stage_roll(const float2* st, int N, int shift, float2* out) {
    const int m = blockIdx.y, n = blockIdx.x * kStageThreads + threadIdx.x;
    if (n >= N) return;
    const int q = (((n - shift) % N) + N) % N;
    out[static_cast<size_t>(m) * N + n] = st[static_cast<size_t>(m) * N + q];
}

// Phase-compensated cut (phase_adjustment.cu:324): out[m, w] = x[m, w] phi[m].
__global__ __launch_bounds__(kStageThreads) void k_stage_rowphase(const float2* x, const float2* phi, int W,
                                                                  float2* out) {
    const int m = blockIdx.y, w = blockIdx.x * kStageThreads + threadIdx.x;
    if (w >= W) return;
    const size_t i = static_cast<size_t>(m) * W + w;
    out[i] = cmul(x[i], phi[m]);
}

// out (C x R) = scale * in (R x C) transposed, with the input's row and column indices rotated by
// sr and sc: out[c, r] = scale in[(r + sr) mod R, (c + sc) mod C]. One thread per output element.
__global__ __launch_bounds__(kStageThreads) void k_stage_transpose(const float2* in, int R, int C, int sr, int sc,
                                                                   float scale, float2* out) {
    const int c = blockIdx.y, r = blockIdx.x * kStageThreads + threadIdx.x;
    if (r >= R) return;
    const int rr = (r + sr) % R, cc = (c + sc) % C;
    out[static_cast<size_t>(c) * R + r] = cscale(in[static_cast<size_t>(rr) * C + cc], scale);
}

// Last index shuffles of MTRC (mtrc.cu:63-75): ifftshift of the rows, ifftshift of the columns,
// and for IFDS fliplr: out[k, c] = a[(k + M/2) mod M, (c' + W/2) mod W], c' = W-1-c for IFDS.
__global__ __launch_bounds__(kStageThreads) void k_stage_mtrc_out(const float2* a, int M, int W, int ifds,
                                                                  float2* out) {
    const int k = blockIdx.y, c = blockIdx.x * kStageThreads + threadIdx.x;
    if (c >= W) return;
    const int cf = ifds ? W - 1 - c : c;
    out[static_cast<size_t>(k) * W + c] = a[static_cast<size_t>((k + M / 2) % M) * W + (cf + W / 2) % W];
}

namespace {
dim3 rows_grid(int len, int rows) { return dim3((len + kStageThreads - 1) / kStageThreads, rows); }
}  // namespace

void launch_stage_hpc(const float2* echo, const double* vel, double hpc_k, int M, int N, float2* out, hipStream_t s) {
    ISAR_LAUNCH(k_stage_hpc, rows_grid(N, M), dim3(kStageThreads), 0, s, echo, vel, hpc_k, N, out);
}

void launch_stage_roll(const float2* st, int M, int N, int shift, float2* out, hipStream_t s) {
    ISAR_LAUNCH(k_stage_roll, rows_grid(N, M), dim3(kStageThreads), 0, s, st, N, shift, out);
}

void launch_stage_rowphase(const float2* x, const float2* phi, int M, int W, float2* out, hipStream_t s) {
    ISAR_LAUNCH(k_stage_rowphase, rows_grid(W, M), dim3(kStageThreads), 0, s, x, phi, W, out);
}

void launch_stage_transpose(const float2* in, int R, int C, int sr, int sc, float scale, float2* out, hipStream_t s) {
    ISAR_LAUNCH(k_stage_transpose, rows_grid(R, C), dim3(kStageThreads), 0, s, in, R, C, sr, sc, scale, out);
}

void launch_stage_mtrc_out(const float2* a, int M, int W, int ifds, float2* out, hipStream_t s) {
    ISAR_LAUNCH(k_stage_mtrc_out, rows_grid(W, M), dim3(kStageThreads), 0, s, a, M, W, ifds, out);
}

}  // namespace isar
