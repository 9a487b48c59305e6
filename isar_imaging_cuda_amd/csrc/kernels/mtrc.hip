// mtrc.hip — MTRC (mtrc.cu:4-192) and azimuth compression (isar_main.cu:191-207) for gfx950.
//
// The reference chain after the autofocus is, with P = half swap, F = DFT, L/Fl = flips:
//   rows:  St = F_W^-1 P_W (X diag(phi)) / W          (IFDS: F_W P_W without the scale)
//   cols:  C  = CZT_alpha_n(St[:, n])                   (Bluestein, alpha_n = 1 - c0 + posa n)
//   cols:  P_M, F_M^-1 / M, P_M ; rows: F_W, P_W ; cols: P_M ; [IFDS rows: L]
//   Q1 window on the first W elements ; cols: F_M, P_M ; [IFDS cols: Fl] ; |.|
// The column operations after the CZT are P F Q F^-1 P with P P = I, and every row operation
// commutes with them. With the reference's window (only row 0, element j scaled by
// c_j = hamming_M[j / M]) the chain collapses to
//   img[k, j] = | R[k', j] + delta_j |,  R = [L] P_W F_W C (rows),  k' = [M-1-k] k,
//   delta = (c - 1) / M * ([L] P_W F_W sum_k C[k, :])
// and with the intended full slow-time window it is R computed from C' = P F diag(h) F^-1 P C / M.
#include <algorithm>

#include "../dev_common.hpp"
#include "../launchers.hpp"

namespace isar {

namespace {
constexpr int kRowE = 16;
constexpr int kRowsPerBlock = 8;   // W = 512 rows: 32 threads per row transform

// Chirp-z column tiles: an F-point transform per range bin (F = 2^k with 16 elements per thread,
// or 3 * 2^k with 12), NT_ threads per workgroup, CT = NT_ / TPF bins per workgroup.
template <int F, int NT_ = 256>
struct CztCfg {
    static constexpr int E = (F % 3 == 0) ? 12 : 16;
    static constexpr int TPF = F / E;
    static constexpr int NT = NT_;
    static constexpr int CT = NT / TPF;
    static_assert(F % E == 0 && NT % TPF == 0, "chirp-z tile");
    static constexpr int R0 = FftPlan<F, E>::R[0];
    static constexpr int RL = FftPlan<F, E>::R[FftPlan<F, E>::NP - 1];
    __device__ static int in(int lane, int e) { return fft_in_pos<F, E, R0>(lane, e); }
    __device__ static int out(int lane, int e) { return fft_out_pos<F, E, RL>(lane, e); }
};

template <int W>
__device__ __forceinline__ int row_in(int lane, int e) { return fft_in_pos<W, kRowE, FftPlan<W, kRowE>::R[0]>(lane, e); }
template <int W>
__device__ __forceinline__ int row_out(int lane, int e) {
    return fft_out_pos<W, kRowE, FftPlan<W, kRowE>::R[FftPlan<W, kRowE>::NP - 1]>(lane, e);
}
}  // namespace

// Bluestein filter spectra: table[n][:] = FFT_F(v), v[t] = exp(+j pi alpha_n t^2 / M) for
// |t| < M placed circularly, zero elsewhere. Phases reduced in FP64. One range bin per block.
template <int F>
__global__ __launch_bounds__(CztCfg<F>::TPF) void k_czt_table(int M, int W, double c0, double posa, float2* table) {
    HIP_DYNAMIC_SHARED(float2, lds);
    using C = CztCfg<F>;
    const int n = blockIdx.x, lane = threadIdx.x;
    const double alpha = 1.0 - c0 + posa * n;
    float2 x[C::E];
#pragma unroll
    for (int e = 0; e < C::E; ++e) {
        const int pos = C::in(lane, e);
        int t = -1;
        if (pos < M) t = pos;
        else if (pos > F - M) t = F - pos;
        x[e] = t >= 0 ? expjpi_reduced(alpha * static_cast<double>(t) * t / M) : make_float2(0.0f, 0.0f);
    }
    fft<F, C::E, -1>(x, lds, lane);
#pragma unroll
    for (int e = 0; e < C::E; ++e) table[static_cast<size_t>(n) * F + C::out(lane, e)] = x[e];
}

// Input and output chirps of the chirp-z transform per range bin n (mtrc.cu:125-192), phases
// reduced in FP64: pre[n][m] = exp(j pi a_n m (M - m) / M), post[n][k] = exp(-j pi a_n k^2 / M) / F.
__global__ void k_czt_chirps(int M, int W, int F, double c0, double posa, float2* pre, float2* post) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M * W) return;
    const int n = i / M, m = i % M;
    const double alpha = 1.0 - c0 + posa * n;
    pre[i] = expjpi_reduced(alpha * static_cast<double>(m) * (M - m) / M);
    post[i] = cscale(expjpi_reduced(-alpha * static_cast<double>(m) * m / M), 1.0f / static_cast<float>(F));
}

// Range-frequency rows (mtrc.cu:14-40) with the autofocus phase applied (phase_adjustment.cu:324).
template <int W>
__global__ __launch_bounds__(kRowsPerBlock * W / kRowE) void k_mtrc_rows(Job J) {
    HIP_DYNAMIC_SHARED(float2, sm);
    constexpr int TPF = W / kRowE;
    const int tid = threadIdx.x, f = blockIdx.y;
    const int g = tid / TPF, lane = tid % TPF;
    const int M = J.M;
    const int m = blockIdx.x * kRowsPerBlock + g;
    const bool valid = m < M;
    const int mm = valid ? m : 0;
    const float2 ph = J.phi[static_cast<size_t>(f) * M + mm];
    const float2* xr = J.x + (static_cast<size_t>(f) * M + mm) * W;
    float2 x[kRowE];
#pragma unroll
    for (int e = 0; e < kRowE; ++e) x[e] = cmul(xr[(row_in<W>(lane, e) + W / 2) & (W - 1)], ph);
    float2* buf = sm + g * lds_padded_size(W);
    if (J.data_type == 1) {
        fft<W, kRowE, -1, true>(x, buf, lane);
    } else {
        fft<W, kRowE, +1, true>(x, buf, lane);
        const float inv = 1.0f / W;
#pragma unroll
        for (int e = 0; e < kRowE; ++e) x[e] = cscale(x[e], inv);
    }
    if (valid) {
        float2* dst = J.stm + (static_cast<size_t>(f) * M + m) * W;
#pragma unroll
        for (int e = 0; e < kRowE; ++e) dst[row_out<W>(lane, e)] = x[e];
    }
}

// Chirp-z transform along slow time for a tile of range bins (mtrc.cu:42-52, 125-192):
// X[k] = exp(-j pi a k^2/M) / F * IFFT_F( FFT_F(y) * table_n )[k],
// y[m] = x[m] exp(j pi a m (M - m) / M) for m < M. Also writes sum_k X[k] per bin.
template <int F, int NTW>
__global__ __launch_bounds__(NTW) void k_mtrc_czt(Job J) {
    using C = CztCfg<F, NTW>;
    constexpr int CT = C::CT, TPF = C::TPF, E = C::E, NT = C::NT;
    HIP_DYNAMIC_SHARED(float2, sm);
    const int tile = xcd_tile(blockIdx.x, gridDim.x), f = blockIdx.y, tid = threadIdx.x;
    const int M = J.M, W = J.W;
    const int g = tid / TPF, lane = tid % TPF;
    const int n0 = tile * CT, n = n0 + g;
    const float2* S = J.stm + static_cast<size_t>(f) * M * W;
    for (int idx = tid; idx < M * CT; idx += NT) {
        const int m = idx / CT, c = idx % CT;
        sm[m * (CT + 1) + c] = S[static_cast<size_t>(m) * W + n0 + c];
    }
    __syncthreads();
    const float2* pre = J.czt_pre + static_cast<size_t>(n) * M;
    const float2* post = J.czt_post + static_cast<size_t>(n) * M;
    float2 x[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int pos = C::in(lane, e);
        x[e] = pos < M ? cmul(sm[pos * (CT + 1) + g], pre[pos]) : make_float2(0.0f, 0.0f);
    }
    __syncthreads();
    // a transform within one wave orders its passes by the wave alone (the transpose back below
    // waits for every wave with a barrier of its own)
    constexpr bool kWave = TPF <= 64;
    float2* buf = sm + g * lds_padded_size(F);
    const FftTwiddles<F, E> tw = fft_twiddles<F, E>(lane);
    fft<F, E, -1, kWave>(x, buf, lane, tw);
    const float2* tab = J.czt_table + static_cast<size_t>(n) * F;
#pragma unroll
    for (int e = 0; e < E; ++e) x[e] = cmul(x[e], tab[C::out(lane, e)]);
    relayout_out_to_in<F, E, kWave>(x, buf, lane);
    fft<F, E, +1, kWave>(x, buf, lane, tw);
    float2 csum = make_float2(0.0f, 0.0f);
    float2 y[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int k = C::out(lane, e);
        y[e] = make_float2(0.0f, 0.0f);
        if (k < M) {
            y[e] = cmul(x[e], post[k]);
            csum = cadd(csum, y[e]);
        }
    }
    if constexpr (TPF <= 64) {
        csum = group_sum2<TPF>(csum);
        if (lane == 0) J.colsum[static_cast<size_t>(f) * W + n] = csum;
        __syncthreads();   // every wave's last transform loads are done before the transpose
    } else {
        // one transform spans TPF / 64 waves: wave sums first, then the transform's waves in LDS.
        // The partials sit past the transpose area (M (CT + 1) entries); czt_lds counts them.
        float2* wave_part = sm + static_cast<size_t>(M) * (CT + 1);
        csum = group_sum2<64>(csum);
        if ((tid & 63) == 0) wave_part[tid / 64] = csum;
        __syncthreads();
        if (lane == 0) {
            float2 t = make_float2(0.0f, 0.0f);
            for (int w = 0; w < TPF / 64; ++w) t = cadd(t, wave_part[g * (TPF / 64) + w]);
            J.colsum[static_cast<size_t>(f) * W + n] = t;
        }
    }
    // transpose back through LDS for row-contiguous stores
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int k = C::out(lane, e);
        if (k < M) sm[k * (CT + 1) + g] = y[e];
    }
    __syncthreads();
    float2* D = J.czt + static_cast<size_t>(f) * M * W;
    for (int idx = tid; idx < M * CT; idx += NT) {
        const int m = idx / CT, c = idx % CT;
        D[static_cast<size_t>(m) * W + n0 + c] = sm[m * (CT + 1) + c];
    }
}

// Intended slow-time window (quirk off): C' = P F diag(hamming_M) F^-1 P C / M in place.
template <int M>
__global__ __launch_bounds__(256) void k_col_window(Job J) {
    using C = ColCfg<M>;
    constexpr int CT = C::CT, TPF = C::TPF, E = C::E, NT = C::NT;
    HIP_DYNAMIC_SHARED(float2, sm);
    const int tile = xcd_tile(blockIdx.x, gridDim.x), f = blockIdx.y, tid = threadIdx.x;
    const int W = J.W;
    const int g = tid / TPF, lane = tid % TPF;
    const int n0 = tile * CT;
    float2* D = J.czt + static_cast<size_t>(f) * M * W;
    for (int idx = tid; idx < M * CT; idx += NT) {
        const int m = idx / CT, c = idx % CT;
        sm[m * (CT + 1) + c] = D[static_cast<size_t>(m) * W + n0 + c];
    }
    __syncthreads();
    float2 x[E];
#pragma unroll
    for (int e = 0; e < E; ++e) x[e] = sm[((C::in(lane, e) + M / 2) % M) * (CT + 1) + g];
    __syncthreads();
    float2* buf = sm + g * lds_padded_size(M);
    const FftTwiddles<M, E> tw = fft_twiddles<M, E>(lane);
    fft<M, E, +1>(x, buf, lane, tw);
    const float invM = 1.0f / M;
#pragma unroll
    for (int e = 0; e < E; ++e) x[e] = cscale(x[e], invM * hamming(C::out(lane, e), M));
    relayout_out_to_in<M, E>(x, buf, lane);
    fft<M, E, -1>(x, buf, lane, tw);
#pragma unroll
    for (int e = 0; e < E; ++e) sm[((C::out(lane, e) + M / 2) % M) * (CT + 1) + g] = x[e];
    __syncthreads();
    for (int idx = tid; idx < M * CT; idx += NT) {
        const int m = idx / CT, c = idx % CT;
        D[static_cast<size_t>(m) * W + n0 + c] = sm[m * (CT + 1) + c];
    }
}

// Slow-time window quirk term, once per frame: delta = (c - 1) / M * ([L] P_W F_W colsum),
// stored at the last-pass positions of the row transform.
template <int W>
__global__ __launch_bounds__(W / kRowE) void k_mtrc_delta(Job J) {
    HIP_DYNAMIC_SHARED(float2, buf);
    const int lane = threadIdx.x, f = blockIdx.x;
    const int M = J.M;
    const bool ifds = J.data_type == 1;
    const float2* cs = J.colsum + static_cast<size_t>(f) * W;
    float2 d[kRowE];
#pragma unroll
    for (int e = 0; e < kRowE; ++e) d[e] = cs[row_in<W>(lane, e)];
    fft<W, kRowE, -1>(d, buf, lane);
    float2* out = J.delta + static_cast<size_t>(f) * W;
#pragma unroll
    for (int e = 0; e < kRowE; ++e) {
        const int p = row_out<W>(lane, e);
        int j = (p + W / 2) & (W - 1);
        if (ifds) j = W - 1 - j;
        out[p] = cscale(d[e], (hamming(j / M, M) - 1.0f) / static_cast<float>(M));
    }
}

// Final rows: img[k, j] = | ([L] P_W F_W C[k'])[j] + delta_j |.
template <int W>
__global__ __launch_bounds__(kRowsPerBlock * W / kRowE) void k_mtrc_image(Job J) {
    HIP_DYNAMIC_SHARED(float2, sm);
    constexpr int TPF = W / kRowE;
    const int tid = threadIdx.x, f = blockIdx.y;
    const int g = tid / TPF, lane = tid % TPF;
    const int M = J.M;
    const bool ifds = J.data_type == 1;
    const int k = blockIdx.x * kRowsPerBlock + g;
    const bool valid = k < M;
    float2* buf = sm + g * lds_padded_size(W);
    float2 d[kRowE];
    const bool quirk = J.quirk_q1 != 0;
    if (quirk) {
        const float2* dl = J.delta + static_cast<size_t>(f) * W;
#pragma unroll
        for (int e = 0; e < kRowE; ++e) d[e] = dl[row_out<W>(lane, e)];
    }
    const int kk = valid ? (ifds ? M - 1 - k : k) : 0;
    const float2* cr = J.czt + (static_cast<size_t>(f) * M + kk) * W;
    float2 x[kRowE];
#pragma unroll
    for (int e = 0; e < kRowE; ++e) x[e] = cr[row_in<W>(lane, e)];
    fft<W, kRowE, -1, true>(x, buf, lane);
    if (valid) {
        float* dst = image_of(J, f) + static_cast<size_t>(k) * W;
#pragma unroll
        for (int e = 0; e < kRowE; ++e) {
            int j = (row_out<W>(lane, e) + W / 2) & (W - 1);
            if (ifds) j = W - 1 - j;
            const float2 v = quirk ? cadd(x[e], d[e]) : x[e];
            dst[j] = cabsf_(v);
        }
    }
}

// No MTRC (if_mtrc = 0): img = | [Fl] P_M F_M window(X diag(phi)) | per column tile.
template <int M>
__global__ __launch_bounds__(256) void k_post_nomtrc(Job J) {
    using C = ColCfg<M>;
    constexpr int CT = C::CT, TPF = C::TPF, E = C::E, NT = C::NT;
    HIP_DYNAMIC_SHARED(float2, sm);
    const int tile = xcd_tile(blockIdx.x, gridDim.x), f = blockIdx.y, tid = threadIdx.x;
    const int W = J.W;
    const int g = tid / TPF, lane = tid % TPF;
    const int n0 = tile * CT, n = n0 + g;
    const float2* X = J.x + static_cast<size_t>(f) * M * W;
    const float2* ph = J.phi + static_cast<size_t>(f) * M;
    for (int idx = tid; idx < M * CT; idx += NT) {
        const int m = idx / CT, c = idx % CT;
        sm[m * (CT + 1) + c] = cmul(X[static_cast<size_t>(m) * W + n0 + c], ph[m]);
    }
    __syncthreads();
    float2 x[E];
    const bool quirk = J.quirk_q1 != 0;
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int pos = C::in(lane, e);
        float w = 1.0f;
        if (quirk) { if (pos == 0) w = hamming(n / M, M); }
        else w = hamming(pos, M);
        x[e] = cscale(sm[pos * (CT + 1) + g], w);
    }
    __syncthreads();
    float2* buf = sm + g * lds_padded_size(M);
    fft<M, E, -1>(x, buf, lane);
    float* fo = reinterpret_cast<float*>(sm);
#pragma unroll
    for (int e = 0; e < E; ++e) {
        int r = (C::out(lane, e) + M / 2) % M;
        if (J.data_type == 1) r = M - 1 - r;
        fo[r * (CT + 1) + g] = cabsf_(x[e]);
    }
    __syncthreads();
    float* D = image_of(J, f);
    for (int idx = tid; idx < M * CT; idx += NT) {
        const int m = idx / CT, c = idx % CT;
        D[static_cast<size_t>(m) * W + n0 + c] = fo[m * (CT + 1) + c];
    }
}

// ------------------------------------------------------------------------------------------
namespace {
template <int F, int NTW>
size_t czt_lds(int M) {
    using C = CztCfg<F, NTW>;
    return std::max(static_cast<size_t>(M) * (C::CT + 1) + NTW / 64, static_cast<size_t>(C::CT) * lds_padded_size(F)) *
           sizeof(float2);
}
template <int M>
size_t col_lds() {
    using C = ColCfg<M>;
    return std::max(static_cast<size_t>(M) * (C::CT + 1), static_cast<size_t>(C::CT) * lds_padded_size(M)) * sizeof(float2);
}
size_t rows_lds() { return kRowsPerBlock * lds_padded_size(kRangeCut) * sizeof(float2); }

#define ISAR_F_SWITCH(F, BODY)                         \
    switch (F) {                                       \
    case 128: { constexpr int FF = 128; BODY; } break;   \
    case 256: { constexpr int FF = 256; BODY; } break;   \
    case 384: { constexpr int FF = 384; BODY; } break;   \
    case 512: { constexpr int FF = 512; BODY; } break;   \
    case 768: { constexpr int FF = 768; BODY; } break;   \
    case 1024: { constexpr int FF = 1024; BODY; } break; \
    case 2048: { constexpr int FF = 2048; BODY; } break; \
    default: break;                                    \
    }
#define ISAR_PULSE_SWITCH2(M, BODY)                      \
    switch (M) {                                         \
    case 64: { constexpr int MM = 64; BODY; } break;     \
    case 128: { constexpr int MM = 128; BODY; } break;   \
    case 192: { constexpr int MM = 192; BODY; } break;   \
    case 256: { constexpr int MM = 256; BODY; } break;   \
    case 384: { constexpr int MM = 384; BODY; } break;   \
    case 512: { constexpr int MM = 512; BODY; } break;   \
    case 1024: { constexpr int MM = 1024; BODY; } break; \
    default: break;                                      \
    }
}  // namespace

hipError_t configure_mtrc_kernels() {
    hipError_t err = hipSuccess;
#ifndef ISAR_EMU
    auto set = [&](const void* fn, size_t bytes) {
        hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(bytes));
        if (e != hipSuccess) {
            err = e;
            report_config_error(__FILE__, bytes);
        }
    };
#define ISAR_SET_F(FF)                                                                          \
    set(reinterpret_cast<const void*>(&k_czt_table<FF>), lds_padded_size(FF) * sizeof(float2)); \
    set(reinterpret_cast<const void*>(&k_mtrc_czt<FF, 256>), czt_lds<FF, 256>(FF / 2));         \
    set(reinterpret_cast<const void*>(&k_mtrc_czt<FF, 512>), czt_lds<FF, 512>(FF / 2));
    ISAR_SET_F(128) ISAR_SET_F(256) ISAR_SET_F(384) ISAR_SET_F(512) ISAR_SET_F(768) ISAR_SET_F(1024) ISAR_SET_F(2048)
#undef ISAR_SET_F
#define ISAR_SET_M(MM)                                                         \
    set(reinterpret_cast<const void*>(&k_col_window<MM>), col_lds<MM>());     \
    set(reinterpret_cast<const void*>(&k_post_nomtrc<MM>), col_lds<MM>());
    ISAR_SET_M(64) ISAR_SET_M(128) ISAR_SET_M(192) ISAR_SET_M(256) ISAR_SET_M(384) ISAR_SET_M(512) ISAR_SET_M(1024)
#undef ISAR_SET_M
    set(reinterpret_cast<const void*>(&k_mtrc_rows<kRangeCut>), rows_lds());
    set(reinterpret_cast<const void*>(&k_mtrc_image<kRangeCut>), rows_lds());
#endif
    return err;
}

void launch_czt_chirps(int M, int W, int F, double c0, double posa, float2* pre, float2* post, hipStream_t s) {
    ISAR_LAUNCH(k_czt_chirps, dim3((M * W + 255) / 256), dim3(256), 0, s, M, W, F, c0, posa, pre, post);
}
void launch_mtrc_delta(const Job& J, hipStream_t s) {
    ISAR_LAUNCH(k_mtrc_delta<kRangeCut>, dim3(J.K), dim3(kRangeCut / kRowE), lds_padded_size(kRangeCut) * sizeof(float2),
                s, J);
}
void launch_czt_table(int M, int W, int F, double c0, double posa, float2* table, hipStream_t s) {
    ISAR_F_SWITCH(F, ISAR_LAUNCH(k_czt_table<FF>, dim3(W), dim3(CztCfg<FF>::TPF), lds_padded_size(FF) * sizeof(float2),
                                 s, M, W, c0, posa, table));
}
void launch_mtrc_rows(const Job& J, hipStream_t s) {
    ISAR_LAUNCH(k_mtrc_rows<kRangeCut>, dim3((J.M + kRowsPerBlock - 1) / kRowsPerBlock, J.K),
                dim3(kRowsPerBlock * kRangeCut / kRowE), rows_lds(), s, J);
}
void launch_mtrc_czt(const Job& J, hipStream_t s) {
    if (J.czt_threads == 512) {
        ISAR_F_SWITCH(J.F, ISAR_LAUNCH((k_mtrc_czt<FF, 512>), dim3(J.W / CztCfg<FF, 512>::CT, J.K), dim3(512),
                                       (czt_lds<FF, 512>(J.M)), s, J));
    } else {
        ISAR_F_SWITCH(J.F, ISAR_LAUNCH((k_mtrc_czt<FF, 256>), dim3(J.W / CztCfg<FF, 256>::CT, J.K), dim3(256),
                                       (czt_lds<FF, 256>(J.M)), s, J));
    }
}
void launch_col_window(const Job& J, hipStream_t s) {
    ISAR_PULSE_SWITCH2(J.M, ISAR_LAUNCH(k_col_window<MM>, dim3(J.W / ColCfg<MM>::CT, J.K), dim3(ColCfg<MM>::NT),
                                        col_lds<MM>(), s, J));
}
void launch_mtrc_image(const Job& J, hipStream_t s) {
    ISAR_LAUNCH(k_mtrc_image<kRangeCut>, dim3((J.M + kRowsPerBlock - 1) / kRowsPerBlock, J.K),
                dim3(kRowsPerBlock * kRangeCut / kRowE), rows_lds(), s, J);
}
void launch_post_nomtrc(const Job& J, hipStream_t s) {
    ISAR_PULSE_SWITCH2(J.M, ISAR_LAUNCH(k_post_nomtrc<MM>, dim3(J.W / ColCfg<MM>::CT, J.K), dim3(ColCfg<MM>::NT),
                                        col_lds<MM>(), s, J));
}

}  // namespace isar
