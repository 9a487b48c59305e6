// ra.hip — range alignment (range_alignment.cu:4-201) for gfx950.
//
// The reference inverse-transforms all M rows at every merge level. Here each row's spectrum is
// re-read from the raw echo with high-speed compensation, Hamming window, (-1)^n centring and
// its cumulative lag ramp applied on the fly, and only rows whose lag changed are transformed
// again. Group energies sum_{rows} |IFFT(row)|^2 / N^2 are kept per group head row and merged
// level by level:  E^{l+1}_g = E^l_{2g} + E~^l_{2g+1}  (E~ = second group after its shift).
#include "../dev_common.hpp"
#include "../launchers.hpp"

namespace isar {

namespace {

constexpr int kE = 16;   // elements per thread for the row transforms

template <int N>
__device__ __forceinline__ int in_pos(int tid, int e) { return fft_in_pos<N, kE, FftPlan<N, kE>::R[0]>(tid, e); }
template <int N>
__device__ __forceinline__ int out_pos(int tid, int e) {
    return fft_out_pos<N, kE, FftPlan<N, kE>::R[FftPlan<N, kE>::NP - 1]>(tid, e);
}

// Spectrum of row m of frame f with HPC (FP64 phase), window, centring and the lag ramp
// exp(-j 2 pi n lag / N) applied, in the first-pass layout.
template <int N>
__device__ __forceinline__ void load_row(float2 (&x)[kE], const Job& J, int f, int m, double lag, int tid) {
    const size_t off = (static_cast<size_t>(f) * J.M + m) * N;
    const float2* row = (J.data_type == 1 ? J.pre : J.echo) + off;
    const bool hpc = J.if_hpc && J.data_type == 2;
    const double vk = hpc ? J.hpc_k * J.velocity[static_cast<size_t>(f) * J.M + m] : 0.0;
    const bool hrrp_window = J.if_hrrp && J.data_type == 2;
    const double lk = -2.0 * lag / N;
    const bool phase = (vk != 0.0) || (lag != 0.0);
#pragma unroll
    for (int e = 0; e < kE; ++e) {
        const int n = in_pos<N>(tid, e);
        float2 v = row[n];
        if (phase) {
            const double dn = static_cast<double>(n);
            v = cmul(v, expjpi_reduced(vk * dn * dn + lk * dn));
        }
        float w = hamming(n, N);
        if (hrrp_window) w *= w;
        if (n & 1) w = -w;
        x[e] = cscale(v, w);
    }
}

// Energy profile of group g at merge stride s (level l = log2 s), in the first-pass layout.
template <int N>
__device__ __forceinline__ void group_energy(float (&a)[kE], const Job& J, int f, int g, int s, int tid) {
    const int M = J.M;
    const float* base = J.egrp + (static_cast<size_t>(f) * M + static_cast<size_t>(g) * s) * N;
#pragma unroll
    for (int e = 0; e < kE; ++e) a[e] = base[in_pos<N>(tid, e)];
    if (s > 1) {
        const int s2 = s / 2;
        const int r0 = (2 * g + 1) * s2;
        if (r0 < M) {
            const int r1 = min(r0 + s2, M);
            const int R = min(s2, kRaChunkRows);
            for (int h = r0; h < r1; h += R) {
                const float* pp = J.epart + (static_cast<size_t>(f) * M + h) * N;
#pragma unroll
                for (int e = 0; e < kE; ++e) a[e] += pp[in_pos<N>(tid, e)];
            }
        }
    }
}

}  // namespace

// IFDS pre-stage (isar_main.cu:15-27): ifftshift, IFFT / N, ifftshift of every row.
template <int N>
__global__ __launch_bounds__(N / kE) void k_ifds_pre(Job J) {
    HIP_DYNAMIC_SHARED(float2, lds);
    const int m = blockIdx.x, f = blockIdx.y, tid = threadIdx.x;
    const size_t off = (static_cast<size_t>(f) * J.M + m) * N;
    float2 x[kE];
#pragma unroll
    for (int e = 0; e < kE; ++e) x[e] = J.echo[off + ((in_pos<N>(tid, e) + N / 2) & (N - 1))];
    fft<N, kE, +1>(x, lds, tid);
    const float inv = 1.0f / N;
#pragma unroll
    for (int e = 0; e < kE; ++e) J.pre[off + ((out_pos<N>(tid, e) + N / 2) & (N - 1))] = cscale(x[e], inv);
}

// Level-0 energies of every row (range_alignment.cu:13-18 and :38-42 at stride 1).
template <int N>
__global__ __launch_bounds__(N / kE) void k_ra_init(Job J) {
    HIP_DYNAMIC_SHARED(float2, lds);
    const int m = blockIdx.x, f = blockIdx.y, tid = threadIdx.x;
    float2 x[kE];
    load_row<N>(x, J, f, m, 0.0, tid);
    fft<N, kE, +1>(x, lds, tid);
    const float inv = 1.0f / (static_cast<float>(N) * static_cast<float>(N));
    float* dst = J.egrp + (static_cast<size_t>(f) * J.M + m) * N;
#pragma unroll
    for (int e = 0; e < kE; ++e) dst[out_pos<N>(tid, e)] = cabs2(x[e]) * inv;
    if (tid == 0) J.lagsum[static_cast<size_t>(f) * J.M + m] = 0.0;
}

// One merge level: correlation of the two group profiles of pair p, argmax of |c| after
// ifftshift (lowest index on ties), parabolic refinement and lag (range_alignment.cu:45-61,
// 108-168). The real correlation comes from one complex FFT of gA + j gB.
template <int N>
__global__ __launch_bounds__(N / kE) void k_ra_corr(Job J, int s, int pair_base) {
    HIP_DYNAMIC_SHARED(float2, lds);
    constexpr int NT = N / kE;
    float* redv = reinterpret_cast<float*>(lds + lds_padded_size(N));
    int* redi = reinterpret_cast<int*>(redv + 16);
    const int p = blockIdx.x, f = blockIdx.y, tid = threadIdx.x;
    const int M = J.M;
    const int ngroups = (M + s - 1) / s, npairs = ngroups / 2;
    const bool unpaired = (p == npairs);
    const int gA = unpaired ? ngroups - 1 : 2 * p;
    float a[kE], b[kE];
    group_energy<N>(a, J, f, gA, s, tid);
    {   // keep the first group's energy for the next level (same head row, same thread)
        float* dst = J.egrp + (static_cast<size_t>(f) * M + static_cast<size_t>(gA) * s) * N;
#pragma unroll
        for (int e = 0; e < kE; ++e) dst[in_pos<N>(tid, e)] = a[e];
    }
    if (unpaired) return;
    group_energy<N>(b, J, f, gA + 1, s, tid);
    float2 x[kE];
#pragma unroll
    for (int e = 0; e < kE; ++e) x[e] = make_float2(sqrtf(a[e]), sqrtf(b[e]));
    fft<N, kE, -1>(x, lds, tid);
#pragma unroll
    for (int e = 0; e < kE; ++e) lds[lds_pad(out_pos<N>(tid, e))] = x[e];
    __syncthreads();
#pragma unroll
    for (int e = 0; e < kE; ++e) {
        const int k = in_pos<N>(tid, e);
        const float2 zk = lds[lds_pad(k)], zn = lds[lds_pad((N - k) & (N - 1))];
        const float2 ga = make_float2(0.5f * (zk.x + zn.x), 0.5f * (zk.y - zn.y));
        const float2 gb = make_float2(0.5f * (zk.y + zn.y), -0.5f * (zk.x - zn.x));
        x[e] = cmulc(ga, gb);
    }
    __syncthreads();
    fft<N, kE, +1>(x, lds, tid);
    float best = -1.0f;
    int bi = 0;
#pragma unroll
    for (int e = 0; e < kE; ++e) {
        const int i = (out_pos<N>(tid, e) + N / 2) & (N - 1);
        const float v = fabsf(x[e].x);
        if (v > best || (v == best && i < bi)) { best = v; bi = i; }
    }
    float* cl = reinterpret_cast<float*>(lds);
#pragma unroll
    for (int e = 0; e < kE; ++e) cl[(out_pos<N>(tid, e) + N / 2) & (N - 1)] = x[e].x;
    block_argmax<NT>(best, bi, redv, redi);
    if (tid == 0) {
        float delta = 0.0f;
        if (bi > 0 && bi < N - 1) {
            const float f1 = cl[bi - 1], f2 = cl[bi], f3 = cl[bi + 1];
            const float fa = (f1 + f3 - 2.0f * f2) / 2.0f, fb = (f3 - f1) / 2.0f;
            if (fa != 0.0f) delta = -fb / (2.0f * fa);
            if (!(delta == delta) || fabsf(delta) > 1e30f) delta = 0.0f;
        }
        const size_t idx = static_cast<size_t>(f) * (M - 1) + pair_base + p;
        J.lags[idx] = static_cast<float>(bi) + delta - static_cast<float>(N) / 2.0f;
        J.imax[idx] = bi;
    }
}

// Shift the rows of the second group of pair p by the pair's lag (range_alignment.cu:66-69)
// and accumulate their new energies; each workgroup handles up to kRaChunkRows rows.
template <int N>
__global__ __launch_bounds__(N / kE) void k_ra_upd(Job J, int s, int pair_base) {
    HIP_DYNAMIC_SHARED(float2, lds);
    const int b = blockIdx.x, f = blockIdx.y, tid = threadIdx.x;
    const int M = J.M;
    const int R = min(s, kRaChunkRows);
    const int cpp = (s + R - 1) / R;
    const int p = b / cpp, i = b % cpp;
    const int gend = min((2 * p + 2) * s, M);
    const int r0 = (2 * p + 1) * s + i * R;
    const int r1 = min(r0 + R, gend);
    if (r0 >= r1) return;
    const float lag = J.lags[static_cast<size_t>(f) * (M - 1) + pair_base + p];
    const float inv = 1.0f / (static_cast<float>(N) * static_cast<float>(N));
    float acc[kE];
#pragma unroll
    for (int e = 0; e < kE; ++e) acc[e] = 0.0f;
    float2 x[kE];
    for (int m = r0; m < r1; ++m) {
        const double L = J.lagsum[static_cast<size_t>(f) * M + m] + static_cast<double>(lag);
        load_row<N>(x, J, f, m, L, tid);
        fft<N, kE, +1>(x, lds, tid);
#pragma unroll
        for (int e = 0; e < kE; ++e) acc[e] += cabs2(x[e]) * inv;
    }
    __syncthreads();
    if (tid < r1 - r0) J.lagsum[static_cast<size_t>(f) * M + r0 + tid] += static_cast<double>(lag);
    float* dst = J.epart + (static_cast<size_t>(f) * M + r0) * N;
#pragma unroll
    for (int e = 0; e < kE; ++e) dst[out_pos<N>(tid, e)] = acc[e];
}

// Final pass (range_alignment.cu:76-77): aligned range profiles IFFT(S) / N, plus the column
// sums of |S| and the maximum |S| per row chunk for HRRP centring (range_alignment.cu:381-399).
template <int N>
__global__ __launch_bounds__(N / kE) void k_ra_final(Job J) {
    HIP_DYNAMIC_SHARED(float2, lds);
    constexpr int NT = N / kE;
    float* red = reinterpret_cast<float*>(lds + lds_padded_size(N));
    const int c = blockIdx.x, f = blockIdx.y, tid = threadIdx.x;
    const int M = J.M;
    const int nchunks = (M + kRaFinalRows - 1) / kRaFinalRows;
    const int r0 = c * kRaFinalRows, r1 = min(r0 + kRaFinalRows, M);
    const float inv = 1.0f / N;
    float acc[kE];
    float mx = 0.0f;
#pragma unroll
    for (int e = 0; e < kE; ++e) acc[e] = 0.0f;
    float2 x[kE];
    for (int m = r0; m < r1; ++m) {
        load_row<N>(x, J, f, m, J.lagsum[static_cast<size_t>(f) * M + m], tid);
        fft<N, kE, +1>(x, lds, tid);
        float2* dst = J.st + (static_cast<size_t>(f) * M + m) * N;
#pragma unroll
        for (int e = 0; e < kE; ++e) {
            const float2 v = cscale(x[e], inv);
            dst[out_pos<N>(tid, e)] = v;
            const float a = cabsf_(v);
            acc[e] += a;
            mx = fmaxf(mx, a);
        }
    }
    float* ap = J.arp_part + (static_cast<size_t>(f) * nchunks + c) * N;
#pragma unroll
    for (int e = 0; e < kE; ++e) ap[out_pos<N>(tid, e)] = acc[e];
    mx = block_max<NT>(mx, red);
    if (tid == 0) J.max_part[static_cast<size_t>(f) * nchunks + c] = mx;
}

// ------------------------------------------------------------------------------------------
// host launchers
// ------------------------------------------------------------------------------------------
namespace {
template <int N>
size_t ra_lds() { return lds_padded_size(N) * sizeof(float2) + 32 * sizeof(float); }

#define ISAR_RANGE_SWITCH(N, BODY)                     \
    switch (N) {                                       \
    case 1024: { constexpr int NN = 1024; BODY; } break;   \
    case 2048: { constexpr int NN = 2048; BODY; } break;   \
    case 4096: { constexpr int NN = 4096; BODY; } break;   \
    case 8192: { constexpr int NN = 8192; BODY; } break;   \
    case 16384: { constexpr int NN = 16384; BODY; } break; \
    default: break;                                    \
    }
}  // namespace

bool range_size_supported(int N) {
    return N == 1024 || N == 2048 || N == 4096 || N == 8192 || N == 16384;
}

hipError_t configure_ra_kernels() {
    hipError_t err = hipSuccess;
#ifndef ISAR_EMU
    auto set = [&](const void* fn, size_t bytes) {
        hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(bytes));
        if (e != hipSuccess) {
            err = e;
            report_config_error(__FILE__, bytes);
        }
    };
#define ISAR_SET_RA(NN)                                                        \
    set(reinterpret_cast<const void*>(&k_ifds_pre<NN>), ra_lds<NN>());         \
    set(reinterpret_cast<const void*>(&k_ra_init<NN>), ra_lds<NN>());          \
    set(reinterpret_cast<const void*>(&k_ra_corr<NN>), ra_lds<NN>());          \
    set(reinterpret_cast<const void*>(&k_ra_upd<NN>), ra_lds<NN>());           \
    set(reinterpret_cast<const void*>(&k_ra_final<NN>), ra_lds<NN>());
    ISAR_SET_RA(1024) ISAR_SET_RA(2048) ISAR_SET_RA(4096) ISAR_SET_RA(8192) ISAR_SET_RA(16384)
#undef ISAR_SET_RA
#endif
    return err;
}

void launch_ifds_pre(const Job& J, hipStream_t s) {
    ISAR_RANGE_SWITCH(J.N, ISAR_LAUNCH(k_ifds_pre<NN>, dim3(J.M, J.K), dim3(NN / kE), ra_lds<NN>(), s, J));
}
void launch_ra_init(const Job& J, hipStream_t s) {
    ISAR_RANGE_SWITCH(J.N, ISAR_LAUNCH(k_ra_init<NN>, dim3(J.M, J.K), dim3(NN / kE), ra_lds<NN>(), s, J));
}
void launch_ra_corr(const Job& J, int stride, int pair_base, int nblocks, hipStream_t s) {
    ISAR_RANGE_SWITCH(J.N, ISAR_LAUNCH(k_ra_corr<NN>, dim3(nblocks, J.K), dim3(NN / kE), ra_lds<NN>(), s, J,
                                       stride, pair_base));
}
void launch_ra_upd(const Job& J, int stride, int pair_base, int nblocks, hipStream_t s) {
    ISAR_RANGE_SWITCH(J.N, ISAR_LAUNCH(k_ra_upd<NN>, dim3(nblocks, J.K), dim3(NN / kE), ra_lds<NN>(), s, J,
                                       stride, pair_base));
}
void launch_ra_final(const Job& J, hipStream_t s) {
    const int nchunks = (J.M + kRaFinalRows - 1) / kRaFinalRows;
    ISAR_RANGE_SWITCH(J.N, ISAR_LAUNCH(k_ra_final<NN>, dim3(nchunks, J.K), dim3(NN / kE), ra_lds<NN>(), s, J));
}

}  // namespace isar
