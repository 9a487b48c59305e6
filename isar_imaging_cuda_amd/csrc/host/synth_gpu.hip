// synth_gpu.hip — the frames of isar_synth_frame (synth.cpp) generated on the GPU, for training
// sets and benchmarks that need thousands of frames.
//
// The host draws each frame's scene, phase error and range walk with the CPU generator's seeded
// streams (synth_model.hpp). The device forms every echo sample as the CPU does, in FP64: the sum
// over scatterers of amp exp(j (phi0 + phi1 n)), each term advanced by a complex rotation from an
// anchor (every 16 samples here, every 256 on the CPU), times the quadratic term exp(j phi2 n^2).
// Noise comes from the same per-row SplitMix64 streams; SplitMix64 is counter-based (draw k of a
// stream is the mixer applied to its seed + (k + 1) * gamma), so every sample's draws are formed
// independently. Scaling to the ADC peak and int16 rounding follow the CPU generator. Agreement
// with the CPU frames is therefore exact except where FP64 rounding differences move a value
// across a float or integer rounding boundary.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../dev_common.hpp"
#include "pipeline.hpp"
#include "isar_synth.h"
#include "synth_model.hpp"

namespace {

constexpr double kPiD = 3.14159265358979323846;
constexpr double kLight = 3e8;
constexpr int kSpt = 16;                 // samples per thread
constexpr int kSigThreads = 64;          // one wave per block: 1024 samples of one row
constexpr uint64_t kGamma = 0x9E3779B97F4A7C15ULL;

struct SynthConsts {
    int M, N, S, K;
    double f_hi, K_chirp, fs, prf, T, cel, sel, v, a, omega;
    double nsig_scale;   // 1 / (10^(snr/10) * 2), or 0 without noise
    double peak_adc;
    int quantize, noisy;
};

__device__ __forceinline__ uint64_t splitmix(uint64_t state_after_add) {
    uint64_t z = state_after_add;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
__device__ __forceinline__ double uniform01(uint64_t seed, uint64_t k) {   // draw k of the stream
    return static_cast<double>(splitmix(seed + (k + 1) * kGamma) >> 11) * (1.0 / 9007199254740992.0);
}

// One block per (1024-sample span, pulse, frame). scat: per frame S x (x, y, z, amp re, amp im);
// pulse: per frame M x (walk, phase error). Writes the noise-free samples as floats and the
// span's FP64 power into rowpow[(f M + m) spans + span].
__global__ __launch_bounds__(kSigThreads) void k_synth_signal(SynthConsts c, const double* __restrict__ scat,
                                                            const double* __restrict__ pulse, float2* __restrict__ echo,
                                                            double* __restrict__ velocity, double* __restrict__ rowpow) {
    __shared__ double sp[6][kSigThreads];   // phi0, phi1, cos phi1, sin phi1, amp re, amp im
    const int span = blockIdx.x, m = blockIdx.y, f = blockIdx.z, t = threadIdx.x;
    const int spans = gridDim.x;
    const int n0 = span * kSigThreads * kSpt + t * kSpt;
    const double tm = m / c.prf;
    const double vm = c.v + c.a * tm;
    const double th = c.omega * (tm - c.T / 2);
    const double ct = cos(th), st = sin(th);
    const double walk = pulse[(static_cast<size_t>(f) * c.M + m) * 2];
    const double perr = pulse[(static_cast<size_t>(f) * c.M + m) * 2 + 1];
    const double* sc = scat + static_cast<size_t>(f) * c.S * 5;
    double ar[kSpt], ai[kSpt];
#pragma unroll
    for (int q = 0; q < kSpt; ++q) ar[q] = ai[q] = 0.0;
    for (int i0 = 0; i0 < c.S; i0 += kSigThreads) {
        __syncthreads();
        const int i = i0 + t;
        if (i < c.S) {
            const double* s = sc + static_cast<size_t>(i) * 5;
            const double dR = walk + (s[1] * ct - s[0] * st) * c.cel + s[2] * c.sel;
            const double phi0 = -(4 * kPiD / kLight) * c.f_hi * dR + perr;
            const double phi1 = (-(4 * kPiD / kLight) * c.f_hi * vm + (4 * kPiD * c.K_chirp / kLight) * dR) / c.fs;
            sp[0][t] = phi0;
            sp[1][t] = phi1;
            sp[2][t] = cos(phi1);
            sp[3][t] = sin(phi1);
            sp[4][t] = s[3];
            sp[5][t] = s[4];
        }
        __syncthreads();
        const int cnt = min(kSigThreads, c.S - i0);
        for (int j = 0; j < cnt; ++j) {
            const double a0 = sp[0][j] + sp[1][j] * n0;
            const double ca = cos(a0), sa = sin(a0);
            double zr = sp[4][j] * ca - sp[5][j] * sa, zi = sp[4][j] * sa + sp[5][j] * ca;
            const double rr = sp[2][j], ri = sp[3][j];
#pragma unroll
            for (int q = 0; q < kSpt; ++q) {
                ar[q] += zr;
                ai[q] += zi;
                const double nr = zr * rr - zi * ri;
                zi = zr * ri + zi * rr;
                zr = nr;
            }
        }
    }
    const double phi2 = (4 * kPiD * c.K_chirp / kLight) * vm / (c.fs * c.fs);
    double p = 0.0;
    float2* row = echo + (static_cast<size_t>(f) * c.M + m) * c.N;
#pragma unroll
    for (int q = 0; q < kSpt; ++q) {
        const int n = n0 + q;
        if (n < c.N) {
            const double a2 = phi2 * static_cast<double>(n) * n;
            const double c2 = cos(a2), s2 = sin(a2);
            const double xr = ar[q] * c2 - ai[q] * s2, xi = ar[q] * s2 + ai[q] * c2;
            p += xr * xr + xi * xi;
            row[n] = make_float2(static_cast<float>(xr), static_cast<float>(xi));
        }
    }
    p = isar::wave_sum_d(p);
    if (t == 0) rowpow[(static_cast<size_t>(f) * c.M + m) * spans + span] = p;
    if (t == 0 && span == 0) velocity[static_cast<size_t>(f) * c.M + m] = vm;
}

// Per frame: the noise level from the mean sample power (FP64), one block of 256 per frame.
__global__ __launch_bounds__(256) void k_synth_level(SynthConsts c, int spans, const double* __restrict__ rowpow,
                                                   double* __restrict__ nsig) {
    __shared__ double red[4];
    const int f = blockIdx.x, t = threadIdx.x;
    double s = 0.0;
    for (int i = t; i < c.M * spans; i += 256) s += rowpow[static_cast<size_t>(f) * c.M * spans + i];
    s = isar::wave_sum_d(s);
    if ((t & 63) == 0) red[t >> 6] = s;
    __syncthreads();
    if (t == 0) {
        const double psig = (red[0] + red[1] + red[2] + red[3]) / (static_cast<double>(c.M) * c.N);
        nsig[f] = (c.noisy && psig > 0) ? sqrt(psig * c.nsig_scale) : 0.0;
    }
}

// Noise: sample q of row m (q < 2N, re and im interleaved) gets nsig * gauss, gauss from draws
// 2q and 2q + 1 of the row's stream (Box-Muller as the CPU generator); then |value| maxima per
// block for the ADC scaling.
__global__ __launch_bounds__(256) void k_synth_noise(SynthConsts c, const uint64_t* __restrict__ row_seed,
                                                   const double* __restrict__ nsig, float* __restrict__ echo,
                                                   float* __restrict__ peak_part) {
    __shared__ float red[4];
    const int f = blockIdx.y, t = threadIdx.x;
    const long long i = static_cast<long long>(blockIdx.x) * 256 + t;
    const long long total = 2LL * c.M * c.N;
    float mx = 0.0f;
    if (i < total) {
        float* e = echo + static_cast<size_t>(f) * total + i;
        float v = *e;
        const double ns = nsig[f];
        if (ns > 0.0) {
            const int m = static_cast<int>(i / (2LL * c.N));
            const long long q = i % (2LL * c.N);
            const uint64_t seed = row_seed[static_cast<size_t>(f) * c.M + m];
            double u1 = uniform01(seed, 2 * q), u2 = uniform01(seed, 2 * q + 1);
            if (u1 < 1e-300) u1 = 1e-300;
            const double g = sqrt(-2.0 * log(u1)) * cos(2.0 * kPiD * u2);
            v += static_cast<float>(ns * g);
            *e = v;
        }
        mx = fabsf(v);
    }
    for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o, 64));
    if ((t & 63) == 0) red[t >> 6] = mx;
    __syncthreads();
    if (t == 0) peak_part[static_cast<size_t>(f) * gridDim.x + blockIdx.x] = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
}

// ADC scaling to the frame's peak (and int16 rounding), as the CPU generator does it.
__global__ __launch_bounds__(256) void k_synth_scale(SynthConsts c, int parts, const float* __restrict__ peak_part,
                                                   float* __restrict__ echo) {
    __shared__ float red[4];
    __shared__ double gain;
    const int f = blockIdx.y, t = threadIdx.x;
    float mx = 0.0f;
    for (int k = t; k < parts; k += 256) mx = fmaxf(mx, peak_part[static_cast<size_t>(f) * parts + k]);
    for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o, 64));
    if ((t & 63) == 0) red[t >> 6] = mx;
    __syncthreads();
    if (t == 0) {
        const float pk = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
        gain = pk > 0 ? c.peak_adc / pk : 1.0;
    }
    __syncthreads();
    const long long total = 2LL * c.M * c.N;
    for (long long i = static_cast<long long>(blockIdx.x) * 256 + t; i < total; i += static_cast<long long>(gridDim.x) * 256) {
        float* e = echo + static_cast<size_t>(f) * total + i;
        double v = static_cast<double>(*e) * gain;
        if (c.quantize) v = fmax(-32768.0, fmin(32767.0, rint(v)));
        *e = static_cast<float>(v);
    }
}

}  // namespace

extern "C" {

int isar_synth_frames_device(const isar_scene_config* cfg, int frames, const uint64_t* seeds, void* d_echo,
                             double* d_velocity) {
    using isar::fail;
    if (!cfg || frames < 1 || !seeds || !d_echo || !d_velocity || cfg->echo_num < 1 || cfg->range_num < 1 ||
        cfg->n_scatterers < 0)
        return fail(ISAR_ERR_INVALID, "isar_synth_frames_device: bad arguments");
    const int M = cfg->echo_num, N = cfg->range_num, S = cfg->n_scatterers, K = frames;
    if (2LL * M * N >= (1LL << 31)) return fail(ISAR_ERR_INVALID, "isar_synth_frames_device: frame too large");
    SynthConsts c{};
    c.M = M; c.N = N; c.S = S; c.K = K;
    c.f_hi = cfg->fc + cfg->band_width / 2;
    c.K_chirp = cfg->band_width / cfg->tp;
    c.fs = cfg->fs; c.prf = cfg->prf; c.T = M / cfg->prf;
    c.cel = std::cos(cfg->elevation); c.sel = std::sin(cfg->elevation);
    c.v = cfg->v; c.a = cfg->a; c.omega = cfg->omega;
    c.noisy = cfg->snr_db < 200 ? 1 : 0;
    c.nsig_scale = c.noisy ? 1.0 / (std::pow(10.0, cfg->snr_db / 10.0) * 2.0) : 0.0;
    c.peak_adc = cfg->peak_adc;
    c.quantize = cfg->quantize;
    // the seeded draws of every frame, on the host
    std::vector<double> scat(static_cast<size_t>(K) * std::max(S, 1) * 5), pulse(static_cast<size_t>(K) * M * 2);
    std::vector<uint64_t> rseed(static_cast<size_t>(K) * M);
    for (int f = 0; f < K; ++f) {
        isar_scene_config cf = *cfg;
        cf.seed = seeds[f];
        isar_synth_model_t md;
        isar_synth_model(&cf, &md);
        for (int i = 0; i < S; ++i) {
            double* s = &scat[(static_cast<size_t>(f) * S + i) * 5];
            s[0] = md.sx[i]; s[1] = md.sy[i]; s[2] = md.sz[i]; s[3] = md.amp[i].real(); s[4] = md.amp[i].imag();
        }
        for (int m = 0; m < M; ++m) {
            pulse[(static_cast<size_t>(f) * M + m) * 2] = md.walk[m];
            pulse[(static_cast<size_t>(f) * M + m) * 2 + 1] = md.perr[m];
            rseed[static_cast<size_t>(f) * M + m] = isar_synth_noise_seed(seeds[f], m);
        }
    }
    const int spans = (N + kSigThreads * kSpt - 1) / (kSigThreads * kSpt);
    const long long total = 2LL * M * N;
    const int parts = static_cast<int>((total + 255) / 256);
    double *d_scat = nullptr, *d_pulse = nullptr, *d_rowpow = nullptr, *d_nsig = nullptr;
    uint64_t* d_rseed = nullptr;
    float* d_peak = nullptr;
    bool ok = hipMalloc(&d_scat, scat.size() * sizeof(double)) == hipSuccess &&
              hipMalloc(&d_pulse, pulse.size() * sizeof(double)) == hipSuccess &&
              hipMalloc(&d_rseed, rseed.size() * sizeof(uint64_t)) == hipSuccess &&
              hipMalloc(&d_rowpow, static_cast<size_t>(K) * M * spans * sizeof(double)) == hipSuccess &&
              hipMalloc(&d_nsig, K * sizeof(double)) == hipSuccess &&
              hipMalloc(&d_peak, static_cast<size_t>(K) * parts * sizeof(float)) == hipSuccess;
    hipError_t e = ok ? hipSuccess : hipErrorOutOfMemory;
    if (e == hipSuccess) e = hipMemcpy(d_scat, scat.data(), scat.size() * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_pulse, pulse.data(), pulse.size() * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d_rseed, rseed.data(), rseed.size() * sizeof(uint64_t), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_synth_signal, dim3(spans, M, K), dim3(kSigThreads), 0, nullptr, c, d_scat, d_pulse,
                           static_cast<float2*>(d_echo), d_velocity, d_rowpow);
        hipLaunchKernelGGL(k_synth_level, dim3(K), dim3(256), 0, nullptr, c, spans, d_rowpow, d_nsig);
        hipLaunchKernelGGL(k_synth_noise, dim3(parts, K), dim3(256), 0, nullptr, c, d_rseed, d_nsig,
                           static_cast<float*>(d_echo), d_peak);
        if (c.peak_adc > 0)
            hipLaunchKernelGGL(k_synth_scale, dim3(std::min(parts, 1024), K), dim3(256), 0, nullptr, c, parts, d_peak,
                               static_cast<float*>(d_echo));
        e = hipGetLastError();
        if (e == hipSuccess) e = hipDeviceSynchronize();
    }
    for (void* p : {static_cast<void*>(d_scat), static_cast<void*>(d_pulse), static_cast<void*>(d_rseed),
                    static_cast<void*>(d_rowpow), static_cast<void*>(d_nsig), static_cast<void*>(d_peak)})
        if (p) (void)hipFree(p);
    if (e != hipSuccess) return fail(ISAR_ERR_DEVICE, std::string("isar_synth_frames_device: ") + hipGetErrorString(e));
    return ISAR_OK;
}

}  // extern "C"
