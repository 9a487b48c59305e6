// fft_debug.hip — batched standalone transforms of the device FFT, exposed for tests.
#include "../dev_common.hpp"
#include "../launchers.hpp"

namespace isar {

template <int N, int E, int DIR>
__global__ __launch_bounds__(N / E) void k_debug_fft(const float2* in, float2* out) {
    HIP_DYNAMIC_SHARED(float2, lds);
    constexpr int R0 = FftPlan<N, E>::R[0];
    constexpr int RL = FftPlan<N, E>::R[FftPlan<N, E>::NP - 1];
    const size_t base = static_cast<size_t>(blockIdx.x) * N;
    const int tid = threadIdx.x;
    float2 x[E];
#pragma unroll
    for (int e = 0; e < E; ++e) x[e] = in[base + fft_in_pos<N, E, R0>(tid, e)];
    fft<N, E, DIR>(x, lds, tid);
#pragma unroll
    for (int e = 0; e < E; ++e) out[base + fft_out_pos<N, E, RL>(tid, e)] = x[e];
}

// 256 / (N / E) transforms per workgroup, each on N / E threads of one wave with its own LDS
// buffer and its passes ordered by the wave alone (fft<..., true>): the arrangement of the
// autofocus, chirp-z and MTRC row transforms. Transforms past `batch` compute on zeros and store
// nothing.
template <int N, int E, int DIR>
__global__ __launch_bounds__(256) void k_debug_fft_grouped(const float2* in, float2* out, int batch) {
    HIP_DYNAMIC_SHARED(float2, lds);
    constexpr int TPF = N / E, G = 256 / TPF;
    constexpr int R0 = FftPlan<N, E>::R[0];
    constexpr int RL = FftPlan<N, E>::R[FftPlan<N, E>::NP - 1];
    const int g = threadIdx.x / TPF, lane = threadIdx.x % TPF;
    const int t = blockIdx.x * G + g;
    const bool valid = t < batch;
    const size_t base = static_cast<size_t>(valid ? t : 0) * N;
    float2 x[E];
#pragma unroll
    for (int e = 0; e < E; ++e) x[e] = valid ? in[base + fft_in_pos<N, E, R0>(lane, e)] : make_float2(0.0f, 0.0f);
    fft<N, E, DIR, true>(x, lds + g * lds_padded_size(N), lane);
    if (valid) {
#pragma unroll
        for (int e = 0; e < E; ++e) out[base + fft_out_pos<N, E, RL>(lane, e)] = x[e];
    }
}

namespace {
template <int N, int E>
void run_fft_grouped(int dir, const float2* in, float2* out, int batch, hipStream_t s) {
    constexpr int G = 256 / (N / E);
    const size_t lds = G * lds_padded_size(N) * sizeof(float2);
    const dim3 grid((batch + G - 1) / G);
    if (dir < 0) ISAR_LAUNCH((k_debug_fft_grouped<N, E, -1>), grid, dim3(256), lds, s, in, out, batch);
    else ISAR_LAUNCH((k_debug_fft_grouped<N, E, +1>), grid, dim3(256), lds, s, in, out, batch);
}

template <int N, int E>
void run_fft(int dir, const float2* in, float2* out, int batch, hipStream_t s) {
    const size_t lds = lds_padded_size(N) * sizeof(float2);
    if (dir < 0) ISAR_LAUNCH((k_debug_fft<N, E, -1>), dim3(batch), dim3(N / E), lds, s, in, out);
    else ISAR_LAUNCH((k_debug_fft<N, E, +1>), dim3(batch), dim3(N / E), lds, s, in, out);
}
}  // namespace

hipError_t configure_debug_kernels() {
    hipError_t err = hipSuccess;
#ifndef ISAR_EMU
    auto set = [&](const void* fn, size_t bytes) {
        hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(bytes));
        if (e != hipSuccess) {
            err = e;
            report_config_error(__FILE__, bytes);
        }
    };
    set(reinterpret_cast<const void*>(&k_debug_fft<16384, 16, -1>), lds_padded_size(16384) * sizeof(float2));
    set(reinterpret_cast<const void*>(&k_debug_fft<16384, 16, +1>), lds_padded_size(16384) * sizeof(float2));
    set(reinterpret_cast<const void*>(&k_debug_fft<8192, 16, -1>), lds_padded_size(8192) * sizeof(float2));
    set(reinterpret_cast<const void*>(&k_debug_fft<8192, 16, +1>), lds_padded_size(8192) * sizeof(float2));
#endif
    return err;
}

bool debug_fft(int n, int dir, const float2* in, float2* out, int batch, hipStream_t s) {
    switch (n) {
    case 64: run_fft<64, 16>(dir, in, out, batch, s); return true;
    case 128: run_fft<128, 16>(dir, in, out, batch, s); return true;
    case 192: run_fft<192, 12>(dir, in, out, batch, s); return true;
    case 256: run_fft<256, 16>(dir, in, out, batch, s); return true;
    case 384: run_fft<384, 12>(dir, in, out, batch, s); return true;
    case 512: run_fft<512, 16>(dir, in, out, batch, s); return true;
    case 1024: run_fft<1024, 16>(dir, in, out, batch, s); return true;
    case 2048: run_fft<2048, 16>(dir, in, out, batch, s); return true;
    case 4096: run_fft<4096, 16>(dir, in, out, batch, s); return true;
    case 8192: run_fft<8192, 16>(dir, in, out, batch, s); return true;
    case 16384: run_fft<16384, 16>(dir, in, out, batch, s); return true;
    default: return false;
    }
}

bool debug_fft_grouped(int n, int dir, const float2* in, float2* out, int batch, hipStream_t s) {
    switch (n) {
    case 64: run_fft_grouped<64, 16>(dir, in, out, batch, s); return true;
    case 128: run_fft_grouped<128, 16>(dir, in, out, batch, s); return true;
    case 192: run_fft_grouped<192, 12>(dir, in, out, batch, s); return true;
    case 256: run_fft_grouped<256, 16>(dir, in, out, batch, s); return true;
    case 384: run_fft_grouped<384, 12>(dir, in, out, batch, s); return true;
    case 512: run_fft_grouped<512, 16>(dir, in, out, batch, s); return true;
    case 1024: run_fft_grouped<1024, 16>(dir, in, out, batch, s); return true;
    default: return false;
    }
}

}  // namespace isar
