// ifds.hip — IFDS pulse compression (common.cu:583-604, 635-746) for gfx950: one workgroup per
// echo forms the velocity-dependent reference chirp, matched-filters the raw echo with three
// NFFT-point transforms in registers and LDS, and cuts RANGE_NUM_IFDS_PC samples around the
// strongest return (cutRangeProfile, common.cu:494-518).
#include "../dev_common.hpp"
#include "../launchers.hpp"

namespace isar {

namespace {
constexpr int kPcE = 16;
template <int NF>
__device__ __forceinline__ int pc_in(int tid, int e) { return fft_in_pos<NF, kPcE, FftPlan<NF, kPcE>::R[0]>(tid, e); }
template <int NF>
__device__ __forceinline__ int pc_out(int tid, int e) {
    return fft_out_pos<NF, kPcE, FftPlan<NF, kPcE>::R[FftPlan<NF, kPcE>::NP - 1]>(tid, e);
}
}  // namespace

// iq: echoes x len int16 I/Q pairs; out: echoes x cut complex. Phases of the reference chirp in FP64.
template <int NF>
__global__ __launch_bounds__(NF / kPcE) void k_pulse_compress(PcJob P) {
    HIP_DYNAMIC_SHARED(float2, lds);
    constexpr int NT = NF / kPcE;
    float* redv = reinterpret_cast<float*>(lds + lds_padded_size(NF));
    int* redi = reinterpret_cast<int*>(redv + 32);
    const int echo = blockIdx.x, tid = threadIdx.x;
    // reference chirp (genTkPulseCompression / genRefPulseCompression, genHammingVec)
    const double v2 = 2.0 * P.velocity[echo] / 3e8;
    const double v2_ = 1.0 - v2;
    const double c1 = -2.0 * 3.14159265358979323846 * P.fc * v2 / v2_;
    const double c2 = 3.14159265358979323846 * P.band_width / P.tp;
    float2 r[kPcE];
#pragma unroll
    for (int e = 0; e < kPcE; ++e) {
        const int k = pc_in<NF>(tid, e);
        r[e] = make_float2(0.0f, 0.0f);
        if (k < P.S) {
            const double t = v2_ * (-P.tp / 2.0 + k * P.tp / (static_cast<double>(P.S) - 1.0));
            const double ph = (c1 + c2 * t) * t;
            r[e] = cscale(expjpi_reduced(ph * 0.31830988618379067154), hamming(k, P.S));
        }
    }
    fft<NF, kPcE, -1>(r, lds, tid);
    // the raw echo, zero-padded to NF
    const int16_t* src = P.iq + static_cast<size_t>(echo) * P.len * 2;
    float2 x[kPcE];
#pragma unroll
    for (int e = 0; e < kPcE; ++e) {
        const int k = pc_in<NF>(tid, e);
        x[e] = k < P.len ? make_float2(static_cast<float>(src[2 * k]), static_cast<float>(src[2 * k + 1]))
                         : make_float2(0.0f, 0.0f);
    }
    fft<NF, kPcE, -1>(x, lds, tid);
#pragma unroll
    for (int e = 0; e < kPcE; ++e) x[e] = cconjmul(r[e], x[e]);   // elementwiseMultiplyConjA
    relayout_out_to_in<NF, kPcE>(x, lds, tid);
    fft<NF, kPcE, +1>(x, lds, tid);
    const float inv = 1.0f / NF;
    float best = -1.0f;
    int bi = 0;
#pragma unroll
    for (int e = 0; e < kPcE; ++e) {
        x[e] = cscale(x[e], inv);
        const int k = pc_out<NF>(tid, e);
        const float a = fabsf(x[e].x) + fabsf(x[e].y);   // cublasIcamax: |Re| + |Im|, first index
        if (a > best || (a == best && k < bi)) { best = a; bi = k; }
    }
    block_argmax<NT>(best, bi, redv, redi);
    int off = max(bi - P.cut / 2, 0);
    off = min(off, NF - P.cut);
    float2* dst = P.out + static_cast<size_t>(echo) * P.cut;
#pragma unroll
    for (int e = 0; e < kPcE; ++e) {
        const int k = pc_out<NF>(tid, e) - off;
        if (k >= 0 && k < P.cut) dst[k] = x[e];
    }
    if (tid == 0 && P.offsets) P.offsets[echo] = off;
}

namespace {
template <int NF>
size_t pc_lds() { return lds_padded_size(NF) * sizeof(float2) + 64 * sizeof(float); }
}  // namespace

bool pulse_compress_supported(int nfft) { return nfft == 4096 || nfft == 8192 || nfft == 16384; }

hipError_t configure_ifds_kernels() {
    hipError_t err = hipSuccess;
#ifndef ISAR_EMU
    auto set = [&](const void* fn, size_t bytes) {
        hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(bytes));
        if (e != hipSuccess) {
            err = e;
            report_config_error(__FILE__, bytes);
        }
    };
    set(reinterpret_cast<const void*>(&k_pulse_compress<4096>), pc_lds<4096>());
    set(reinterpret_cast<const void*>(&k_pulse_compress<8192>), pc_lds<8192>());
    set(reinterpret_cast<const void*>(&k_pulse_compress<16384>), pc_lds<16384>());
#endif
    return err;
}

void launch_pulse_compress(const PcJob& P, int nfft, hipStream_t s) {
    switch (nfft) {
    case 4096: ISAR_LAUNCH(k_pulse_compress<4096>, dim3(P.echoes), dim3(4096 / kPcE), pc_lds<4096>(), s, P); break;
    case 8192: ISAR_LAUNCH(k_pulse_compress<8192>, dim3(P.echoes), dim3(8192 / kPcE), pc_lds<8192>(), s, P); break;
    case 16384: ISAR_LAUNCH(k_pulse_compress<16384>, dim3(P.echoes), dim3(16384 / kPcE), pc_lds<16384>(), s, P); break;
    default: break;
    }
}

}  // namespace isar
