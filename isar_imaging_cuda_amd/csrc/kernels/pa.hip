// pa.hip — phase adjustment (phase_adjustment.cu:4-41, 167-346) for gfx950:
// Doppler-centroid tracking, column energy/entropy statistics, bin selection and the
// fast minimum-entropy autofocus with the selected bins kept resident per frame.
#include <algorithm>

#include "../dev_common.hpp"
#include "../launchers.hpp"

namespace isar {

namespace {

// phi[0] = 0, phi[m+1] = exp(-j theta_m), theta = inclusive scan of arg(r_m) (accumulated in
// FP64 and reduced before the FP32 sincos). r: M-1 adjacent-pulse products (LDS or global).
template <int NT>
__device__ void doppler_phase(const float2* r, int M, float2* phi, double* scratch, double* red) {
    for (int m = threadIdx.x; m < M - 1; m += NT) scratch[m] = static_cast<double>(atan2f(r[m].y, r[m].x));
    __syncthreads();
    block_inclusive_scan_d<NT>(scratch, M - 1, red);
    for (int m = threadIdx.x; m < M; m += NT) {
        if (m == 0) phi[0] = make_float2(0.0f, 0.0f);   // phase_adjustment.cu:29
        else phi[m] = expjpi_reduced(-scratch[m - 1] * 0.31830988618379067154);
    }
    __syncthreads();
}

// Ascending bitonic sort of NT (key, tag) pairs, one per thread, ties ordered by tag (stable for
// unique tags), NaN keys last: stage (k, j) orders each pair i < i ^ j ascending where i & k == 0,
// descending otherwise. The stages whose partner lies in the same wave (j < 64) exchange through
// shuffles in registers; only those across waves go through LDS, with a barrier each (6 of the 45
// stages at 512). key and tag are in LDS, filled before a barrier; every thread calls it.
template <int NT>
__device__ void bitonic_sort_kv(float* key, int* tag) {
    static_assert(NT >= 64 && (NT & (NT - 1)) == 0, "one wave or more, a power of two");
    const int t = threadIdx.x;
    auto greater = [](float a, float b, int ta, int tb) { return (a > b) || (a == b && ta > tb) || (a != a && b == b); };
    float kv = 0.0f;
    int tv = 0;
    bool in_regs = false;
    for (int k = 2; k <= NT; k <<= 1) {
        int j = k >> 1;
        for (; j >= 64; j >>= 1) {   // partner in another wave
            if (in_regs) {
                key[t] = kv;
                tag[t] = tv;
                __syncthreads();
                in_regs = false;
            }
            const int l = t ^ j;
            if (l > t) {
                const float a = key[t], b = key[l];
                const int ta = tag[t], tb = tag[l];
                if (greater(a, b, ta, tb) == ((t & k) == 0)) { key[t] = b; key[l] = a; tag[t] = tb; tag[l] = ta; }
            }
            __syncthreads();
        }
        if (!in_regs) {   // after the last LDS stage's barrier (or the caller's)
            kv = key[t];
            tv = tag[t];
            in_regs = true;
        }
        for (; j >= 1; j >>= 1) {   // partner lane t ^ j of the same wave
            const float pk = __shfl_xor(kv, j, 64);
            const int pt = __shfl_xor(tv, j, 64);
            const bool lower = (t & j) == 0;
            const bool gt = lower ? greater(kv, pk, tv, pt) : greater(pk, kv, pt, tv);
            if (gt == ((t & k) == 0)) {
                kv = pk;
                tv = pt;
            }
        }
    }
    key[t] = kv;   // the last stage (j = 1) left the pairs in registers
    tag[t] = tv;
    __syncthreads();
}

}  // namespace

// Doppler-centroid phase of the cut (dopplerTracking, phase_adjustment.cu:4-41) from the
// adjacent-pulse products k_cut formed: once per frame, for every column tile of k_pa_colfft.
template <int M>
__global__ __launch_bounds__(256) void k_pa_doppler(Job J) {
    HIP_DYNAMIC_SHARED(double, scr);   // M (scan) + 256 (scan partials)
    const int f = blockIdx.x;
    doppler_phase<256>(J.rowdop + static_cast<size_t>(f) * M, M, J.dopphi + static_cast<size_t>(f) * M, scr,
                       scr + M);
}

// Per column tile: Doppler-tracked image column FFT, energy E_col, entropy H_col and the
// column maxima of |X| (phase_adjustment.cu:175-246).
template <int M>
__global__ __launch_bounds__(256) void k_pa_colfft(Job J) {
    using C = ColCfg<M>;
    constexpr int NT = C::NT, CT = C::CT, TPF = C::TPF, E = C::E;
    HIP_DYNAMIC_SHARED(float2, sm);
    float2* phi = sm;                                    // M
    float2* work = phi + M;                              // tile / transform buffers
    const int tile = xcd_tile(blockIdx.x, gridDim.x), f = blockIdx.y, tid = threadIdx.x;
    const int W = J.W;
    const int g = tid / TPF, lane = tid % TPF;
    const int n0 = tile * CT, n = n0 + g;

    for (int m = tid; m < M; m += NT) phi[m] = J.dopphi[static_cast<size_t>(f) * M + m];

    const float2* X = J.x + static_cast<size_t>(f) * M * W;
    for (int idx = tid; idx < M * CT; idx += NT) {
        const int m = idx / CT, c = idx % CT;
        work[m * (CT + 1) + c] = X[static_cast<size_t>(m) * W + n0 + c];
    }
    __syncthreads();
    float2 x[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int pos = C::in(lane, e);
        x[e] = cmul(phi[pos], work[pos * (CT + 1) + g]);
    }
    __syncthreads();
    float2* buf = work + g * lds_padded_size(M);
    fft<M, E, -1, true>(x, buf, lane);
    float p2[E];
    float esum = 0.0f;
#pragma unroll
    for (int e = 0; e < E; ++e) { p2[e] = cabs2(x[e]); esum += p2[e]; }
    const float Ecol = group_sum<TPF>(esum);
    float h = 0.0f;
    if (Ecol > 0.0f) {
        const float inv = 1.0f / Ecol;
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const float p = p2[e] * inv;
            if (p > 0.0f) h -= p * logf(p);
        }
    }
    const float H = group_sum<TPF>(h);
    if (lane == 0) {
        const int nb = (M + kCutBandRows - 1) / kCutBandRows;
        float cm = 0.0f;
        for (int b = 0; b < nb; ++b) cm = fmaxf(cm, J.colmax_part[(static_cast<size_t>(f) * nb + b) * W + n]);
        J.colmax[static_cast<size_t>(f) * W + n] = cm;
        J.colE[static_cast<size_t>(f) * W + n] = Ecol;
        J.colH[static_cast<size_t>(f) * W + n] = Ecol > 0.0f ? H : __int_as_float(0x7f800000);
    }
}

// Bin selection (phase_adjustment.cu:186-264): int-truncated mean of the column maxima,
// tgt_num bins above 1.48 * mean, the tgt_num most energetic (stable, descending), the
// unit2_num = nextPow2(tgt_num / 2) of those with the lowest entropy (stable, ascending),
// sorted by index. One workgroup of W threads per frame.
template <int W>
__global__ __launch_bounds__(W) void k_pa_select(Job J) {
    HIP_DYNAMIC_SHARED(float, sm_sel);
    float* key = sm_sel;                              // W
    int* tag = reinterpret_cast<int*>(key + W);       // W
    int* order1 = tag + W;                            // W
    float* red = reinterpret_cast<float*>(order1 + W);   // 32
    int* ired = reinterpret_cast<int*>(red + 32);        // 32
    const int f = blockIdx.x, n = threadIdx.x;
    const float mx = J.colmax[static_cast<size_t>(f) * W + n];
    const float E = J.colE[static_cast<size_t>(f) * W + n];
    // thrust::reduce(..., 0, plus<int>()) (phase_adjustment.cu:186) converts every maximum to int
    // as the device does (truncation, clamped) and adds them in 32-bit arithmetic, which wraps
    unsigned utot = 0;
    (void)block_exclusive_scan<W, unsigned>(static_cast<unsigned>(trunc_sat_i32(mx)),
                                            reinterpret_cast<unsigned*>(ired), &utot);
    const int tot = static_cast<int>(utot);   // two's complement
    const float mean_val = static_cast<float>(tot) / static_cast<float>(W);
    const float threshold = 1.48f * mean_val;
    int tgt = 0;
    (void)block_exclusive_scan<W>(mx > threshold ? 1 : 0, ired, &tgt);
    int half = tgt / 2, U = 1;
    while (half >> 1) { U <<= 1; half >>= 1; }
    U <<= 1;   // nextPow2 (common.cu:521-529)
    // sort 1: energy descending, index ascending on ties
    key[n] = (E == E) ? -E : __int_as_float(0x7f800000);
    tag[n] = n;
    __syncthreads();
    bitonic_sort_kv<W>(key, tag);
    order1[n] = tag[n];
    __syncthreads();
    // sort 2: entropy ascending among the first tgt, rank ascending on ties
    key[n] = n < tgt ? J.colH[static_cast<size_t>(f) * W + order1[n]] : __int_as_float(0x7f800000);
    tag[n] = n < tgt ? n : W + n;
    __syncthreads();
    bitonic_sort_kv<W>(key, tag);
    // first U of them (bin 0 fills positions beyond tgt, as std::vector growth does)
    int bin = 0x7fffffff;
    if (n < U) bin = (n < tgt) ? order1[tag[n]] : 0;
    __syncthreads();
    key[n] = static_cast<float>(bin == 0x7fffffff ? 1 << 30 : bin);
    tag[n] = bin;
    __syncthreads();
    bitonic_sort_kv<W>(key, tag);
    if (n < U) J.sel[static_cast<size_t>(f) * W + n] = tag[n];
    if (n == 0) {
        J.dec[f].tgt_num = tgt;
        J.dec[f].unit2_num = U;
        J.dec[f].pa_threshold = threshold;
    }
}

// Fast minimum-entropy autofocus (phase_adjustment.cu:270-324): one workgroup per frame; the
// selected bins are gathered once into a column-major copy, then J.iterations fixed-point
// updates phi_m <- unit(sum_u conj(S[m,u]) IFFT(G |G|)[m,u] / M), G = FFT(diag(phi) S).
// CF columns are transformed at once (TPF threads each); when all U selected bins fit in one
// chunk they stay in registers for every iteration.
template <int M>
struct AfCfg {
    static constexpr int E = ColCfg<M>::E;
    static constexpr int TPF = ColCfg<M>::TPF;
    static constexpr int NT = M <= 384 ? 1024 : 512;
    static constexpr int CF = NT / TPF;                       // columns per chunk
    static constexpr int RPT = (M + NT - 1) / NT;             // rows per thread when M > NT
    static constexpr int Q = M <= NT ? (NT / M >= 16 ? 16 : NT / M >= 8 ? 8 : NT / M >= 4 ? 4 : NT / M >= 2 ? 2 : 1) : 1;
    // Summation order of the per-pulse sums: the selected bins in sub-chunks of SUB columns, each
    // summed column by column, the sub-chunk sums added in order. The one-workgroup form and the
    // split form (k_af_cols / k_af_phase) both follow it.
    static constexpr int SUB = CF / Q;
    static_assert(CF % Q == 0, "sub-chunks must tile a chunk");
};

// One term conj(S) IFFT(G |G|) / M of the per-pulse sums. Only the unit phasor of the sum is
// kept, so for a power-of-two M the 1/M is left out: scaling by 2^-k commutes with every rounding
// of the sum, its magnitude and the division, and the phases are the same bit for bit.
template <int M>
__device__ __forceinline__ float2 af_term(float2 s, float2 x, float invM) {
    if constexpr ((M & (M - 1)) == 0) return cconjmul(s, x);
    else return cscale(cconjmul(s, x), invM);
}

// The scale exponent of a frame's selected bins: 2^-e takes the largest selected |S| into [1, 2)
// (e = 0 for no finite positive maximum). cm: the workgroup's maximum of their column maxima.
__device__ __forceinline__ int af_scale_exponent(float cm) {
    return cm > 0.0f && cm <= 3.40282347e38f ? ilogbf(cm) : 0;
}

// PRE = true: the selected bins are in su and their Doppler products in rowdop already
// (k_af_gather, k_af_bin_doppler), and only the start phase is formed (no iterations).
template <int M, bool PRE = false>
__global__ __launch_bounds__(AfCfg<M>::NT) void k_autofocus(Job J) {
    using A = AfCfg<M>;
    using C = ColCfg<M>;
    constexpr int NT = A::NT, E = A::E, TPF = A::TPF, CF = A::CF, RPT = A::RPT, Q = A::Q;
    HIP_DYNAMIC_SHARED(float2, sm);
    float2* phi = sm;                                     // M
    float2* rdop = phi + M;                               // M
    double* scr = reinterpret_cast<double*>(rdop + M);    // M
    double* red = scr + M;                                // NT
    float2* bufs = reinterpret_cast<float2*>(red + NT);   // CF * lds_padded_size(M)
    const int f = blockIdx.x, tid = threadIdx.x;
    const int W = J.W;
    const int U = J.dec[f].unit2_num;
    const int* sel = J.sel + static_cast<size_t>(f) * W;
    float2* su = J.su + static_cast<size_t>(f) * W * M;
    const float2* X = J.x + static_cast<size_t>(f) * M * W;
    if (tid == 0) {   // the resident form's exchange epoch (the frame's run count) and hand-off counter
        J.af_sync[2 * f] += 1;
        J.af_sync[2 * f + 1] = 0;
    }

    // gather the selected bins, column-major (selectRangeBins, :335-346), scaled by 2^-e so that
    // the largest selected |S| lies in [1, 2): the update takes the unit phasor of sums cubic in
    // S, and a power of two scales every product exactly, so the phases are those of the unscaled
    // bins while conj(S) IFFT(G |G|) stays far from FP32 overflow at any input level
    if constexpr (!PRE) {
        float cm = 0.0f;
        for (int u = tid; u < U; u += NT) cm = fmaxf(cm, J.colmax[static_cast<size_t>(f) * W + sel[u]]);
        cm = block_max<NT>(cm, reinterpret_cast<float*>(red));
        const int e = af_scale_exponent(cm);
        {   // the selected bins' indices in LDS (the transform buffers are free here), then the
            // gather with eight loads in flight per thread
            int* lsel = reinterpret_cast<int*>(bufs);
            for (int u = tid; u < U; u += NT) lsel[u] = sel[u];
            __syncthreads();
            for (int idx0 = tid; idx0 < U * M; idx0 += 8 * NT) {
                float2 v[8];
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    const int idx = idx0 + i * NT;
                    if (idx < U * M) v[i] = X[static_cast<size_t>(idx / U) * W + lsel[idx % U]];
                }
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    const int idx = idx0 + i * NT;
                    if (idx < U * M)
                        su[static_cast<size_t>(idx % U) * M + idx / U] = make_float2(ldexpf(v[i].x, -e), ldexpf(v[i].y, -e));
                }
            }
        }
        __syncthreads();
    }
    if (J.phi0_given) {
        for (int m = tid; m < M; m += NT) phi[m] = J.phi0_in[static_cast<size_t>(f) * M + m];
        __syncthreads();
    } else if constexpr (PRE) {
        doppler_phase<NT>(J.rowdop + static_cast<size_t>(f) * M, M, phi, scr, red);
    } else {
        // Doppler tracking of the selected bins without compensation (:282), the products added
        // in bin order with eight bins' loads in flight
        for (int m = tid; m < M - 1; m += NT) {
            float2 acc = make_float2(0.0f, 0.0f);
            for (int u0 = 0; u0 < U; u0 += 8) {
                float2 a[8], b[8];
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    if (u0 + i < U) {
                        a[i] = su[static_cast<size_t>(u0 + i) * M + m];
                        b[i] = su[static_cast<size_t>(u0 + i) * M + m + 1];
                    }
                }
#pragma unroll
                for (int i = 0; i < 8; ++i)
                    if (u0 + i < U) acc = cadd(acc, cconjmul(a[i], b[i]));
            }
            rdop[m] = acc;
        }
        __syncthreads();
        doppler_phase<NT>(rdop, M, phi, scr, red);
    }
    for (int m = tid; m < M; m += NT) J.phi0[static_cast<size_t>(f) * M + m] = phi[m];
    if constexpr (PRE) {
        for (int m = tid; m < M; m += NT) J.phi[static_cast<size_t>(f) * M + m] = phi[m];
        return;
    }

    const int g = tid / TPF, lane = tid % TPF;
    float2* buf = bufs + g * lds_padded_size(M);
    const float invM = 1.0f / static_cast<float>(M);
    const bool resident = U <= CF;   // one chunk: the columns stay in registers
    // the twiddle powers stay in registers for all iterations' transforms
    const FftTwiddlePowers<M, E> tw = fft_twiddle_powers<M, E>(fft_twiddles<M, E>(lane));
    float2 s[E];
    for (int it = 0; it < J.iterations; ++it) {
        float2 acc[RPT];
#pragma unroll
        for (int i = 0; i < RPT; ++i) acc[i] = make_float2(0.0f, 0.0f);
        for (int c0 = 0; c0 < U; c0 += CF) {
            const int u = c0 + g;
            const bool valid = u < U;
            const float2* col = su + static_cast<size_t>(valid ? u : 0) * M;
            if (!resident || it == 0) {
#pragma unroll
                for (int e = 0; e < E; ++e) s[e] = valid ? col[C::in(lane, e)] : make_float2(0.0f, 0.0f);
            }
            float2 x[E];
#pragma unroll
            for (int e = 0; e < E; ++e) x[e] = cmul(phi[C::in(lane, e)], s[e]);
            fft<M, E, -1, true>(x, buf, lane, tw);
#pragma unroll
            for (int e = 0; e < E; ++e) x[e] = cscale(x[e], cabs_hw(x[e]));
            relayout_out_to_in<M, E, true>(x, buf, lane);
            fft<M, E, +1, true>(x, buf, lane, tw);
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const int pos = C::out(lane, e);
                const float2 sv = C::same_layout ? s[e] : (valid ? col[pos] : make_float2(0.0f, 0.0f));
                buf[pos] = af_term<M>(sv, x[e], invM);
            }
            __syncthreads();
            const int ncol = min(CF, U - c0);
            if constexpr (Q > 1) {
                // Q lanes per row: lane q sums sub-chunk q column by column, then every lane adds
                // the Q sub-chunk sums in order
                const int m = tid / Q, q = tid % Q;
                float2 t = make_float2(0.0f, 0.0f);
                if (m < M) {
                    const int k1 = min((q + 1) * A::SUB, ncol);
                    for (int k = q * A::SUB; k < k1; ++k) t = cadd(t, bufs[k * lds_padded_size(M) + m]);
                }
#pragma unroll
                for (int j = 0; j < Q; ++j) acc[0] = cadd(acc[0], make_float2(__shfl(t.x, j, Q), __shfl(t.y, j, Q)));
            } else {
#pragma unroll
                for (int i = 0; i < RPT; ++i) {
                    const int m = tid + i * NT;
                    if (m < M) {
                        float2 t = make_float2(0.0f, 0.0f);
                        for (int k = 0; k < ncol; ++k) t = cadd(t, bufs[k * lds_padded_size(M) + m]);
                        acc[i] = cadd(acc[i], t);
                    }
                }
            }
            __syncthreads();
        }
        if constexpr (Q > 1) {
            const int m = tid / Q;
            if (m < M && tid % Q == 0) {
                const float a = cabsf_(acc[0]);
                phi[m] = a > 0.0f ? cscale(acc[0], 1.0f / a) : make_float2(0.0f, 0.0f);
            }
        } else {
#pragma unroll
            for (int i = 0; i < RPT; ++i) {
                const int m = tid + i * NT;
                if (m < M) {
                    const float a = cabsf_(acc[i]);
                    phi[m] = a > 0.0f ? cscale(acc[i], 1.0f / a) : make_float2(0.0f, 0.0f);
                }
            }
        }
        __syncthreads();
    }
    for (int m = tid; m < M; m += NT) J.phi[static_cast<size_t>(f) * M + m] = phi[m];
}

// Start of the split and resident forms with the frame's work spread over many workgroups (a
// batch of a few frames leaves most CUs idle under one workgroup per frame): k_af_gather copies
// the selected bins into su as k_autofocus does (same scale exponent, same values), kAfGatherGroups
// workgroups per frame each taking a contiguous share of the (pulse, bin) elements; then
// k_af_bin_doppler forms each pulse's Doppler product of the selected bins (:282) with the
// products added in bin order as k_autofocus adds them, into rowdop, which is free once the bins
// are selected; k_autofocus<M, true> then forms the start phase from it.
constexpr int kAfGatherGroups = 32;

template <int M>
__global__ __launch_bounds__(256) void k_af_gather(Job J) {
    HIP_DYNAMIC_SHARED(int, lsel);                                 // W
    float* red = reinterpret_cast<float*>(lsel + J.W);             // 32
    const int f = blockIdx.y, tid = threadIdx.x;
    const int W = J.W;
    const int U = J.dec[f].unit2_num;
    const int* sel = J.sel + static_cast<size_t>(f) * W;
    float cm = 0.0f;
    for (int u = tid; u < U; u += 256) cm = fmaxf(cm, J.colmax[static_cast<size_t>(f) * W + sel[u]]);
    cm = block_max<256>(cm, red);
    const int e = af_scale_exponent(cm);
    for (int u = tid; u < U; u += 256) lsel[u] = sel[u];
    __syncthreads();
    const float2* X = J.x + static_cast<size_t>(f) * M * W;
    float2* su = J.su + static_cast<size_t>(f) * W * M;
    const int total = U * M, share = (total + gridDim.x - 1) / gridDim.x;
    const int i0 = blockIdx.x * share, i1 = min(i0 + share, total);
    for (int idx0 = i0 + tid; idx0 < i1; idx0 += 8 * 256) {
        float2 v[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int idx = idx0 + i * 256;
            if (idx < i1) v[i] = X[static_cast<size_t>(idx / U) * W + lsel[idx % U]];
        }
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int idx = idx0 + i * 256;
            if (idx < i1) su[static_cast<size_t>(idx % U) * M + idx / U] = make_float2(ldexpf(v[i].x, -e), ldexpf(v[i].y, -e));
        }
    }
}

template <int M>
__global__ __launch_bounds__(256) void k_af_bin_doppler(Job J) {
    const int f = blockIdx.y, m = blockIdx.x * 256 + threadIdx.x;
    if (m >= M - 1) return;
    const int U = J.dec[f].unit2_num;
    const float2* su = J.su + static_cast<size_t>(f) * J.W * M;
    float2 acc = make_float2(0.0f, 0.0f);
    for (int u0 = 0; u0 < U; u0 += 8) {
        float2 a[8], b[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            if (u0 + i < U) {
                a[i] = su[static_cast<size_t>(u0 + i) * M + m];
                b[i] = su[static_cast<size_t>(u0 + i) * M + m + 1];
            }
        }
#pragma unroll
        for (int i = 0; i < 8; ++i)
            if (u0 + i < U) acc = cadd(acc, cconjmul(a[i], b[i]));
    }
    J.rowdop[static_cast<size_t>(f) * M + m] = acc;
}

// The same fixed-point update with a frame spread over several workgroups, for batches too small
// to fill the GPU with one workgroup per frame. k_autofocus with no iterations gathers the
// selected bins and forms the starting phase; then iteration `it` is a launch of k_af_cols:
// workgroup (j, f) takes the frame's current phase, transforms sub-chunk j of the selected bins
// and writes its per-pulse sums of conj(S) IFFT(G |G|) (/ M unless M is a power of two, af_term)
// into the iteration's half of af_part.
// The current phase is either formed by every workgroup alike from the previous iteration's sums
// (FOLD: one launch per iteration) or read from J.phi, which k_af_phase writes after each
// iteration (two launches per iteration, no repeated work). The transforms and the summation
// order are those of k_autofocus, so the phases agree bit for bit.
// Unit phasor of pulse m's sum over the nsub sub-chunk sums part[q M + m], added in order;
// load(p) reads one sum.
template <int M, typename Load>
__device__ __forceinline__ float2 af_phase_from(const float2* part, int nsub, int m, Load load) {
    float2 acc = make_float2(0.0f, 0.0f);
    for (int q = 0; q < nsub; q += 8) {   // up to eight loads in flight, added in order
        float2 v[8];
#pragma unroll
        for (int i = 0; i < 8; ++i)
            if (q + i < nsub) v[i] = load(part + static_cast<size_t>(q + i) * M + m);
#pragma unroll
        for (int i = 0; i < 8; ++i)
            if (q + i < nsub) acc = cadd(acc, v[i]);
    }
    const float a = cabsf_(acc);
    return a > 0.0f ? cscale(acc, 1.0f / a) : make_float2(0.0f, 0.0f);
}
template <int M>
__device__ __forceinline__ float2 af_phase_from(const float2* part, int nsub, int m) {
    return af_phase_from<M>(part, nsub, m, [](const float2* p) { return *p; });
}

// the half of af_part that iteration `it` writes; frame f's sub-chunk sums in it
template <int M>
__device__ __forceinline__ float2* af_part_of(const Job& J, int f, int it) {
    const size_t per_frame = static_cast<size_t>(J.W / AfCfg<M>::SUB) * M;
    return J.af_part + (static_cast<size_t>(it & 1) * J.K + f) * per_frame;
}

// One iteration's work on one sub-chunk of a frame's selected bins: the SUB columns (column
// c0 + tid / TPF in s, first-pass layout, zero where invalid) are multiplied by the phase phi
// (M values), transformed, scaled by |G|, transformed back, multiplied by conj(S) (af_term) and
// left in bufs; then the per-pulse sums over the sub-chunk's ncol columns, each added column by
// column, are handed to put(m, sum). Every thread of the workgroup calls it.
template <int M, typename Put>
__device__ __forceinline__ void af_subchunk(const float2 (&s)[AfCfg<M>::E], const float2* phi, const float2* col,
                                            bool valid, int ncol, float2* bufs, int tid,
                                            const FftTwiddles<M, AfCfg<M>::E>& tw, Put put) {
    using A = AfCfg<M>;
    using C = ColCfg<M>;
    constexpr int E = A::E, TPF = A::TPF, NTS = A::SUB * TPF;
    const int g = tid / TPF, lane = tid % TPF;
    float2* buf = bufs + g * lds_padded_size(M);
    const float invM = 1.0f / static_cast<float>(M);
    float2 x[E];
#pragma unroll
    for (int e = 0; e < E; ++e) x[e] = cmul(phi[C::in(lane, e)], s[e]);
    fft<M, E, -1, true>(x, buf, lane, tw);
#pragma unroll
    for (int e = 0; e < E; ++e) x[e] = cscale(x[e], cabs_hw(x[e]));
    relayout_out_to_in<M, E, true>(x, buf, lane);
    fft<M, E, +1, true>(x, buf, lane, tw);
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int pos = C::out(lane, e);
        const float2 sv = C::same_layout ? s[e] : (valid ? col[pos] : make_float2(0.0f, 0.0f));
        buf[pos] = af_term<M>(sv, x[e], invM);
    }
    __syncthreads();
    for (int m = tid; m < M; m += NTS) {
        float2 t = make_float2(0.0f, 0.0f);
        for (int k = 0; k < ncol; ++k) t = cadd(t, bufs[k * lds_padded_size(M) + m]);
        put(m, t);
    }
}

template <int M, bool FOLD>
__global__ __launch_bounds__(AfCfg<M>::SUB * AfCfg<M>::TPF) void k_af_cols(Job J, int it) {
    using A = AfCfg<M>;
    using C = ColCfg<M>;
    constexpr int E = A::E, TPF = A::TPF, SUB = A::SUB, NTS = SUB * TPF;
    HIP_DYNAMIC_SHARED(float2, bufs);   // SUB * lds_padded_size(M), then (FOLD) the phase (M)
    const int j = blockIdx.x, f = blockIdx.y, tid = threadIdx.x;
    const int U = J.dec[f].unit2_num;
    const int nsub = (U + SUB - 1) / SUB;
    if (j >= nsub) return;   // the whole workgroup leaves before any barrier
    const int c0 = j * SUB;
    const int W = J.W;
    const float2* phi = J.phi + static_cast<size_t>(f) * M;
    if (FOLD && it > 0) {
        float2* formed = bufs + SUB * lds_padded_size(M);
        const float2* prev = af_part_of<M>(J, f, it - 1);
        for (int m = tid; m < M; m += NTS) formed[m] = af_phase_from<M>(prev, nsub, m);
        __syncthreads();
        phi = formed;
    }
    const int u = c0 + tid / TPF, lane = tid % TPF;
    const bool valid = u < U;
    const float2* col = J.su + static_cast<size_t>(f) * W * M + static_cast<size_t>(valid ? u : 0) * M;
    const FftTwiddles<M, E> tw = fft_twiddles<M, E>(lane);
    float2 s[E];
#pragma unroll
    for (int e = 0; e < E; ++e) s[e] = valid ? col[C::in(lane, e)] : make_float2(0.0f, 0.0f);
    float2* part = af_part_of<M>(J, f, it) + static_cast<size_t>(j) * M;
    af_subchunk<M>(s, phi, col, valid, min(SUB, U - c0), bufs, tid, tw, [&](int m, float2 t) { part[m] = t; });
}

// The exchange words of frame f in the resident form: per iteration parity, sub-chunk q < W/SUB
// and pulse m, the real and the imaginary part of the sub-chunk's per-pulse sum as tagged words
// (xwg_word). Frame f owns 4 (W/8) M words of af_part, the buffer being sized for it.
template <int M>
__device__ __forceinline__ uint64_t* af_words_of(const Job& J, int f, int it) {
    const size_t nq = static_cast<size_t>(J.W / AfCfg<M>::SUB);
    uint64_t* frame = reinterpret_cast<uint64_t*>(J.af_part) + static_cast<size_t>(f) * 4 * (J.W / 8) * M;
    return frame + static_cast<size_t>(it & 1) * nq * M * 2;
}

// af_phase_from over tagged words (af_words_of): pulse m's sub-chunk sums are loaded eight at a
// time, each word polled until it carries `tag`, and added in sub-chunk order, so the phase is
// af_phase_from's bit for bit. A thread that has polled more than `limit` times gives up for the
// rest of the launch (waiting = false) and sets status bit 2.
template <int M>
__device__ __forceinline__ float2 af_phase_tagged(const uint64_t* words, int nsub, int m, unsigned tag, int limit,
                                                  bool& waiting, int32_t& status) {
    float2 acc = make_float2(0.0f, 0.0f);
    int polls = 0;
    for (int q0 = 0; q0 < nsub; q0 += 8) {
        uint64_t re[8], im[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            if (q0 + i < nsub) {
                const uint64_t* w = words + (static_cast<size_t>(q0 + i) * M + m) * 2;
                re[i] = xwg_load(w);
                im[i] = xwg_load(w + 1);
            }
        }
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            if (q0 + i >= nsub) break;
            const uint64_t* w = words + (static_cast<size_t>(q0 + i) * M + m) * 2;
            while (waiting && (xwg_tag(re[i]) != tag || xwg_tag(im[i]) != tag)) {
                if (polls++ >= limit) {
                    waiting = false;
                    atomicOr(&status, 2);
                    break;
                }
                xwg_pause();
                re[i] = xwg_load(w);
                im[i] = xwg_load(w + 1);
            }
            acc = cadd(acc, make_float2(xwg_value(re[i]), xwg_value(im[i])));
        }
    }
    const float a = cabsf_(acc);
    return a > 0.0f ? cscale(acc, 1.0f / a) : make_float2(0.0f, 0.0f);
}

// The fixed-point iterations of k_af_cols (FOLD) in one launch: workgroup (j, f) keeps sub-chunk
// j of frame f's selected bins in registers for every iteration and hands its per-pulse sums to
// the frame's other workgroups; every workgroup then forms the next phase from all nsub sums, in
// sub-chunk order. The transforms, sums and phases are those of k_af_cols bit for bit.
// Up to 512 pulses the sums travel as tagged words: each word is polled until it carries
// iteration it's tag (epoch << 7) | it, where the epoch is the frame's run count that k_autofocus
// advances before this launch, so words of an earlier run never pass (at most 128 iterations;
// the host runs the split form above that). At 1024 pulses, where a frame has more sub-chunks and
// tagged words would double the bytes every workgroup reads per iteration, the sums travel as
// plain 64-bit words: each thread waits for its stores to be acknowledged, the workgroup counts
// itself in on the frame's counter (cleared by k_autofocus), and thread 0 polls the counter until
// all nsub workgroups have counted in for the iteration.
// Either way a workgroup writes iteration it + 2 into the words of iteration it only after it has
// read every workgroup's iteration it + 1, which each wrote after reading iteration it.
// All of a frame's workgroups must be resident together; the dispatcher places a grid's
// workgroups in order, so every frame but the last one placed is complete, and that one
// completes as earlier frames finish. A wait that exceeds J.af_wait_polls polls (about a second by
// default; a negative bound gives up at once) sets status bit 2 of the frame and stops that
// thread waiting; the frame's phase is then not valid. Workgroup 0 of each frame writes the final
// phase to J.phi.
template <int M>
__global__ __launch_bounds__(AfCfg<M>::SUB * AfCfg<M>::TPF) void k_af_resident(Job J) {
    using A = AfCfg<M>;
    using C = ColCfg<M>;
    constexpr int E = A::E, TPF = A::TPF, SUB = A::SUB, NTS = SUB * TPF;
    constexpr bool kTagged = M <= 512;
    HIP_DYNAMIC_SHARED(float2, bufs);   // SUB * lds_padded_size(M), then the phase (M)
    float2* phi = bufs + SUB * lds_padded_size(M);
    const int j = blockIdx.x, f = blockIdx.y, tid = threadIdx.x;
    const int U = J.dec[f].unit2_num;
    const int nsub = (U + SUB - 1) / SUB;
    if (j >= nsub) return;   // the whole workgroup leaves before any barrier
    const int c0 = j * SUB;
    const int W = J.W;
    const int u = c0 + tid / TPF, lane = tid % TPF;
    const bool valid = u < U;
    const float2* col = J.su + static_cast<size_t>(f) * W * M + static_cast<size_t>(valid ? u : 0) * M;
    const FftTwiddles<M, E> tw = fft_twiddles<M, E>(lane);
    float2 s[E];
#pragma unroll
    for (int e = 0; e < E; ++e) s[e] = valid ? col[C::in(lane, e)] : make_float2(0.0f, 0.0f);
    for (int m = tid; m < M; m += NTS) phi[m] = J.phi[static_cast<size_t>(f) * M + m];
    __syncthreads();
    const unsigned epoch = J.af_sync[2 * f];
    unsigned* const count = J.af_sync + 2 * f + 1;
    bool waiting = J.af_wait_polls >= 0;   // false once this thread's polls have hit their bound
    if (!waiting && tid == 0) atomicOr(&J.dec[f].status, 2);   // a negative bound: every wait gives up
    for (int it = 0; it < J.iterations; ++it) {
        uint64_t* words = af_words_of<M>(J, f, it);
        if constexpr (kTagged) {
            const unsigned tag = (epoch << 7) | static_cast<unsigned>(it);
            af_subchunk<M>(s, phi, col, valid, min(SUB, U - c0), bufs, tid, tw, [&](int m, float2 t) {
                uint64_t* w = words + (static_cast<size_t>(j) * M + m) * 2;
                xwg_store(w, xwg_word(t.x, tag));
                xwg_store(w + 1, xwg_word(t.y, tag));
            });
            for (int m = tid; m < M; m += NTS)
                phi[m] = af_phase_tagged<M>(words, nsub, m, tag, J.af_wait_polls, waiting, J.dec[f].status);
        } else {
            af_subchunk<M>(s, phi, col, valid, min(SUB, U - c0), bufs, tid, tw, [&](int m, float2 t) {
                uint64_t v;
                __builtin_memcpy(&v, &t, sizeof v);
                xwg_store(words + static_cast<size_t>(j) * M + m, v);
            });
            xwg_stores_done();
            __syncthreads();
            if (tid == 0) {
                xwg_arrive(count);
                if (waiting && !xwg_wait(count, static_cast<unsigned>(it + 1) * static_cast<unsigned>(nsub), J.af_wait_polls)) {
                    waiting = false;
                    atomicOr(&J.dec[f].status, 2);
                }
            }
            __syncthreads();
            for (int m = tid; m < M; m += NTS)
                phi[m] = af_phase_from<M>(reinterpret_cast<const float2*>(words), nsub, m, [](const float2* p) {
                    const uint64_t v = xwg_load(reinterpret_cast<const uint64_t*>(p));
                    float2 t;
                    __builtin_memcpy(&t, &v, sizeof t);
                    return t;
                });
        }
        __syncthreads();
    }
    if (j == 0)
        for (int m = tid; m < M; m += NTS) J.phi[static_cast<size_t>(f) * M + m] = phi[m];
}

template <int M>
__global__ __launch_bounds__(256) void k_af_phase(Job J, int it) {
    using A = AfCfg<M>;
    const int f = blockIdx.y, m = blockIdx.x * 256 + threadIdx.x;
    if (m >= M) return;
    const int nsub = (J.dec[f].unit2_num + A::SUB - 1) / A::SUB;
    J.phi[static_cast<size_t>(f) * M + m] = af_phase_from<M>(af_part_of<M>(J, f, it), nsub, m);
}

// ------------------------------------------------------------------------------------------
namespace {
template <int M>
size_t af_cols_lds(bool fold) {
    return (static_cast<size_t>(AfCfg<M>::SUB) * lds_padded_size(M) + (fold ? M : 0)) * sizeof(float2);
}
template <int M>
size_t colfft_lds() {
    using C = ColCfg<M>;
    const size_t work = std::max(static_cast<size_t>(M) * (C::CT + 1), static_cast<size_t>(C::CT) * lds_padded_size(M));
    return M * sizeof(float2) + work * sizeof(float2);
}
template <int M>
size_t autofocus_lds() {
    using A = AfCfg<M>;
    return 2 * M * sizeof(float2) + (M + A::NT) * sizeof(double) + static_cast<size_t>(A::CF) * lds_padded_size(M) * sizeof(float2);
}
size_t select_lds(int W) { return 3 * W * sizeof(int) + 64 * sizeof(float); }
}  // namespace

#define ISAR_PULSE_SWITCH(M, BODY)                       \
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

bool pulse_count_supported(int M) {
    return M == 64 || M == 128 || M == 192 || M == 256 || M == 384 || M == 512 || M == 1024;
}

hipError_t configure_pa_kernels() {
    hipError_t err = hipSuccess;
#ifndef ISAR_EMU
    auto set = [&](const void* fn, size_t bytes) {
        hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(bytes));
        if (e != hipSuccess) {
            err = e;
            report_config_error(__FILE__, bytes);
        }
    };
#define ISAR_SET_PA(MM)                                                          \
    set(reinterpret_cast<const void*>(&k_pa_colfft<MM>), colfft_lds<MM>());      \
    set(reinterpret_cast<const void*>(&k_autofocus<MM>), autofocus_lds<MM>());   \
    set(reinterpret_cast<const void*>(&k_autofocus<MM, true>), autofocus_lds<MM>());   \
    set(reinterpret_cast<const void*>(&k_af_cols<MM, true>), af_cols_lds<MM>(true));   \
    set(reinterpret_cast<const void*>(&k_af_cols<MM, false>), af_cols_lds<MM>(false));   \
    set(reinterpret_cast<const void*>(&k_af_resident<MM>), af_cols_lds<MM>(true));
    ISAR_SET_PA(64) ISAR_SET_PA(128) ISAR_SET_PA(192) ISAR_SET_PA(256) ISAR_SET_PA(384) ISAR_SET_PA(512)
    ISAR_SET_PA(1024)
#undef ISAR_SET_PA
#endif
    return err;
}

void launch_pa_colfft(const Job& J, hipStream_t s) {
    ISAR_PULSE_SWITCH(J.M, ISAR_LAUNCH(k_pa_doppler<MM>, dim3(J.K), dim3(256), (MM + 256) * sizeof(double), s, J));
    ISAR_PULSE_SWITCH(J.M, ISAR_LAUNCH(k_pa_colfft<MM>, dim3(J.W / ColCfg<MM>::CT, J.K), dim3(ColCfg<MM>::NT),
                                       colfft_lds<MM>(), s, J));
}
void launch_pa_select(const Job& J, hipStream_t s) {
    if (J.W == kRangeCut) ISAR_LAUNCH(k_pa_select<kRangeCut>, dim3(J.K), dim3(kRangeCut), select_lds(J.W), s, J);
}
void launch_autofocus(const Job& J, hipStream_t s) {
    ISAR_PULSE_SWITCH(J.M, ISAR_LAUNCH(k_autofocus<MM>, dim3(J.K), dim3(AfCfg<MM>::NT), autofocus_lds<MM>(), s, J));
}
namespace {
// The start phase and the gathered bins of the split and resident forms (k_af_gather,
// k_af_bin_doppler, k_autofocus<MM, true>).
template <int MM>
void autofocus_start_spread(const Job& J, hipStream_t s) {
    ISAR_LAUNCH(k_af_gather<MM>, dim3(kAfGatherGroups, J.K), dim3(256), J.W * sizeof(int) + 32 * sizeof(float), s, J);
    if (!J.phi0_given) ISAR_LAUNCH(k_af_bin_doppler<MM>, dim3((MM - 1 + 255) / 256, J.K), dim3(256), 0, s, J);
    Job J0 = J;
    J0.iterations = 0;
    ISAR_LAUNCH((k_autofocus<MM, true>), dim3(J.K), dim3(AfCfg<MM>::NT), autofocus_lds<MM>(), s, J0);
}

template <int MM>
void autofocus_split(const Job& J, bool fold, hipStream_t s) {
    autofocus_start_spread<MM>(J, s);
    constexpr int SUB = AfCfg<MM>::SUB;
    const dim3 cols_grid(J.W / SUB, J.K);
    const dim3 cols_block(SUB * AfCfg<MM>::TPF);
    const dim3 phase_grid((MM + 255) / 256, J.K);
    for (int it = 0; it < J.iterations; ++it) {
        if (fold) {
            ISAR_LAUNCH((k_af_cols<MM, true>), cols_grid, cols_block, af_cols_lds<MM>(true), s, J, it);
        } else {
            ISAR_LAUNCH((k_af_cols<MM, false>), cols_grid, cols_block, af_cols_lds<MM>(false), s, J, it);
            ISAR_LAUNCH(k_af_phase<MM>, phase_grid, dim3(256), 0, s, J, it);
        }
    }
    if (fold && J.iterations > 0) ISAR_LAUNCH(k_af_phase<MM>, phase_grid, dim3(256), 0, s, J, J.iterations - 1);
}
}  // namespace

void launch_autofocus_split(const Job& J, bool fold, hipStream_t s) {
    ISAR_PULSE_SWITCH(J.M, autofocus_split<MM>(J, fold, s));
}

namespace {
template <int MM>
void autofocus_resident(const Job& J, hipStream_t s) {
    autofocus_start_spread<MM>(J, s);
    if (J.iterations > 0) {
        constexpr int SUB = AfCfg<MM>::SUB;
        ISAR_LAUNCH_RESIDENT(k_af_resident<MM>, dim3(J.W / SUB, J.K), dim3(SUB * AfCfg<MM>::TPF), af_cols_lds<MM>(true), s,
                             J);
    }
}
}  // namespace

void launch_autofocus_resident(const Job& J, hipStream_t s) {
    ISAR_PULSE_SWITCH(J.M, autofocus_resident<MM>(J, s));
}

}  // namespace isar
