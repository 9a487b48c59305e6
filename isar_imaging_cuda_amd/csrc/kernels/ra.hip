// ra.hip — range alignment (range_alignment.cu:4-201) for gfx950.
//
// The reference inverse-transforms all M rows at every merge level and once more at the end.
// Here each row's spectrum S0 (HPC, Hamming window and (-1)^n centring applied) is formed once
// and stored; a row is transformed again only when its cumulative lag changes, by re-reading S0
// and applying its lag ramp on the fly. Group energies sum_{rows} |IFFT(row)|^2 / N^2 are kept per
// group head row and merged level by level, E^{l+1}_g = E^l_{2g} + E~^l_{2g+1} (E~ = the second
// group after its shift). Row m moves for the last time at level floor(log2 m) (it moves at level
// l exactly when bit l of m is set), so its aligned profile is written then, and no final pass
// over all rows is needed.
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

// Dynamic LDS of the row kernels: the transform buffer (lds_padded_size(N) float2), 64 floats of
// reductions, then the 256 twiddles of the stride-16 radix-16 pass (fft_tw16_table_fill).
template <int N>
__device__ __forceinline__ float2* tw16_table(float2* lds) {
    return reinterpret_cast<float2*>(reinterpret_cast<char*>(lds + lds_padded_size(N)) + 64 * sizeof(float));
}

// Spectrum of row m of frame f with HPC (FP64 phase), window and (-1)^n centring applied, in
// the first-pass layout (range_alignment.cu:13-18, high_speed_compensation.cu:4-54). With
// if_hrrp the STRETCH rows carry the window twice (getHRRP windows them in place first,
// common.cu:563-566) unless ONE_WINDOW asks for the rows as getHRRP transforms them.
//
// The HPC phase exp(j pi vk n^2) at the first-pass positions n = tid + e S, S = N / 16, is
// factorised as exp(j pi vk tid^2) * d^e * exp(j pi vk S^2 e^2) with d = exp(j pi 2 vk S tid):
// three FP64-reduced evaluations per thread (tid^2, d, d^4) and one per lane e < 16 of each wave
// (the e^2 factor, read by the other lanes), instead of sixteen. exp(j pi vk tid^2) d^e is walked
// along each group of four elements by one product with d, and the group base steps by d^4, so
// no factor is more than six products from an FP64-reduced value and only the running factor,
// the group base, d, d^4 and the lane's e^2 factor stay live beside the row. Every FP64 argument
// is a part of vk n^2 and is reduced at least as exactly as vk n^2 itself.
template <int N, bool ONE_WINDOW = false>
__device__ __forceinline__ void load_row_raw(float2 (&x)[kE], const Job& J, int f, int m, int tid_) {
    static_assert(FftPlan<N, kE>::R[0] == 16, "factorised phase assumes a radix-16 first pass");
    const int tid = opaque(tid_);
    const long long erow = echo_first_row(J, f) + m;
    const float2* row = J.data_type == 1 ? J.pre + (static_cast<size_t>(f) * J.M + m) * N
                                         : J.echo + static_cast<size_t>(erow) * N;
    const bool hpc = J.if_hpc && J.data_type == 2;
    const bool hrrp_window = !ONE_WINDOW && J.if_hrrp && J.data_type == 2;
    auto window = [&](int n) {
        float w = elem_u32(J.win_range, n);
        if (hrrp_window) w *= w;
        return (n & 1) ? -w : w;
    };
    if (!hpc) {
#pragma unroll
        for (int e = 0; e < kE; ++e) {
            const int n = in_pos<N>(tid, e);
            x[e] = cscale(elem_u32(row, n), window(n));
        }
        return;
    }
    constexpr double S = N / kE;
    const double vk = J.hpc_k * J.velocity[erow];
    const double t = static_cast<double>(tid);
    const double lin = 2.0 * S * vk * t;
    const float2 d1 = expjpi_reduced(lin);
    const float2 d4 = expjpi_reduced(4.0 * lin);
    float2 base = expjpi_reduced(vk * t * t);
    const double e0 = static_cast<double>(tid & 15);
    const float2 c = expjpi_reduced((S * S) * vk * (e0 * e0));
#pragma unroll
    for (int q = 0; q < kE / 4; ++q) {
        float2 h = base;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int e = 4 * q + r;
            const int n = in_pos<N>(tid, e);
            const float2 ce = make_float2(wave_read_lane(c.x, e), wave_read_lane(c.y, e));
            x[e] = cscale(cmul(elem_u32(row, n), cmul(h, ce)), window(n));
            if (r < 3) h = cmul(h, d1);
        }
        if (q + 1 < kE / 4) base = cmul(base, d4);
    }
}

// Multiply a first-pass-layout row by the lag ramp exp(-j 2 pi n lag / N). With n = tid + e N/16
// the ramp factors into exp(j pi b tid) * u_e, u_e = exp(j pi b e N/16), b = -2 lag / N. The u_e
// are the same for every thread: lane e of each wave evaluates u_e and the other lanes read it.
// Both arguments are reduced in FP64.
template <int N>
__device__ __forceinline__ void apply_ramp(float2 (&x)[kE], double lag, int tid) {
    static_assert(FftPlan<N, kE>::R[0] == 16, "factorised ramp assumes a radix-16 first pass");
    if (lag == 0.0) return;
    const double b = -2.0 * lag / N;
    const float2 w0 = expjpi_reduced(b * tid);
    const float2 u = expjpi_reduced(b * static_cast<double>((N / 16) * (tid & 15)));
#pragma unroll
    for (int e = 0; e < kE; ++e) {
        const float2 ue = make_float2(wave_read_lane(u.x, e), wave_read_lane(u.y, e));
        x[e] = cmul(cmul(x[e], ue), w0);
    }
}

// Stored spectrum S0 of row m times its lag ramp.
template <int N>
__device__ __forceinline__ void load_row_shifted(float2 (&x)[kE], const Job& J, int f, int m, double lag, int tid_) {
    const int tid = opaque(tid_);
    const float2* row = J.pre + (static_cast<size_t>(f) * J.M + m) * N;
#pragma unroll
    for (int e = 0; e < kE; ++e) x[e] = elem_u32(row, static_cast<unsigned>(in_pos<N>(tid, e)));
    apply_ramp<N>(x, lag, tid);
}

// Energy profile of group g at merge stride s >= 2, in the first-pass layout. Level 1 (s = 2)
// energies are complete (written by the fused level-0 pass); from level 2 on, the shifted second
// subgroup of the previous level is added from its row-chunk partial sums.
template <int N>
__device__ __forceinline__ void group_energy(float (&a)[kE], const Job& J, int f, int g, int s, int tid) {
    const int M = J.M;
    const float* base = J.egrp + (static_cast<size_t>(f) * M + static_cast<size_t>(g) * s) * N;
#pragma unroll
    for (int e = 0; e < kE; ++e) a[e] = elem_u32(base, in_pos<N>(tid, e));
    if (s > 2) {
        const int s2 = s / 2;
        const int r0 = (2 * g + 1) * s2;
        if (r0 < M) {
            const int r1 = min(r0 + s2, M);
            const int R = min(s2, J.ra_chunk);
            for (int h = r0; h < r1; h += R) {
                const float* pp = J.epart + (static_cast<size_t>(f) * M + h) * N;
#pragma unroll
                for (int e = 0; e < kE; ++e) a[e] += elem_u32(pp, in_pos<N>(tid, e));
            }
        }
    }
}

// Correlation of two non-negative profiles given as energies (first-pass layout): the argmax of
// |c| after ifftshift, lowest index on ties (range_alignment.cu:45-61, 108-168). The real
// correlation comes from one complex FFT of gA + j gB. Every thread gets the argmax.
template <int N>
__device__ int correlation_argmax(const float (&a)[kE], const float (&b)[kE], float2* lds, float* redv, int* redi,
                                  int tid, const FftTwiddles<N, kE>& tw, const float2* tab) {
    constexpr int NT = N / kE;
    float2 x[kE];
#pragma unroll
    for (int e = 0; e < kE; ++e) x[e] = make_float2(sqrt_hw(a[e]), sqrt_hw(b[e]));
    fft<N, kE, -1>(x, lds, tid, tw, tab);
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
    fft<N, kE, +1>(x, lds, tid, tw, tab);
    float best = -1.0f;
    int bi = 0;
#pragma unroll
    for (int e = 0; e < kE; ++e) {
        const int i = (out_pos<N>(tid, e) + N / 2) & (N - 1);
        const float v = fabsf(x[e].x);
        if (v > best || (v == best && i < bi)) { best = v; bi = i; }
    }
    // the transform's last LDS reads precede the reduction's barriers
    block_argmax<NT>(best, bi, redv, redi);
    return bi;
}

// sqrt(x) to about 1e-14 relative for x >= 0: the hardware FP32 root (within 1 ulp) refined by one
// Newton step in FP64, s + (x - s^2) / (2 s); the step's own factor 1 / (2 s) needs only the
// hardware reciprocal's accuracy (1 ulp). A few FP64 operations instead of the long correctly rounded FP64 square root.
__device__ __forceinline__ double sqrt_refined(float x) {
    const float s = sqrt_hw(x);
    const double sd = static_cast<double>(s);
    const double d = fma(-sd, sd, static_cast<double>(x));
    return s > 0.0f ? fma(d, static_cast<double>(0.5f * rcp_hw(s)), sd) : 0.0;
}

// Parabolic lag around the argmax bi (binomialFixDevice, range_alignment.cu:121-131). The three
// samples c[bi-1], c[bi], c[bi+1] are formed again as direct sums sum_n gA[n] gB[n - l] in FP64,
// not read from the FP32 transform: near a broad peak the second
// difference f1 + f3 - 2 f2 cancels most of its digits, and the transform's rounding alone moved
// the vertex by up to 2.4e-4 cell at 1024 x 16384, which 120 autofocus iterations amplify into a
// 1e-3 image difference. The vertex and the lag are FP64. a, b: the two energy profiles, element
// e of thread tid at range cell pos(tid, e), where pos is the first-pass layout (OUT = false) or
// the last-pass layout (OUT = true). Must follow correlation_argmax (its barriers end the LDS
// use). Thread 0 records the lag and argmax; every thread gets the lag.
template <int N, bool OUT = false>
__device__ double parabola_lag(const float (&a)[kE], const float (&b)[kE], int bi, float2* lds, int tid,
                               double* lag_out, int32_t* imax_out) {
    constexpr int NT = N / kE;
    // cl[i] = c[i - N/2] and c[l] = sum_n gA[n] gB[n - l] (mod N): the samples at bi - 1, bi,
    // bi + 1 are lags l0 - 1, l0, l0 + 1 with l0 = bi - N/2. gB goes to LDS by range cell (N
    // doubles fit the transform's padded N float2); each thread sums over its own cells. The
    // per-wave sums follow gB in LDS and are added in wave order, so the result does not depend
    // on scheduling.
    // The profiles are the square roots of the FP32 energies to FP64 accuracy (the hardware FP32
    // root alone is off by up to 1 ulp, which the vertex would see). The energies pass through
    // opaque() so that the roots are not shared with the transform input above and kept live
    // across its transforms.
    double* gb = reinterpret_cast<double*>(lds);
    double* red = gb + N;
    auto pos = [&](int e) { return OUT ? out_pos<N>(tid, e) : in_pos<N>(tid, e); };
#pragma unroll
    for (int e = 0; e < kE; ++e) gb[pos(e)] = sqrt_refined(opaque(b[e]));
    __syncthreads();
    const int l0 = bi - N / 2;
    double s1 = 0.0, s2 = 0.0, s3 = 0.0;
#pragma unroll
    for (int e = 0; e < kE; ++e) {
        const int n = pos(e);
        const double ga = sqrt_refined(opaque(a[e]));
        s1 = fma(ga, gb[(n - l0 + 1) & (N - 1)], s1);
        s2 = fma(ga, gb[(n - l0) & (N - 1)], s2);
        s3 = fma(ga, gb[(n - l0 - 1) & (N - 1)], s3);
    }
    s1 = wave_sum_d(s1);
    s2 = wave_sum_d(s2);
    s3 = wave_sum_d(s3);
    const int w = tid >> 6;
    if ((tid & 63) == 0) {
        red[3 * w] = s1;
        red[3 * w + 1] = s2;
        red[3 * w + 2] = s3;
    }
    __syncthreads();
    if (tid == 0) {
        double f1 = 0.0, f2 = 0.0, f3 = 0.0;
#pragma unroll
        for (int k = 0; k < (NT + 63) / 64; ++k) {
            f1 += red[3 * k];
            f2 += red[3 * k + 1];
            f3 += red[3 * k + 2];
        }
        double delta = 0.0;
        if (bi > 0 && bi < N - 1) {   // Q9a: no refinement at the row edge
            const double fa = (f1 + f3 - 2.0 * f2) / 2.0, fb = (f3 - f1) / 2.0;
            if (fa != 0.0) delta = -fb / (2.0 * fa);
            if (!(delta == delta) || fabs(delta) > 1e30) delta = 0.0;
        }
        const double lag = static_cast<double>(l0) + delta;
        *lag_out = lag;
        *imax_out = bi;
        red[0] = lag;
    }
    __syncthreads();
    const double lag = red[0];
    __syncthreads();
    return lag;
}

}  // namespace

// IFDS pre-stage (isar_main.cu:15-27): ifftshift, IFFT / N, ifftshift of every row.
template <int N>
__global__ __launch_bounds__(N / kE) void k_ifds_pre(Job J) {
    HIP_DYNAMIC_SHARED(float2, lds);
    const int m = blockIdx.x, f = blockIdx.y, tid = threadIdx.x;
    const size_t off = (static_cast<size_t>(f) * J.M + m) * N;
    const float2* src = J.echo + static_cast<size_t>(echo_first_row(J, f) + m) * N;
    float2 x[kE];
#pragma unroll
    for (int e = 0; e < kE; ++e) x[e] = src[(in_pos<N>(tid, e) + N / 2) & (N - 1)];
    // getHRRP for IFDS (common.cu:546-558) undoes this stage, ifftshift(FFT(ifftshift(pre))) =
    // the input row, so the profile is the row itself
    if (J.hrrp) {
#pragma unroll
        for (int e = 0; e < kE; ++e) J.hrrp[off + ((in_pos<N>(tid, e) + N / 2) & (N - 1))] = x[e];
    }
    fft<N, kE, +1>(x, lds, tid);
    const float inv = 1.0f / N;
#pragma unroll
    for (int e = 0; e < kE; ++e) J.pre[off + ((out_pos<N>(tid, e) + N / 2) & (N - 1))] = cscale(x[e], inv);
}

// getHRRP for STRETCH (common.cu:561-571): ifftshift(FFT(hamming_N * row)) of the compensated
// rows, one workgroup per row. The ifftshift is the (-1)^n modulation load_row_raw applies.
template <int N>
__global__ __launch_bounds__(N / kE) void k_hrrp(Job J) {
    HIP_DYNAMIC_SHARED(float2, lds);
    const int m = blockIdx.x, f = blockIdx.y, tid = threadIdx.x;
    float2 x[kE];
    load_row_raw<N, true>(x, J, f, m, tid);
    fft<N, kE, -1>(x, lds, tid);
    float2* dst = J.hrrp + (static_cast<size_t>(f) * J.M + m) * N;
#pragma unroll
    for (int e = 0; e < kE; ++e) dst[out_pos<N>(tid, e)] = x[e];
}

// Merge level 0 fused with the spectrum preparation (range_alignment.cu:13-18 and :34-73 at
// stride 1), one workgroup per row pair (2p, 2p+1): S0 of both rows (stored for the later
// levels), both energy profiles, their correlation and lag, the shift of row 2p+1 and the
// merged level-1 energy. Rows 0 and 1 never move again: their aligned profiles and column sums
// are written here.
template <int N>
__global__ __launch_bounds__(N / kE, 4) void k_ra_pair0(Job J) {
    HIP_DYNAMIC_SHARED(float2, lds);
    constexpr int NT = N / kE;
    float* redv = reinterpret_cast<float*>(lds + lds_padded_size(N));
    int* redi = reinterpret_cast<int*>(redv + 16);
    const int p = blockIdx.x, f = blockIdx.y, tid = threadIdx.x;
    const int M = J.M;
    const int ra = 2 * p, rb = 2 * p + 1;
    const float inv2 = 1.0f / (static_cast<float>(N) * static_cast<float>(N));
    const float invN = 1.0f / N;
    const bool final_rows = (p == 0);
    float2* s0a = J.pre + (static_cast<size_t>(f) * M + ra) * N;
    float2* s0b = J.pre + (static_cast<size_t>(f) * M + rb) * N;
    float* dst = J.egrp + (static_cast<size_t>(f) * M + ra) * N;
    float* ap = J.arp_part + static_cast<size_t>(f) * (M / 2) * N;
    const FftTwiddles<N, kE> tw = fft_twiddles<N, kE>(tid);
    float2* const tab = tw16_table<N>(lds);
    fft_tw16_table_fill(tab, tid, NT);
    float2 x[kE];
    float ea[kE];
    float mx = 0.0f;
    // Every global location below is re-read only by the thread that wrote it, and the stores
    // into the S0 rows (which st aliases) are ordered by the barriers inside each transform.
    // row 2p: S0, profile, energy (written now as the first term of the merged energy)
    load_row_raw<N>(x, J, f, ra, tid);
    {
        const int t = opaque(tid);
#pragma unroll
        for (int e = 0; e < kE; ++e) elem_u32(s0a, in_pos<N>(t, e)) = x[e];
    }
    fft<N, kE, +1>(x, lds, tid, tw, tab);
    {
        const int t = opaque(tid);
#pragma unroll
        for (int e = 0; e < kE; ++e) {
            ea[e] = cabs2(x[e]) * inv2;
            elem_u32(dst, out_pos<N>(t, e)) = ea[e];
        }
    }
    if (final_rows) {   // row 0 is the reference of the merge tree and never moves
        const int t = opaque(tid);
#pragma unroll
        for (int e = 0; e < kE; ++e) {
            const float2 v = cscale(x[e], invN);
            const int pos = out_pos<N>(t, e);
            elem_u32(s0a, pos) = v;
            const float av = cabs_hw(v);
            elem_u32(ap, pos) = av;
            mx = fmaxf(mx, av);
        }
    }
    // row 2p+1: S0 and energy
    load_row_raw<N>(x, J, f, rb, tid);
    {
        const int t = opaque(tid);
#pragma unroll
        for (int e = 0; e < kE; ++e) elem_u32(s0b, in_pos<N>(t, e)) = x[e];
    }
    fft<N, kE, +1>(x, lds, tid, tw, tab);
    // Above 4096 cells the two energy rows do not stay in registers across the correlation's
    // transforms: row 2p's is in its merged-energy row already, and row 2p+1's goes to that
    // row's energy slot, which no group at stride >= 2 uses (group heads are even rows). Each
    // thread reads back only what it wrote.
    constexpr bool kStash = N > 4096;
    float* eb_row = J.egrp + (static_cast<size_t>(f) * M + rb) * N;
    if constexpr (kStash) {
        const int t = opaque(tid);
#pragma unroll
        for (int e = 0; e < kE; ++e) elem_u32(eb_row, out_pos<N>(t, e)) = cabs2(x[e]) * inv2;
    }
    // the correlation takes the energies in the first-pass layout
#pragma unroll
    for (int e = 0; e < kE; ++e) x[e] = make_float2(ea[e], cabs2(x[e]) * inv2);
    relayout_out_to_in<N, kE>(x, lds, tid);
    float ta[kE], tb[kE];
#pragma unroll
    for (int e = 0; e < kE; ++e) { ta[e] = x[e].x; tb[e] = x[e].y; }
    const size_t idx = static_cast<size_t>(f) * (M - 1) + p;
    const int bi = correlation_argmax<N>(ta, tb, lds, redv, redi, tid, tw, tab);
    double lag;
    if constexpr (kStash) {
        const int t = opaque(tid);
#pragma unroll
        for (int e = 0; e < kE; ++e) {
            ta[e] = elem_u32(dst, out_pos<N>(t, e));
            tb[e] = elem_u32(eb_row, out_pos<N>(t, e));
        }
        lag = parabola_lag<N, true>(ta, tb, bi, lds, tid, J.lags + idx, J.imax + idx);
    } else {
        lag = parabola_lag<N>(ta, tb, bi, lds, tid, J.lags + idx, J.imax + idx);
    }
    // shift row 2p+1 by the lag and merge (range_alignment.cu:66-69)
    load_row_shifted<N>(x, J, f, rb, lag, tid);
    fft<N, kE, +1>(x, lds, tid, tw, tab);
    {
        const int t = opaque(tid);
#pragma unroll
        for (int e = 0; e < kE; ++e) elem_u32(dst, out_pos<N>(t, e)) += cabs2(x[e]) * inv2;
    }
    if (tid == 0) {
        J.lagsum[static_cast<size_t>(f) * M + ra] = 0.0;
        J.lagsum[static_cast<size_t>(f) * M + rb] = lag;
    }
    if (final_rows) {   // row 1 never moves again either
        const int t = opaque(tid);
#pragma unroll
        for (int e = 0; e < kE; ++e) {
            const float2 v = cscale(x[e], invN);
            const int pos = out_pos<N>(t, e);
            elem_u32(s0b, pos) = v;
            const float av = cabs_hw(v);
            elem_u32(ap, pos) += av;
            mx = fmaxf(mx, av);
        }
        mx = block_max<NT>(mx, redv);
        if (tid == 0) J.max_part[static_cast<size_t>(f) * (M / 2)] = mx;
    }
}

// One merge level s >= 2: group energies of pair p and their correlation (range_alignment.cu:
// 45-61). The first group's energy is kept for the next level.
template <int N>
__global__ __launch_bounds__(N / kE) void k_ra_corr(Job J, int s, int pair_base) {
    HIP_DYNAMIC_SHARED(float2, lds);
    float* redv = reinterpret_cast<float*>(lds + lds_padded_size(N));
    int* redi = reinterpret_cast<int*>(redv + 16);
    const int p = blockIdx.x, f = blockIdx.y, tid = threadIdx.x;
    const int M = J.M;
    const int ngroups = (M + s - 1) / s, npairs = ngroups / 2;
    const bool unpaired = (p == npairs);
    const int gA = unpaired ? ngroups - 1 : 2 * p;
    float a[kE], b[kE];
    group_energy<N>(a, J, f, gA, s, tid);
    const FftTwiddles<N, kE> tw = fft_twiddles<N, kE>(tid);
    float2* const tab = tw16_table<N>(lds);
    fft_tw16_table_fill(tab, tid, N / kE);
    {   // keep the first group's energy for the next level (same head row, same thread)
        float* dst = J.egrp + (static_cast<size_t>(f) * M + static_cast<size_t>(gA) * s) * N;
#pragma unroll
        for (int e = 0; e < kE; ++e) elem_u32(dst, in_pos<N>(tid, e)) = a[e];
    }
    if (unpaired) return;
    group_energy<N>(b, J, f, gA + 1, s, tid);
    const size_t idx = static_cast<size_t>(f) * (M - 1) + pair_base + p;
    const int bi = correlation_argmax<N>(a, b, lds, redv, redi, tid, tw, tab);
    // read both energies again rather than hold them across the transforms (the first from the
    // row this thread just wrote)
    {
        const float* src = J.egrp + (static_cast<size_t>(f) * M + static_cast<size_t>(gA) * s) * N;
        const int t = opaque(tid);
#pragma unroll
        for (int e = 0; e < kE; ++e) a[e] = elem_u32(src, in_pos<N>(t, e));
    }
    group_energy<N>(b, J, f, gA + 1, s, opaque(tid));
    (void)parabola_lag<N>(a, b, bi, lds, tid, J.lags + idx, J.imax + idx);
}

// Shift the rows of the second group of pair p by the pair's lag (range_alignment.cu:66-69)
// and accumulate their new energies; each workgroup handles up to J.ra_chunk rows. The rows of
// pair 0 move for the last time at this level: their aligned profiles replace their spectra and
// their column sums feed HRRP centring.
// PART selects the grid: kUpdAll, the chunks of every pair (final rows where p = 0, one launch
// per level); kUpdFinal, the chunks of pair 0 only; kUpdRest, those of pairs 1, 2, ... Apart
// from the final-row work the three run the same code; split into kUpdFinal + kUpdRest, the
// other pairs, most of the rows, do not carry that work's registers, at the cost of a second
// launch per level.
constexpr int kUpdAll = 0, kUpdFinal = 1, kUpdRest = 2;
template <int N, int PART>
__global__ __launch_bounds__(N / kE, PART == kUpdFinal && N <= 4096 ? 3 : 4) void k_ra_upd(Job J, int s, int pair_base) {
    HIP_DYNAMIC_SHARED(float2, lds);
    const int b = blockIdx.x, f = blockIdx.y, tid = threadIdx.x;
    const int M = J.M;
    const int R = min(s, J.ra_chunk);
    const int cpp = (s + R - 1) / R;
    const int p = b / cpp + (PART == kUpdRest ? 1 : 0), i = b % cpp;
    const int gend = min((2 * p + 2) * s, M);
    const int r0 = (2 * p + 1) * s + i * R;
    const int r1 = min(r0 + R, gend);
    if (r0 >= r1) return;
    const double lag = J.lags[static_cast<size_t>(f) * (M - 1) + pair_base + p];
    const float inv = 1.0f / (static_cast<float>(N) * static_cast<float>(N));
    const float invN = 1.0f / N;
    const bool final_rows = PART == kUpdFinal || (PART == kUpdAll && p == 0);
    float acc[kE];
    float mx = 0.0f;
#pragma unroll
    for (int e = 0; e < kE; ++e) acc[e] = 0.0f;
    float2 x[kE];
    const FftTwiddles<N, kE> tw = fft_twiddles<N, kE>(tid);
    float2* const tab = tw16_table<N>(lds);
    fft_tw16_table_fill(tab, tid, N / kE);
    // column sums of |profile| over the final rows of this chunk: kept in registers by the
    // final-row grid up to 4096 cells (above, its workgroup size leaves no room) and stored once,
    // otherwise accumulated in place (each location is read back only by the thread that wrote
    // it); the same additions in the same order either way
    constexpr bool kArpInRegs = PART == kUpdFinal && N <= 4096;
    float* ap = J.arp_part + (static_cast<size_t>(f) * (M / 2) + r0 / 2) * N;
    float arp[kArpInRegs ? kE : 1];
#pragma unroll
    for (int e = 0; e < (kArpInRegs ? kE : 1); ++e) arp[e] = 0.0f;
    for (int m = r0; m < r1; ++m) {
        const double L = J.lagsum[static_cast<size_t>(f) * M + m] + lag;
        load_row_shifted<N>(x, J, f, m, L, tid);
        fft<N, kE, +1>(x, lds, tid, tw, tab);
#pragma unroll
        for (int e = 0; e < kE; ++e) acc[e] += cabs2(x[e]) * inv;
        if (final_rows) {   // S0 of row m was read before the transform's first barrier
            float2* st = J.st + (static_cast<size_t>(f) * M + m) * N;
            const int t = opaque(tid);
#pragma unroll
            for (int e = 0; e < kE; ++e) {
                const float2 v = cscale(x[e], invN);
                const int pos = out_pos<N>(t, e);
                elem_u32(st, pos) = v;
                const float av = cabs_hw(v);
                if constexpr (kArpInRegs) arp[e] += av;
                else elem_u32(ap, pos) = (m == r0 ? 0.0f : elem_u32(ap, pos)) + av;
                mx = fmaxf(mx, av);
            }
        }
    }
    if (final_rows) {
        if constexpr (kArpInRegs) {
            const int t = opaque(tid);
#pragma unroll
            for (int e = 0; e < kE; ++e) elem_u32(ap, out_pos<N>(t, e)) = arp[e];
        }
        float* red = reinterpret_cast<float*>(lds + lds_padded_size(N));
        mx = block_max<N / kE>(mx, red);
        if (tid == 0) J.max_part[static_cast<size_t>(f) * (M / 2) + r0 / 2] = mx;
    }
    __syncthreads();
    if (tid < r1 - r0) J.lagsum[static_cast<size_t>(f) * M + r0 + tid] += lag;
    float* dst = J.epart + (static_cast<size_t>(f) * M + r0) * N;
#pragma unroll
    for (int e = 0; e < kE; ++e) elem_u32(dst, out_pos<N>(tid, e)) = acc[e];
}

// Range Hamming window 0.54 - 0.46 cos(2 pi n / (N - 1)) (genHammingVec, common.cu:260-266),
// evaluated in FP64 and rounded once to FP32.
__global__ __launch_bounds__(256) void k_range_window(int N, float* win) {
    const int n = blockIdx.x * 256 + threadIdx.x;
    if (n < N) win[n] = static_cast<float>(0.54 - 0.46 * cos(2.0 * 3.14159265358979323846 * n / (N - 1)));
}

// ------------------------------------------------------------------------------------------
// host launchers
// ------------------------------------------------------------------------------------------
namespace {
template <int N>
size_t ra_lds() { return lds_padded_size(N) * sizeof(float2) + 64 * sizeof(float) + 256 * sizeof(float2); }

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
    set(reinterpret_cast<const void*>(&k_hrrp<NN>), ra_lds<NN>());             \
    set(reinterpret_cast<const void*>(&k_ra_pair0<NN>), ra_lds<NN>());         \
    set(reinterpret_cast<const void*>(&k_ra_corr<NN>), ra_lds<NN>());          \
    set(reinterpret_cast<const void*>(&k_ra_upd<NN, kUpdAll>), ra_lds<NN>());   \
    set(reinterpret_cast<const void*>(&k_ra_upd<NN, kUpdFinal>), ra_lds<NN>()); \
    set(reinterpret_cast<const void*>(&k_ra_upd<NN, kUpdRest>), ra_lds<NN>());
    ISAR_SET_RA(1024) ISAR_SET_RA(2048) ISAR_SET_RA(4096) ISAR_SET_RA(8192) ISAR_SET_RA(16384)
#undef ISAR_SET_RA
#endif
    return err;
}

void launch_range_window(int N, float* win, hipStream_t s) {
    ISAR_LAUNCH(k_range_window, dim3((N + 255) / 256), dim3(256), 0, s, N, win);
}

void launch_ifds_pre(const Job& J, hipStream_t s) {
    ISAR_RANGE_SWITCH(J.N, ISAR_LAUNCH(k_ifds_pre<NN>, dim3(J.M, J.K), dim3(NN / kE), ra_lds<NN>(), s, J));
}

void launch_hrrp(const Job& J, hipStream_t s) {
    ISAR_RANGE_SWITCH(J.N, ISAR_LAUNCH(k_hrrp<NN>, dim3(J.M, J.K), dim3(NN / kE), ra_lds<NN>(), s, J));
}
void launch_ra_pair0(const Job& J, hipStream_t s) {
    ISAR_RANGE_SWITCH(J.N, ISAR_LAUNCH(k_ra_pair0<NN>, dim3(J.M / 2, J.K), dim3(NN / kE), ra_lds<NN>(), s, J));
}
void launch_ra_corr(const Job& J, int stride, int pair_base, int nblocks, hipStream_t s) {
    ISAR_RANGE_SWITCH(J.N, ISAR_LAUNCH(k_ra_corr<NN>, dim3(nblocks, J.K), dim3(NN / kE), ra_lds<NN>(), s, J,
                                       stride, pair_base));
}
void launch_ra_upd(const Job& J, int stride, int pair_base, int nblocks, hipStream_t s) {
    // nblocks = pairs x chunks per pair. Large batches launch pair 0 (final rows) and the other
    // pairs apart; small ones, where the second launch per level costs more than it saves, in one.
    const bool split = J.ra_split == 2 || (J.ra_split == 0 && static_cast<long long>(J.K) * J.M * J.N >= kRaSplitCells);
    if (!split) {
        ISAR_RANGE_SWITCH(J.N, ISAR_LAUNCH((k_ra_upd<NN, kUpdAll>), dim3(nblocks, J.K), dim3(NN / kE), ra_lds<NN>(),
                                           s, J, stride, pair_base));
        return;
    }
    const int R = std::min(stride, J.ra_chunk), cpp = (stride + R - 1) / R;
    const int rest = nblocks - cpp;
    ISAR_RANGE_SWITCH(J.N, ISAR_LAUNCH((k_ra_upd<NN, kUpdFinal>), dim3(cpp, J.K), dim3(NN / kE), ra_lds<NN>(), s, J,
                                       stride, pair_base));
    if (rest > 0)
        ISAR_RANGE_SWITCH(J.N, ISAR_LAUNCH((k_ra_upd<NN, kUpdRest>), dim3(rest, J.K), dim3(NN / kE), ra_lds<NN>(), s,
                                           J, stride, pair_base));
}

}  // namespace isar
