// center.hip — HRRP centring (range_alignment.cu:371-465) and range-window cut
// (common.cu:494-518) with every decision taken on the device (no host round trips).
#include "../dev_common.hpp"
#include "../launchers.hpp"

namespace isar {

namespace {
constexpr int kCenterThreads = 1024;
constexpr int kInterLength = 30;   // isar_main.cu:98
constexpr int kArpWindow = 8;      // range_alignment.cu:434

// ascending bitonic sort of n (power of two) floats in LDS by the whole workgroup
template <int NT>
__device__ void bitonic_sort_f(float* a, int n) {
    for (int k = 2; k <= n; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < n; i += NT) {
                const int l = i ^ j;
                if (l > i) {
                    const float x = a[i], y = a[l];
                    const bool up = (i & k) == 0;
                    if ((x > y) == up) { a[i] = y; a[l] = x; }
                }
            }
            __syncthreads();
        }
    }
}

// The same compare-exchange network (stage (k, j) orders each pair i < i ^ j ascending where
// i & k == 0, descending otherwise, swapping when (a[i] > a[i ^ j]) == ascending), with the stages
// inside a wave taken out of LDS: thread t holds elements t E ... t E + E - 1 in registers, stages
// j < E exchange within the thread, stages E <= j < 64 E with the partner thread t ^ (j / E) of
// the same wave by shuffles, and only the stages j >= 64 E go through LDS with a barrier each
// (10 of the 105 stages at 16384 elements). Every pair meets the same comparator in the same
// stage as in bitonic_sort_f, so the sorted array is the same bit for bit, NaNs included.
// n = NT E; every thread of the workgroup calls it, with a in LDS and a barrier after its fill.
template <int NT, int E>
__device__ void bitonic_sort_f_waves(float* a) {
    constexpr int n = NT * E, WE = 64 * E;
    const int t = threadIdx.x, i0 = t * E;
    float v[E];
    bool in_regs = false;
    for (int k = 2; k <= n; k <<= 1) {
        int j = k >> 1;
        for (; j >= WE; j >>= 1) {   // partner in another wave
            if (in_regs) {
#pragma unroll
                for (int e = 0; e < E; ++e) a[i0 + e] = v[e];
                __syncthreads();
                in_regs = false;
            }
            for (int i = t; i < n; i += NT) {
                const int l = i ^ j;
                if (l > i) {
                    const float x = a[i], y = a[l];
                    const bool up = (i & k) == 0;
                    if ((x > y) == up) { a[i] = y; a[l] = x; }
                }
            }
            __syncthreads();
        }
        if (!in_regs) {   // the last LDS stage (or the caller's fill) ended with a barrier
#pragma unroll
            for (int e = 0; e < E; ++e) v[e] = a[i0 + e];
            in_regs = true;
        }
        for (; j >= E; j >>= 1) {   // partner thread t ^ (j / E) of the same wave
            const bool lower = (i0 & j) == 0;
            const bool up = (i0 & k) == 0;
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const float p = __shfl_xor(v[e], j / E, 64);
                // x = the lower index's value, y = the higher one's
                const float x = lower ? v[e] : p, y = lower ? p : v[e];
                const bool sw = (x > y) == up;
                v[e] = sw ? p : v[e];
            }
        }
#pragma unroll
        for (int jj = E / 2; jj >= 1; jj >>= 1) {   // within the thread (jj < k)
            if (jj < k) {
#pragma unroll
                for (int e = 0; e < E; ++e) {
                    if ((e & jj) == 0) {
                        const float x = v[e], y = v[e | jj];
                        const bool up = ((i0 + e) & k) == 0;
                        if ((x > y) == up) { v[e] = y; v[e | jj] = x; }
                    }
                }
            }
        }
    }
#pragma unroll
    for (int e = 0; e < E; ++e) a[i0 + e] = v[e];
    __syncthreads();
}

template <int NT>
__device__ void sort_arp(float* a, int n) {
    switch (n) {
    case 1024: bitonic_sort_f_waves<NT, 1024 / NT>(a); break;
    case 2048: bitonic_sort_f_waves<NT, 2048 / NT>(a); break;
    case 4096: bitonic_sort_f_waves<NT, 4096 / NT>(a); break;
    case 8192: bitonic_sort_f_waves<NT, 8192 / NT>(a); break;
    case 16384: bitonic_sort_f_waves<NT, 16384 / NT>(a); break;
    default: bitonic_sort_f<NT>(a, n); break;
    }
}
}  // namespace

// Rows of arp_part / max_part holding partials, in summation order: row 0 (rows 0-1) and, at
// every merge stride s >= 2, the chunks of min(s, ra_chunk) rows starting at s (the second group
// of pair 0), stored at head row / 2. At most M/2 rows.
__device__ int arp_partial_rows(const Job& J, int* rows) {
    int c = 0;
    rows[c++] = 0;
    for (int s = 2; s < J.M; s *= 2)
        for (int h = s; h < min(2 * s, J.M); h += min(s, J.ra_chunk)) rows[c++] = h / 2;
    return c;
}

// Column sums of |S| over all M rows (the ARP numerator, range_alignment.cu:381-399) from the
// partials range alignment wrote as rows finished moving, added in the order of their rows; one
// thread per range cell, eight independent loads in flight. The sums go to the frame's first egrp
// row, free once range alignment has finished.
__global__ __launch_bounds__(256) void k_arp(Job J) {
    HIP_DYNAMIC_SHARED(int, rows);   // M/2 + 1
    const int f = blockIdx.y, n = blockIdx.x * 256 + threadIdx.x;
    const int M = J.M, N = J.N;
    if (threadIdx.x == 0) rows[M / 2] = arp_partial_rows(J, rows);
    __syncthreads();
    if (n >= N) return;
    const float* ap = J.arp_part + static_cast<size_t>(f) * (M / 2) * N + n;
    const int P = rows[M / 2];
    float acc = 0.0f;
    int i = 0;
    for (; i + 8 <= P; i += 8) {
        float v[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) v[q] = ap[static_cast<size_t>(rows[i + q]) * N];
#pragma unroll
        for (int q = 0; q < 8; ++q) acc += v[q];
    }
    for (; i < P; ++i) acc += ap[static_cast<size_t>(rows[i]) * N];
    J.egrp[static_cast<size_t>(f) * M * N + n] = acc;
}

// One workgroup per frame. LDS: arp[N], sorted[N], reduction scratch, partial-row list.
__global__ __launch_bounds__(kCenterThreads) void k_center(Job J) {
    HIP_DYNAMIC_SHARED(float, sm);
    constexpr int NT = kCenterThreads;
    const int f = blockIdx.x, tid = threadIdx.x;
    const int M = J.M, N = J.N, W = J.W;
    float* arp = sm;
    float* srt = sm + N;
    float* redv = srt + N;
    int* redi = reinterpret_cast<int*>(redv + 32);
    int* rows = redi + 32;   // M/2 + 1
    // max |S| arrives as partials per chunk of rows that finished moving (arp_partial_rows), the
    // column sums of |S| from k_arp
    const float* mp = J.max_part + static_cast<size_t>(f) * (M / 2);
    const float* colsum = J.egrp + static_cast<size_t>(f) * M * N;
    if (tid == 0) rows[M / 2] = arp_partial_rows(J, rows);
    __syncthreads();

    // H = |S| / max|S| ; ARP = column sums / M  (range_alignment.cu:381-399, all M rows)
    float gm = 0.0f;
    for (int i = tid; i < rows[M / 2]; i += NT) gm = fmaxf(gm, mp[rows[i]]);
    gm = block_max<NT>(gm, redv);
    const float inv = gm > 0.0f ? 1.0f / gm : 0.0f;
    const float invM = 1.0f / static_cast<float>(M);
    for (int n = tid; n < N; n += NT) {
        const float v = colsum[n] * inv * invM;
        arp[n] = v;
        srt[n] = v;
    }
    __syncthreads();
    sort_arp<NT>(srt, N);

    // threshold from the sorted ARP (range_alignment.cu:402-423)
    const float amax = fabsf(srt[N - 1]);
    const float extra = static_cast<float>(kInterLength) * amax / static_cast<float>(N);
    float dmin = 3.4e38f;
    int imin = 0x7fffffff;
    for (int i = tid; i < N - kInterLength; i += NT) {
        const float d = fabsf(srt[i + kInterLength] - srt[i] - extra);
        if (d < dmin || (d == dmin && i < imin)) { dmin = d; imin = i; }
    }
    block_argmin<NT>(dmin, imin, redv, redi);
    int status = 0;
    if (imin < 0 || imin >= N - kInterLength) {   // every difference was NaN
        imin = 0;
        status |= 1;
    }
    const float thr = srt[imin + kInterLength / 2];

    // idx_1 = find(arp > thr); windowed means (getARPMean, :524-550); idx_2 = find(mean < thr)
    int c1 = 0, c2 = 0;
    long long keep_sum = 0;
    for (int n = tid; n < N; n += NT) {
        if (arp[n] > thr) {
            float acc = 0.0f;
            float mean;
            if (n - kArpWindow / 2 >= 0 && n + kArpWindow / 2 <= N - 1) {
                for (int q = n - kArpWindow / 2; q <= n + kArpWindow / 2; ++q) acc += arp[q];
                mean = acc / static_cast<float>(kArpWindow + 1);
            } else if (n - kArpWindow / 2 < 0) {
                for (int q = 0; q <= n + kArpWindow / 2; ++q) acc += arp[q];
                mean = acc / (static_cast<float>(kArpWindow) / 2.0f + static_cast<float>(n) + 1.0f);
            } else {
                for (int q = n - kArpWindow / 2; q < N; ++q) acc += arp[q];
                mean = acc / (static_cast<float>(N) - static_cast<float>(n) - static_cast<float>(kArpWindow) / 2.0f - 1.0f);
            }
            ++c1;
            if (mean < thr) ++c2;
            else keep_sum += n;
        }
    }
    const int len1 = static_cast<int>(block_sum<NT>(static_cast<float>(c1), redv) + 0.5f);
    const int len2 = static_cast<int>(block_sum<NT>(static_cast<float>(c2), redv) + 0.5f);
    // the kept-index sum can exceed FP32 precision: reduce it as two exact halves
    const float ks_hi = block_sum<NT>(static_cast<float>(keep_sum >> 12), redv);
    const float ks_lo = block_sum<NT>(static_cast<float>(keep_sum & 4095), redv);
    const long long ksum = (static_cast<long long>(ks_hi + 0.5f) << 12) + static_cast<long long>(ks_lo + 0.5f);
    int shift = 0, applied = 0;
    if (len2 != len1) {
        const int mean_idx = static_cast<int>(ksum / (len1 - len2));
        shift = -(mean_idx - N / 2);
        applied = 1;
    }

    // cut offset: cublasIcamax over row 0 after the circular shift (|Re| + |Im|, first index)
    const float2* row0 = J.st + static_cast<size_t>(f) * M * N;
    float best = -1.0f;
    int bi = 0x7fffffff;
    for (int n = tid; n < N; n += NT) {
        const float2 v = row0[((n - shift) % N + N) % N];
        const float key = fabsf(v.x) + fabsf(v.y);
        if (key > best || (key == best && n < bi)) { best = key; bi = n; }
    }
    block_argmax<NT>(best, bi, redv, redi);
    int off = bi - W / 2;
    if (off < 0) off = 0;
    if (off > N - W) off = N - W;
    if (tid == 0) {
        FrameDecisions& d = J.dec[f];
        d.center_shift = shift;
        d.center_applied = applied;
        d.idx1_len = len1;
        d.idx2_len = len2;
        d.diff_imin = imin;
        d.cut_offset = off;
        d.center_thr = thr;
        d.gmax = gm;
        d.status = status;
    }
}

// Cut: X[m, j] = S[m, (off + j - shift) mod N] for a band of rows; column maxima of |X| per
// band (fastEntropy :180-184) and the adjacent-pulse products sum_j conj(X[m,j]) X[m+1,j]
// (dopplerTracking :12-20). Block = W threads, one per range bin of the window. Each wave's sum
// of a row's products goes to LDS; one barrier after the band, then thread r adds row r's wave
// sums in wave order.
template <int W>
__global__ __launch_bounds__(W) void k_cut(Job J) {
    HIP_DYNAMIC_SHARED(float2, wred);   // (W/64) x kCutBandRows wave sums
    const int band = blockIdx.x, f = blockIdx.y, j = threadIdx.x;
    const int M = J.M, N = J.N;
    const int nbands = (M + kCutBandRows - 1) / kCutBandRows;
    const FrameDecisions d = J.dec[f];
    const int q = (((d.cut_offset + j - d.center_shift) % N) + N) % N;
    const float2* src = J.st + static_cast<size_t>(f) * M * N + q;
    float2* xo = J.x + static_cast<size_t>(f) * M * W + j;
    const int m0 = band * kCutBandRows, m1 = min(m0 + kCutBandRows, M);
    float cmax = 0.0f;
    float2 cur = src[static_cast<size_t>(m0) * N];
    for (int m = m0; m < m1; ++m) {
        xo[static_cast<size_t>(m) * W] = cur;
        cmax = fmaxf(cmax, cabsf_(cur));
        float2 prod = make_float2(0.0f, 0.0f);
        float2 nxt = make_float2(0.0f, 0.0f);
        if (m + 1 < M) {
            nxt = src[static_cast<size_t>(m + 1) * N];
            prod = cconjmul(cur, nxt);
        }
        const float2 sum = wave_sum2(prod);
        if ((j & 63) == 0) wred[(j >> 6) * kCutBandRows + (m - m0)] = sum;
        cur = nxt;
    }
    J.colmax_part[(static_cast<size_t>(f) * nbands + band) * W + j] = cmax;
    __syncthreads();
    if (j < m1 - m0 && m0 + j + 1 < M) {
        float2 t = make_float2(0.0f, 0.0f);
        for (int k = 0; k < W / 64; ++k) {
            const float2 v = wred[k * kCutBandRows + j];
            t.x += v.x;
            t.y += v.y;
        }
        J.rowdop[static_cast<size_t>(f) * M + m0 + j] = t;
    }
}

hipError_t configure_center_kernels() {
#ifndef ISAR_EMU
    const size_t bytes = (2 * 16384 + 64) * sizeof(float) + (1024 / 2 + 1) * sizeof(int);
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_center), hipFuncAttributeMaxDynamicSharedMemorySize,
                                       static_cast<int>(bytes));
    if (e != hipSuccess) report_config_error(__FILE__, bytes);
    return e;
#else
    return hipSuccess;
#endif
}

void launch_center(const Job& J, hipStream_t s) {
    const size_t rows = (static_cast<size_t>(J.M) / 2 + 1) * sizeof(int);
    ISAR_LAUNCH(k_arp, dim3((J.N + 255) / 256, J.K), dim3(256), rows, s, J);
    const size_t lds = (2 * static_cast<size_t>(J.N) + 64) * sizeof(float) + rows;
    ISAR_LAUNCH(k_center, dim3(J.K), dim3(kCenterThreads), lds, s, J);
}

void launch_cut(const Job& J, hipStream_t s) {
    const int nbands = (J.M + kCutBandRows - 1) / kCutBandRows;
    if (J.W == kRangeCut)
        ISAR_LAUNCH(k_cut<kRangeCut>, dim3(nbands, J.K), dim3(kRangeCut), (kRangeCut / 64) * kCutBandRows * sizeof(float2),
                    s, J);
}

}  // namespace isar
