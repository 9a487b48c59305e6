// isar_oracle.cpp — FP64 CPU restatement of the reference imaging chain (test infrastructure).
//
// The HIP pipeline is tested against this oracle, and bench.py times it as the CPU baseline. It
// is never linked into the product library. Each function cites the reference lines it restates;
// paths are relative to the reference repository root.
//
// Deviations from the reference (each one resolves undefined or NaN-producing behaviour; all
// deterministic reference behaviour is reproduced):
//   Q2  conjMulAveProfile launched with N blocks writes past the half spectrum: only k < N/2+1.
//   Q3  cublasDger accumulates into uninitialised memory: phase = coef * v (x) tk^2 exactly.
//   Q4  CZT chirp padding uninitialised; FP32 thrust::pow chirps: exact FP64 chirp-z transform.
//   Q5  sumCols/maxCols launched with 256 threads: reductions run over all M rows.
//   Q6  getARPMean has no bounds guard: one mean per idx_1 entry (quirky edge divisors kept).
//   Q8  host dereference of a device pointer: entropies of the selected bins are gathered.
//   Q9a parabola at the row edge or with zero curvature: offset 0.
//   Q10 non-power-of-two M: the last odd group is carried to the next level unpaired.
//   FP32 phase ramps exp(-j 2 pi k s / N) with k*s up to ~1e7: argument reduced exactly.
//   circshiftFreq (FFT, ramp, IFFT) with an integer shift: exact circular shift.
//   cos(PI_FLT * n) in the centring vector: exact (-1)^n.
//   argmax/argmin ties: first index (cuBLAS I?amax/I?amin semantics); the block-tree argmax of
//   maxRowsIdxABS is resolved the same way.
//   NaN safety: 0 * log(0) := 0; a zero column energy gives entropy +inf (sorted last);
//   x / |x| with x == 0 gives 0.
#include "isar_oracle.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <map>
#include <numeric>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

using cd = std::complex<double>;
constexpr double kPi = 3.14159265358979323846;
constexpr double kLightSpeed = 3e8;  // LIGHT_SPEED, common.cuh:69
constexpr int kInterLength = 30;     // isar_main.cu:98
constexpr int kArpWindow = 8;        // range_alignment.cu:434

// ------------------------------------------------------------------------------------------
// FP64 FFT: iterative radix-2 for powers of two, Bluestein otherwise. Unnormalised both ways.
// ------------------------------------------------------------------------------------------
bool is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

struct Pow2Plan {
    int n = 0;
    std::vector<cd> tw;          // exp(-j 2 pi k / n), k < n/2
    std::vector<int> rev;
};

const Pow2Plan& pow2_plan(int n) {
    thread_local std::map<int, Pow2Plan> cache;
    auto it = cache.find(n);
    if (it != cache.end()) return it->second;
    Pow2Plan p;
    p.n = n;
    p.tw.resize(std::max(1, n / 2));
    for (int k = 0; k < n / 2; ++k) {
        double a = -2.0 * kPi * static_cast<double>(k) / n;
        p.tw[k] = cd(std::cos(a), std::sin(a));
    }
    p.rev.resize(n);
    int lg = 0;
    while ((1 << lg) < n) ++lg;
    for (int i = 0; i < n; ++i) {
        int r = 0;
        for (int b = 0; b < lg; ++b)
            if (i & (1 << b)) r |= 1 << (lg - 1 - b);
        p.rev[i] = r;
    }
    return cache.emplace(n, std::move(p)).first->second;
}

void fft_pow2(cd* x, int n, int dir) {
    if (n <= 1) return;
    const Pow2Plan& p = pow2_plan(n);
    for (int i = 0; i < n; ++i)
        if (i < p.rev[i]) std::swap(x[i], x[p.rev[i]]);
    for (int len = 2; len <= n; len <<= 1) {
        int half = len >> 1, step = n / len;
        for (int i = 0; i < n; i += len) {
            for (int j = 0; j < half; ++j) {
                cd w = p.tw[j * step];
                if (dir > 0) w = std::conj(w);
                cd u = x[i + j], v = x[i + j + half] * w;
                x[i + j] = u + v;
                x[i + j + half] = u - v;
            }
        }
    }
}

// exp(-j pi k^2 / n) with k^2 reduced modulo 2n in integers (exact argument).
cd chirp(long long k, int n) {
    long long r = (k * k) % (2LL * n);
    double a = -kPi * static_cast<double>(r) / n;
    return cd(std::cos(a), std::sin(a));
}

void fft_any(cd* x, int n, int dir) {
    if (n <= 1) return;
    if (is_pow2(n)) {
        fft_pow2(x, n, dir);
        return;
    }
    int L = 1;
    while (L < 2 * n - 1) L <<= 1;
    std::vector<cd> a(L, cd(0, 0)), b(L, cd(0, 0)), c(n);
    for (int k = 0; k < n; ++k) {
        c[k] = chirp(k, n);
        if (dir > 0) c[k] = std::conj(c[k]);
    }
    for (int k = 0; k < n; ++k) a[k] = x[k] * c[k];
    b[0] = std::conj(c[0]);
    for (int k = 1; k < n; ++k) b[k] = b[L - k] = std::conj(c[k]);
    fft_pow2(a.data(), L, -1);
    fft_pow2(b.data(), L, -1);
    for (int k = 0; k < L; ++k) a[k] *= b[k];
    fft_pow2(a.data(), L, +1);
    for (int k = 0; k < n; ++k) x[k] = a[k] * c[k] / static_cast<double>(L);
}

// Column FFT of a rows x cols row-major matrix (cuFFT plan with stride = cols, batch = cols).
void fft_cols(cd* x, int rows, int cols, int dir) {
#pragma omp parallel
    {
        std::vector<cd> col(rows);
#pragma omp for schedule(static)
        for (int c = 0; c < cols; ++c) {
            for (int r = 0; r < rows; ++r) col[r] = x[static_cast<size_t>(r) * cols + c];
            fft_any(col.data(), rows, dir);
            for (int r = 0; r < rows; ++r) x[static_cast<size_t>(r) * cols + c] = col[r];
        }
    }
}

void fft_rows(cd* x, int rows, int cols, int dir) {
#pragma omp parallel for schedule(static)
    for (int r = 0; r < rows; ++r) fft_any(x + static_cast<size_t>(r) * cols, cols, dir);
}

// ifftshiftRows / ifftshiftCols (common.cuh:208-244): swap the two halves (floor(len/2)).
void swap_halves_rows(cd* x, int rows, int cols) {
    int h = cols >> 1;
    for (int r = 0; r < rows; ++r) {
        cd* row = x + static_cast<size_t>(r) * cols;
        for (int i = 0; i < h; ++i) std::swap(row[i], row[i + h]);
    }
}
void swap_halves_cols(cd* x, int rows, int cols) {
    int h = rows >> 1;
    for (int r = 0; r < h; ++r)
        for (int c = 0; c < cols; ++c)
            std::swap(x[static_cast<size_t>(r) * cols + c], x[static_cast<size_t>(r + h) * cols + c]);
}
// fliplr / flipud (common.cuh:255-296)
void flip_rows(cd* x, int rows, int cols) {
    for (int r = 0; r < rows; ++r) std::reverse(x + static_cast<size_t>(r) * cols, x + static_cast<size_t>(r + 1) * cols);
}
void flip_cols(cd* x, int rows, int cols) {
    for (int r = 0; r < rows / 2; ++r)
        for (int c = 0; c < cols; ++c)
            std::swap(x[static_cast<size_t>(r) * cols + c], x[static_cast<size_t>(rows - 1 - r) * cols + c]);
}

// genHammingVec (common.cu:260-266)
std::vector<double> hamming(int len) {
    std::vector<double> h(len);
    for (int n = 0; n < len; ++n) h[n] = 0.54 - 0.46 * std::cos(2.0 * kPi * n / (len - 1));
    return h;
}

// exp(-j 2 pi k s / n) with the argument k*s/n reduced to [0, 1) before scaling.
cd ramp(long long k, double s, int n) {
    double t = static_cast<double>(k) * s / n;
    t -= std::floor(t);
    double a = -2.0 * kPi * t;
    return cd(std::cos(a), std::sin(a));
}

void put_c(double* dst, cd v) { dst[0] = v.real(); dst[1] = v.imag(); }

// x / |x| with 0 -> 0 (the reference divides 0/0 and propagates NaN).
cd unit(cd v) {
    double a = std::abs(v);
    return a > 0 ? v / a : cd(0, 0);
}

}  // namespace

extern "C" {

int orc_next_pow2(int N) {
    // common.cu:521-529: returns 2^(floor(log2 N)+1); 2n for exact powers of two, 2 for 0 and 1.
    int n = 1;
    while (N >> 1) {
        n <<= 1;
        N >>= 1;
    }
    return n << 1;
}

void orc_fft(double* x, int n, int dir) { fft_any(reinterpret_cast<cd*>(x), n, dir); }

void orc_czt(const double* x, int m, int k, double w_re, double w_im, double a_re, double a_im,
             double* out) {
    // Direct evaluation of X[k] = sum_m x[m] a^-m w^(mk): the transform cztRange computes with
    // Bluestein's algorithm (mtrc.cu:125-172).
    const cd* xc = reinterpret_cast<const cd*>(x);
    cd w(w_re, w_im), a(a_re, a_im);
    double lw = std::arg(w), la = std::arg(a);
    double rw = std::abs(w), ra = std::abs(a);
    for (int kk = 0; kk < k; ++kk) {
        cd acc(0, 0);
        for (int mm = 0; mm < m; ++mm) {
            double ph = lw * static_cast<double>(mm) * kk - la * mm;
            double mag = std::pow(rw, static_cast<double>(mm) * kk) * std::pow(ra, -static_cast<double>(mm));
            acc += xc[mm] * std::polar(mag, ph);
        }
        put_c(out + 2 * kk, acc);
    }
}

void orc_czt_alpha(const double* x, int m, double alpha, double* out) {
    // w = exp(-j 2 pi alpha / M), a = exp(-j pi alpha)  (mtrc.cu:175-192), so
    // X[k] = sum_m x[m] exp(-j 2 pi alpha m (k - M/2) / M).  The phase alpha*m*(2k - M)/M is
    // formed in FP64 from the exact integer m*(2k - M) and reduced modulo 2 before sincos.
    const cd* xc = reinterpret_cast<const cd*>(x);
    for (int k = 0; k < m; ++k) {
        cd acc(0, 0);
        for (int mm = 0; mm < m; ++mm) {
            double t = alpha * static_cast<double>(static_cast<long long>(mm) * (2LL * k - m)) / m;
            t -= 2.0 * std::floor(t * 0.5);
            double a = -kPi * t;
            acc += xc[mm] * cd(std::cos(a), std::sin(a));
        }
        put_c(out + 2 * k, acc);
    }
}

void orc_hpc(double* s, const double* velocity, const orc_params* p) {
    // high_speed_compensation.cu:4-54 (Q3: the phase matrix is formed exactly).
    const int M = p->echo_num, N = p->range_num;
    const double chirp_rate = static_cast<double>(p->band_width) / p->Tp;
    const double coef = -4.0 * kPi * chirp_rate / kLightSpeed;
    const double Fs = static_cast<double>(p->Fs);
    cd* x = reinterpret_cast<cd*>(s);
#pragma omp parallel for schedule(static)
    for (int m = 0; m < M; ++m) {
        for (int n = 0; n < N; ++n) {
            double tk = static_cast<double>(n) / Fs;
            double ph = coef * velocity[m] * tk * tk;
            x[static_cast<size_t>(m) * N + n] *= cd(std::cos(ph), std::sin(ph));
        }
    }
}

}  // extern "C"

namespace {
// FP32 radix-2 transform with twiddles rounded once from FP64, for the range_fft32 option.
thread_local int t_range_fft32 = 0;
void fft_pow2_f32(cd* x, int n, int dir) {
    std::vector<std::complex<float>> v(n);
    for (int i = 0; i < n; ++i) v[i] = std::complex<float>(static_cast<float>(x[i].real()), static_cast<float>(x[i].imag()));
    for (int i = 1, j = 0; i < n; ++i) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) std::swap(v[i], v[j]);
    }
    for (int len = 2; len <= n; len <<= 1)
        for (int i = 0; i < n; i += len)
            for (int k = 0; k < len / 2; ++k) {
                const double a = dir * 2.0 * kPi * k / len;
                const std::complex<float> w(static_cast<float>(std::cos(a)), static_cast<float>(std::sin(a)));
                const std::complex<float> u = v[i + k], t = v[i + k + len / 2] * w;
                v[i + k] = u + t;
                v[i + k + len / 2] = u - t;
            }
    for (int i = 0; i < n; ++i) x[i] = cd(v[i].real(), v[i].imag());
}
}  // namespace

extern "C" {

void orc_range_alignment(double* s_, const orc_params* p, orc_trace* tr) {
    // range_alignment.cu:4-78, with the group structure generalised to any M (Q10).
    const int M = p->echo_num, N = p->range_num;
    cd* S = reinterpret_cast<cd*>(s_);
    std::vector<double> ham = hamming(N);
    // :13-18 frequency centring: hamming[n] * cos(pi n) == hamming[n] * (-1)^n
    for (int m = 0; m < M; ++m)
        for (int n = 0; n < N; ++n) S[static_cast<size_t>(m) * N + n] *= ham[n] * ((n & 1) ? -1.0 : 1.0);

    std::vector<double> P(static_cast<size_t>(M) * N);
    std::vector<double> g(static_cast<size_t>(M) * N);
    const int fft32 = t_range_fft32;   // read before the parallel regions (thread-local)
    int pair_index = 0;
    for (int s = 1; s < M; s *= 2) {
        // :38-42 profile magnitude of every row: |IFFT(S)| / N
#pragma omp parallel
        {
            std::vector<cd> row(N);
#pragma omp for schedule(static)
            for (int m = 0; m < M; ++m) {
                std::copy(S + static_cast<size_t>(m) * N, S + static_cast<size_t>(m + 1) * N, row.begin());
                if (fft32) fft_pow2_f32(row.data(), N, +1);
                else fft_pow2(row.data(), N, +1);
                for (int n = 0; n < N; ++n) P[static_cast<size_t>(m) * N + n] = std::abs(row[n]) / N;
            }
        }
        // :45 group profile: sqrt of the sum of squares over the rows of each group
        const int ngroups = (M + s - 1) / s;
        for (int j = 0; j < ngroups; ++j) {
            int r0 = j * s, r1 = std::min(M, r0 + s);
            for (int n = 0; n < N; ++n) {
                double acc = 0;
                for (int r = r0; r < r1; ++r) acc += P[static_cast<size_t>(r) * N + n] * P[static_cast<size_t>(r) * N + n];
                g[static_cast<size_t>(r0) * N + n] = std::sqrt(acc);
            }
        }
        // :50-61 correlation of each pair of groups, argmax and parabolic refinement
        const int npairs = ngroups / 2;
        std::vector<double> lag(npairs);
#pragma omp parallel
        {
            std::vector<cd> a(N), b(N);
            std::vector<double> c(N);
#pragma omp for schedule(static)
            for (int pp = 0; pp < npairs; ++pp) {
                const double* g1 = &g[static_cast<size_t>(2 * pp * s) * N];
                const double* g2 = &g[static_cast<size_t>((2 * pp + 1) * s) * N];
                for (int n = 0; n < N; ++n) { a[n] = cd(g1[n], 0); b[n] = cd(g2[n], 0); }
                fft_pow2(a.data(), N, -1);   // R2C (full spectrum; Hermitian half used by ref)
                fft_pow2(b.data(), N, -1);
                for (int k = 0; k < N; ++k) a[k] *= std::conj(b[k]);   // conjMulAveProfile
                fft_pow2(a.data(), N, +1);   // C2R, unnormalised
                for (int i = 0; i < N; ++i) c[i] = a[(i + N / 2) % N].real();   // ifftshiftRows
                // maxRowsIdxABS (:134-168): first index of max |c|
                int imax = 0;
                for (int i = 1; i < N; ++i) if (std::fabs(c[imax]) < std::fabs(c[i])) imax = i;
                double second = -1.0;
                for (int i = 0; i < N; ++i) if (i != imax) second = std::max(second, std::fabs(c[i]));
                // binomialFixDevice (:121-131); Q9a: no refinement at the edge or for fa == 0
                double delta = 0.0;
                if (imax > 0 && imax < N - 1) {
                    double f1 = c[imax - 1], f2 = c[imax], f3 = c[imax + 1];
                    double fa = (f1 + f3 - 2 * f2) / 2, fb = (f3 - f1) / 2;
                    if (fa != 0.0) delta = -fb / (2 * fa);
                    if (!std::isfinite(delta)) delta = 0.0;
                }
                lag[pp] = imax + delta - static_cast<double>(N) / 2;
                if (tr) {
                    int idx = pair_index + pp;
                    if (tr->ra_imax) tr->ra_imax[idx] = imax;
                    if (tr->ra_lag) tr->ra_lag[idx] = lag[pp];
                    double cm = std::fabs(c[imax]);
                    if (tr->ra_gap) tr->ra_gap[idx] = cm > 0 ? (cm - second) / cm : 0.0;
                }
            }
        }
        // :66-69 shift the second group of every pair by its lag (frequency-domain ramp)
#pragma omp parallel for schedule(static)
        for (int pp = 0; pp < npairs; ++pp) {
            int r0 = (2 * pp + 1) * s, r1 = std::min(M, r0 + s);
            for (int r = r0; r < r1; ++r)
                for (int k = 0; k < N; ++k) S[static_cast<size_t>(r) * N + k] *= ramp(k, lag[pp], N);
        }
        pair_index += npairs;
    }
    if (tr) tr->n_pairs = pair_index;
    // :76-77 back to range profiles
    fft_rows(S, M, N, +1);
    for (size_t i = 0; i < static_cast<size_t>(M) * N; ++i) S[i] /= static_cast<double>(N);
}

void orc_hrrp_center(double* s_, const orc_params* p, orc_trace* tr) {
    // range_alignment.cu:371-465 with inter_length = 30 (isar_main.cu:98).
    const int M = p->echo_num, N = p->range_num;
    cd* S = reinterpret_cast<cd*>(s_);
    const size_t MN = static_cast<size_t>(M) * N;
    // :381-389 H = |S| / max|S|
    double gmax = 0;
    for (size_t i = 0; i < MN; ++i) gmax = std::max(gmax, std::abs(S[i]));
    double inv = gmax > 0 ? 1.0 / gmax : 0.0;
    // :392-399 ARP = column sums / M  (Q5: all rows)
    std::vector<double> arp(N, 0.0);
    for (int m = 0; m < M; ++m)
        for (int n = 0; n < N; ++n) arp[n] += std::abs(S[static_cast<size_t>(m) * N + n]) * inv;
    for (int n = 0; n < N; ++n) arp[n] /= M;
    // :402-423 noise threshold from the sorted ARP
    std::vector<double> arp1(arp);
    std::stable_sort(arp1.begin(), arp1.end());
    int amax = 0;  // cublasIsamax: first index of the largest |x|
    for (int i = 1; i < N; ++i) if (std::fabs(arp1[amax]) < std::fabs(arp1[i])) amax = i;
    const double extra = static_cast<double>(kInterLength) * std::fabs(arp1[amax]) / N;
    const int diff_len = N - kInterLength;
    int imin = 0;
    double dmin = 0;
    for (int i = 0; i < diff_len; ++i) {  // cublasIsamin: first index of the smallest |x|
        double d = std::fabs(arp1[i + kInterLength] - arp1[i] - extra);
        if (i == 0 || d < dmin) { dmin = d; imin = i; }
    }
    const double thr = arp1[imin + kInterLength / 2];
    // :425-432 idx_1 = find(arp > thr)
    std::vector<int> idx1;
    for (int n = 0; n < N; ++n) if (arp[n] > thr) idx1.push_back(n);
    const int len1 = static_cast<int>(idx1.size());
    // :434-440 getARPMean (:524-550), one mean per idx_1 entry (Q6), edge divisors as in the ref
    std::vector<double> mean(len1);
    for (int t = 0; t < len1; ++t) {
        const int i = idx1[t];
        double acc = 0;
        if (i - kArpWindow / 2 >= 0 && i + kArpWindow / 2 <= N - 1) {
            for (int q = i - kArpWindow / 2; q <= i + kArpWindow / 2; ++q) acc += arp[q];
            acc /= kArpWindow + 1;
        } else if (i - kArpWindow / 2 < 0) {
            for (int q = 0; q <= i + kArpWindow / 2; ++q) acc += arp[q];
            acc /= (static_cast<double>(kArpWindow) / 2 + i + 1);
        } else {
            for (int q = i - kArpWindow / 2; q < N; ++q) acc += arp[q];
            acc /= (static_cast<double>(N) - i - static_cast<double>(kArpWindow) / 2 - 1);
        }
        mean[t] = acc;
    }
    // :442-449 idx_2 = find(mean < thr)
    std::vector<int> idx2;
    for (int t = 0; t < len1; ++t) if (mean[t] < thr) idx2.push_back(t);
    const int len2 = static_cast<int>(idx2.size());
    int shift = 0, applied = 0;
    if (len2 != len1) {
        // :451-461 zero the rejected indices, integer centroid, circular shift
        std::vector<int> kept(idx1);
        for (int t : idx2) kept[t] = 0;
        long long sum = 0;
        for (int v : kept) sum += v;
        const int mean_idx = static_cast<int>(sum / (len1 - len2));
        shift = -(mean_idx - N / 2);
        applied = 1;
        // circshiftFreq (:496-521) with an integer shift == circular shift by `shift`
        std::vector<cd> row(N);
        for (int m = 0; m < M; ++m) {
            cd* r = S + static_cast<size_t>(m) * N;
            for (int n = 0; n < N; ++n) row[((n + shift) % N + N) % N] = r[n];
            std::copy(row.begin(), row.end(), r);
        }
    }
    if (tr) {
        tr->center_applied = applied;
        tr->center_shift = shift;
        tr->idx1_len = len1;
        tr->idx2_len = len2;
        tr->diff_imin = imin;
        tr->center_thr = thr;
        if (tr->arp) std::copy(arp.begin(), arp.end(), tr->arp);
        if (tr->arp_mean) std::copy(mean.begin(), mean.end(), tr->arp_mean);
    }
}

void orc_cut_range_profile(const double* s_, const orc_params* p, double* cut_, orc_trace* tr) {
    // common.cu:494-518: cublasIcamax over row 0 (|Re|+|Im|, first index), window of W cells.
    const int M = p->echo_num, N = p->range_num, W = p->range_num_cut;
    const cd* S = reinterpret_cast<const cd*>(s_);
    cd* cut = reinterpret_cast<cd*>(cut_);
    int imax = 0;
    double best = -1;
    for (int n = 0; n < N; ++n) {
        double key = std::fabs(S[n].real()) + std::fabs(S[n].imag());
        if (tr && tr->cut_key) tr->cut_key[n] = key;
        if (key > best) { best = key; imax = n; }
    }
    int offset = std::max(imax - W / 2, 0);
    offset = std::min(offset, N - W);
    for (int m = 0; m < M; ++m)
        for (int j = 0; j < W; ++j) cut[static_cast<size_t>(m) * W + j] = S[static_cast<size_t>(m) * N + offset + j];
    if (tr) tr->cut_offset = offset;
}

void orc_doppler_tracking(double* comp_, double* phase_, const double* data_, int rows, int cols,
                          int if_comp) {
    // phase_adjustment.cu:4-41
    const cd* X = reinterpret_cast<const cd*>(data_);
    cd* phase = reinterpret_cast<cd*>(phase_);
    std::vector<double> angle(rows > 1 ? rows - 1 : 0);
    for (int m = 0; m + 1 < rows; ++m) {
        cd acc(0, 0);   // sumRows of conj(X[m]) * X[m+1]
        for (int n = 0; n < cols; ++n)
            acc += std::conj(X[static_cast<size_t>(m) * cols + n]) * X[static_cast<size_t>(m + 1) * cols + n];
        angle[m] = std::arg(acc);
    }
    for (size_t m = 1; m < angle.size(); ++m) angle[m] += angle[m - 1];   // inclusive_scan
    phase[0] = cd(0, 0);                                                  // :29 (Q9b kept)
    for (int m = 0; m + 1 < rows; ++m) phase[m + 1] = std::exp(cd(0, -angle[m]));
    if (if_comp) {
        cd* out = reinterpret_cast<cd*>(comp_);
        for (int m = 0; m < rows; ++m)
            for (int n = 0; n < cols; ++n)
                out[static_cast<size_t>(m) * cols + n] = phase[m] * X[static_cast<size_t>(m) * cols + n];
    }
}

void orc_fast_entropy(double* x_, int M, int W, int iterations, int phi0_given, orc_trace* tr) {
    // phase_adjustment.cu:167-332
    cd* X = reinterpret_cast<cd*>(x_);
    const size_t MW = static_cast<size_t>(M) * W;
    // :175-188 column maxima of |X| and the int-truncated mean (Q7 kept)
    std::vector<double> mx(W, 0.0);
    for (int m = 0; m < M; ++m)
        for (int n = 0; n < W; ++n) mx[n] = std::max(mx[n], std::abs(X[static_cast<size_t>(m) * W + n]));
    // :186 thrust::reduce(..., 0, plus<int>()): each maximum becomes an int as CUDA device code
    // converts it (truncation toward zero, clamped to the int range), and the ints add in 32-bit
    // arithmetic, which wraps modulo 2^32
    uint32_t usum = 0;
    for (int n = 0; n < W; ++n) {
        const double v = mx[n];
        const int32_t t = !(v == v) ? 0 : v >= 2147483648.0 ? INT32_MAX : static_cast<int32_t>(v);
        usum += static_cast<uint32_t>(t);
    }
    const float mean_val = static_cast<float>(static_cast<int32_t>(usum)) / static_cast<float>(W);
    const float threshold = 1.48f * mean_val;
    // :190-194 tgt_num
    int tgt_num = 0;
    for (int n = 0; n < W; ++n) if (mx[n] > static_cast<double>(threshold)) ++tgt_num;
    // :197-201 Doppler-tracked copy
    std::vector<cd> comp(MW), phase(M);
    orc_doppler_tracking(reinterpret_cast<double*>(comp.data()), reinterpret_cast<double*>(phase.data()),
                         x_, M, W, 1);
    // :216-246 image, column energy and column entropy
    fft_cols(comp.data(), M, W, -1);
    std::vector<double> E(W, 0.0), H(W, 0.0);
    for (int m = 0; m < M; ++m)
        for (int n = 0; n < W; ++n) E[n] += std::norm(comp[static_cast<size_t>(m) * W + n]);
    for (int n = 0; n < W; ++n) {
        if (!(E[n] > 0)) { H[n] = INFINITY; continue; }
        double h = 0;
        for (int m = 0; m < M; ++m) {
            double pr = std::norm(comp[static_cast<size_t>(m) * W + n]) / E[n];
            if (pr > 0) h -= pr * std::log(pr);
        }
        H[n] = h;
    }
    // :248-264 bin selection
    const int unit1 = tgt_num;
    const int unit2 = orc_next_pow2(tgt_num / 2);
    std::vector<int> order(W);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return E[a] > E[b]; });
    order.resize(unit1);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return H[a] < H[b]; });
    order.resize(unit2, 0);   // std::vector growth value-initialises to bin 0, as the ref does
    std::sort(order.begin(), order.end());
    // :270-278 selectRangeBins
    const int U = unit2;
    std::vector<cd> SU(static_cast<size_t>(M) * U);
    for (int m = 0; m < M; ++m)
        for (int u = 0; u < U; ++u) SU[static_cast<size_t>(m) * U + u] = X[static_cast<size_t>(m) * W + order[u]];
    // :281-282 starting phase from Doppler tracking of the selected bins
    std::vector<cd> phi(M);
    if (phi0_given && tr && tr->phi0) {
        for (int m = 0; m < M; ++m) phi[m] = cd(tr->phi0[2 * m], tr->phi0[2 * m + 1]);
    } else {
        orc_doppler_tracking(nullptr, reinterpret_cast<double*>(phi.data()),
                             reinterpret_cast<const double*>(SU.data()), M, U, 0);
        if (tr && tr->phi0)
            for (int m = 0; m < M; ++m) put_c(tr->phi0 + 2 * m, phi[m]);
    }
    // :302-322 fixed-count minimum-entropy iteration
    std::vector<cd> T(static_cast<size_t>(M) * U);
    for (int it = 0; it < iterations; ++it) {
        for (int m = 0; m < M; ++m)
            for (int u = 0; u < U; ++u) T[static_cast<size_t>(m) * U + u] = phi[m] * SU[static_cast<size_t>(m) * U + u];
        fft_cols(T.data(), M, U, -1);
        for (auto& v : T) v *= std::abs(v);
        fft_cols(T.data(), M, U, +1);
        for (int m = 0; m < M; ++m) {
            cd acc(0, 0);
            for (int u = 0; u < U; ++u)
                acc += std::conj(SU[static_cast<size_t>(m) * U + u]) * (T[static_cast<size_t>(m) * U + u] / static_cast<double>(M));
            phi[m] = unit(acc);
        }
    }
    // :324 apply
    for (int m = 0; m < M; ++m)
        for (int n = 0; n < W; ++n) X[static_cast<size_t>(m) * W + n] *= phi[m];
    if (tr) {
        tr->tgt_num = tgt_num;
        tr->unit2_num = U;
        tr->pa_threshold = threshold;
        if (tr->pa_colmax) std::copy(mx.begin(), mx.end(), tr->pa_colmax);
        if (tr->pa_colE) std::copy(E.begin(), E.end(), tr->pa_colE);
        if (tr->pa_colH) std::copy(H.begin(), H.end(), tr->pa_colH);
        if (tr->select_bin) std::copy(order.begin(), order.end(), tr->select_bin);
        if (tr->phi)
            for (int m = 0; m < M; ++m) put_c(tr->phi + 2 * m, phi[m]);
    }
}

void orc_mtrc(double* x_, const orc_params* p, int data_type) {
    // mtrc.cu:4-76
    const int M = p->echo_num, W = p->range_num_cut;
    cd* X = reinterpret_cast<cd*>(x_);
    const double B = static_cast<double>(p->band_width), fc = static_cast<double>(p->fc);
    const double posa = B / (fc * (W - 1.0));     // K * Tp / (fc * (W - 1)), :12
    const double constant = B / (2.0 * fc);       // K * Tp / (2 fc), :47
    // :14-40 range-frequency domain
    std::vector<cd> st(X, X + static_cast<size_t>(M) * W);
    swap_halves_rows(st.data(), M, W);
    if (data_type == 1) {
        fft_rows(st.data(), M, W, -1);
    } else {
        fft_rows(st.data(), M, W, +1);
        for (auto& v : st) v /= static_cast<double>(W);
    }
    // :42-52 chirp-z transform of every column n with alpha_n = 1 - constant + posa * n
#pragma omp parallel
    {
        std::vector<cd> col(M), out(M);
#pragma omp for schedule(static)
        for (int n = 0; n < W; ++n) {
            for (int m = 0; m < M; ++m) col[m] = st[static_cast<size_t>(m) * W + n];
            orc_czt_alpha(reinterpret_cast<const double*>(col.data()), M, 1.0 - constant + posa * n,
                          reinterpret_cast<double*>(out.data()));
            for (int m = 0; m < M; ++m) X[static_cast<size_t>(m) * W + n] = out[m];
        }
    }
    // :53-70
    swap_halves_cols(X, M, W);
    fft_cols(X, M, W, +1);
    for (size_t i = 0; i < static_cast<size_t>(M) * W; ++i) X[i] /= static_cast<double>(M);
    swap_halves_cols(X, M, W);
    fft_rows(X, M, W, -1);
    swap_halves_rows(X, M, W);
    swap_halves_cols(X, M, W);
    if (data_type == 1) flip_rows(X, M, W);   // :72-75 fliplr
}

// IFDS pulse compression (common.cu:583-604 multiThreadIFDS, 635-746 pulseCompression): the
// reference chirp with the echo's velocity, FFT matched filter, and the cut around the peak
// (cutRangeProfile, common.cu:494-518, with cols = NFFT).
int orc_pulse_compress(const float* iq, int len, double velocity, const orc_params* p, int cut, double* out) {
    const int nfft = orc_next_pow2(len);
    const int S = static_cast<int>(static_cast<double>(p->Fs) * p->Tp);
    if (len < 1 || cut < 1 || cut > nfft || S < 2 || S > nfft) return -1;
    const double c = 3e8;   // LIGHT_SPEED (common.cuh:69)
    const double v2 = 2.0 * velocity / c;
    const double v2_ = 1.0 - v2;
    const double c1 = -2.0 * M_PI * static_cast<double>(p->fc) * v2 / v2_;
    const double c2 = M_PI * static_cast<double>(p->band_width) / p->Tp;
    std::vector<double> ham = hamming(S);
    std::vector<cd> ref(nfft, cd(0.0, 0.0)), dat(nfft, cd(0.0, 0.0));
    for (int k = 0; k < S; ++k) {
        const double t = v2_ * (-p->Tp / 2.0 + k * p->Tp / (static_cast<double>(S) - 1.0));   // genTkPulseCompression
        const double ph = (c1 + c2 * t) * t;                                                  // genRefPulseCompression
        ref[k] = ham[k] * cd(std::cos(ph), std::sin(ph));
    }
    for (int k = 0; k < len; ++k) dat[k] = cd(iq[2 * k], iq[2 * k + 1]);
    fft_any(ref.data(), nfft, -1);
    fft_any(dat.data(), nfft, -1);
    for (int k = 0; k < nfft; ++k) dat[k] = std::conj(ref[k]) * dat[k];   // elementwiseMultiplyConjA
    fft_any(dat.data(), nfft, +1);
    int imax = 0;
    double best = -1.0;
    for (int k = 0; k < nfft; ++k) {
        dat[k] /= static_cast<double>(nfft);
        const double a = std::abs(dat[k].real()) + std::abs(dat[k].imag());   // Icamax: first max of |Re|+|Im|
        if (a > best) { best = a; imax = k; }
    }
    int off = std::max(imax - cut / 2, 0);
    off = std::min(off, nfft - cut);
    for (int j = 0; j < cut; ++j) {
        out[2 * j] = dat[off + j].real();
        out[2 * j + 1] = dat[off + j].imag();
    }
    return off;
}

double orc_image_entropy(const float* img, int n) {
    double tot = 0;
    for (int i = 0; i < n; ++i) tot += static_cast<double>(img[i]) * img[i];
    if (!(tot > 0)) return 0.0;
    double h = 0;
    for (int i = 0; i < n; ++i) {
        double pr = static_cast<double>(img[i]) * img[i] / tot;
        if (pr > 0) h -= pr * std::log(pr);
    }
    return h;
}

namespace {
// The rows as range alignment receives them, before getHRRP: the IFDS pre-stage
// (isar_main.cu:15-27) or, for STRETCH, the HPC (:31-48).
void front_rows(std::vector<cd>& S, const float* echo, const double* velocity, const orc_params* p,
                const orc_options* opt) {
    const int M = p->echo_num, N = p->range_num;
    const size_t MN = static_cast<size_t>(M) * N;
    for (size_t i = 0; i < MN; ++i) S[i] = cd(echo[2 * i], echo[2 * i + 1]);
    if (opt->data_type == 1) {
        swap_halves_rows(S.data(), M, N);
        fft_rows(S.data(), M, N, +1);
        for (auto& v : S) v /= static_cast<double>(N);
        swap_halves_rows(S.data(), M, N);
    }
    if (opt->if_hpc && opt->data_type == 2) orc_hpc(reinterpret_cast<double*>(S.data()), velocity, p);
}
}  // namespace

int orc_hrrp(const float* echo, const double* velocity, const orc_params* p, const orc_options* opt,
             double* hrrp) {
    // getHRRP (common.cu:541-580) on the rows of isar_main.cu:57-67
    const int M = p->echo_num, N = p->range_num;
    if (M < 1 || N < 2 || !is_pow2(N) || (opt->data_type != 1 && opt->data_type != 2)) return 1;
    std::vector<cd> S(static_cast<size_t>(M) * N);
    front_rows(S, echo, velocity, p, opt);
    if (opt->data_type == 1) {
        // IFDS: ifftshift, forward FFT, ifftshift of a copy (common.cu:546-558)
        swap_halves_rows(S.data(), M, N);
        fft_rows(S.data(), M, N, -1);
        swap_halves_rows(S.data(), M, N);
    } else {
        // STRETCH: Hamming window over range, forward FFT, ifftshift (common.cu:561-571)
        const std::vector<double> ham = hamming(N);
        for (int m = 0; m < M; ++m)
            for (int n = 0; n < N; ++n) S[static_cast<size_t>(m) * N + n] *= ham[n];
        fft_rows(S.data(), M, N, -1);
        swap_halves_rows(S.data(), M, N);
    }
    std::memcpy(hrrp, S.data(), S.size() * sizeof(cd));
    return 0;
}

int orc_isar_main(const float* echo, const double* velocity, const orc_params* p,
                  const orc_options* opt, float* img, orc_trace* tr) {
    return orc_isar_stage_times(echo, velocity, p, opt, img, tr, nullptr);
}

int orc_isar_stage_times(const float* echo, const double* velocity, const orc_params* p,
                         const orc_options* opt, float* img, orc_trace* tr, double* stage_ms) {
    // isar_main.cu:3-227 (isarMainSingle, isar_imaging_export.cpp:208-212); stage_ms (optional)
    // gets the wall time of the reference's SEPARATE_TIMING_ stages (isar_main.cu:34-218)
    const int M = p->echo_num, N = p->range_num, W = p->range_num_cut;
    // imagingMemInit validation (isar_imaging_export.cpp:144-155)
    if (M < 2 || M > 1024 || N < W || N < M || W < 2 || !is_pow2(N)) return 1;
    const size_t MN = static_cast<size_t>(M) * N, MW = static_cast<size_t>(M) * W;
    auto clock = [] { return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count(); };
    double t0 = clock();
    auto lap = [&](int stage) {
        const double t = clock();
        if (stage_ms) stage_ms[stage] = t - t0;
        t0 = t;
    };
    std::vector<cd> S(MN);
    front_rows(S, echo, velocity, p, opt);
    lap(0);
    if (opt->if_hrrp && opt->data_type == 2) {
        // getHRRP, common.cu:563-566: the STRETCH branch windows d_data in place (Q11 kept);
        // the HRRP itself is a side output that the image does not depend on.
        std::vector<double> ham = hamming(N);
        for (int m = 0; m < M; ++m)
            for (int n = 0; n < N; ++n) S[static_cast<size_t>(m) * N + n] *= ham[n];
    }
    t_range_fft32 = opt->range_fft32;
    orc_range_alignment(reinterpret_cast<double*>(S.data()), p, tr);
    t_range_fft32 = 0;
    lap(1);
    orc_hrrp_center(reinterpret_cast<double*>(S.data()), p, tr);
    std::vector<cd> X(MW);
    orc_cut_range_profile(reinterpret_cast<const double*>(S.data()), p, reinterpret_cast<double*>(X.data()), tr);
    lap(2);
    orc_fast_entropy(reinterpret_cast<double*>(X.data()), M, W, opt->iterations, opt->phi0_given, tr);
    lap(3);
    if (opt->if_mtrc) orc_mtrc(reinterpret_cast<double*>(X.data()), p, opt->data_type);
    lap(4);
    orc_post_process(reinterpret_cast<double*>(X.data()), M, W, opt->quirk_q1, opt->data_type, img);
    lap(5);
    if (tr) tr->entropy = orc_image_entropy(img, static_cast<int>(MW));
    return 0;
}

void orc_post_process(double* x_, int M, int W, int quirk_q1, int data_type, float* img) {
    // isar_main.cu:191-211
    cd* X = reinterpret_cast<cd*>(x_);
    const size_t MW = static_cast<size_t>(M) * W;
    // :191-193 slow-time window
    std::vector<double> hm = hamming(M);
    if (quirk_q1) {
        // diagMulMat(d_hamming_echoes, X, X, cols = echo_num, len = range_num_cut): only the
        // first W elements, element i scaled by hamming_M[i / M].
        for (int i = 0; i < W; ++i) X[i] *= hm[i / M];
    } else {
        for (int m = 0; m < M; ++m)
            for (int n = 0; n < W; ++n) X[static_cast<size_t>(m) * W + n] *= hm[m];
    }
    // :195-207 azimuth FFT, shift, [flipud], abs
    fft_cols(X, M, W, -1);
    swap_halves_cols(X, M, W);
    if (data_type == 1) flip_cols(X, M, W);
    for (size_t i = 0; i < MW; ++i) img[i] = static_cast<float>(std::abs(X[i]));
}

int orc_isar_batch(int K, const float* echo, const double* velocity, const orc_params* p,
                   const orc_options* opt, float* img) {
    const size_t MN = static_cast<size_t>(p->echo_num) * p->range_num;
    const size_t MW = static_cast<size_t>(p->echo_num) * p->range_num_cut;
    int threads = 1;
#pragma omp parallel
    {
#pragma omp single
        {
#ifdef _OPENMP
            threads = omp_get_num_threads();
#endif
        }
#pragma omp for schedule(dynamic, 1)
        for (int k = 0; k < K; ++k)
            orc_isar_main(echo + 2 * MN * k, velocity + static_cast<size_t>(p->echo_num) * k, p, opt,
                          img + MW * k, nullptr);
    }
    return threads;
}

}  // extern "C"
