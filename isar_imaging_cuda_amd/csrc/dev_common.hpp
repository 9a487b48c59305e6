// dev_common.hpp — device primitives for the gfx950 imaging kernels: complex arithmetic,
// wave64/workgroup reductions, exact phase generation and an LDS-resident Stockham FFT.
#pragma once

#include "pipeline_types.hpp"

namespace isar {

// ------------------------------------------------------------------------------------------
// complex float helpers (float2 = re, im)
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {
    return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
// conj(a) * b
__device__ __forceinline__ float2 cconjmul(float2 a, float2 b) {
    return make_float2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__device__ __forceinline__ float cabs2(float2 a) { return a.x * a.x + a.y * a.y; }
__device__ __forceinline__ float cabsf_(float2 a) { return sqrtf(a.x * a.x + a.y * a.y); }
// Square root by the hardware instruction alone (within 1 ulp). sqrtf compiles to that
// instruction plus about 15 more for correct rounding; where a square root is taken per element
// inside FP32 transforms, that expansion costs more issue slots than it buys accuracy.
__device__ __forceinline__ float sqrt_hw(float v) {
#ifdef ISAR_EMU
    return sqrtf(v);
#else
    return __builtin_amdgcn_sqrtf(v);
#endif
}
__device__ __forceinline__ float cabs_hw(float2 a) { return sqrt_hw(a.x * a.x + a.y * a.y); }
// float -> int as a CUDA device conversion performs it: truncation toward zero, clamped to the int
// range, NaN -> 0 (the reference relies on it where Thrust accumulates floats as ints)
__device__ __forceinline__ int trunc_sat_i32(float v) {
    if (!(v == v)) return 0;
    if (v >= 2147483648.0f) return 2147483647;
    if (v <= -2147483648.0f) return -2147483647 - 1;
    return static_cast<int>(v);
}
// multiply by (dir * i): dir = -1 -> -i, dir = +1 -> +i
template <int DIR>
__device__ __forceinline__ float2 cmul_i(float2 a) {
    return DIR < 0 ? make_float2(a.y, -a.x) : make_float2(-a.y, a.x);
}

__device__ __forceinline__ float2 expjpi_twiddle(float t);

// exp(j pi t) for a phase given in half-turns; t is reduced to [-1, 1] in FP64 before the FP32
// evaluation, so arguments carrying n^2 or k*s growth keep their digits.
__device__ __forceinline__ float2 expjpi_reduced(double t) {
    t -= 2.0 * rint(0.5 * t);
    return expjpi_twiddle(static_cast<float>(t));
}
// exp(j pi t) evaluated in FP64 and rounded once to FP32 (within half an ulp per component).
__device__ __forceinline__ float2 expjpi_rounded(double t) {
    double s, c;
    sincospi(t, &s, &c);
    return make_float2(static_cast<float>(c), static_cast<float>(s));
}
// exp(j pi t) for |t| small enough that FP32 holds it exactly (dyadic FFT twiddles).
__device__ __forceinline__ float2 expjpi(float t) {
    float s, c;
    sincospif(t, &s, &c);
    return make_float2(c, s);
}

// exp(j pi t) for t in [-2, 2] given exactly in FP32 (FFT twiddle arguments): octant reduction
// t = q/4 + r, |r| <= 1/8, then Taylor polynomials of sin(pi r) (degree 9) and cos(pi r)
// (degree 10), whose truncation errors (< 3e-11) are far below FP32 rounding, and a rotation by
// exp(j pi q / 4).
__device__ __forceinline__ float2 expjpi_twiddle(float t) {
    const float qf = rintf(4.0f * t);
    const int q = static_cast<int>(qf) & 7;
    const float r = t - 0.25f * qf;   // exact for dyadic t
    const float x = 3.14159265358979323846f * r;
    const float x2 = x * x;
    const float sn = x * (1.0f + x2 * (-1.6666667163e-1f + x2 * (8.3333337680e-3f + x2 * (-1.9841270114e-4f + x2 * 2.7557319224e-6f))));
    const float cs = 1.0f + x2 * (-0.5f + x2 * (4.1666667908e-2f + x2 * (-1.3888889225e-3f + x2 * (2.4801587642e-5f - x2 * 2.7557319224e-7f))));
    // rotation by exp(j pi q / 4) with selects only: the odd octant first (times (1 + j)/sqrt 2),
    // then the quadrant (times j^(q/2))
    constexpr float h = 0.70710678118654752440f;
    const bool odd = q & 1;
    float c = odd ? h * (cs - sn) : cs;
    float s = odd ? h * (cs + sn) : sn;
    const bool q1 = q & 2, q2 = q & 4;
    const float c1 = q1 ? -s : c, s1 = q1 ? c : s;
    c = q2 ? -c1 : c1;
    s = q2 ? -s1 : s1;
    return make_float2(c, s);
}

// ------------------------------------------------------------------------------------------
// reductions
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ float2 wave_sum2(float2 v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        v.x += __shfl_xor(v.x, o, 64);
        v.y += __shfl_xor(v.y, o, 64);
    }
    return v;
}
// (value, index) argmax with lowest index on ties
__device__ __forceinline__ void wave_argmax(float& v, int& i) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        float ov = __shfl_xor(v, o, 64);
        int oi = __shfl_xor(i, o, 64);
        if (ov > v || (ov == v && oi < i)) { v = ov; i = oi; }
    }
}
// (value, index) argmin with lowest index on ties
__device__ __forceinline__ void wave_argmin(float& v, int& i) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        float ov = __shfl_xor(v, o, 64);
        int oi = __shfl_xor(i, o, 64);
        if (ov < v || (ov == v && oi < i)) { v = ov; i = oi; }
    }
}

// Workgroup-wide reductions through a small LDS scratch (>= 16 slots). All threads get the
// result. Must be called by every thread of the workgroup.
template <int NT>
__device__ __forceinline__ float block_sum(float v, float* red) {
    v = wave_sum(v);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    __syncthreads();
    if (l == 0) red[w] = v;
    __syncthreads();
    float r = 0.f;
#pragma unroll
    for (int i = 0; i < NT / 64; ++i) r += red[i];
    return r;
}
template <int NT>
__device__ __forceinline__ float block_max(float v, float* red) {
    v = wave_max(v);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    __syncthreads();
    if (l == 0) red[w] = v;
    __syncthreads();
    float r = red[0];
#pragma unroll
    for (int i = 1; i < NT / 64; ++i) r = fmaxf(r, red[i]);
    return r;
}
template <int NT>
__device__ __forceinline__ void block_argmax(float& v, int& i, float* redv, int* redi) {
    wave_argmax(v, i);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    __syncthreads();
    if (l == 0) { redv[w] = v; redi[w] = i; }
    __syncthreads();
    v = redv[0];
    i = redi[0];
#pragma unroll
    for (int k = 1; k < NT / 64; ++k) {
        float ov = redv[k];
        int oi = redi[k];
        if (ov > v || (ov == v && oi < i)) { v = ov; i = oi; }
    }
}
template <int NT>
__device__ __forceinline__ void block_argmin(float& v, int& i, float* redv, int* redi) {
    wave_argmin(v, i);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    __syncthreads();
    if (l == 0) { redv[w] = v; redi[w] = i; }
    __syncthreads();
    v = redv[0];
    i = redi[0];
#pragma unroll
    for (int k = 1; k < NT / 64; ++k) {
        float ov = redv[k];
        int oi = redi[k];
        if (ov < v || (ov == v && oi < i)) { v = ov; i = oi; }
    }
}
// Exclusive prefix sum of integers over the workgroup; returns the exclusive prefix, total in
// *tot. With T = unsigned the sums wrap modulo 2^32.
template <int NT, typename T = int>
__device__ __forceinline__ T block_exclusive_scan(T v, T* red, T* tot) {
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    T x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        T y = __shfl_up(x, o, 64);
        if (l >= o) x += y;
    }
    __syncthreads();
    if (l == 63) red[w] = x;
    __syncthreads();
    T base = 0, all = 0;
#pragma unroll
    for (int k = 0; k < NT / 64; ++k) {
        T r = red[k];
        if (k < w) base += r;
        all += r;
    }
    *tot = all;
    return base + x - v;
}

// ------------------------------------------------------------------------------------------
// Small DFTs in registers. DIR = -1 forward (exp(-j...)), +1 inverse. Output in natural order.
// ------------------------------------------------------------------------------------------
template <int DIR>
__device__ __forceinline__ void dft2(float2& a, float2& b) {
    float2 t = a;
    a = cadd(t, b);
    b = csub(t, b);
}
template <int DIR>
__device__ __forceinline__ void dft4(float2& a0, float2& a1, float2& a2, float2& a3) {
    float2 t0 = cadd(a0, a2), t1 = csub(a0, a2);
    float2 t2 = cadd(a1, a3), t3 = cmul_i<DIR>(csub(a1, a3));
    a0 = cadd(t0, t2);
    a2 = csub(t0, t2);
    a1 = cadd(t1, t3);
    a3 = csub(t1, t3);
}
template <int DIR>
__device__ __forceinline__ void dft3(float2& a0, float2& a1, float2& a2) {
    constexpr float h = 0.86602540378443864676f;   // sqrt(3)/2
    float2 t = cadd(a1, a2), d = csub(a1, a2);
    float2 m = make_float2(a0.x - 0.5f * t.x, a0.y - 0.5f * t.y);
    float2 r = cmul_i<DIR>(make_float2(h * d.x, h * d.y));
    a0 = cadd(a0, t);
    a1 = cadd(m, r);
    a2 = csub(m, r);
}

template <int R, int DIR>
struct Dft;

template <int DIR>
struct Dft<2, DIR> {
    __device__ __forceinline__ static void run(float2* v) { dft2<DIR>(v[0], v[1]); }
};
template <int DIR>
struct Dft<3, DIR> {
    __device__ __forceinline__ static void run(float2* v) { dft3<DIR>(v[0], v[1], v[2]); }
};
template <int DIR>
struct Dft<4, DIR> {
    __device__ __forceinline__ static void run(float2* v) { dft4<DIR>(v[0], v[1], v[2], v[3]); }
};
template <int DIR>
struct Dft<8, DIR> {
    // 8 = 4 (inner, over n1 with stride 2) x 2
    __device__ __forceinline__ static void run(float2* v) {
        constexpr float r = 0.70710678118654752440f;
        float2 e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
        float2 o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
        dft4<DIR>(e0, e1, e2, e3);
        dft4<DIR>(o0, o1, o2, o3);
        // twiddle o_k *= W8^k, W8 = exp(DIR j pi/4)
        o1 = cmul(o1, make_float2(r, DIR * r));
        o2 = cmul_i<DIR>(o2);
        o3 = cmul(o3, make_float2(-r, DIR * r));
        v[0] = cadd(e0, o0); v[4] = csub(e0, o0);
        v[1] = cadd(e1, o1); v[5] = csub(e1, o1);
        v[2] = cadd(e2, o2); v[6] = csub(e2, o2);
        v[3] = cadd(e3, o3); v[7] = csub(e3, o3);
    }
};
template <int DIR>
struct Dft<16, DIR> {
    // 16 = 4 x 4: n = 4 n1 + n2, k = k1 + 4 k2
    __device__ __forceinline__ static void run(float2* v) {
        constexpr float c1 = 0.92387953251128675613f, s1 = 0.38268343236508977173f;
        constexpr float r = 0.70710678118654752440f;
        float2 b[4][4];
#pragma unroll
        for (int n2 = 0; n2 < 4; ++n2) {
            b[n2][0] = v[n2]; b[n2][1] = v[4 + n2]; b[n2][2] = v[8 + n2]; b[n2][3] = v[12 + n2];
            dft4<DIR>(b[n2][0], b[n2][1], b[n2][2], b[n2][3]);
        }
        // twiddles W16^(n2 k1)
        b[1][1] = cmul(b[1][1], make_float2(c1, DIR * s1));
        b[1][2] = cmul(b[1][2], make_float2(r, DIR * r));
        b[1][3] = cmul(b[1][3], make_float2(s1, DIR * c1));
        b[2][1] = cmul(b[2][1], make_float2(r, DIR * r));
        b[2][2] = cmul_i<DIR>(b[2][2]);
        b[2][3] = cmul(b[2][3], make_float2(-r, DIR * r));
        b[3][1] = cmul(b[3][1], make_float2(s1, DIR * c1));
        b[3][2] = cmul(b[3][2], make_float2(-r, DIR * r));
        b[3][3] = cmul(b[3][3], make_float2(-c1, -DIR * s1));
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) {
            float2 a0 = b[0][k1], a1 = b[1][k1], a2 = b[2][k1], a3 = b[3][k1];
            dft4<DIR>(a0, a1, a2, a3);
            v[k1] = a0; v[k1 + 4] = a1; v[k1 + 8] = a2; v[k1 + 12] = a3;
        }
    }
};

// ------------------------------------------------------------------------------------------
// LDS-resident Stockham FFT (radix-2/3/4/8/16 passes).
//
// A transform of length N is computed by T = N / E threads, each holding E elements in
// registers. Pass p with radix R processes N/R butterflies; thread t owns butterflies
// j = t + b*T, b < E/R, and holds x[j + r*N/R] for r < R on entry. After the last pass the
// thread holds the outputs at the same positions j + r*N/R_last, so the first-pass load and
// the last-pass store are both unit-stride across threads (coalesced against global memory).
// Between passes the data goes through an LDS buffer with one padding slot every 16 elements.
// All threads of the workgroup must call the transform together (it uses __syncthreads).
// With WAVE = true the passes are ordered by fft_sync<true> instead: for transforms of at most
// 64 threads, aligned within a wave, that touch only their own LDS buffer (N/E <= 64 and
// 64 % (N/E) == 0); the caller orders other waves' accesses to that buffer with barriers of its own.
// ------------------------------------------------------------------------------------------
// A copy of v that the compiler cannot see through. Index and address arithmetic derived from it
// is recomputed where it is used instead of being kept live in registers across transforms.
__device__ __forceinline__ int opaque(int v) {
#ifndef ISAR_EMU
    asm volatile("" : "+v"(v));
#endif
    return v;
}
// Element i of p for 0 <= i < 2^28 / sizeof(T) (p typically wave-uniform): the byte offset is
// formed in 32 bits, so the access can take the base in scalar registers and a 32-bit per-lane
// offset instead of a 64-bit address per lane.
template <typename T>
__device__ __forceinline__ T& elem_u32(T* p, unsigned i) {
    using B = std::conditional_t<std::is_const_v<T>, const char, char>;
    return *reinterpret_cast<T*>(reinterpret_cast<B*>(p) + i * static_cast<unsigned>(sizeof(T)));
}

// Value of v held by lane `lane` (a compile-time constant in practice) of the calling wave.
__device__ __forceinline__ float wave_read_lane(float v, int lane) {
#ifdef ISAR_EMU
    return __shfl(v, lane, 64);
#else
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
#endif
}

// 1 / v to about 1 ulp (v_rcp_f32), without the correctly rounded division sequence.
__device__ __forceinline__ float rcp_hw(float v) {
#ifdef ISAR_EMU
    return 1.0f / v;
#else
    return __builtin_amdgcn_rcpf(v);
#endif
}

__device__ __forceinline__ float opaque(float v) {
#ifndef ISAR_EMU
    asm volatile("" : "+v"(v));
#endif
    return v;
}

__device__ __forceinline__ float2 opaque(float2 v) {
#ifndef ISAR_EMU
    asm volatile("" : "+v"(v.x), "+v"(v.y));
#endif
    return v;
}

// Column tile of workgroup b among the `tiles` tiles (gridDim.x) of a frame, ordered so that one
// XCD runs adjacent tiles. Workgroups go to the 8 XCDs round-robin by linear id (b + tiles f), so
// with tiles % 8 == 0 XCD x takes tiles [x tiles/8, (x+1) tiles/8) of every frame: tiles narrower
// than a 128-byte line then share its fetch in one L2 instead of each XCD fetching it.
__device__ __forceinline__ int xcd_tile(int b, int tiles) {
    return (tiles & 7) ? b : (b & 7) * (tiles >> 3) + (b >> 3);
}

// Orders the LDS accesses of an FFT's passes. WAVE = false: a workgroup barrier. WAVE = true: the
// transform lives in one wave, whose LDS instructions the hardware performs in issue order, so
// only the compiler has to be kept from moving accesses across (a wavefront-scope fence needs no
// wait and no barrier instruction).
template <bool WAVE>
__device__ __forceinline__ void fft_sync() {
#ifdef ISAR_EMU
    __syncthreads();   // every thread of the block calls the transform together
#else
    if constexpr (WAVE) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    } else {
        __syncthreads();
    }
#endif
}
template <int N, int E, bool WAVE>
__device__ __forceinline__ constexpr void fft_check_wave() {
    static_assert(!WAVE || (N / E <= 64 && 64 % (N / E) == 0), "a wave-ordered transform must lie in one wave");
}

__device__ __forceinline__ int lds_pad(int i) { return i + (i >> 4); }
__host__ __device__ constexpr int lds_padded_size(int n) { return n + (n >> 4) + 1; }

// w[r] = exp(DIR j pi base r), r < R: w^1 and w^4 from expjpi_twiddle, every other power by at
// most three complex products (relative error a few FP32 ulp).
template <int R, int DIR>
__device__ __forceinline__ void twiddle_powers(float2 (&w)[R], float base) {
    w[0] = make_float2(1.f, 0.f);
    w[1] = expjpi_twiddle(DIR * base);
    if constexpr (R == 3) {
        w[2] = cmul(w[1], w[1]);
    } else if constexpr (R >= 4) {
        w[2] = cmul(w[1], w[1]);
        w[3] = cmul(w[2], w[1]);
        if constexpr (R >= 8) {
            w[4] = expjpi_twiddle(DIR * base * 4.0f);
            w[5] = cmul(w[4], w[1]);
            w[6] = cmul(w[4], w[2]);
            w[7] = cmul(w[4], w[3]);
        }
        if constexpr (R >= 16) {
            const float2 w8 = cmul(w[4], w[4]);
            w[8] = w8;
#pragma unroll
            for (int r = 9; r < 16; ++r) w[r] = cmul(w8, w[r - 8]);
        }
    }
}

// w[r] for r < R from given exp(DIR j pi base) and exp(DIR j pi 4 base).
template <int R>
__device__ __forceinline__ void twiddle_powers_from(float2 (&w)[R], float2 w1, float2 w4) {
    w[0] = make_float2(1.f, 0.f);
    w[1] = w1;
    if constexpr (R >= 3) w[2] = cmul(w1, w1);
    if constexpr (R >= 4) w[3] = cmul(w[2], w1);
    if constexpr (R >= 8) {
        w[4] = w4;
        w[5] = cmul(w4, w1);
        w[6] = cmul(w4, w[2]);
        w[7] = cmul(w4, w[3]);
    }
    if constexpr (R >= 16) {
        const float2 w8 = cmul(w4, w4);
        w[8] = w8;
#pragma unroll
        for (int r = 9; r < 16; ++r) w[r] = cmul(w8, w[r - 8]);
    }
}

// Twiddle bases of every pass of a transform plan for one thread, computed once and reused by
// all transforms of a kernel: w1[p] = exp(+j pi b), w4[p] = exp(+j pi 4 b), b = 2k / (Ns R) with
// k = tid mod Ns, for the passes that hold one radix group per thread (the others compute
// their twiddles in place).
template <int N, int E>
struct FftTwiddles {
    float2 w1[5], w4[5];
};

template <int N, int E, int DIR, int Ns, int R>
__device__ __forceinline__ void stockham_twiddle_cached(float2* v, float2 w1, float2 w4) {
    // the powers are formed here, at each use, rather than once per kernel and kept live
    w1 = opaque(w1);
    w4 = opaque(w4);
    if constexpr (DIR < 0) {
        w1.y = -w1.y;
        w4.y = -w4.y;
    }
    float2 w[R];
    twiddle_powers_from<R>(w, w1, w4);
#pragma unroll
    for (int r = 1; r < R; ++r) v[r] = cmul(v[r], w[r]);
}

template <int N, int E, int R, int DIR, int Ns>
__device__ __forceinline__ void stockham_compute(float2 (&x)[E], int tid) {
    constexpr int T = N / E;
    constexpr int B = E / R;
    static_assert(E % R == 0, "radix must divide the elements per thread");
    // The twiddles depend only on the thread index, so the compiler would otherwise compute
    // those of every pass (and of every transform in a loop over rows) once and keep them live,
    // which costs far more registers than recomputing them here.
    const int t = Ns > 1 ? opaque(tid) : tid;
#pragma unroll
    for (int b = 0; b < B; ++b) {
        const int j = t + b * T;
        float2* v = &x[b * R];
        if constexpr (Ns > 1) {
            const int k = j % Ns;
            // twiddles exp(DIR j 2 pi k r / (Ns R)); 2 k r / (Ns R) is exact in FP32. Even
            // powers up to 8 come straight from sincospi, the rest are one or two products.
            const float base = static_cast<float>(2 * k) / static_cast<float>(Ns * R);
            float2 w[R];
            twiddle_powers<R, DIR>(w, base);
#pragma unroll
            for (int r = 1; r < R; ++r) v[r] = cmul(v[r], w[r]);
        }
        Dft<R, DIR>::run(v);
    }
}

template <int N, int E, int R, int Ns>
__device__ __forceinline__ void stockham_store(const float2 (&x)[E], float2* lds, int tid) {
    constexpr int T = N / E;
    constexpr int B = E / R;
#pragma unroll
    for (int b = 0; b < B; ++b) {
        const int j = tid + b * T;
        const int k = j % Ns;
        const int d = (j / Ns) * Ns * R + k;
#pragma unroll
        for (int r = 0; r < R; ++r) lds[lds_pad(d + r * Ns)] = x[b * R + r];
    }
}

template <int N, int E, int R>
__device__ __forceinline__ void stockham_load(float2 (&x)[E], const float2* lds, int tid) {
    constexpr int T = N / E;
    constexpr int B = E / R;
#pragma unroll
    for (int b = 0; b < B; ++b) {
        const int j = tid + b * T;
#pragma unroll
        for (int r = 0; r < R; ++r) x[b * R + r] = lds[lds_pad(j + r * (N / R))];
    }
}

// Element position held in register slot e after the transform (for the last radix RL).
template <int N, int E, int RL>
__device__ __forceinline__ int fft_out_pos(int tid, int e) {
    constexpr int T = N / E;
    return (tid + (e / RL) * T) + (e % RL) * (N / RL);
}
// Element position expected in register slot e before the transform (first radix R0).
template <int N, int E, int R0>
__device__ __forceinline__ int fft_in_pos(int tid, int e) {
    constexpr int T = N / E;
    return (tid + (e / R0) * T) + (e % R0) * (N / R0);
}

// Radix sequence per transform length for E elements per thread.
template <int N, int E>
struct FftPlan;

// E = 16 plans
template <> struct FftPlan<16384, 16> { static constexpr int NP = 4; static constexpr int R[4] = {16, 16, 16, 4}; };
template <> struct FftPlan<8192, 16>  { static constexpr int NP = 4; static constexpr int R[4] = {16, 16, 16, 2}; };
template <> struct FftPlan<4096, 16>  { static constexpr int NP = 3; static constexpr int R[3] = {16, 16, 16}; };
template <> struct FftPlan<2048, 16>  { static constexpr int NP = 3; static constexpr int R[3] = {16, 16, 8}; };
template <> struct FftPlan<1024, 16>  { static constexpr int NP = 3; static constexpr int R[3] = {16, 16, 4}; };
template <> struct FftPlan<512, 16>   { static constexpr int NP = 3; static constexpr int R[3] = {16, 16, 2}; };
template <> struct FftPlan<256, 16>   { static constexpr int NP = 2; static constexpr int R[2] = {16, 16}; };
template <> struct FftPlan<128, 16>   { static constexpr int NP = 2; static constexpr int R[2] = {16, 8}; };
template <> struct FftPlan<64, 16>    { static constexpr int NP = 2; static constexpr int R[2] = {16, 4}; };
// E = 8 plans
template <> struct FftPlan<512, 8>    { static constexpr int NP = 3; static constexpr int R[3] = {8, 8, 8}; };
template <> struct FftPlan<4096, 8>   { static constexpr int NP = 4; static constexpr int R[4] = {8, 8, 8, 8}; };
// E = 12 plans (mixed radix for 3 * 2^k lengths)
template <> struct FftPlan<192, 12>   { static constexpr int NP = 4; static constexpr int R[4] = {4, 4, 4, 3}; };
template <> struct FftPlan<384, 12>   { static constexpr int NP = 5; static constexpr int R[5] = {4, 4, 4, 2, 3}; };
template <> struct FftPlan<768, 12>   { static constexpr int NP = 5; static constexpr int R[5] = {4, 4, 4, 4, 3}; };

template <int N, int E, int DIR, int P, int Ns, bool WAVE = false>
struct FftPasses {
    using Plan = FftPlan<N, E>;
    // runs passes P..NP-1; x holds the layout of pass P on entry
    __device__ __forceinline__ static void run(float2 (&x)[E], float2* lds, int tid) {
        fft_check_wave<N, E, WAVE>();
        constexpr int R = Plan::R[P];
        stockham_compute<N, E, R, DIR, Ns>(x, tid);
        if constexpr (P + 1 < Plan::NP) {
            constexpr int RN = Plan::R[P + 1];
            stockham_store<N, E, R, Ns>(x, lds, tid);
            fft_sync<WAVE>();
            stockham_load<N, E, RN>(x, lds, tid);
            fft_sync<WAVE>();
            FftPasses<N, E, DIR, P + 1, Ns * R, WAVE>::run(x, lds, tid);
        }
    }
};

// Same passes with the twiddle bases taken from a per-thread cache (one group per thread).
// The radix-16 pass at stride 16 needs only the 256 twiddles exp(j 2 pi m / 256): a kernel that
// runs many transforms may keep them in LDS (fft_tw16_table_fill) and read them there
// (TAB = true) instead of forming the 15 powers from the cached bases at every transform.
template <int N, int E, int DIR, int P, int Ns, bool TAB = false, bool WAVE = false>
struct FftPassesCached {
    using Plan = FftPlan<N, E>;
    __device__ __forceinline__ static void run(float2 (&x)[E], float2* lds, int tid, const FftTwiddles<N, E>& tw,
                                               const float2* tab = nullptr) {
        fft_check_wave<N, E, WAVE>();
        // the LDS table is filled by the whole workgroup and ordered by the first workgroup barrier
        static_assert(!(TAB && WAVE), "table twiddles need workgroup-ordered passes");
        constexpr int R = Plan::R[P];
        if constexpr (E == R) {   // one radix group per thread: twiddles from the cache
            if constexpr (TAB && Ns == 16 && R == 16) {
                const int k = opaque(tid) & 15;
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    const float2 w = tab[k * r];
                    x[r] = cmul(x[r], DIR < 0 ? make_float2(w.x, -w.y) : w);
                }
            } else if constexpr (Ns > 1) {
                stockham_twiddle_cached<N, E, DIR, Ns, R>(x, tw.w1[P], tw.w4[P]);
            }
            Dft<R, DIR>::run(x);
        } else {
            stockham_compute<N, E, R, DIR, Ns>(x, tid);
        }
        if constexpr (P + 1 < Plan::NP) {
            constexpr int RN = Plan::R[P + 1];
            stockham_store<N, E, R, Ns>(x, lds, tid);
            fft_sync<WAVE>();
            stockham_load<N, E, RN>(x, lds, tid);
            fft_sync<WAVE>();
            FftPassesCached<N, E, DIR, P + 1, Ns * R, TAB, WAVE>::run(x, lds, tid, tw, tab);
        }
    }
};

// tab[m] = exp(j 2 pi m / 256), m < 256, each from an FP64 sincospi rounded once; every thread of
// the workgroup calls it, and the transforms' first LDS barrier orders it before any use.
__device__ __forceinline__ void fft_tw16_table_fill(float2* tab, int tid, int nthreads) {
    for (int m = tid; m < 256; m += nthreads) tab[m] = expjpi_rounded(static_cast<double>(m) / 128.0);
}

// All twiddle powers w^r (r < E) of the passes that hold one radix group per thread, formed once
// per thread from the cached bases (for the + direction; the - direction uses their conjugates,
// which are bit for bit the powers of the conjugated bases). For kernels that run many
// transforms and have the registers to keep E complex values per such pass live across them.
template <int N, int E>
struct FftTwiddlePowers {
    float2 w[5][E];   // rows of the other passes are never formed or read
};

template <int N, int E, int DIR, int P, int Ns, bool WAVE = false>
struct FftPassesPow {
    using Plan = FftPlan<N, E>;
    __device__ __forceinline__ static void run(float2 (&x)[E], float2* lds, int tid, const FftTwiddlePowers<N, E>& pw) {
        fft_check_wave<N, E, WAVE>();
        constexpr int R = Plan::R[P];
        if constexpr (E == R) {
            if constexpr (Ns > 1) {
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    const float2 w = pw.w[P][r];
                    x[r] = cmul(x[r], DIR < 0 ? make_float2(w.x, -w.y) : w);
                }
            }
            Dft<R, DIR>::run(x);
        } else {
            stockham_compute<N, E, R, DIR, Ns>(x, tid);
        }
        if constexpr (P + 1 < Plan::NP) {
            constexpr int RN = Plan::R[P + 1];
            stockham_store<N, E, R, Ns>(x, lds, tid);
            fft_sync<WAVE>();
            stockham_load<N, E, RN>(x, lds, tid);
            fft_sync<WAVE>();
            FftPassesPow<N, E, DIR, P + 1, Ns * R, WAVE>::run(x, lds, tid, pw);
        }
    }
};

template <int N, int E, int P, int Ns>
__device__ __forceinline__ void fft_twiddle_powers_fill(FftTwiddlePowers<N, E>& pw, const FftTwiddles<N, E>& tw) {
    using Plan = FftPlan<N, E>;
    if constexpr (P < Plan::NP) {
        constexpr int R = Plan::R[P];
        if constexpr (Ns > 1 && R == E) twiddle_powers_from<R>(pw.w[P], tw.w1[P], tw.w4[P]);
        fft_twiddle_powers_fill<N, E, P + 1, Ns * R>(pw, tw);
    }
}

template <int N, int E, int P, int Ns>
__device__ __forceinline__ void fft_twiddles_fill(FftTwiddles<N, E>& tw, int tid) {
    using Plan = FftPlan<N, E>;
    if constexpr (P < Plan::NP) {
        constexpr int R = Plan::R[P];
        if constexpr (Ns > 1 && R == E) {
            // exp(j pi b) and exp(j pi 4 b), b = 2k / (Ns R), from an FP64 sincospi rounded once:
            // once per thread and kernel, so every transform of the kernel starts from bases
            // within half an FP32 ulp (the powers are products of these)
            const double base = static_cast<double>(2 * (tid % Ns)) / static_cast<double>(Ns * R);
            tw.w1[P] = expjpi_rounded(base);
            tw.w4[P] = R >= 8 ? expjpi_rounded(4.0 * base) : make_float2(1.0f, 0.0f);
        } else {
            tw.w1[P] = make_float2(1.0f, 0.0f);
            tw.w4[P] = make_float2(1.0f, 0.0f);
        }
        fft_twiddles_fill<N, E, P + 1, Ns * R>(tw, tid);
    }
}

template <int N, int E>
__device__ __forceinline__ FftTwiddles<N, E> fft_twiddles(int tid) {
    FftTwiddles<N, E> tw;
    fft_twiddles_fill<N, E, 0, 1>(tw, tid);
    return tw;
}

// In-register/LDS transform. On entry x[e] = in[fft_in_pos(tid, e)], on exit
// x[e] = out[fft_out_pos(tid, e)]. lds must hold lds_padded_size(N) float2 for this transform.
template <int N, int E, int DIR, bool WAVE = false>
__device__ __forceinline__ void fft(float2 (&x)[E], float2* lds, int tid) {
    FftPasses<N, E, DIR, 0, 1, WAVE>::run(x, lds, tid);
}
// The same transform with twiddle bases from fft_twiddles<N, E>(tid).
template <int N, int E, int DIR, bool WAVE = false>
__device__ __forceinline__ void fft(float2 (&x)[E], float2* lds, int tid, const FftTwiddles<N, E>& tw) {
    FftPassesCached<N, E, DIR, 0, 1, false, WAVE>::run(x, lds, tid, tw);
}
// ... and the stride-16 radix-16 pass's twiddles from an LDS table (fft_tw16_table_fill).
template <int N, int E, int DIR>
__device__ __forceinline__ void fft(float2 (&x)[E], float2* lds, int tid, const FftTwiddles<N, E>& tw,
                                    const float2* tab) {
    FftPassesCached<N, E, DIR, 0, 1, true>::run(x, lds, tid, tw, tab);
}
// The same transform with all twiddle powers precomputed (fft_twiddle_powers).
template <int N, int E>
__device__ __forceinline__ FftTwiddlePowers<N, E> fft_twiddle_powers(const FftTwiddles<N, E>& tw) {
    FftTwiddlePowers<N, E> pw;
    fft_twiddle_powers_fill<N, E, 0, 1>(pw, tw);
    return pw;
}
template <int N, int E, int DIR, bool WAVE = false>
__device__ __forceinline__ void fft(float2 (&x)[E], float2* lds, int tid, const FftTwiddlePowers<N, E>& pw) {
    FftPassesPow<N, E, DIR, 0, 1, WAVE>::run(x, lds, tid, pw);
}

template <int N, int E>
__device__ __forceinline__ constexpr int fft_first_radix() { return FftPlan<N, E>::R[0]; }
template <int N, int E>
__device__ __forceinline__ constexpr int fft_last_radix() { return FftPlan<N, E>::R[FftPlan<N, E>::NP - 1]; }

// Hamming window 0.54 - 0.46 cos(2 pi n / (len - 1))  (genHammingVec, common.cu:260-266)
__device__ __forceinline__ float hamming(int n, int len) {
    return 0.54f - 0.46f * cospif(2.0f * static_cast<float>(n) / static_cast<float>(len - 1));
}

}  // namespace isar

namespace isar {

// Sum over aligned groups of WIDTH lanes (WIDTH a power of two <= 64).
template <int WIDTH>
__device__ __forceinline__ float group_sum(float v) {
    static_assert(WIDTH <= 64, "shuffle groups cannot span waves");
#pragma unroll
    for (int o = WIDTH / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, WIDTH);
    return v;
}
template <int WIDTH>
__device__ __forceinline__ float2 group_sum2(float2 v) {
    static_assert(WIDTH <= 64, "shuffle groups cannot span waves");
#pragma unroll
    for (int o = WIDTH / 2; o > 0; o >>= 1) {
        v.x += __shfl_xor(v.x, o, WIDTH);
        v.y += __shfl_xor(v.y, o, WIDTH);
    }
    return v;
}

// Inclusive prefix sum in FP64 of n values held as v[i] (i < n) in LDS, in place, by the
// whole workgroup (NT threads). Each thread scans a contiguous segment sequentially.
template <int NT>
__device__ __forceinline__ void block_inclusive_scan_d(double* v, int n, double* red) {
    const int tid = threadIdx.x;
    const int seg = (n + NT - 1) / NT;
    const int b = tid * seg, e = min(b + seg, n);
    double run = 0.0;
    for (int i = b; i < e; ++i) { run += v[i]; v[i] = run; }
    __syncthreads();
    red[tid] = run;
    __syncthreads();
    if (tid == 0) {
        double acc = 0.0;
        for (int i = 0; i < NT; ++i) { const double t = red[i]; red[i] = acc; acc += t; }
    }
    __syncthreads();
    const double base = red[tid];
    for (int i = b; i < e; ++i) v[i] += base;
    __syncthreads();
}

// Column-transform configuration for M-point slow-time transforms inside a 256-thread block.
template <int M>
struct ColCfg {
    static constexpr int E = (M == 192 || M == 384 || M == 768) ? 12 : 16;
    static constexpr int TPF = M / E;                  // threads per transform
    static constexpr int NT = 256;
    static constexpr int CT = NT / TPF;                // transforms (columns) per block
    static constexpr int R0 = FftPlan<M, E>::R[0];
    static constexpr int RL = FftPlan<M, E>::R[FftPlan<M, E>::NP - 1];
    static constexpr bool same_layout = (R0 == RL);
    __device__ static int in(int lane, int e) { return fft_in_pos<M, E, R0>(lane, e); }
    __device__ static int out(int lane, int e) { return fft_out_pos<M, E, RL>(lane, e); }
};

// Move a transform result from the last-pass layout to the first-pass layout through LDS
// (WAVE: ordered as the transforms with fft_sync<WAVE>).
template <int M, int E, bool WAVE = false>
__device__ __forceinline__ void relayout_out_to_in(float2 (&x)[E], float2* buf, int lane) {
    fft_check_wave<M, E, WAVE>();
    constexpr int R0 = FftPlan<M, E>::R[0];
    constexpr int RL = FftPlan<M, E>::R[FftPlan<M, E>::NP - 1];
    if constexpr (R0 != RL) {
#pragma unroll
        for (int e = 0; e < E; ++e) buf[lds_pad(fft_out_pos<M, E, RL>(lane, e))] = x[e];
        fft_sync<WAVE>();
#pragma unroll
        for (int e = 0; e < E; ++e) x[e] = buf[lds_pad(fft_in_pos<M, E, R0>(lane, e))];
        fft_sync<WAVE>();
    }
}

// ------------------------------------------------------------------------------------------
// Hand-off between the workgroups of one launch (the resident autofocus). Values pass through
// memory as 64-bit words stored and loaded as agent-scope relaxed atomics: global accesses with
// sc1 set, which are written through to the device's point of coherence and read from it, never
// from a stale line of another XCD's L2. Either a word carries a 32-bit value and a 32-bit tag,
// and a reader polls it until the tag is the one it expects, so value and tag arrive together;
// or each writer waits for its stores to be acknowledged (s_waitcnt 0) before its workgroup
// counts itself in on a counter, and readers poll the counter. Neither involves a fence, an L2
// write-back or an invalidation.
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ uint64_t xwg_word(float v, unsigned tag) {
    return (static_cast<uint64_t>(tag) << 32) | static_cast<uint64_t>(__float_as_uint(v));
}
__device__ __forceinline__ unsigned xwg_tag(uint64_t w) { return static_cast<unsigned>(w >> 32); }
__device__ __forceinline__ float xwg_value(uint64_t w) { return __uint_as_float(static_cast<unsigned>(w)); }
__device__ __forceinline__ void xwg_store(uint64_t* p, uint64_t w) {
#ifdef ISAR_EMU
    __atomic_store_n(p, w, __ATOMIC_RELAXED);
#else
    __hip_atomic_store(p, w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#endif
}
__device__ __forceinline__ uint64_t xwg_load(const uint64_t* p) {
#ifdef ISAR_EMU
    return __atomic_load_n(p, __ATOMIC_RELAXED);
#else
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#endif
}
// Waits until this thread's stores are acknowledged.
__device__ __forceinline__ void xwg_stores_done() {
#ifdef ISAR_EMU
    __atomic_thread_fence(__ATOMIC_SEQ_CST);
#else
    __builtin_amdgcn_s_waitcnt(0);
#endif
}
__device__ __forceinline__ void xwg_arrive(unsigned* c) {
#ifdef ISAR_EMU
    __atomic_fetch_add(c, 1u, __ATOMIC_SEQ_CST);
#else
    __hip_atomic_fetch_add(c, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#endif
}
__device__ __forceinline__ void xwg_pause() {
#ifdef ISAR_EMU
    std::this_thread::yield();
#else
    __builtin_amdgcn_s_sleep(2);
#endif
}
// Polls *c until it reaches target; false when `limit` polls pass first, and at once when limit
// is negative.
__device__ __forceinline__ bool xwg_wait(const unsigned* c, unsigned target, int limit) {
    if (limit < 0) return false;
    for (int n = 0;; ++n) {
#ifdef ISAR_EMU
        if (__atomic_load_n(c, __ATOMIC_SEQ_CST) >= target) return true;
#else
        if (__hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= target) return true;
#endif
        if (n >= limit) return false;
        xwg_pause();
    }
}

}  // namespace isar
