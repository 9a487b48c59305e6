// nn.hip — kernels of the learned autofocus warm start: a small multilayer perceptron with BF16
// weights and FP32 accumulation on the matrix cores (gfx950's v_mfma_f32_16x16x32_bf16), its backward
// pass and Adam update, and the per-pulse phase features it reads and writes.
//
// There is no counterpart in the reference: the network seeds fast-entropy autofocus
// (phase_adjustment.cu:282-324) with a predicted per-pulse phase; the plain iteration still runs
// from that start and remains the judge of the result.
#include "../dev_common.hpp"
#include "../launchers.hpp"

namespace isar {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// float -> bfloat16 bits, round to nearest even (finite inputs)
__device__ __forceinline__ short to_bf16(float f) {
    unsigned u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return static_cast<short>(u >> 16);
}
__device__ __forceinline__ __bf16 bits_bf16(short s) { return __builtin_bit_cast(__bf16, s); }
__device__ __forceinline__ __bf16 load_bf16(const float* p) { return bits_bf16(to_bf16(*p)); }
__device__ __forceinline__ __bf16 load_bf16(const uint16_t* p) { return bits_bf16(static_cast<short>(*p)); }

// D = A B + C for one 16 x 16 x 32 tile. Lane l holds A[l % 16][8 (l / 16) + t] and
// B[8 (l / 16) + t][l % 16] for t < 8, and D[4 (l / 16) + t][l % 16] for t < 4.
__device__ __forceinline__ f32x4 mfma16x32(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

constexpr int kGemmTile = 32;   // output rows and columns per wave (2 x 2 MFMA tiles)

}  // namespace

// C[m][n] = sum_k A(m, k) B(k, n) with A(m, k) = a[m sam + k sak], B(k, n) = b[k sbk + n sbn],
// operands rounded to BF16, products accumulated in FP32. One wave per 32 x 32 block of C.
// Epilogue: 0 store; 1 + bias[n], then max(0, .); 2 + bias[n]; 3 times (mask[m][n] > 0).
template <typename TA, typename TB, int EPI>
__global__ __launch_bounds__(64) void k_gemm_bf16(int m, int n, int k, const TA* __restrict__ a, long sam, long sak,
                                                  const TB* __restrict__ b, long sbk, long sbn, float* __restrict__ c,
                                                  long scm, const float* __restrict__ bias,
                                                  const float* __restrict__ mask, long smm) {
    const int l = threadIdx.x;
    const int bm = blockIdx.y * kGemmTile, bn = blockIdx.x * kGemmTile;
    const int li = l % 16, lk = 8 * (l / 16), lr = 4 * (l / 16);
    f32x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    const __bf16 zero = bits_bf16(0);
    for (int k0 = 0; k0 < k; k0 += 32) {
        bf16x8 fa[2], fb[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const int row = bm + 16 * i + li;
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                const int kk = k0 + lk + t;
                fa[i][t] = (row < m && kk < k) ? load_bf16(a + row * sam + kk * sak) : zero;
            }
        }
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int col = bn + 16 * j + li;
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                const int kk = k0 + lk + t;
                fb[j][t] = (col < n && kk < k) ? load_bf16(b + kk * sbk + col * sbn) : zero;
            }
        }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) acc[i][j] = mfma16x32(fa[i], fb[j], acc[i][j]);
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int col = bn + 16 * j + li;
            if (col >= n) continue;
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int row = bm + 16 * i + lr + t;
                if (row >= m) continue;
                float v = acc[i][j][t];
                if constexpr (EPI == 1) v = fmaxf(v + bias[col], 0.0f);
                if constexpr (EPI == 2) v += bias[col];
                if constexpr (EPI == 3) v = mask[row * smm + col] > 0.0f ? v : 0.0f;
                c[row * scm + col] = v;
            }
        }
}

// bias gradient: db[j] = sum_b dz[b][j]
__global__ void k_colsum(int rows, int cols, const float* __restrict__ dz, float* __restrict__ db) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= cols) return;
    float s = 0.0f;
    for (int r = 0; r < rows; ++r) s += dz[static_cast<size_t>(r) * cols + j];
    db[j] = s;
}

// Adam (bias-corrected) on FP32 master parameters; the BF16 copy of each weight follows.
__global__ void k_adam(int n, float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m1,
                       float* __restrict__ m2, uint16_t* __restrict__ pb, const int* __restrict__ bf_index,
                       float lr, float beta1, float beta2, float eps, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float a = beta1 * m1[i] + (1.0f - beta1) * gi;
    const float b = beta2 * m2[i] + (1.0f - beta2) * gi * gi;
    m1[i] = a;
    m2[i] = b;
    const float v = p[i] - lr * (a / c1) / (sqrtf(b / c2) + eps);
    p[i] = v;
    const int bi = bf_index[i];
    if (bi >= 0) pb[bi] = static_cast<uint16_t>(to_bf16(v));
}

// BF16 copies of the weights from the FP32 master parameters
__global__ void k_to_bf16(int n, const float* __restrict__ p, uint16_t* __restrict__ pb,
                          const int* __restrict__ bf_index) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int bi = bf_index[i];
    if (bi >= 0) pb[bi] = static_cast<uint16_t>(to_bf16(p[i]));
}

// mean squared error over rows x cols and its gradient 2 (pred - target) / (rows cols)
__global__ void k_mse(int rows, int cols, const float* __restrict__ pred, const float* __restrict__ target,
                      float* __restrict__ grad, double* __restrict__ loss) {
    __shared__ double red[256];
    const int r = blockIdx.x, tid = threadIdx.x;
    const float scale = 2.0f / (static_cast<float>(rows) * static_cast<float>(cols));
    double s = 0.0;
    for (int j = tid; j < cols; j += blockDim.x) {
        const size_t i = static_cast<size_t>(r) * cols + j;
        const float d = pred[i] - target[i];
        if (grad) grad[i] = scale * d;
        s += static_cast<double>(d) * d;
    }
    red[tid] = s;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if (tid < o) red[tid] += red[tid + o];
        __syncthreads();
    }
    if (tid == 0) loss[r] = red[0] / (static_cast<double>(rows) * cols);
}

// Per-pulse phase change between two autofocus phases: out[f stride + 2m] + j out[f stride + 2m+1]
// = phi_b[f][m] conj(phi_a[f][m]).
__global__ void k_phase_change(int frames, int M, const float2* __restrict__ phi_a, const float2* __restrict__ phi_b,
                               float* __restrict__ out, int stride) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= frames * M) return;
    const int f = i / M, m = i % M;
    const float2 v = cmulc(phi_b[i], phi_a[i]);
    float* o = out + static_cast<size_t>(f) * stride + 2 * m;
    o[0] = v.x;
    o[1] = v.y;
}

// Warm start phi_w = phi_b * unit(pred); a zero prediction keeps phi_b.
__global__ void k_compose_phase(int frames, int M, const float2* __restrict__ phi_b, const float* __restrict__ pred,
                                float2* __restrict__ phi_w) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= frames * M) return;
    const float2 q = make_float2(pred[2 * static_cast<size_t>(i)], pred[2 * static_cast<size_t>(i) + 1]);
    const float a = cabsf_(q);
    phi_w[i] = a > 0.0f ? cmul(phi_b[i], cscale(q, 1.0f / a)) : phi_b[i];
}

// Per frame, the starting phase whose image entropy is lower: out = ent_b < ent_a ? phi_b : phi_a
// (ties keep phi_a); ent_out (optional) receives the entropy of the phase kept.
__global__ void k_select_phase(int frames, int M, const float2* __restrict__ phi_a, const double* __restrict__ ent_a,
                               const float2* __restrict__ phi_b, const double* __restrict__ ent_b,
                               float2* __restrict__ out, double* __restrict__ ent_out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= frames * M) return;
    const int f = i / M;
    const bool take_b = ent_b[f] < ent_a[f];
    out[i] = take_b ? phi_b[i] : phi_a[i];
    if (ent_out && i % M == 0) ent_out[f] = take_b ? ent_b[f] : ent_a[f];
}

// The final judgement of a learned result: per frame, out = img_b (the learned image) when its
// entropy is no worse than img_a's (the plain result) within rel_tol, else img_a; accepted[f]
// records the choice. out may alias either input.
__global__ void k_accept_image(int frames, int n, const float* __restrict__ img_a, const double* __restrict__ ent_a,
                               const float* img_b, const double* __restrict__ ent_b, double rel_tol, float* out,
                               int32_t* __restrict__ accepted) {
    const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= static_cast<long long>(frames) * n) return;
    const int f = static_cast<int>(i / n);
    const bool take_b = ent_b[f] <= ent_a[f] * (1.0 + rel_tol);
    out[i] = take_b ? img_b[i] : img_a[i];
    if (accepted && i % n == 0) accepted[f] = take_b ? 1 : 0;
}

// Half step from phi_a toward phi_b: out = phi_a * sqrt(unit(phi_b conj(phi_a))), the square root
// taken on the shorter arc (unit(1 + c)); an exactly opposite pair turns by +pi/2.
__global__ void k_half_phase(int frames, int M, const float2* __restrict__ phi_a, const float2* __restrict__ phi_b,
                             float2* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= frames * M) return;
    const float2 a = phi_a[i];
    float2 c = cconjmul(a, phi_b[i]);                  // phi_b conj(phi_a)
    const float cm = cabsf_(c);
    c = cm > 0.0f ? cscale(c, 1.0f / cm) : make_float2(1.0f, 0.0f);
    float2 h = make_float2(1.0f + c.x, c.y);
    const float hm = cabsf_(h);
    h = hm > 1e-20f ? cscale(h, 1.0f / hm) : make_float2(0.0f, 1.0f);
    out[i] = cmul(a, h);
}

// Image entropy -sum p ln p, p = img^2 / sum img^2, one workgroup per image, FP64 sums.
__global__ void k_image_entropy(int n, const float* __restrict__ img, double* __restrict__ out) {
    __shared__ double red[256];
    const int f = blockIdx.x, tid = threadIdx.x;
    const float* x = img + static_cast<size_t>(f) * n;
    double s = 0.0;
    for (int i = tid; i < n; i += blockDim.x) s += static_cast<double>(x[i]) * x[i];
    red[tid] = s;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if (tid < o) red[tid] += red[tid + o];
        __syncthreads();
    }
    const double total = red[0];
    __syncthreads();
    double h = 0.0;
    if (total > 0.0)
        for (int i = tid; i < n; i += blockDim.x) {
            const double p = static_cast<double>(x[i]) * x[i] / total;
            if (p > 0.0) h -= p * log(p);
        }
    red[tid] = h;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if (tid < o) red[tid] += red[tid + o];
        __syncthreads();
    }
    if (tid == 0) out[f] = red[0];
}

// ------------------------------------------------------------------------------------------
// host launchers
// ------------------------------------------------------------------------------------------
namespace {
dim3 gemm_grid(int m, int n) { return dim3((n + kGemmTile - 1) / kGemmTile, (m + kGemmTile - 1) / kGemmTile); }
}  // namespace

void launch_gemm_f32_bf16w(int m, int n, int k, const float* a, long sam, long sak, const uint16_t* w, long sbk,
                           long sbn, float* c, long scm, const float* bias, int relu, hipStream_t s) {
    if (relu)
        ISAR_LAUNCH((k_gemm_bf16<float, uint16_t, 1>), gemm_grid(m, n), dim3(64), 0, s, m, n, k, a, sam, sak, w, sbk,
                    sbn, c, scm, bias, static_cast<const float*>(nullptr), 0L);
    else
        ISAR_LAUNCH((k_gemm_bf16<float, uint16_t, 2>), gemm_grid(m, n), dim3(64), 0, s, m, n, k, a, sam, sak, w, sbk,
                    sbn, c, scm, bias, static_cast<const float*>(nullptr), 0L);
}
void launch_gemm_f32_bf16w_masked(int m, int n, int k, const float* a, long sam, long sak, const uint16_t* w,
                                  long sbk, long sbn, float* c, long scm, const float* mask, long smm, hipStream_t s) {
    ISAR_LAUNCH((k_gemm_bf16<float, uint16_t, 3>), gemm_grid(m, n), dim3(64), 0, s, m, n, k, a, sam, sak, w, sbk, sbn,
                c, scm, static_cast<const float*>(nullptr), mask, smm);
}
void launch_gemm_f32_f32(int m, int n, int k, const float* a, long sam, long sak, const float* b, long sbk, long sbn,
                         float* c, long scm, hipStream_t s) {
    ISAR_LAUNCH((k_gemm_bf16<float, float, 0>), gemm_grid(m, n), dim3(64), 0, s, m, n, k, a, sam, sak, b, sbk, sbn, c,
                scm, static_cast<const float*>(nullptr), static_cast<const float*>(nullptr), 0L);
}
void launch_colsum(int rows, int cols, const float* dz, float* db, hipStream_t s) {
    ISAR_LAUNCH(k_colsum, dim3((cols + 255) / 256), dim3(256), 0, s, rows, cols, dz, db);
}
void launch_adam(int n, float* p, const float* g, float* m1, float* m2, uint16_t* pb, const int* bf_index, float lr,
                 float beta1, float beta2, float eps, float c1, float c2, hipStream_t s) {
    ISAR_LAUNCH(k_adam, dim3((n + 255) / 256), dim3(256), 0, s, n, p, g, m1, m2, pb, bf_index, lr, beta1, beta2, eps,
                c1, c2);
}
void launch_to_bf16(int n, const float* p, uint16_t* pb, const int* bf_index, hipStream_t s) {
    ISAR_LAUNCH(k_to_bf16, dim3((n + 255) / 256), dim3(256), 0, s, n, p, pb, bf_index);
}
void launch_mse(int rows, int cols, const float* pred, const float* target, float* grad, double* loss, hipStream_t s) {
    ISAR_LAUNCH(k_mse, dim3(rows), dim3(256), 0, s, rows, cols, pred, target, grad, loss);
}
void launch_phase_change(int frames, int M, const float2* a, const float2* b, float* out, int stride, hipStream_t s) {
    ISAR_LAUNCH(k_phase_change, dim3((frames * M + 255) / 256), dim3(256), 0, s, frames, M, a, b, out, stride);
}
void launch_compose_phase(int frames, int M, const float2* phi_b, const float* pred, float2* phi_w, hipStream_t s) {
    ISAR_LAUNCH(k_compose_phase, dim3((frames * M + 255) / 256), dim3(256), 0, s, frames, M, phi_b, pred, phi_w);
}
void launch_select_phase(int frames, int M, const float2* phi_a, const double* ent_a, const float2* phi_b,
                         const double* ent_b, float2* out, double* ent_out, hipStream_t s) {
    ISAR_LAUNCH(k_select_phase, dim3((frames * M + 255) / 256), dim3(256), 0, s, frames, M, phi_a, ent_a, phi_b, ent_b,
                out, ent_out);
}
void launch_half_phase(int frames, int M, const float2* phi_a, const float2* phi_b, float2* out, hipStream_t s) {
    ISAR_LAUNCH(k_half_phase, dim3((frames * M + 255) / 256), dim3(256), 0, s, frames, M, phi_a, phi_b, out);
}
void launch_accept_image(int frames, int n, const float* img_a, const double* ent_a, const float* img_b,
                         const double* ent_b, double rel_tol, float* out, int32_t* accepted, hipStream_t s) {
    const long long total = static_cast<long long>(frames) * n;
    ISAR_LAUNCH(k_accept_image, dim3(static_cast<unsigned>((total + 255) / 256)), dim3(256), 0, s, frames, n, img_a,
                ent_a, img_b, ent_b, rel_tol, out, accepted);
}
void launch_image_entropy(int frames, int n, const float* img, double* out, hipStream_t s) {
    ISAR_LAUNCH(k_image_entropy, dim3(frames), dim3(256), 0, s, n, img, out);
}

}  // namespace isar
