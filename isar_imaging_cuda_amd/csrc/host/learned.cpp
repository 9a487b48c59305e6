// learned.cpp — the warm-start network object behind isar_learned.h: FP32 master parameters,
// BF16 weight copies, Adam moments and per-layer activations, all allocated at creation.
#include <climits>
#include <cmath>
#include <string>
#include <vector>

#include "isar_learned.h"
#include "pipeline.hpp"

namespace isar {

struct Mlp {
    std::vector<int> widths;          // n_layers + 1
    int max_batch = 0;
    int n_params = 0;
    std::vector<size_t> w_off, b_off;  // offsets of W_l and b_l in the parameter vector
    std::vector<size_t> wb_off;        // offsets of the BF16 copy of W_l
    float* params = nullptr;
    float* grads = nullptr;
    float* m1 = nullptr;
    float* m2 = nullptr;
    uint16_t* wbf = nullptr;
    int* bf_index = nullptr;           // parameter -> BF16 slot (-1 for biases)
    std::vector<float*> act;           // act[l], l = 1..n: batch x widths[l]
    float* dz[2] = {nullptr, nullptr};
    const float* input = nullptr;      // input of the last forward
    int last_batch = 0;
    std::vector<void*> allocations;

    int layers() const { return static_cast<int>(widths.size()) - 1; }
    ~Mlp() {
        for (void* p : allocations) (void)hipFree(p);
    }
};

}  // namespace isar

struct isar_mlp : isar::Mlp {};

using isar::fail;

namespace {
// The element-wise phase kernels index frames x M, and the GEMMs batch x width, with 32-bit ints
// (grid size included): larger requests are refused rather than wrapped.
bool int_indexable(long long a, long long b) { return a * b <= static_cast<long long>(INT_MAX) - 1024; }
}  // namespace

extern "C" {

ISAR_API int isar_mlp_create(int n_layers, const int* widths, int max_batch, isar_mlp** out) {
    if (!out || !widths || n_layers < 1 || n_layers > 16 || max_batch < 1)
        return fail(ISAR_ERR_INVALID, "isar_mlp_create: bad arguments");
    for (int l = 0; l <= n_layers; ++l) {
        if (widths[l] < 1 || widths[l] > 8192) return fail(ISAR_ERR_INVALID, "isar_mlp_create: width out of range");
        if (!int_indexable(max_batch, widths[l])) return fail(ISAR_ERR_INVALID, "isar_mlp_create: max_batch x width too large");
    }
    auto* net = new isar_mlp();
    net->widths.assign(widths, widths + n_layers + 1);
    net->max_batch = max_batch;
    size_t np = 0, nw = 0, maxw = 0;
    for (int l = 0; l < n_layers; ++l) {
        const size_t in = widths[l], o = widths[l + 1];
        net->w_off.push_back(np);
        np += in * o;
        net->b_off.push_back(np);
        np += o;
        net->wb_off.push_back(nw);
        nw += in * o;
        maxw = std::max(maxw, std::max(in, o));
    }
    net->n_params = static_cast<int>(np);
    bool ok = true;
    auto alloc = [&](size_t bytes) -> void* {
        void* p = nullptr;
        if (!ok || hipMalloc(&p, bytes ? bytes : 16) != hipSuccess) {
            ok = false;
            return nullptr;
        }
        net->allocations.push_back(p);
        return p;
    };
    net->params = static_cast<float*>(alloc(np * sizeof(float)));
    net->grads = static_cast<float*>(alloc(np * sizeof(float)));
    net->m1 = static_cast<float*>(alloc(np * sizeof(float)));
    net->m2 = static_cast<float*>(alloc(np * sizeof(float)));
    net->wbf = static_cast<uint16_t*>(alloc(nw * sizeof(uint16_t)));
    net->bf_index = static_cast<int*>(alloc(np * sizeof(int)));
    net->act.assign(n_layers + 1, nullptr);
    for (int l = 1; l <= n_layers; ++l)
        net->act[l] = static_cast<float*>(alloc(static_cast<size_t>(max_batch) * widths[l] * sizeof(float)));
    net->dz[0] = static_cast<float*>(alloc(static_cast<size_t>(max_batch) * maxw * sizeof(float)));
    net->dz[1] = static_cast<float*>(alloc(static_cast<size_t>(max_batch) * maxw * sizeof(float)));
    if (!ok) {
        delete net;
        return fail(ISAR_ERR_MEMORY, "isar_mlp_create: device allocation failed");
    }
    std::vector<int> idx(np, -1);
    for (int l = 0; l < n_layers; ++l) {
        const size_t n = static_cast<size_t>(widths[l]) * widths[l + 1];
        for (size_t i = 0; i < n; ++i) idx[net->w_off[l] + i] = static_cast<int>(net->wb_off[l] + i);
    }
    if (hipMemcpy(net->bf_index, idx.data(), np * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(net->params, 0, np * sizeof(float)) != hipSuccess ||
        hipMemset(net->grads, 0, np * sizeof(float)) != hipSuccess ||
        hipMemset(net->m1, 0, np * sizeof(float)) != hipSuccess ||
        hipMemset(net->m2, 0, np * sizeof(float)) != hipSuccess ||
        hipMemset(net->wbf, 0, nw * sizeof(uint16_t)) != hipSuccess) {
        delete net;
        return fail(ISAR_ERR_DEVICE, "isar_mlp_create: initialisation failed");
    }
    *out = net;
    return ISAR_OK;
}

ISAR_API int isar_mlp_destroy(isar_mlp* net) {
    delete net;
    return ISAR_OK;
}

ISAR_API int isar_mlp_param_count(const isar_mlp* net) { return net ? net->n_params : 0; }

ISAR_API int isar_mlp_set_params(isar_mlp* net, const float* h_params) {
    if (!net || !h_params) return fail(ISAR_ERR_INVALID, "isar_mlp_set_params: null argument");
    const size_t bytes = static_cast<size_t>(net->n_params) * sizeof(float);
    if (hipMemcpy(net->params, h_params, bytes, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(net->m1, 0, bytes) != hipSuccess || hipMemset(net->m2, 0, bytes) != hipSuccess)
        return fail(ISAR_ERR_DEVICE, "isar_mlp_set_params: copy failed");
    isar::launch_to_bf16(net->n_params, net->params, net->wbf, net->bf_index, nullptr);
    if (hipDeviceSynchronize() != hipSuccess) return fail(ISAR_ERR_DEVICE, "isar_mlp_set_params: conversion failed");
    return ISAR_OK;
}

ISAR_API int isar_mlp_get_params(const isar_mlp* net, float* h_params) {
    if (!net || !h_params) return fail(ISAR_ERR_INVALID, "isar_mlp_get_params: null argument");
    if (hipMemcpy(h_params, net->params, static_cast<size_t>(net->n_params) * sizeof(float), hipMemcpyDeviceToHost) !=
        hipSuccess)
        return fail(ISAR_ERR_DEVICE, "isar_mlp_get_params: copy failed");
    return ISAR_OK;
}

ISAR_API int isar_mlp_get_grads(const isar_mlp* net, float* h_grads) {
    if (!net || !h_grads) return fail(ISAR_ERR_INVALID, "isar_mlp_get_grads: null argument");
    if (hipMemcpy(h_grads, net->grads, static_cast<size_t>(net->n_params) * sizeof(float), hipMemcpyDeviceToHost) !=
        hipSuccess)
        return fail(ISAR_ERR_DEVICE, "isar_mlp_get_grads: copy failed");
    return ISAR_OK;
}

ISAR_API int isar_mlp_forward(isar_mlp* net, int batch, const float* d_in, float* d_out, void* stream) {
    if (!net || !d_in || !d_out) return fail(ISAR_ERR_INVALID, "isar_mlp_forward: null argument");
    if (batch < 1 || batch > net->max_batch) return fail(ISAR_ERR_INVALID, "isar_mlp_forward: batch out of range");
    const hipStream_t s = static_cast<hipStream_t>(stream);
    const int L = net->layers();
    const float* a = d_in;
    for (int l = 0; l < L; ++l) {
        const int in = net->widths[l], o = net->widths[l + 1];
        float* z = net->act[l + 1];
        isar::launch_gemm_f32_bf16w(batch, o, in, a, in, 1, net->wbf + net->wb_off[l], o, 1, z, o,
                                    net->params + net->b_off[l], l + 1 < L ? 1 : 0, s);
        a = z;
    }
    const size_t bytes = static_cast<size_t>(batch) * net->widths[L] * sizeof(float);
    if (hipMemcpyAsync(d_out, net->act[L], bytes, hipMemcpyDeviceToDevice, s) != hipSuccess)
        return fail(ISAR_ERR_DEVICE, "isar_mlp_forward: output copy failed");
    if (hipGetLastError() != hipSuccess) return fail(ISAR_ERR_DEVICE, "isar_mlp_forward: launch failed");
    net->input = d_in;
    net->last_batch = batch;
    return ISAR_OK;
}

ISAR_API int isar_mlp_backward(isar_mlp* net, int batch, const float* d_grad_out, void* stream) {
    if (!net || !d_grad_out) return fail(ISAR_ERR_INVALID, "isar_mlp_backward: null argument");
    if (batch != net->last_batch || !net->input)
        return fail(ISAR_ERR_STATE, "isar_mlp_backward: needs a forward pass of the same batch first");
    const hipStream_t s = static_cast<hipStream_t>(stream);
    const int L = net->layers();
    const float* dz = d_grad_out;
    int ping = 0;
    for (int l = L - 1; l >= 0; --l) {
        const int in = net->widths[l], o = net->widths[l + 1];
        const float* a_prev = l == 0 ? net->input : net->act[l];
        // dW = A_prev^T dZ (in x o), db = column sums of dZ
        isar::launch_gemm_f32_f32(in, o, batch, a_prev, 1, in, dz, o, 1, net->grads + net->w_off[l], o, s);
        isar::launch_colsum(batch, o, dz, net->grads + net->b_off[l], s);
        if (l > 0) {
            // dA_prev = dZ W^T, then through the ReLU of layer l - 1 (its output is A_prev)
            float* next = net->dz[ping];
            isar::launch_gemm_f32_bf16w_masked(batch, in, o, dz, o, 1, net->wbf + net->wb_off[l], 1, o, next, in,
                                               a_prev, in, s);
            dz = next;
            ping ^= 1;
        }
    }
    if (hipGetLastError() != hipSuccess) return fail(ISAR_ERR_DEVICE, "isar_mlp_backward: launch failed");
    return ISAR_OK;
}

ISAR_API int isar_mlp_adam_step(isar_mlp* net, float lr, float beta1, float beta2, float eps, int step,
                                void* stream) {
    if (!net || step < 1) return fail(ISAR_ERR_INVALID, "isar_mlp_adam_step: bad arguments");
    const float c1 = 1.0f - std::pow(beta1, static_cast<float>(step));
    const float c2 = 1.0f - std::pow(beta2, static_cast<float>(step));
    isar::launch_adam(net->n_params, net->params, net->grads, net->m1, net->m2, net->wbf, net->bf_index, lr, beta1,
                      beta2, eps, c1, c2, static_cast<hipStream_t>(stream));
    if (hipGetLastError() != hipSuccess) return fail(ISAR_ERR_DEVICE, "isar_mlp_adam_step: launch failed");
    return ISAR_OK;
}

ISAR_API int isar_mse(int rows, int cols, const float* d_pred, const float* d_target, float* d_grad,
                      double* d_loss_rows, void* stream) {
    if (rows < 1 || cols < 1 || !d_pred || !d_target || !d_loss_rows)
        return fail(ISAR_ERR_INVALID, "isar_mse: bad arguments");
    isar::launch_mse(rows, cols, d_pred, d_target, d_grad, d_loss_rows, static_cast<hipStream_t>(stream));
    return hipGetLastError() == hipSuccess ? ISAR_OK : fail(ISAR_ERR_DEVICE, "isar_mse: launch failed");
}

ISAR_API int isar_phase_change(int frames, int M, const void* d_phi_a, const void* d_phi_b, float* d_out,
                               int out_stride, void* stream) {
    if (frames < 1 || M < 1 || !int_indexable(frames, M) || !d_phi_a || !d_phi_b || !d_out || out_stride < 2LL * M)
        return fail(ISAR_ERR_INVALID, "isar_phase_change: bad arguments");
    isar::launch_phase_change(frames, M, static_cast<const float2*>(d_phi_a), static_cast<const float2*>(d_phi_b),
                              d_out, out_stride, static_cast<hipStream_t>(stream));
    return hipGetLastError() == hipSuccess ? ISAR_OK : fail(ISAR_ERR_DEVICE, "isar_phase_change: launch failed");
}

ISAR_API int isar_compose_phase(int frames, int M, const void* d_phi_b, const float* d_pred, void* d_phi_w,
                                void* stream) {
    if (frames < 1 || M < 1 || !int_indexable(frames, M) || !d_phi_b || !d_pred || !d_phi_w)
        return fail(ISAR_ERR_INVALID, "isar_compose_phase: bad arguments");
    isar::launch_compose_phase(frames, M, static_cast<const float2*>(d_phi_b), d_pred, static_cast<float2*>(d_phi_w),
                               static_cast<hipStream_t>(stream));
    return hipGetLastError() == hipSuccess ? ISAR_OK : fail(ISAR_ERR_DEVICE, "isar_compose_phase: launch failed");
}

ISAR_API int isar_select_phase(int frames, int M, const void* d_phi_a, const double* d_ent_a, const void* d_phi_b,
                               const double* d_ent_b, void* d_out, double* d_ent_out, void* stream) {
    if (frames < 1 || M < 1 || !int_indexable(frames, M) || !d_phi_a || !d_ent_a || !d_phi_b || !d_ent_b || !d_out)
        return fail(ISAR_ERR_INVALID, "isar_select_phase: bad arguments");
    isar::launch_select_phase(frames, M, static_cast<const float2*>(d_phi_a), d_ent_a, static_cast<const float2*>(d_phi_b),
                              d_ent_b, static_cast<float2*>(d_out), d_ent_out, static_cast<hipStream_t>(stream));
    return hipGetLastError() == hipSuccess ? ISAR_OK : fail(ISAR_ERR_DEVICE, "isar_select_phase: launch failed");
}

ISAR_API int isar_half_phase(int frames, int M, const void* d_phi_a, const void* d_phi_b, void* d_out, void* stream) {
    if (frames < 1 || M < 1 || !int_indexable(frames, M) || !d_phi_a || !d_phi_b || !d_out) return fail(ISAR_ERR_INVALID, "isar_half_phase: bad arguments");
    isar::launch_half_phase(frames, M, static_cast<const float2*>(d_phi_a), static_cast<const float2*>(d_phi_b),
                            static_cast<float2*>(d_out), static_cast<hipStream_t>(stream));
    return hipGetLastError() == hipSuccess ? ISAR_OK : fail(ISAR_ERR_DEVICE, "isar_half_phase: launch failed");
}

ISAR_API int isar_accept_image(int frames, int n, const float* d_img_plain, const double* d_ent_plain,
                               const float* d_img_learned, const double* d_ent_learned, double rel_tol, float* d_out,
                               int32_t* d_accepted, void* stream) {
    if (frames < 1 || n < 1 || !d_img_plain || !d_ent_plain || !d_img_learned || !d_ent_learned || !d_out ||
        !(rel_tol >= 0.0))
        return fail(ISAR_ERR_INVALID, "isar_accept_image: bad arguments");
    isar::launch_accept_image(frames, n, d_img_plain, d_ent_plain, d_img_learned, d_ent_learned, rel_tol, d_out,
                              d_accepted, static_cast<hipStream_t>(stream));
    return hipGetLastError() == hipSuccess ? ISAR_OK : fail(ISAR_ERR_DEVICE, "isar_accept_image: launch failed");
}

ISAR_API int isar_image_entropy(int frames, int n, const float* d_img, double* d_out, void* stream) {
    if (frames < 1 || n < 1 || !d_img || !d_out) return fail(ISAR_ERR_INVALID, "isar_image_entropy: bad arguments");
    isar::launch_image_entropy(frames, n, d_img, d_out, static_cast<hipStream_t>(stream));
    return hipGetLastError() == hipSuccess ? ISAR_OK : fail(ISAR_ERR_DEVICE, "isar_image_entropy: launch failed");
}

}  // extern "C"
