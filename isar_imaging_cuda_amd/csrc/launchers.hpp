// launchers.hpp — host entry points of the kernel families (one per stage).
#pragma once

#include "pipeline_types.hpp"

namespace isar {

// Range sizes N and pulse counts M the kernels are instantiated for.
bool range_size_supported(int N);
bool pulse_count_supported(int M);

// Sets the dynamic-LDS limits of every kernel instantiation (call once per process/device).
hipError_t configure_kernels();
// Records which kernel family could not be configured (for isar_last_error).
void report_config_error(const char* file, size_t lds_bytes);

// range alignment (range_alignment.cu:4-78)
void launch_ifds_pre(const Job& J, hipStream_t s);
void launch_hrrp(const Job& J, hipStream_t s);
void launch_ra_pair0(const Job& J, hipStream_t s);
void launch_ra_corr(const Job& J, int stride, int pair_base, int nblocks, hipStream_t s);
void launch_ra_upd(const Job& J, int stride, int pair_base, int nblocks, hipStream_t s);

// HRRP centring, cut and range-window statistics (range_alignment.cu:371-465, common.cu:494-518)
void launch_center(const Job& J, hipStream_t s);
void launch_cut(const Job& J, hipStream_t s);

// phase adjustment (phase_adjustment.cu:4-41, 167-346)
void launch_pa_colfft(const Job& J, hipStream_t s);
void launch_pa_select(const Job& J, hipStream_t s);
void launch_autofocus(const Job& J, hipStream_t s);
void launch_autofocus_split(const Job& J, bool fold, hipStream_t s);
// all iterations in one launch, each frame over several workgroups that stay resident together
void launch_autofocus_resident(const Job& J, hipStream_t s);

// MTRC and azimuth compression (mtrc.cu:4-192, isar_main.cu:191-207)
void launch_czt_table(int M, int W, int F, double c0, double posa, float2* table, hipStream_t s);
void launch_range_window(int N, float* win, hipStream_t s);
void launch_czt_chirps(int M, int W, int F, double c0, double posa, float2* pre, float2* post, hipStream_t s);
void launch_mtrc_delta(const Job& J, hipStream_t s);
void launch_mtrc_rows(const Job& J, hipStream_t s);
void launch_mtrc_czt(const Job& J, hipStream_t s);
void launch_col_window(const Job& J, hipStream_t s);
void launch_mtrc_image(const Job& J, hipStream_t s);
void launch_post_nomtrc(const Job& J, hipStream_t s);

// IFDS pulse compression (common.cu:583-604, 635-746)
bool pulse_compress_supported(int nfft);
void launch_pulse_compress(const PcJob& P, int nfft, hipStream_t s);

// learned warm start: BF16 matrix-core GEMMs with FP32 accumulation and their helpers
void launch_gemm_f32_bf16w(int m, int n, int k, const float* a, long sam, long sak, const uint16_t* w, long sbk,
                           long sbn, float* c, long scm, const float* bias, int relu, hipStream_t s);
void launch_gemm_f32_bf16w_masked(int m, int n, int k, const float* a, long sam, long sak, const uint16_t* w,
                                  long sbk, long sbn, float* c, long scm, const float* mask, long smm, hipStream_t s);
void launch_gemm_f32_f32(int m, int n, int k, const float* a, long sam, long sak, const float* b, long sbk, long sbn,
                         float* c, long scm, hipStream_t s);
void launch_colsum(int rows, int cols, const float* dz, float* db, hipStream_t s);
void launch_adam(int n, float* p, const float* g, float* m1, float* m2, uint16_t* pb, const int* bf_index, float lr,
                 float beta1, float beta2, float eps, float c1, float c2, hipStream_t s);
void launch_to_bf16(int n, const float* p, uint16_t* pb, const int* bf_index, hipStream_t s);
void launch_mse(int rows, int cols, const float* pred, const float* target, float* grad, double* loss, hipStream_t s);
void launch_phase_change(int frames, int M, const float2* a, const float2* b, float* out, int stride, hipStream_t s);
void launch_compose_phase(int frames, int M, const float2* phi_b, const float* pred, float2* phi_w, hipStream_t s);
void launch_image_entropy(int frames, int n, const float* img, double* out, hipStream_t s);
void launch_select_phase(int frames, int M, const float2* phi_a, const double* ent_a, const float2* phi_b,
                         const double* ent_b, float2* out, double* ent_out, hipStream_t s);
void launch_half_phase(int frames, int M, const float2* phi_a, const float2* phi_b, float2* out, hipStream_t s);
void launch_accept_image(int frames, int n, const float* img_a, const double* ent_a, const float* img_b,
                         const double* ent_b, double rel_tol, float* out, int32_t* accepted, hipStream_t s);

// standalone batched FFT used by the tests: n-point transforms of `batch` contiguous rows
bool debug_fft(int n, int dir, const float2* in, float2* out, int batch, hipStream_t s);
bool debug_fft_grouped(int n, int dir, const float2* in, float2* out, int batch, hipStream_t s);

// stage.hip: intermediate arrays of a finished run, rebuilt for inspection
void launch_stage_hpc(const float2* echo, const double* vel, double hpc_k, int M, int N, float2* out, hipStream_t s);
void launch_stage_roll(const float2* st, int M, int N, int shift, float2* out, hipStream_t s);
void launch_stage_rowphase(const float2* x, const float2* phi, int M, int W, float2* out, hipStream_t s);
void launch_stage_transpose(const float2* in, int R, int C, int sr, int sc, float scale, float2* out, hipStream_t s);
void launch_stage_mtrc_out(const float2* a, int M, int W, int ifds, float2* out, hipStream_t s);

}  // namespace isar
