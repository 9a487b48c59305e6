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

// MTRC and azimuth compression (mtrc.cu:4-192, isar_main.cu:191-207)
void launch_czt_table(int M, int W, int F, double c0, double posa, float2* table, hipStream_t s);
void launch_mtrc_rows(const Job& J, hipStream_t s);
void launch_mtrc_czt(const Job& J, hipStream_t s);
void launch_col_window(const Job& J, hipStream_t s);
void launch_mtrc_image(const Job& J, hipStream_t s);
void launch_post_nomtrc(const Job& J, hipStream_t s);

// standalone batched FFT used by the tests: n-point transforms of `batch` contiguous rows
bool debug_fft(int n, int dir, const float2* in, float2* out, int batch, hipStream_t s);

}  // namespace isar
