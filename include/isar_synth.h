/*
 * isar_synth.h — seeded synthetic STRETCH (dechirped) ISAR echo generator.
 *
 * The reference ships no data; this generator produces frames with the layout the reference
 * reads from .wbd files (common.cu:1091-1100): M pulses x N range samples of complex float,
 * row-major, plus the per-pulse radial velocity the reference passes to high-speed
 * compensation (common.cu:1106).
 *
 * Signal model (down-chirp, dechirp-on-receive, plane-wave):
 *   s[m,n] = exp(j phi2_m n^2) * sum_i A_i exp(j (phi0_im + phi1_im n)) * exp(j e_m) + noise
 *   phi0_im = -(4 pi / c) (fc + B/2) dR_im
 *   phi1_im = ( -(4 pi / c) (fc + B/2) v_m + (4 pi K / c) dR_im ) / Fs
 *   phi2_m  =  (4 pi K / c) v_m / Fs^2                (removed exactly by HPC)
 *   dR_im   = walk_m + (y_i cos th_m - x_i sin th_m) cos(el) + z_i sin(el)
 *   th_m = omega (t_m - T/2),  t_m = m / PRF,  v_m = v + a t_m,  K = B / Tp
 * The residual-video-phase term is omitted. All randomness comes from SplitMix64 streams
 * keyed by (seed, purpose, index), so output does not depend on the thread count.
 */
#ifndef ISAR_SYNTH_H_
#define ISAR_SYNTH_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum isar_phase_error_model {
    ISAR_PHASE_NONE = 0,
    ISAR_PHASE_UNIFORM = 1,    /* i.i.d. uniform in [-pi, pi] per pulse                      */
    ISAR_PHASE_POLY = 2        /* sum_{k=2..6} c_k tau^k, c_k ~ N(0, poly_sigma^2), tau in
                                  [-1, 1], plus white N(0, white_sigma^2)                     */
};

typedef struct isar_scene_config {
    int32_t  echo_num;         /* M                                                  */
    int32_t  range_num;        /* N                                                  */
    int32_t  n_scatterers;
    int32_t  phase_error_model;
    double   box_x, box_y, box_z;   /* scene extent: cross-range, range, height (m) */
    double   fc, band_width, prf, fs, tp;
    double   omega;            /* rotation rate (rad/s)                              */
    double   v, a;             /* radial velocity (m/s) at t=0 and acceleration      */
    double   walk_cells;       /* residual range walk amplitude after coarse tracking
                                  (range cells); negative: untracked, walk = v t + a t^2/2 */
    double   snr_db;           /* per-sample SNR; > 200 disables noise               */
    double   elevation;        /* line-of-sight elevation (rad)                      */
    double   poly_sigma, white_sigma;
    double   peak_adc;         /* scale so max(|Re|,|Im|) equals this; 0 = no scaling */
    int32_t  quantize;         /* 1: round to integers (int16 ADC samples)           */
    int32_t  reserved;
    uint64_t seed;
} isar_scene_config;

/* Fill `cfg` with a preset: 1..5 are the five benchmark configurations (sizes, scene, motion,
 * phase error and SNR), 4 uses the range-aligned 256 x 2048 learned-autofocus frames, 6 is an
 * IFDS (raw chirp) configuration for isar_synth_ifds.
 * Returns 0, or -1 for an unknown id. */
int isar_synth_preset(int config_id, isar_scene_config* cfg);

/* Generate one frame. echo: 2*M*N floats (interleaved), velocity: M doubles.
 * truth_phase (M) and truth_walk (M, metres) may be NULL. Returns 0 or -1. */
int isar_synth_frame(const isar_scene_config* cfg, float* echo, double* velocity,
                     double* truth_phase, double* truth_walk);

/* Raw IFDS echoes (intermediate-frequency direct sampling): M pulses of `len` complex samples
 * (iq: 2*M*len floats, int16-valued when quantised) at cfg->fs, each the sum over scatterers of
 * the linear chirp of cfg->tp seconds and cfg->band_width sweep that the reference's pulse
 * compression correlates against (common.cu:635-746, with the pulse's velocity), delayed to the
 * scatterer's range (nearest sample, c / (2 fs) metres per sample) and rotated by its carrier
 * phase -4 pi fc dR / c plus the per-pulse phase error. velocity: M doubles. Preset 6 is an
 * IFDS configuration (range_num = len = 6000). Returns 0 or -1. */
int isar_synth_ifds(const isar_scene_config* cfg, int len, float* iq, double* velocity);

/* Radar parameters that go with a config, in the reference's integer header units
 * (MHz and microseconds, common.cu:809-826). */
void isar_synth_radar_header(const isar_scene_config* cfg, int64_t* band_width_hz,
                             int64_t* fc_hz, double* tp_s, int64_t* fs_hz);

/* The frames isar_synth_frame makes for the seeds seeds[0 .. frames-1] (each replacing cfg's
 * seed), generated on the GPU of the calling thread: d_echo frames x M x N complex float and
 * d_velocity frames x M doubles, device buffers. Same scene, phase-error, walk and noise draws
 * and FP64 sums as the CPU generator, so the samples agree with it except where rounding
 * differences cross a float or integer rounding boundary. Synchronous; allocates and frees a few
 * small device buffers. Returns 0 or a negative isar_status code (libisar_imaging only). */
int isar_synth_frames_device(const isar_scene_config* cfg, int frames, const uint64_t* seeds, void* d_echo,
                             double* d_velocity);

/* Write frames to a STRETCH .wbd file in the layout ioOperation reads (common.cu:796-899,
 * 1082-1100): frames of 256-byte header + int16 I/Q pairs. Samples are rounded and
 * saturated to int16. range/velocity/azimuth/pitch are per frame (may be NULL -> 0).
 * Returns 0 or -1. */
int isar_write_wbd(const char* path, const float* echo, int frames, int range_num,
                   int64_t fc_hz, int64_t band_width_hz, double tp_s, int64_t fs_hz,
                   const double* range_m, const double* velocity_mps,
                   const double* azimuth_deg, const double* pitch_deg);

#ifdef __cplusplus
}
#endif

#endif /* ISAR_SYNTH_H_ */
