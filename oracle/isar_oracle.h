/*
 * isar_oracle.h — FP64 CPU restatement of the reference ISAR range-Doppler imaging chain.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the oracle that the HIP pipeline is compared
 * against, and the CPU baseline that bench.py times. Product code never links, loads or calls it.
 *
 * Every stage follows the reference source literally (same stage order, same data flow, same
 * integer decisions); the only deliberate differences are resolutions of undefined behaviour in
 * the reference (uninitialised buffers, out-of-bounds reads, NaN-producing divisions) to their
 * intended semantics, each marked with its quirk label (Q2 ... Q13) where it is resolved.
 *
 * Reference files (paths relative to the reference repository):
 *   isar_imaging/isar_main.cu:3-227            pipeline order            -> orc_isar_main
 *   isar_imaging/high_speed_compensation.cu:4-81                          -> orc_hpc
 *   isar_imaging/range_alignment.cu:4-201     binary-merge alignment     -> orc_range_alignment
 *   isar_imaging/range_alignment.cu:371-550   HRRP centring              -> orc_hrrp_center
 *   isar_imaging/common.cu:494-518            cutRangeProfile            -> orc_cut_range_profile
 *   isar_imaging/phase_adjustment.cu:4-41     dopplerTracking            -> orc_doppler_tracking
 *   isar_imaging/phase_adjustment.cu:167-346  fastEntropy                -> orc_fast_entropy
 *   isar_imaging/mtrc.cu:4-192                MTRC (chirp-z)             -> orc_mtrc, orc_czt
 *   isar_imaging/common.cu:583-746            IFDS pulse compression     -> orc_pulse_compress
 */
#ifndef ISAR_ORACLE_H_
#define ISAR_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Mirrors RadarParameters (common.cuh:96-117). */
typedef struct orc_params {
    int32_t echo_num;      /* M: pulses (slow time, rows)            */
    int32_t range_num;     /* N: range cells (fast time, columns)    */
    int32_t range_num_cut; /* W: cut width, 512 in the reference     */
    int32_t reserved;
    int64_t band_width;    /* B, Hz                                  */
    int64_t fc;            /* carrier, Hz                            */
    double  Tp;            /* pulse width, s                         */
    int64_t Fs;            /* sampling rate, Hz                      */
} orc_params;

/* Runtime switches of ISAR_RD_Imaging_Main_Ku (isar_main.cu:3-4) plus the two documented
 * deviations that can be switched back. */
typedef struct orc_options {
    int32_t data_type;       /* 1 = IFDS, 2 = STRETCH (common.cuh:84-89)                  */
    int32_t if_hpc;
    int32_t if_hrrp;
    int32_t if_mtrc;
    int32_t quirk_q1;        /* 1: reproduce isar_main.cu:192 (only the first W elements
                                scaled); 0: apply the slow-time Hamming window to all rows  */
    int32_t iterations;      /* fast-entropy iterations, 120 in the reference (common.cuh:73) */
    int32_t phi0_given;      /* 1: use trace->phi0 as the autofocus start (warm start seam) */
    int32_t range_fft32;     /* 1: the range-alignment row transforms in FP32 arithmetic (radix 2,
                                correctly rounded twiddles), everything else FP64: a sensitivity
                                probe of how far FP32 range profiles alone move the result      */
} orc_options;

/* Optional outputs. Any pointer may be NULL. Array sizes in brackets. */
typedef struct orc_trace {
    /* range alignment: one entry per merged pair, level-major (level 0 first), pair-minor */
    int32_t  n_pairs;          /* out: number of pairs = M - 1                              */
    int32_t  reserved0;
    int32_t* ra_imax;          /* [M-1] integer argmax of |c| after ifftshift               */
    double*  ra_lag;           /* [M-1] imax + parabolic offset - N/2                        */
    double*  ra_gap;           /* [M-1] (c[imax] - max_{i != imax} c[i]) / c[imax]           */
    /* HRRP centring */
    int32_t  center_applied;   /* 1 when the shift branch ran (idx2_len != idx1_len)         */
    int32_t  center_shift;     /* circular shift applied to every row                        */
    int32_t  idx1_len;
    int32_t  idx2_len;
    int32_t  diff_imin;        /* argmin of the sorted-ARP difference                        */
    int32_t  cut_offset;
    double   center_thr;       /* low_threshold_gray                                          */
    double*  arp;              /* [N] normalised average range profile                       */
    double*  arp_mean;         /* [N] windowed means for idx1 entries (first idx1_len valid) */
    double*  cut_key;          /* [N] |Re|+|Im| of row 0 after centring                      */
    /* fast entropy */
    int32_t  tgt_num;
    int32_t  unit2_num;
    double   pa_threshold;
    double*  pa_colmax;        /* [W] column maxima of |X|                                   */
    double*  pa_colE;          /* [W] column energies of the Doppler-tracked image           */
    double*  pa_colH;          /* [W] column entropies                                       */
    int32_t* select_bin;       /* [W] selected range bins, first unit2_num valid              */
    double*  phi0;             /* [2M] autofocus start phase (interleaved re, im)            */
    double*  phi;              /* [2M] autofocus result                                       */
    /* result */
    double   entropy;          /* image entropy -sum p ln p, p = img^2 / sum img^2           */
    double   reserved1;
} orc_trace;

/* Full imaging chain for one frame (isar_main.cu:3-227).
 * echo: M*N interleaved complex float (re, im), row-major, rows = pulses.
 * velocity: M doubles (per-pulse radial velocity, used by HPC for STRETCH data).
 * img: M*W floats, output.  Returns 0 on success, nonzero on invalid sizes. */
int orc_isar_main(const float* echo, const double* velocity, const orc_params* p,
                  const orc_options* opt, float* img, orc_trace* trace);

/* orc_isar_main with the wall time (ms) of the reference's SEPARATE_TIMING_ stages
 * (isar_main.cu:34-218) in stage_ms[6] (may be NULL): 0 HPC (with the IFDS pre-stage), 1 range
 * alignment, 2 HRRP centring and cut, 3 phase compensation (Doppler tracking, bin selection,
 * fast entropy), 4 MTRC, 5 post-processing (window, azimuth FFT, |.|). */
int orc_isar_stage_times(const float* echo, const double* velocity, const orc_params* p,
                         const orc_options* opt, float* img, orc_trace* trace, double* stage_ms);

/* K frames in parallel (one OpenMP thread per frame). echo: K*M*N complex, velocity: K*M,
 * img: K*M*W. Traces are not collected. Returns the number of threads used. */
int orc_isar_batch(int K, const float* echo, const double* velocity, const orc_params* p,
                   const orc_options* opt, float* img);

/* ---- stage-level entry points (complex arrays are interleaved doubles) ---- */

/* Unnormalised complex DFT of length n; dir = -1 forward (exp(-j...)), +1 inverse. */
void orc_fft(double* x, int n, int dir);

/* Chirp-z transform X[k] = sum_m x[m] a^-m w^(mk), k < K (mtrc.cu:125-172 semantics). */
void orc_czt(const double* x, int m, int k, double w_re, double w_im, double a_re, double a_im,
             double* out);

/* Chirp-z transform with w = exp(-j 2 pi alpha / m), a = exp(-j pi alpha) (mtrc.cu:175-192). */
void orc_czt_alpha(const double* x, int m, double alpha, double* out);

/* s[m,n] *= exp(j coef v_m (n/Fs)^2), coef = -4 pi (B/Tp) / c  (high_speed_compensation.cu:4-54) */
void orc_hpc(double* s, const double* velocity, const orc_params* p);

/* Range alignment on frequency-domain rows; output time-domain rows (range_alignment.cu:4-78). */
void orc_range_alignment(double* s, const orc_params* p, orc_trace* trace);

/* HRRP centring on time-domain rows, inter_length = 30 (range_alignment.cu:371-465). */
void orc_hrrp_center(double* s, const orc_params* p, orc_trace* trace);

/* Cut W columns around the |Re|+|Im| maximum of row 0 (common.cu:494-518). */
void orc_cut_range_profile(const double* s, const orc_params* p, double* cut, orc_trace* trace);

/* Doppler-centroid tracking (phase_adjustment.cu:4-41). data: rows x cols. phase: rows. */
void orc_doppler_tracking(double* data_comp, double* phase, const double* data, int rows, int cols,
                          int if_compensation);

/* Fast minimum-entropy autofocus on the M x W cut, in place (phase_adjustment.cu:167-332). */
void orc_fast_entropy(double* x, int rows, int cols, int iterations, int phi0_given,
                      orc_trace* trace);

/* MTRC on the M x W cut, in place (mtrc.cu:4-76). */
void orc_mtrc(double* x, const orc_params* p, int data_type);

/* Post-processing on the M x W image data, in place (isar_main.cu:191-211): the slow-time
 * Hamming window (quirk_q1 as in orc_options), the slow-time FFT, the half swap along pulses,
 * flipud for IFDS, and |.| into img (M*W floats). */
void orc_post_process(double* x, int M, int W, int quirk_q1, int data_type, float* img);

/* Synthetic-frame-free helpers used by the tests. */
int orc_next_pow2(int n);                            /* common.cu:521-529 */
double orc_image_entropy(const float* img, int n);   /* -sum p ln p, p = img^2/sum img^2 */

/* IFDS pulse compression of one echo (common.cu:583-604, 635-746): iq = `len` complex samples
 * (interleaved), NFFT = orc_next_pow2(len); the reference chirp of S = (int)(Fs Tp) samples,
 * t_k = (1 - 2v/c)(-Tp/2 + k Tp/(S-1)), ref_k = hamming_S[k] exp(j (c1 + c2 t_k) t_k) with
 * c1 = -2 pi fc v2 / (1 - v2), c2 = pi B / Tp, zero-padded to NFFT; out = IFFT(conj(FFT(ref))
 * FFT(iq)) / NFFT cut to `cut` samples from offset = clamp(i* - cut/2, 0, NFFT - cut), i* the
 * first index of max |Re|+|Im| (cublasIcamax). Returns the offset, or -1 for bad sizes. */
int orc_pulse_compress(const float* iq, int len, double velocity, const orc_params* p, int cut, double* out);

/* High-resolution range profiles of one frame (getHRRP, common.cu:541-580, called at
 * isar_main.cu:57-67): the frame's rows after the IFDS pre-stage or the STRETCH HPC; then for
 * IFDS ifftshift(FFT(ifftshift(row))), for STRETCH ifftshift(FFT(hamming_N * row)). hrrp: M*N
 * interleaved complex doubles. Returns 0, or 1 for bad sizes or data type. */
int orc_hrrp(const float* echo, const double* velocity, const orc_params* p, const orc_options* opt,
             double* hrrp);

#ifdef __cplusplus
}
#endif

#endif /* ISAR_ORACLE_H_ */
