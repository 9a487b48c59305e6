/*
 * isar_imaging_ext.h — plain C ABI of the MI355X imaging pipeline (plain pointers and sizes).
 *
 * This header sits beside the drop-in isar_imaging_export.h. It exposes the same imaging chain
 * (isar_main.cu:3-227 of the reference) as batched, device-resident entry points for callers
 * that bring their own frames (synthetic or ingested), plus the per-frame decisions for
 * verification. All functions return 0 on success and a negative code on failure; the reason
 * is available from isar_last_error(). Nothing here terminates the process.
 */
#ifndef ISAR_IMAGING_EXT_H_
#define ISAR_IMAGING_EXT_H_

#include <stddef.h>
#include <stdint.h>

#if defined(_WIN32)
#define ISAR_API __declspec(dllexport)
#else
#define ISAR_API __attribute__((visibility("default")))
#endif

#ifdef __cplusplus
extern "C" {
#endif

enum isar_status {
    ISAR_OK = 0,
    ISAR_ERR_INVALID = -1,      /* bad argument or unsupported size                */
    ISAR_ERR_DEVICE = -2,       /* HIP runtime error                               */
    ISAR_ERR_MEMORY = -3,       /* device allocation failed                        */
    ISAR_ERR_STATE = -4,        /* call out of order (e.g. run before init)        */
    ISAR_ERR_IO = -5            /* file not found / unreadable                     */
};

/* RadarParameters (common.cuh:96-117). range_num_cut must be 512. */
typedef struct isar_radar_params {
    int32_t echo_num;       /* M pulses; 64, 128, 192, 256, 384, 512 or 1024            */
    int32_t range_num;      /* N range cells; 1024, 2048, 4096, 8192 or 16384           */
    int32_t range_num_cut;  /* W = 512                                                   */
    int32_t reserved;
    int64_t band_width;     /* Hz */
    int64_t fc;             /* Hz */
    double  Tp;             /* s  */
    int64_t Fs;             /* Hz */
} isar_radar_params;

/* Switches of ISAR_RD_Imaging_Main_Ku (isar_main.cu:3-4) plus two additions. */
typedef struct isar_options {
    int32_t data_type;        /* 1 IFDS, 2 STRETCH                                            */
    int32_t option_alignment; /* accepted and ignored, as in the reference                    */
    int32_t option_phase;     /* accepted and ignored, as in the reference                    */
    int32_t if_hrrp;
    int32_t if_hpc;
    int32_t if_mtrc;
    int32_t quirk_q1;         /* 1 (default): slow-time window as isar_main.cu:192 applies it  */
    int32_t iterations;       /* fast-entropy iterations (reference: 120)                      */
    int32_t warm_start;       /* 1: phi0 supplied by the caller seeds the autofocus            */
    int32_t stage_mode;       /* ISAR_RUN_FULL, ISAR_RUN_FRONT or ISAR_RUN_BACK                */
} isar_options;

/* Parts of the chain a run executes. FRONT stops after the autofocus (no image); BACK repeats
 * the autofocus from the cut and selection kept by the previous FRONT or FULL run of the same
 * context and frame count, then forms the images (echo and velocity may be NULL; with a NULL
 * image pointer only the autofocus runs). */
enum isar_stage_mode { ISAR_RUN_FULL = 0, ISAR_RUN_FRONT = 1, ISAR_RUN_BACK = 2 };

/* Per-frame decisions and vectors copied back by isar_get_trace. Pointers are caller-owned
 * and may be NULL. */
typedef struct isar_frame_trace {
    int32_t n_pairs;          /* M - 1 */
    int32_t center_applied;
    int32_t center_shift;
    int32_t idx1_len;
    int32_t idx2_len;
    int32_t diff_imin;
    int32_t cut_offset;
    int32_t tgt_num;
    int32_t unit2_num;
    int32_t status;
    float   center_thr;
    float   pa_threshold;
    double* ra_lag;           /* [M-1] level-major, pair-minor */
    int32_t* ra_imax;         /* [M-1] */
    float*  pa_colmax;        /* [W] */
    float*  pa_colE;          /* [W] */
    float*  pa_colH;          /* [W] */
    int32_t* select_bin;      /* [W], first unit2_num valid */
    float*  phi0;             /* [2M] interleaved */
    float*  phi;              /* [2M] interleaved */
} isar_frame_trace;

typedef struct isar_context isar_context;

ISAR_API void isar_default_options(isar_options* opt);
ISAR_API int isar_create(const isar_radar_params* p, int max_frames, isar_context** ctx);
ISAR_API int isar_destroy(isar_context* ctx);
ISAR_API size_t isar_workspace_bytes(const isar_radar_params* p, int max_frames);

/* Device-resident batch: echo K*M*N complex float, velocity K*M double, phi0 K*M complex
 * float (only with warm_start), img K*M*W float. Asynchronous on `stream` (hipStream_t,
 * NULL = default stream). No device memory is allocated. */
ISAR_API int isar_run_device(isar_context* ctx, int frames, const void* d_echo, const double* d_velocity,
                             const void* d_phi0, float* d_img, const isar_options* opt, void* stream);

/* isar_run_device without a context and without any device allocation. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under stream capture,
 * tight serving loops, hosts that forbid hipMalloc). The workspace is a device array the library
 * holds per GPU until the process ends, isar_static_workspace_bytes() bytes (64 MiB: two frames
 * of 256 x 4096, one of 512 x 4096); a batch needing more (isar_workspace_bytes) is refused
 * with ISAR_ERR_INVALID. The chirp-z and window tables are formed on `stream` at every call.
 * Full runs only (stage_mode 0). Every call on one GPU shares that GPU's array: calls whose
 * work overlaps on one GPU race. The caller passes device inputs and outputs as for
 * isar_run_device. */
ISAR_API int isar_run_device_static(const isar_radar_params* p, int frames, const void* d_echo,
                                    const double* d_velocity, const void* d_phi0, float* d_img,
                                    const isar_options* opt, void* stream);
ISAR_API size_t isar_static_workspace_bytes(void);

/* Device-resident batch whose frames are windows of a longer recording: frame f uses echo rows
 * (and velocities) d_frame_rows[f] .. d_frame_rows[f] + M - 1 of d_echo (N complex floats per
 * row) and writes its M x W image at d_img + d_frame_img[f] floats. Both tables are device
 * arrays of `frames` int64 entries; either may be NULL for contiguous frames. */
ISAR_API int isar_run_frames(isar_context* ctx, int frames, const void* d_echo, const double* d_velocity,
                             const int64_t* d_frame_rows, const void* d_phi0, float* d_img,
                             const int64_t* d_frame_img, const isar_options* opt, void* stream);

/* IFDS pulse compression (common.cu:583-604, 635-746 of the reference): `echoes` raw echoes of
 * `len` int16 I/Q pairs each (d_iq, as read from IFDS .wbd frames after the 256-byte header),
 * one velocity per echo (m/s). Each echo is matched-filtered with the velocity-dependent
 * reference chirp (Fs Tp samples, Hamming-weighted) through NFFT = nextPow2(len)-point
 * transforms (nextPow2 as in common.cu:521-529: 2 len for powers of two; NFFT 4096, 8192 or
 * 16384), and the 4096 samples around the first maximum of |Re| + |Im| are written to
 * d_out + 4096 e (complex float). d_offsets (optional) receives each cut offset.
 * Asynchronous on `stream`. */
ISAR_API int isar_pulse_compress(const isar_radar_params* p, int echoes, int len, const int16_t* d_iq,
                                 const double* d_velocity, void* d_out, int32_t* d_offsets, void* stream);

/* High-resolution range profiles (getHRRP, common.cu:541-580, called at isar_main.cu:57-67; the
 * reference keeps them in the DLL's device buffer d_hrrp, isar_imaging_export.cpp:18, 176-178).
 * After this call, runs with if_hrrp = 1 also write each frame's M x N profiles as complex
 * float, frame f at d_hrrp + f M N:
 *   STRETCH: ifftshift(FFT(hamming_N * row)) of the rows after high-speed compensation;
 *   IFDS: the pulse-compressed rows themselves, which getHRRP's shift-transform-shift returns
 *         unchanged.
 * The buffer holds max_frames * M * N complex floats. NULL stops the output. */
ISAR_API int isar_set_hrrp_output(isar_context* ctx, void* d_hrrp);

/* Drop-in session: after isarMainSingle with if_hrrp (and imagingMemInit with if_hrrp), copies
 * the window's M x N profiles (complex float) to h_hrrp. Synchronous. */
ISAR_API int isar_session_hrrp(float* h_hrrp);

/* Drop-in session: its context (NULL before imagingMemInit), for the setters below. The session
 * images each frame with the resident autofocus; a frame on which a resident wait hit its bound
 * (status bit 2) is imaged again with the split form, with a note on stderr. */
ISAR_API isar_context* isar_session_context(void);

/* Intermediate arrays of one frame of the last run, as the reference writes them with its
 * DATA_WRITE_BACK_* compile switches (isar_main.cu:10-12, 51-53, 76-78, 101-103, 152-154,
 * 178-180, 221-223). The fused kernels never hold most of them whole; they are rebuilt on the
 * device from what the run kept. The last run must have been a full run, and the caller's echo,
 * velocity and image buffers must still hold its data. */
enum isar_dump_stage {
    ISAR_DUMP_DATAW = 0,   /* dataW.dat: the echo rows, M x N complex                          */
    ISAR_DUMP_HPC = 1,     /* hpc.dat: after the IFDS pre-stage or the STRETCH HPC, M x N       */
    ISAR_DUMP_HRRP = 2,    /* hrrp.dat: getHRRP's profiles, M x N complex                       */
    ISAR_DUMP_RA = 3,      /* ra.dat: aligned and centred range profiles, M x N complex         */
    ISAR_DUMP_PC = 4,      /* pc.dat: the cut after Doppler tracking and autofocus, M x W       */
    ISAR_DUMP_MTRC = 5,    /* mtrc.dat: after MTRC (the PC array when MTRC is off), M x W       */
    ISAR_DUMP_FINAL = 6,   /* final.dat: the image, M x W float                                 */
    ISAR_DUMP_COUNT = 7
};
/* Bytes of one frame's array of `stage`; 0 for a bad stage. */
ISAR_API size_t isar_stage_bytes(const isar_context* ctx, int stage);
/* Copies frame `frame`'s array of `stage` to host memory (isar_stage_bytes bytes; complex
 * arrays as interleaved floats). Synchronous: waits for the device first. */
ISAR_API int isar_stage_read(isar_context* ctx, int stage, int frame, void* h_out);
/* Writes the arrays of the stages whose bit is set in `mask` (bit s = stage s) as the reference's
 * text files <dir>/dataW.dat ... final.dat: one value per line, 5 decimals, complex values as
 * the real part and then the imaginary part (ioOperation::writeFile, common.cu:1194-1259). */
ISAR_API int isar_stage_dump(isar_context* ctx, int frame, unsigned mask, const char* dir);
/* Drop-in session: isar_stage_dump of the last isarMainSingle. */
ISAR_API int isar_session_dump(unsigned mask, const char* dir);

/* Form of the fast-entropy autofocus. FUSED runs each frame's 120 iterations in one workgroup,
 * which fills the GPU when a batch has hundreds of frames. SPLIT spreads each frame's selected
 * range bins over several workgroups, for small batches and single frames: one launch per
 * iteration, every workgroup forming the current phase from the previous iteration's sums.
 * SPLIT_PHASE does the same with a separate phase launch per iteration, which repeats no work
 * and suits GPUs shared with other streams. RESIDENT spreads the bins like SPLIT but runs all
 * iterations in one launch: each workgroup keeps its bins in registers and the frame's
 * workgroups exchange their sums through memory after every iteration. Its waits are bounded;
 * a frame whose workgroups could not all be resident together gets status bit 2
 * (isar_get_trace) and no valid phase. AUTO (default) uses SPLIT for batches of at most 32
 * frames and FUSED above. All forms give the same phases bit for bit. */
enum isar_autofocus_mode {
    ISAR_AF_AUTO = 0,
    ISAR_AF_FUSED = 1,
    ISAR_AF_SPLIT = 2,
    ISAR_AF_SPLIT_PHASE = 3,
    ISAR_AF_RESIDENT = 4
};
ISAR_API int isar_set_autofocus_mode(isar_context* ctx, int mode);
ISAR_API int isar_get_autofocus_mode(isar_context* ctx);
/* Bound of each wait of the resident autofocus, in polls of about a microsecond (default 2^20);
 * a negative bound makes every wait give up at once, which forces status bit 2 (for tests of the
 * fallback). */
ISAR_API int isar_set_af_wait_polls(isar_context* ctx, int polls);

/* Graph mode (default off): a full run whose arguments (frame count, pointers, options, tuning)
 * equal the previous full run's replays a HIP graph of its launches instead of launching them one
 * by one; a run with new arguments is captured once first. Runs with probes enabled, and front or
 * back runs, launch directly. The drop-in session turns it on: its buffers never change. */
ISAR_API int isar_set_graph_mode(isar_context* ctx, int on);

/* Rows of the range-alignment update handled per workgroup (2, 4, 8 or 16; default 8). The
 * choice changes the summation order of the group energies, not the algorithm. */
ISAR_API int isar_set_ra_chunk(isar_context* ctx, int rows);
ISAR_API int isar_get_ra_chunk(isar_context* ctx);
/* Grids of the range-alignment update per merge level: 0 (default) by batch size, one grid below
 * 2^25 echo cells per batch and two from there (the rows of pair 0, which move for the last time
 * and carry the aligned-profile output, apart from the other pairs); 1 always one grid; 2 always
 * two. All give the same results bit for bit; the setting exists so that both launch forms can
 * be exercised on small frames. */
ISAR_API int isar_set_ra_split(isar_context* ctx, int mode);
ISAR_API int isar_get_ra_split(isar_context* ctx);
/* Chirp-z transform length F of the MTRC's Bluestein convolution (mtrc.cu:125-192). The default
 * (0) is the reference's nextPow2(2M - 1); for 192 and 384 pulses the shorter 3 * 2^k lengths 384
 * and 768 (mixed-radix transforms) are accepted too. Every length gives the same transform up to
 * FP32 rounding. isar_czt_lengths lists the accepted lengths, the default first, and returns how
 * many (at most max). Setting a new length waits for the device, then re-forms the chirp-z
 * tables. */
ISAR_API int isar_set_czt_length(isar_context* ctx, int length);
ISAR_API int isar_get_czt_length(isar_context* ctx);
ISAR_API int isar_czt_lengths(isar_context* ctx, int32_t* out, int max);
/* Chirp-z column tile: 256 (default) or 512 threads per workgroup, i.e. 256 / (F / E) or twice as
 * many range bins per workgroup. Results are identical bit for bit. */
ISAR_API int isar_set_czt_threads(isar_context* ctx, int threads);
ISAR_API int isar_get_czt_threads(isar_context* ctx);
/* Autotuner: times every range-alignment chunk size on these frames (`repeats` runs each after
 * one warm-up, on `stream`), keeps the fastest for the context and returns it in *chosen; then
 * times the autofocus forms with that chunk size and keeps the fastest (isar_get_autofocus_mode);
 * then, for runs with MTRC, times every chirp-z length and column tile on the chirp-z launches
 * alone (through the kernel probe, which it leaves enabled for the stage it found, with its
 * accumulated time cleared) and keeps the fastest (isar_get_czt_length, isar_get_czt_threads).
 * Synchronous; the image buffer receives the last run's images. */
ISAR_API int isar_autotune(isar_context* ctx, int frames, const void* d_echo, const double* d_velocity, float* d_img,
                           const isar_options* opt, int repeats, int32_t* chosen, void* stream);

/* Host batch: copies in, runs, copies the images out; synchronous. */
ISAR_API int isar_run_host(isar_context* ctx, int frames, const float* h_echo, const double* h_velocity,
                           const float* h_phi0, float* h_img, const isar_options* opt);

/* Autofocus phases of the last run (device to device, asynchronous on `stream`): the starting
 * phase (Doppler tracking, or the warm start) and the final phase, K*M complex float each.
 * Either destination may be NULL. */
ISAR_API int isar_get_phases(isar_context* ctx, int frames, void* d_phi_start, void* d_phi_final, void* stream);

/* After a run has completed: decisions of frame `frame` of the last batch. */
ISAR_API int isar_get_trace(isar_context* ctx, int frame, isar_frame_trace* out);

/* Unnormalised n-point transforms of `batch` contiguous rows on the device (dir -1 forward,
 * +1 inverse); n in {64,128,192,256,384,512,1024,2048,4096,8192,16384}. Synchronous. */
ISAR_API int isar_fft_host(int n, int dir, const float* h_in, float* h_out, int batch);

/* The same transforms arranged as the autofocus, chirp-z and MTRC kernels run them: several per
 * 256-thread workgroup, each within one wave and ordered by that wave alone; n in
 * {64,128,192,256,384,512,1024}. The results equal isar_fft_host's bit for bit. No counterpart in
 * the reference (a test entry point). Synchronous. */
ISAR_API int isar_fft_host_grouped(int n, int dir, const float* h_in, float* h_out, int batch);

/* Kernel families that can be probed with HIP events (isar_probe_enable). */
enum isar_stage {
    ISAR_STAGE_IFDS_PRE = 0, ISAR_STAGE_RA_PAIR0 = 1, ISAR_STAGE_RA_CORR = 2, ISAR_STAGE_RA_UPD = 3,
    ISAR_STAGE_CENTER = 4, ISAR_STAGE_CUT = 5, ISAR_STAGE_PA_COLFFT = 6, ISAR_STAGE_PA_SELECT = 7,
    ISAR_STAGE_AUTOFOCUS = 8, ISAR_STAGE_MTRC_ROWS = 9, ISAR_STAGE_MTRC_CZT = 10, ISAR_STAGE_COL_WINDOW = 11,
    ISAR_STAGE_MTRC_IMAGE = 12, ISAR_STAGE_POST_NOMTRC = 13, ISAR_STAGE_HRRP = 14,
    ISAR_STAGE_COUNT = 15
};

/* Record a HIP event pair around every launch of kernel family `stage` (on the run stream) in
 * the following runs; stage < 0 disables. Events are created here, not during runs. */
ISAR_API int isar_probe_enable(isar_context* ctx, int stage);
/* Waits for the recorded events and returns the summed duration (ms) and the launch count since
 * the last enable/read, then resets the accumulator. */
ISAR_API int isar_probe_read(isar_context* ctx, double* total_ms, int* launches);

/* Message of the last failure on the calling thread (each thread keeps its own, like errno);
 * valid until that thread's next failing call. */
ISAR_API const char* isar_last_error(void);
ISAR_API int isar_device_count(void);

#ifdef __cplusplus
}
#endif

#endif /* ISAR_IMAGING_EXT_H_ */
