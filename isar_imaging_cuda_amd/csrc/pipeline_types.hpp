// pipeline_types.hpp — structures shared by the host orchestrator and the kernels.
#pragma once

#ifdef ISAR_EMU
#include "emu/hip_emu.hpp"
#else
#include <hip/hip_runtime.h>
#define ISAR_LAUNCH(kernel, grid, block, smem, stream, ...) \
    hipLaunchKernelGGL(kernel, dim3(grid), dim3(block), (smem), (stream), __VA_ARGS__)
// a kernel whose workgroups wait for each other (an ordinary launch on the device)
#define ISAR_LAUNCH_RESIDENT(kernel, grid, block, smem, stream, ...) \
    hipLaunchKernelGGL(kernel, dim3(grid), dim3(block), (smem), (stream), __VA_ARGS__)
#endif

#include <stdint.h>

namespace isar {

constexpr int kRangeCut = 512;        // RANGE_NUM_CUT (common.cuh:71)
constexpr int kRangeNumIfdsPc = 4096; // RANGE_NUM_IFDS_PC (common.cuh:72)
constexpr int kRaChunkRowsDefault = 8;   // rows per workgroup in the alignment update kernel
constexpr int kCutBandRows = 16;      // rows per workgroup when cutting the range window
constexpr int kAfSplitFrames = 32;    // batches of at most this many frames split the autofocus
// batches of at least this many echo cells (frames x pulses x range cells; 32 frames of 256 x 4096)
// update the final rows of range alignment in a launch of their own
constexpr long long kRaSplitCells = 1LL << 25;

// Per-frame scalar decisions written by the device (one record per frame).
struct FrameDecisions {
    int32_t center_shift;     // circular shift applied by HRRP centring
    int32_t center_applied;   // 1 if the centring branch ran
    int32_t idx1_len;
    int32_t idx2_len;
    int32_t diff_imin;
    int32_t cut_offset;
    int32_t tgt_num;
    int32_t unit2_num;
    float center_thr;
    float pa_threshold;
    float gmax;
    int32_t status;           // 0 ok; bit flags: 1 centring input all NaN, 2 resident autofocus wait bound hit
};

// Everything a kernel needs to know about one imaging job (identical for all frames of a
// batch). Device pointers are per batch; frame f uses the f-th slice of every buffer.
struct Job {
    int M, N, W, F;               // pulses, range cells, cut width, CZT length
    int K;                        // frames in this batch
    int data_type;                // 1 IFDS, 2 STRETCH
    int if_hpc, if_hrrp, if_mtrc, quirk_q1, iterations;
    int phi0_given;
    int ra_chunk;                 // rows per workgroup in the alignment update (2, 4, 8 or 16)
    int ra_split;                 // alignment update grids per level: 0 by batch size, 1 one, 2 two
    int czt_threads;              // chirp-z workgroup width (256 or 512 threads)
    int af_wait_polls;            // bound of each resident-autofocus wait in polls (< 0: give up at once)
    double hpc_k;                 // HPC phase in half-turns = hpc_k * v * n^2
    double czt_c0, czt_posa;      // alpha_n = 1 - c0 + posa * n
    // inputs
    const float2* echo;           // frame f: rows frame_row[f] .. + M - 1 (f M when frame_row is null), N each
    const double* velocity;       // indexed like the echo rows
    const long long* frame_row;   // K first echo rows, or null
    const long long* frame_img;   // K image offsets in floats, or null (f M W)
    const float2* phi0_in;        // K x M (optional warm start)
    // workspaces
    float2* pre;                  // K x M x N: IFDS pre-stage output, then the row spectra S0
    double* lagsum;               // K x M
    double* lags;                 // K x (M-1)
    int32_t* imax;                // K x (M-1)
    float* egrp;                  // K x M x N
    float* epart;                 // K x M x N
    float2* st;                   // K x M x N (aligned range profiles; aliases pre)
    float* arp_part;              // K x M/2 x N: column sums of |profile| per final-row chunk
    float* max_part;              // K x M/2: max |profile| per final-row chunk
    FrameDecisions* dec;          // K
    float2* x;                    // K x M x W (cut)
    float2* rowdop;               // K x M
    float2* dopphi;               // K x M: Doppler-centroid phase of the cut (k_pa_doppler)
    float* colmax_part;           // K x ceil(M/kCutBandRows) x W
    float* colmax;                // K x W
    float* colE;                  // K x W
    float* colH;                  // K x W
    int32_t* sel;                 // K x W
    float2* su;                   // K x W x M (selected columns, column-major)
    float2* phi0;                 // K x M
    float2* phi;                  // K x M
    float2* stm;                  // K x M x W (MTRC range-frequency rows)
    float2* czt;                  // K x M x W (chirp-z output)
    float2* colsum;               // K x W
    float2* af_part;              // K x 4 (W/8) M: per-pulse sums per sub-chunk of selected bins, by iteration parity (split autofocus: the first half; resident: all, as tagged words)
    unsigned* af_sync;            // 2K: per frame the run count (the resident autofocus's exchange epoch, advanced by k_autofocus) and its hand-off counter (cleared there)
    const float2* czt_table;      // W x F (FFT of the Bluestein chirp per range bin)
    const float2* czt_pre;        // W x M: exp(j pi a_n m (M - m) / M), the input chirp
    const float2* czt_post;       // W x M: exp(-j pi a_n k^2 / M) / F, the output chirp
    float2* delta;                // K x W: transformed column sums for the slow-time window quirk
    const float* win_range;       // N: Hamming window over range cells (genHammingVec, common.cu:260-266)
    // outputs
    float* img;                   // K x M x W
    float2* hrrp;                 // K x M x N high-resolution range profiles (getHRRP), or null
};

// One batch of IFDS echoes to pulse-compress (common.cu:583-604, 635-746).
struct PcJob {
    const int16_t* iq;            // echoes x len I/Q pairs
    const double* velocity;       // echoes
    float2* out;                  // echoes x cut
    int32_t* offsets;             // echoes (cut offsets), or null
    int echoes, len, S, cut;      // S: reference chirp samples (int)(Fs Tp)
    double fc, band_width, tp;
};

// First echo row of frame f and the start of its image.
__host__ __device__ inline long long echo_first_row(const Job& J, int f) {
    return J.frame_row ? J.frame_row[f] : static_cast<long long>(f) * J.M;
}
__host__ __device__ inline float* image_of(const Job& J, int f) {
    return J.img + (J.frame_img ? J.frame_img[f] : static_cast<long long>(f) * J.M * J.W);
}

}  // namespace isar
