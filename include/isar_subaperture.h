/*
 * isar_subaperture.h — many sliding sub-apertures of one recording in one graph launch.
 *
 * No counterpart in the reference, whose caller images one window per isarMainSingle call
 * (calling_dll.cpp:48-86, window_head advanced by imaging_stride). A plan images n windows of
 * uneven pulse counts of one dwell at once: the windows are grouped by pulse count, each group
 * runs as one batch straight out of the dwell (isar_run_frames, no copies), and the launches of
 * all groups are captured once into a HIP graph that later runs replay. An autotuner picks, per
 * pulse count, the range-alignment chunk size (isar_set_ra_chunk) with the shortest measured
 * run. Functions return 0 or a negative isar_status code (isar_last_error() has the reason).
 */
#ifndef ISAR_SUBAPERTURE_H_
#define ISAR_SUBAPERTURE_H_

#include <stddef.h>
#include <stdint.h>

#include "isar_imaging_ext.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct isar_sa_plan isar_sa_plan;

/* p: the radar parameters (range_num, range_num_cut and the radar constants; echo_num is
 * ignored). Window i covers pulses first_pulse[i] .. first_pulse[i] + pulses[i] - 1 of a dwell of
 * dwell_pulses pulses; every pulse count must be one the pipeline supports. */
ISAR_API int isar_sa_plan_create(const isar_radar_params* p, int dwell_pulses, int n_windows,
                                 const int32_t* first_pulse, const int32_t* pulses, isar_sa_plan** out);
ISAR_API int isar_sa_plan_destroy(isar_sa_plan* plan);
/* Floats of the output: window i's pulses[i] x range_num_cut image starts at offsets[i]. */
ISAR_API size_t isar_sa_plan_image_floats(const isar_sa_plan* plan);
ISAR_API int isar_sa_plan_image_offsets(const isar_sa_plan* plan, int64_t* offsets);
/* Image every window: d_dwell is dwell_pulses x range_num complex float, d_velocity one double
 * per pulse. With use_graph the launches are captured into a graph on the first call for these
 * pointers and options and replayed afterwards. Asynchronous on `stream`. */
ISAR_API int isar_sa_plan_run(isar_sa_plan* plan, const void* d_dwell, const double* d_velocity, float* d_img,
                              const isar_options* opt, int use_graph, void* stream);
/* Tunes every pulse-count group like isar_autotune (range-alignment chunk rows, autofocus form,
 * chirp-z length and column tile; `repeats` runs each after one warm-up), then times whole graph
 * replays for a few combinations of the groups' autofocus forms and keeps the fastest; chosen[g]
 * receives the chunk rows of group g (groups in increasing pulse count; pass NULL to skip; the
 * rest is read with isar_sa_plan_group_tuning). Synchronous. Invalidates a captured graph. */
ISAR_API int isar_sa_plan_autotune(isar_sa_plan* plan, const void* d_dwell, const double* d_velocity,
                                   float* d_img, const isar_options* opt, int repeats, int32_t* chosen);
ISAR_API int isar_sa_plan_groups(const isar_sa_plan* plan, int32_t* pulses, int32_t* windows);
/* Tuning of group g (groups in increasing pulse count): tuning[0] range-alignment chunk rows,
 * [1] autofocus form (isar_autofocus_mode), [2] chirp-z length, [3] chirp-z workgroup threads. */
ISAR_API int isar_sa_plan_group_tuning(const isar_sa_plan* plan, int group, int32_t* tuning);
/* Sets the tuning of group g in the same layout (for example one found by isar_sa_plan_autotune
 * earlier): chunk rows 2, 4, 8 or 16; autofocus form 1 to 4 (isar_autofocus_mode); a chirp-z
 * length that isar_czt_lengths lists for the group's pulse count; 256 or 512 threads. Waits for
 * the device when the chirp-z length changes. Invalidates a captured graph. */
ISAR_API int isar_sa_plan_set_group_tuning(isar_sa_plan* plan, int group, const int32_t* tuning);
/* Times each group alone with its current tuning (one warm-up and `repeats` runs, eager, on an
 * internal stream) and places the groups on the graph's branches by those times, longest first
 * onto the least loaded branch. Synchronous. Invalidates a captured graph. */
ISAR_API int isar_sa_plan_calibrate(isar_sa_plan* plan, const void* d_dwell, const double* d_velocity, float* d_img,
                                    const isar_options* opt, int repeats);
/* Graph branch that group g is placed on, or -1 for a bad argument. */
ISAR_API int isar_sa_plan_group_lane(const isar_sa_plan* plan, int group);

#ifdef __cplusplus
}
#endif

#endif /* ISAR_SUBAPERTURE_H_ */
