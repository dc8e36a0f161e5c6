/*
 * gnss_gpu_amd — C ABI of the MI355X (gfx950) particle-filter step and
 * batched WLS library (libgnss_gpu_amd.so).
 *
 * Conventions
 *   - Every function returns 0 on success and a negative code on failure;
 *     gg_last_error() returns the thread-local message of the last failure.
 *   - Host pointers are plain C arrays (row-major, binary64 unless stated).
 *     Functions named *_device take device pointers and a hipStream_t.
 *   - A gg_pf handle owns its device memory and one HIP stream (or uses the
 *     stream given to gg_pf_set_stream).  A handle is not thread-safe.
 *
 * Reference counterparts are cited per function (paths relative to the
 * rsasaki0109/gnss_gpu tree).
 */
#ifndef GNSS_GPU_AMD_H
#define GNSS_GPU_AMD_H

#include <stdint.h>

#include "gnss_gpu_amd_layout.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gg_pf gg_pf;
typedef struct gg_mlp gg_mlp;

/* --- library ----------------------------------------------------------- */
const char* gg_last_error(void);
int gg_abi_version(void);
/* sizeof(gg_step_args) and sizeof(gg_dyn), for binding layout checks */
int gg_struct_sizes(int32_t* step_args, int32_t* dyn, int32_t* filter);
int gg_device_count(int32_t* out);

/* --- particle filter: handle lifecycle ----------------------------------- */
/* Batch of n_filters independent filters sharing one arena.
 * sat_cap[f]: satellite rows reserved for filter f (grown on demand by the
 * single-filter weight call).  n_slots: epoch slots of the satellite table
 * (1 for host-driven use, T to preload a T-epoch trajectory).
 * with_velocity: reserve per-particle velocity and velocity-covariance
 * storage (needed for sigma_vel > 0, non-uniform states, RBPF). */
int gg_pf_create(int32_t n_filters, const int32_t* n_particles, const int32_t* sat_cap,
                 int32_t n_isb, int32_t n_slots, int32_t with_velocity, gg_pf** out);
/* reference: pf_device_create, src/particle_filter/pf_device.cu:1306 */
int gg_pf_device_create(int32_t n_particles, int32_t n_isb, gg_pf** out);
/* reference: pf_device_destroy, pf_device.cu:1445; idempotent on NULL */
int gg_pf_destroy(gg_pf* h);
int gg_pf_set_stream(gg_pf* h, void* hip_stream);
/* reference: pf_device_sync, pf_device.cu:2274 */
int gg_pf_device_sync(gg_pf* h);
int gg_pf_info(gg_pf* h, int32_t* n_filters, int32_t* n_isb, int64_t* arena_slots);

/* --- particle filter: reference-shaped single-filter API (filter 0) ------- */
/* reference: pf_device_initialize, pf_device.cu:1455-1476 (init_isb may be NULL) */
int gg_pf_device_initialize(gg_pf* h, double init_x, double init_y, double init_z,
                            double init_cb, double spread_pos, double spread_cb, uint64_t seed,
                            double init_vx, double init_vy, double init_vz, double spread_vel,
                            double init_vel_sigma, const double* init_isb, double spread_isb);
/* reference: pf_device_predict, pf_device.cu:1482-1510 */
int gg_pf_device_predict(gg_pf* h, double vx, double vy, double vz, double dt, double sigma_pos,
                         double sigma_cb, uint64_t seed, int32_t step, double sigma_vel,
                         double velocity_guide_alpha, int32_t velocity_kf,
                         double velocity_process_noise, double sigma_isb);
/* reference: pf_device_weight, pf_device.cu:1516-1574.  sys (may be NULL):
 * per-satellite system id 0..3 (0 = reference system, k>0 uses ISB state k). */
int gg_pf_device_weight(gg_pf* h, const double* sat_ecef, const double* pseudoranges,
                        const double* weights_sat, const int32_t* sys, int32_t n_sat,
                        double sigma_pr, double nu, double per_particle_nlos_threshold_m,
                        int32_t per_particle_huber, double per_particle_huber_k);
/* reference: pf_device_weight_gmm, src/particle_filter/pf_device.cu:1622-1680 and
 * pfd_weight_gmm_kernel :373-437 (binding _pf_device_bindings.cpp:147-165):
 * overwrites the log-weights with sum_s log(w_los N(e_s; 0, sigma_pr) +
 * (1 - w_los) N(e_s; mu_nlos, sigma_nlos)) (up to the constant -log sqrt(2 pi)),
 * e_s = pr_s - |sat_s - x| - cb, per-satellite weights scaling both exponents. */
int gg_pf_device_weight_gmm(gg_pf* h, const double* sat_ecef, const double* pseudoranges,
                            const double* weights_sat, int32_t n_sat, double sigma_pr,
                            double w_los, double mu_nlos, double sigma_nlos);
/* reference: pf_device_ess, pf_device.cu:1722-1762 */
int gg_pf_device_ess(gg_pf* h, double* out);
/* reference: pf_device_position_spread, pf_device.cu:1764-1807 */
int gg_pf_device_position_spread(gg_pf* h, double cx, double cy, double cz, double* out);
/* reference: pf_device_resample_systematic, pf_device.cu:1813-1878.
 * u0_mode 0 reproduces the reference offset (seed % 1e6) / (1e6 * N) with the
 * reference's int32 product; 1 draws a fresh Philox u0 in [0,1). */
int gg_pf_device_resample_systematic(gg_pf* h, uint64_t seed, int32_t u0_mode);
/* reference: pf_device_resample_megopolis, pf_device.cu:1884-1920 (same
 * proposal/acceptance rule, ancestor form, one gather) */
int gg_pf_device_resample_megopolis(gg_pf* h, int32_t n_iterations, uint64_t seed);
/* Megopolis of Chesser et al. (arXiv:2109.13504); no reference counterpart
 * (the reference's kernel is the slot-weight variant above). */
int gg_pf_device_resample_megopolis_chesser(gg_pf* h, int32_t n_iterations);
/* reference: pf_device_estimate, pf_device.cu:1926-1969; out[4] */
int gg_pf_device_estimate(gg_pf* h, double* out);
/* out[4 + n_isb]: x y z cb isb...  (no reference counterpart: ISB states) */
int gg_pf_device_estimate_full(gg_pf* h, double* out);
/* reference: pf_device_get_particles, pf_device.cu:1975-1993; out[N*4] */
int gg_pf_device_get_particles(gg_pf* h, double* out);
/* out[N*(4+n_isb)] */
int gg_pf_device_get_particles_full(gg_pf* h, double* out);
/* reference: pf_device_get_particle_states / set_particle_states,
 * pf_device.cu:1995-2031; rows [x y z cb vx vy vz S(3x3)] */
int gg_pf_device_get_particle_states(gg_pf* h, double* out);
int gg_pf_device_set_particle_states(gg_pf* h, const double* in);
/* Positions and clock biases as four host arrays px, py, pz, pcb [N] (the
 * layout of the legacy functions of python/gnss_gpu/_particle_filter_bindings.cpp:
 * 10-145, which the reference keeps on the host between calls).  set replaces
 * them (weights reset, ISB offsets kept, velocities back to the uniform one);
 * no reference counterpart in pf_device.cu. */
int gg_pf_device_get_positions(gg_pf* h, double* px, double* py, double* pz, double* pcb);
int gg_pf_device_set_positions(gg_pf* h, const double* px, const double* py, const double* pz,
                               const double* pcb);
/* reference: pf_device_get_log_weights / set_log_weights, pf_device.cu:2033-2051 */
int gg_pf_device_get_log_weights(gg_pf* h, double* out);
int gg_pf_device_set_log_weights(gg_pf* h, const double* in);
/* reference: pf_device_get_resample_ancestors, pf_device.cu:2053-2058 */
int gg_pf_device_get_resample_ancestors(gg_pf* h, int32_t* out);
/* reference: pf_device_shift_clock_bias / shift_position, pf_device.cu:1699-1716 */
int gg_pf_device_shift_clock_bias(gg_pf* h, double shift);
int gg_pf_device_shift_position(gg_pf* h, double dx, double dy, double dz);
/* reference: pf_device_position_update, pf_device.cu:1682-1693 */
int gg_pf_device_position_update(gg_pf* h, double ref_x, double ref_y, double ref_z,
                                 double sigma_pos);

/* --- particle filter: batched, device-driven step -------------------------- */
/* init[f*(4+n_isb) + k]: x y z cb isb...; seeds[f]. */
int gg_pfb_initialize(gg_pf* h, const double* init, const uint64_t* seeds, double spread_pos,
                      double spread_cb, double spread_isb);
/* Host arrays for n_slots epochs:
 *   sat_xyz[s][row][3], pr[s][row], w[s][row], sys[s][row] with the rows of
 *   filter f at [sum_{g<f} sat_cap[g], ...); n_sat[s][f]; vel[s][f][3] (ECEF). */
int gg_pfb_load_epochs(gg_pf* h, int32_t n_slots, const double* sat_xyz, const double* pr,
                       const double* w, const int32_t* sys, const int32_t* n_sat,
                       const double* vel);
/* resampler: 0 systematic, 1 Megopolis (Chesser), 2 Megopolis (reference rule) */
int gg_pfb_step(gg_pf* h, const gg_step_args* args, int32_t resampler, int32_t n_iter);
/* Capture `steps` consecutive steps in one hipGraph (replaced on each call).
 * The capture runs on a private stream; gg_pfb_replay launches the graph on
 * the stream in use (the handle's own or the one given to gg_pf_set_stream). */
int gg_pfb_capture(gg_pf* h, const gg_step_args* args, int32_t resampler, int32_t n_iter,
                   int32_t steps);
int gg_pfb_replay(gg_pf* h, int32_t n_replays);
/* Device time of n_replays graph launches (hipEvents on the handle's stream). */
int gg_pfb_time_replays(gg_pf* h, int32_t n_replays, float* ms_total);
/* Device time of each of n_replays graph launches (ms_each[n_replays]). */
int gg_pfb_time_replays_each(gg_pf* h, int32_t n_replays, float* ms_each);
/* n_steps eager steps, each waited for, every kernel timed at its own start
 * and end (hipExtLaunchKernelGGL events on the handle's stream); a launch
 * group's time runs from its first kernel's start to its last kernel's end.
 * stage_ms[0..2] = mean ms of (predict+weight, then for systematic the one
 * resampling launch (decision, CDF ends, gather, finalize) and 0, for Megopolis
 * the decision and the chains with their gather) over steps in which some
 * filter resampled, stage_ms[3..5] over the other steps; counts[0..1] = the two
 * step counts. */
int gg_pfb_profile_steps(gg_pf* h, const gg_step_args* args, int32_t resampler, int32_t n_iter,
                         int32_t n_steps, float* stage_ms, int32_t* counts);
/* Device time of `reps` launches of one kernel: 0 weight+predict, 1 weight,
 * 2 and 3 systematic resampling (decision, CDF ends, gather and finalize in one
 * launch, forced), 4 Megopolis
 * (Chesser: weights and packed states, chains, state copy and finalize),
 * 5 gather (ancestors, from the packed states), 6 decision (one workgroup per
 * filter). */
int gg_pfb_time_stage(gg_pf* h, const gg_step_args* args, int32_t stage, int32_t n_iter,
                      int32_t reps, float* ms_total);
int gg_pfb_read_dyn(gg_pf* h, gg_dyn* out);
int gg_pfb_set_epoch(gg_pf* h, int32_t epoch);
/* Export filter f's particles as binary64 rows [x y z cb isb...]. */
int gg_pfb_get_particles(gg_pf* h, int32_t f, double* out);
int gg_pfb_get_log_weights(gg_pf* h, int32_t f, double* out);
/* Inspection of the device-driven step (no reference counterpart):
 * gg_pfb_step_part(.., 0) enqueues the step's likelihood launch (predict +
 * weight + tile partials) and (.., 1) the rest (decision, resampling and its
 * finalize); part 0 then part 1 gives exactly gg_pfb_step's results (a whole
 * Megopolis step also keeps its predicted state as packed records between its
 * launches; the parts keep the state arrays, which a caller may read between
 * them).
 * gg_pfb_get_log_weights_raw copies filter f's binary32 log-weight buffer as
 * the last likelihood launch wrote it (no zeroed-weights flag applied);
 * gg_pfb_get_ancestors the ancestor indices of its last resampling. */
int gg_pfb_step_part(gg_pf* h, const gg_step_args* args, int32_t resampler, int32_t n_iter,
                     int32_t part);
int gg_pfb_get_log_weights_raw(gg_pf* h, int32_t f, float* out);
int gg_pfb_get_ancestors(gg_pf* h, int32_t f, int32_t* out);
/* Launch shapes of the device-driven step (no reference counterpart):
 * weight_tpw tiles of 1024 particles per workgroup of the predict + likelihood
 * kernel (1: one tile per workgroup, one round over the batch; > 1: fewer
 * workgroups streaming tiles with the next tile prefetched); mego_depth (1, 2,
 * 4, 8) Megopolis iterations whose proposal loads are issued together.
 * Results do not depend on the shapes; only the speed does.  Graphs captured
 * earlier keep the shapes they were captured with. */
int gg_pfb_set_tuning(gg_pf* h, int32_t weight_tpw, int32_t mego_depth);
int gg_pfb_get_tuning(gg_pf* h, int32_t* weight_tpw, int32_t* mego_depth);
/* Autotuner (no reference counterpart): times graphs of `steps` steps, `reps`
 * replays each, for every candidate shape (weight_tpw in {1, 2, 4} for the
 * analytic likelihood, mego_depth in {1, 2, 4, 8} for resampler 1), every trial
 * starting from a device snapshot of the particles, log-weights, ancestors and
 * filter records, which are restored at the end.  Keeps the fastest shape in
 * the handle; cand[2k], cand[2k+1] = (weight_tpw, mego_depth) of candidate k
 * and ms[k] its device ms per step, for the n_cand <= max_cand candidates. */
int gg_pfb_autotune(gg_pf* h, const gg_step_args* args, int32_t resampler, int32_t n_iter,
                    int32_t steps, int32_t reps, int32_t max_cand, int32_t* cand, float* ms,
                    int32_t* n_cand);

/* --- batched Gauss-Newton WLS ---------------------------------------------- */
/* reference: wls_batch, src/positioning/wls.cu:172-278 (Sagnac range model).
 * n_sat may be NULL (all epochs use n_sat_max rows) or give per-epoch counts;
 * rows beyond an epoch's count are ignored.  Host arrays. */
int gg_wls_batch(const double* sat_ecef, const double* pseudoranges, const double* weights,
                 const int32_t* n_sat, int32_t n_epoch, int32_t n_sat_max, int32_t max_iter,
                 double tol, double* results, int32_t* iters);
/* Device-pointer variant; mode 0 = wls_batch semantics, 1 = wls_position semantics
 * (no Sagnac, singular-matrix early return, -1 for < 4 satellites). */
int gg_wls_batch_device(const double* sat_ecef, const double* pseudoranges,
                        const double* weights, const int32_t* n_sat, int32_t n_epoch,
                        int32_t n_sat_max, int32_t max_iter, double tol, double* results,
                        int32_t* iters, int32_t mode, void* hip_stream);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): gg_wls_batch with the same arguments and results, staged
 * in a static device area of GG_WLS_NOALLOC_BYTES per GPU (held until the
 * process ends) instead of a per-call hipMalloc.  Calls no hipMalloc/hipFree.
 * Sizes needing more than the area (about 40 B per epoch-satellite plus
 * 40 B per epoch) are rejected with an error.  Every call on a GPU shares
 * that GPU's area and leaves its contents behind: calls that overlap on one
 * GPU race (this library serialises its own callers with a process-wide
 * lock).  Synchronous on the null stream, like gg_wls_batch. */
#define GG_WLS_NOALLOC_BYTES (16 * 1024 * 1024)
int gg_wls_batch_noalloc(const double* sat_ecef, const double* pseudoranges,
                         const double* weights, const int32_t* n_sat, int32_t n_epoch,
                         int32_t n_sat_max, int32_t max_iter, double tol, double* results,
                         int32_t* iters);
/* reference: wls_position, src/positioning/wls.cu:10-123 (no Sagnac) */
int gg_wls_position(const double* sat_ecef, const double* pseudoranges, const double* weights,
                    int32_t n_sat, int32_t max_iter, double tol, double* result,
                    int32_t* iters);

/* --- multi-GNSS WLS: position + one clock bias per system ------------------ */
#define GG_MAX_SYSTEMS 5 /* GPS, GLONASS, Galileo, BeiDou, QZSS (multi_gnss.h:5) */

/* reference: wls_multi_gnss_batch, src/positioning/multi_gnss.cu:307-350 (kernel
 * :211-305; binding python/gnss_gpu/_multi_gnss_bindings.cpp:36-66).  Host arrays
 * sat_ecef [n_epoch][n_sat][3], pseudoranges / weights [n_epoch][n_sat], int32
 * system_ids [n_epoch][n_sat] in [0, n_systems) (others add no clock column and
 * use clock 0's bias); results [n_epoch][3 + n_systems] = x, y, z, cb_0..;
 * iters [n_epoch] (may be NULL).  Sagnac range model.  n_systems in [0, 5];
 * n_sat < 3 + n_systems is an error, as in the reference binding.  weights NULL
 * means all ones. */
int gg_wls_multi_gnss_batch(const double* sat_ecef, const double* pseudoranges,
                            const double* weights, const int32_t* system_ids, int32_t n_epoch,
                            int32_t n_sat, int32_t n_systems, int32_t max_iter, double tol,
                            double* results, int32_t* iters);
/* reference: wls_multi_gnss, src/positioning/multi_gnss.cu:50-153 (binding
 * _multi_gnss_bindings.cpp:10-34).  One epoch, result [3 + n_systems]; plain
 * ranges (no Sagnac term, as the reference's CPU solver); fewer than 3 + n_systems
 * satellites or a satellite centroid below 1e-6 m gives zeros and *iters = -1. */
int gg_wls_multi_gnss(const double* sat_ecef, const double* pseudoranges, const double* weights,
                      const int32_t* system_ids, int32_t n_sat, int32_t n_systems,
                      int32_t max_iter, double tol, double* result, int32_t* iters);
/* Device-pointer variant of both, enqueued on hip_stream without synchronising;
 * mode 0 = wls_multi_gnss_batch semantics, 1 = wls_multi_gnss semantics. */
int gg_wls_multi_gnss_batch_device(const double* sat_ecef, const double* pseudoranges,
                                   const double* weights, const int32_t* system_ids,
                                   int32_t n_epoch, int32_t n_sat, int32_t n_systems,
                                   int32_t max_iter, double tol, double* results, int32_t* iters,
                                   int32_t mode, void* hip_stream);

/* --- ray casting through building meshes (binary64 ECEF) ------------------- */
/* Trees use the reference's flat node rows [n_nodes][10]: min_xyz, max_xyz, left,
 * right, tri_start, tri_count (left == -1: leaf; node 0 is the root), with the
 * triangles [n_tri][3][3] in the tree's order.  Every entry point that takes a
 * tree checks it first (a proper tree from node 0, leaf ranges inside the
 * triangle array, depth <= 62) and returns an error otherwise.  An empty tree
 * (n_nodes == 0) blocks nothing. */
typedef struct gg_bvh gg_bvh;

/* reference: bvh_build, src/raytrace/bvh.cu:221-243 (binding _bvh_bindings.cpp:12-49).
 * Binned-SAH build on the host; nodes_flat holds up to 2 * n_tri rows, *n_nodes is
 * the count written, sorted_tri_indices [n_tri] the triangle order of the leaves.
 * The tree may differ from the reference's; every query result depends only on the
 * triangles. */
int gg_bvh_build(const double* triangles, int32_t n_tri, double* nodes_flat, int32_t* n_nodes,
                 int32_t* sorted_tri_indices);
/* Device-resident copy of a tree for repeated queries (no reference counterpart:
 * the reference uploads the tree on every call). */
int gg_bvh_create(const double* nodes_flat, int32_t n_nodes, const double* sorted_tris,
                  int32_t n_tri, gg_bvh** out);
int gg_bvh_destroy(gg_bvh* bvh);

/* reference: raytrace_los_check_bvh, src/raytrace/bvh.cu:798-846 (kernel :293-354).
 * is_los [n_sat]: 1 when no triangle lies strictly between receiver and satellite. */
int gg_raytrace_los_check_bvh(const double* rx_ecef, const double* sat_ecef,
                              const double* nodes_flat, int32_t n_nodes,
                              const double* sorted_tris, int32_t n_tri, int32_t* is_los,
                              int32_t n_sat);
/* reference: raytrace_los_check_bvh_batch, bvh.cu:693-741 (kernel :517-585).
 * rx_ecef [n_epoch][3], sat_ecef [n_epoch][n_sat][3], is_los [n_epoch][n_sat];
 * a non-finite coordinate gives 0. */
int gg_raytrace_los_check_bvh_batch(const double* rx_ecef, const double* sat_ecef,
                                    const double* nodes_flat, int32_t n_nodes,
                                    const double* sorted_tris, int32_t n_tri, int32_t* is_los,
                                    int32_t n_epoch, int32_t n_sat);
/* reference: raytrace_multipath_bvh, bvh.cu:464-515 (kernel :362-462).  Smallest
 * positive excess delay of a first-order specular reflection per satellite
 * (0 and a zero point when there is none); reflection_points [n_sat][3]. */
int gg_raytrace_multipath_bvh(const double* rx_ecef, const double* sat_ecef,
                              const double* nodes_flat, int32_t n_nodes,
                              const double* sorted_tris, int32_t n_tri,
                              double* reflection_points, double* excess_delays, int32_t n_sat);
/* reference: raytrace_multipath_bvh_batch, bvh.cu:743-796 (kernel :587-691); a
 * non-finite coordinate gives 0 and a zero point. */
int gg_raytrace_multipath_bvh_batch(const double* rx_ecef, const double* sat_ecef,
                                    const double* nodes_flat, int32_t n_nodes,
                                    const double* sorted_tris, int32_t n_tri,
                                    double* reflection_points, double* excess_delays,
                                    int32_t n_epoch, int32_t n_sat);
/* reference: raytrace_los_check / raytrace_multipath, src/raytrace/raytrace.cu:195-277
 * (linear scans); same results through a tree built on the call. */
int gg_raytrace_los_check(const double* rx_ecef, const double* sat_ecef, const double* triangles,
                          int32_t n_tri, int32_t* is_los, int32_t n_sat);
int gg_raytrace_multipath(const double* rx_ecef, const double* sat_ecef, const double* triangles,
                          int32_t n_tri, double* reflection_points, double* excess_delays,
                          int32_t n_sat);

/* reference: pf_weight_3d_bvh, src/particle_filter/weight_3d_bvh.cu:225-300 (kernel
 * :127-220; binding _pf3d_bvh_bindings.cpp:18-92).  Host arrays px/py/pz/pcb [N];
 * overwrites log_weights [N] with the sum over satellites of the LOS / NLOS mixture:
 * LOS term -0.5 w r^2 / sigma_los^2, NLOS term with r - nlos_bias for r > 0 and
 * sigma_nlos, mixed with P(NLOS) = blocked_nlos_prob when a triangle lies on the
 * particle-satellite ray (1e-6 m < t < range), clear_nlos_prob otherwise. */
int gg_pf_weight_3d_bvh(const double* px, const double* py, const double* pz, const double* pcb,
                        const double* sat_ecef, const double* pseudoranges,
                        const double* weights_sat, const double* nodes_flat, int32_t n_nodes,
                        const double* sorted_tris, int32_t n_tri, double* log_weights,
                        int32_t n_particles, int32_t n_sat, double sigma_pr_los,
                        double sigma_pr_nlos, double nlos_bias, double blocked_nlos_prob,
                        double clear_nlos_prob);
/* reference: pf_weight_3d, src/particle_filter/weight_3d.cu:184-253 (linear scan); same
 * result through a tree built on the call. */
int gg_pf_weight_3d(const double* px, const double* py, const double* pz, const double* pcb,
                    const double* sat_ecef, const double* pseudoranges, const double* weights_sat,
                    const double* triangles, int32_t n_tri, double* log_weights,
                    int32_t n_particles, int32_t n_sat, double sigma_pr_los,
                    double sigma_pr_nlos, double nlos_bias, double blocked_nlos_prob,
                    double clear_nlos_prob);
/* Device-pointer variant of gg_pf_weight_3d_bvh on a gg_bvh handle, enqueued on
 * hip_stream without synchronising (all arrays device-resident binary64). */
int gg_pf_weight_3d_bvh_device(const gg_bvh* bvh, const double* px, const double* py,
                               const double* pz, const double* pcb, const double* sat_ecef,
                               const double* pseudoranges, const double* weights_sat,
                               double* log_weights, int32_t n_particles, int32_t n_sat,
                               double sigma_pr_los, double sigma_pr_nlos, double nlos_bias,
                               double blocked_nlos_prob, double clear_nlos_prob,
                               void* hip_stream);

/* Ray-traced NLOS weight on a single device filter's particles (the device-filter
 * counterpart of pf_weight_3d_bvh, src/particle_filter/weight_3d_bvh.cu:127-220):
 * overwrites the log-weights with the LOS / NLOS mixture of gg_pf_weight_3d_bvh for
 * each particle's ECEF position and clock bias (plus the inter-system bias of
 * system_ids[k] in 1..n_isb; system_ids may be NULL).  n_sat <= 1024. */
int gg_pf_device_weight_3d_bvh(gg_pf* h, const gg_bvh* bvh, const double* sat_ecef,
                               const double* pseudoranges, const double* weights_sat,
                               const int32_t* system_ids, int32_t n_sat, double sigma_pr_los,
                               double sigma_pr_nlos, double nlos_bias, double blocked_nlos_prob,
                               double clear_nlos_prob);

/* --- backward trajectory sampling over a particle history ----------------- */
/* reference: ffbsi_sample_indices / genealogy_smooth_indices,
 * python/gnss_gpu/particle_ffbsi.py:32-104, 121-151 (NumPy).  n_traj trajectories,
 * one workgroup each, steps t = T-1 .. 0 in one launch.  log_weights [T][N];
 * FFBSi (ancestors == NULL): states [T][N][4] (x, y, z, cb), vel [T][3], dt [T],
 * sigma_pos [T] (row t+1 drives the t -> t+1 transition), sigma_cb; genealogy
 * (ancestors [T][N] given, values in [0, N)): only the last step samples.
 * u [n_traj][T]: the uniform used at step t; forced [n_traj][T]: an index in [0, N)
 * replaces step t's draw (-1: sample; other values are rejected with -2 here).
 * indices [n_traj][T] (int64) out, always in [0, N).
 * status [n_traj]: -1 done; -2 the last step is degenerate (indices not written);
 * -3 (device variant) a forced value outside [-1, N) or a genealogy ancestor
 * outside [0, N) met: the walk stopped there, no out-of-range index written;
 * t >= 0 the first degenerate step met (a NaN or +inf log-weight, or all -inf),
 * where the walk stopped -- the reference draws an integer there, which the caller
 * passes through forced[] in a second call. */
int gg_ffbsi_sample(const double* log_weights, const double* states, const double* vel,
                    const double* dt, const double* sigma_pos, double sigma_cb,
                    const int32_t* ancestors, int32_t T, int32_t N, int32_t n_traj,
                    const double* u, const int32_t* forced, int64_t* indices, int32_t* status);
/* Device-pointer variant of gg_ffbsi_sample (every array device-resident), enqueued on
 * hip_stream without synchronising; no reference counterpart (the reference smoother
 * takes host arrays).  Pairs with gg_pf_device_record, which fills a history in HBM. */
int gg_ffbsi_sample_device(const double* log_weights, const double* states, const double* vel,
                           const double* dt, const double* sigma_pos, double sigma_cb,
                           const int32_t* ancestors, int32_t T, int32_t N, int32_t n_traj,
                           const double* u, const int32_t* forced, int64_t* indices,
                           int32_t* status, void* hip_stream);
/* reference: transition_logpdf, particle_ffbsi.py:14-29.  out [N] = log p(x_next | x_t[i])
 * of one predict step (Gaussian position noise sigma_pos per axis after v*dt, clock
 * noise sigma_cb), up to the normalising constant. */
int gg_transition_logpdf(const double* x_next, const double* x_t, int32_t N, const double* vel,
                         double dt, double sigma_pos, double sigma_cb, double* out);

/* Writes the single filter's particles as ECEF rows [N][4] (x, y, z, cb; the values
 * gg_pf_device_get_particles returns) and its log-weights [N] (binary64) to device
 * memory, enqueued on the handle's stream: one epoch of a device-resident history for
 * the smoother (the reference copies get_particles / get_log_weights to the host,
 * experiments/exp_ffbsi_eval.py:189-190).  No reference counterpart. */
int gg_pf_device_record(gg_pf* h, double* states_dev, double* log_weights_dev);

/* --- double-differenced and carrier-phase updates (single filter) ---------- */
/* reference: pf_device_weight_dd_pseudorange, src/particle_filter/pf_device.cu:1576-1620
 * (kernel :282-368).  Overwrites the log-weights with the DD pseudorange
 * likelihood over n_dd pairs (sat_ecef_k / ref_ecef [n_dd][3] ECEF, dd_pseudorange,
 * base_range_k, base_range_ref, weights_dd [n_dd]); no clock bias.  Gate: a
 * threshold > 0 drops |residual| > threshold unless fewer than min(n_dd, 3)
 * remain; per_particle_huber selects the weighted Huber cost.  n_dd <= 1024. */
int gg_pf_device_weight_dd_pseudorange(gg_pf* h, const double* sat_ecef_k, const double* ref_ecef,
                                       const double* dd_pseudorange, const double* base_range_k,
                                       const double* base_range_ref, const double* weights_dd,
                                       int32_t n_dd, double sigma_pr,
                                       double per_particle_nlos_threshold_m,
                                       int32_t per_particle_huber, double per_particle_huber_k);
/* reference: pf_device_weight_carrier_afv, pf_device.cu:2062-2116 (kernel :439-494).
 * Adds the undifferenced carrier-phase AFV log-likelihood
 * -w afv^2 / (2 sigma^2), afv = x - rint(x), x = phase - (range + cb) / wavelength,
 * to the log-weights (carrier_phase in cycles). */
int gg_pf_device_weight_carrier_afv(gg_pf* h, const double* sat_ecef, const double* carrier_phase,
                                    const double* weights_sat, int32_t n_sat, double wavelength,
                                    double sigma_cycles);
/* reference: pf_device_weight_dd_carrier_afv, pf_device.cu:2122-2170 (kernel :502-604).
 * Adds the DD carrier-phase AFV log-likelihood with per-pair wavelengths; gate in
 * cycles and Huber as for the DD pseudorange update. */
int gg_pf_device_weight_dd_carrier_afv(gg_pf* h, const double* sat_ecef_k, const double* ref_ecef,
                                       const double* dd_carrier, const double* base_range_k,
                                       const double* base_range_ref, const double* weights_dd,
                                       const double* wavelengths_m, int32_t n_dd,
                                       double sigma_cycles,
                                       double per_particle_nlos_threshold_cycles,
                                       int32_t per_particle_huber, double per_particle_huber_k);

/* reference: pf_device_weight_doppler, pf_device.cu:2183-2226 (kernel :610-740).
 * Doppler rows (sat_ecef, sat_vel [n_sat][3] ECEF, doppler_hz, weights_sat):
 * adds the range-rate log-likelihood (clock drift = weighted mean misfit) and,
 * for velocity_update_gain > 0 and n_sat >= 4, moves each particle's velocity
 * toward its Doppler WLS solution (step capped at max_velocity_update_mps).
 * A no-op for n_sat <= 0, sigma_mps <= 0 or wavelength_m <= 0, as in the
 * reference.  Needs a handle with velocity storage; n_sat <= 512. */
int gg_pf_device_weight_doppler(gg_pf* h, const double* sat_ecef, const double* sat_vel,
                                const double* doppler_hz, const double* weights_sat,
                                int32_t n_sat, double wavelength_m, double sigma_mps,
                                double velocity_update_gain, double max_velocity_update_mps);
/* reference: pf_device_doppler_kf_update, pf_device.cu:2228-2270 (kernel :742-880).
 * Per-particle Kalman update of the velocity Gaussian (RBPF), clock drift
 * eliminated by centring; adds the innovation log-likelihood. */
int gg_pf_device_doppler_kf_update(gg_pf* h, const double* sat_ecef, const double* sat_vel,
                                   const double* doppler_hz, const double* weights_sat,
                                   int32_t n_sat, double wavelength_m, double sigma_mps);

/* --- learned likelihood (no reference counterpart) -------------------------- */
/* An 8 -> 64 -> 64 -> 2 MLP per (particle, satellite) pair predicting the NLOS
 * bias mean and log-variance of the pseudorange residual; it replaces the
 * Student-t term of pfd_weight_kernel (src/particle_filter/pf_device.cu:262-266)
 * when a filter is given one.  Parameters: GG_MLP_NPARAMS binary32 values
 * (W1[64][8] b1[64] W2[64][64] b2[64] W3[2][64] b3[2]); bf16 operands with
 * binary32 accumulation on the matrix cores. */
int gg_mlp_create(gg_mlp** out);
int gg_mlp_destroy(gg_mlp* h);
int gg_mlp_set_params(gg_mlp* h, const float* params);
int gg_mlp_get_params(gg_mlp* h, float* params);
/* FP8 weight mode (enable != 0): W1, W2 and W3 are used as OCP FP8 E4M3 values
 * (round to nearest even, saturating) times one power-of-two scale per tensor,
 * the smallest with max|W| / scale <= 448; the binary32 master parameters are
 * kept, so training in this mode updates them through the quantised forward
 * and backward passes.  Off (bf16 weights) by default. */
int gg_mlp_set_fp8(gg_mlp* h, int32_t enable);
/* Pair log-likelihoods for n_groups groups of 16 pairs sharing one satellite:
 * e[n_groups*16] residuals [m], f[n_groups*7] satellite features; outputs
 * ll[n_groups*16] and z3[n_groups*16*2] (network outputs). */
int gg_mlp_eval(gg_mlp* h, int32_t n_groups, const float* e, const float* f, float* ll, float* z3);

/* Training set of n_epochs epochs with up to s_max (<= 16) satellites each:
 * f[E][s_max][7] features, r0[E][s_max] residual at the true position [m],
 * u[E][s_max][3] unit vectors to the satellites, n_sat[E]. */
int gg_mlp_load_data(gg_mlp* h, int32_t n_epochs, int32_t s_max, const float* f, const float* r0,
                     const float* u, const int32_t* n_sat);
/* n_steps Adam steps (batch epochs each, 64 candidate offsets per epoch drawn
 * from N(0, diag(sh^2, sh^2, sv^2, scb^2)), candidate 0 = the truth); loss_out
 * (may be NULL) gets the mean batch loss of every step. */
int gg_mlp_train(gg_mlp* h, int32_t n_steps, int32_t batch, float lr, uint64_t seed,
                 float cand_sigma_h, float cand_sigma_v, float cand_sigma_cb, float* loss_out);
/* Loss and gradient (sum over epochs) of the first n_epochs loaded epochs for
 * given candidates cand[n_epochs][64][4] (dx dy dz dcb); loss[n_epochs]. */
int gg_mlp_loss_grad(gg_mlp* h, int32_t n_epochs, const float* cand, float* grad, float* loss);

/* Attach a learned likelihood to a particle filter (NULL detaches).  The gg_mlp
 * handle must outlive every weight launch that uses it; steps select it with
 * gg_step_args.lik = GG_LIK_LEARNED. */
int gg_pf_set_mlp(gg_pf* h, gg_mlp* m);
/* Per-row satellite inputs of the learned likelihood for n_slots epoch slots:
 * cn0[n_slots][rows] (dB-Hz) and los_flag[n_slots][rows] (1 = line of sight at
 * the particle mean), rows as in gg_pfb_load_epochs. */
int gg_pfb_load_features(gg_pf* h, int32_t n_slots, const float* cn0, const float* los_flag);
/* pf_device_weight (pf_device.cu:1516-1574) with the learned likelihood in
 * place of the Gaussian / Student-t term; cn0[n_sat], los_flag[n_sat].
 * learned_mode 0 evaluates the network per (particle, satellite) pair; 1 first
 * tabulates it per satellite over the bf16 codes of its scaled-residual input
 * (GG_MLP_CODES entries) and looks the pairs up: the same values, fewer
 * network evaluations when particles outnumber the codes. */
int gg_pf_device_weight_learned(gg_pf* h, const double* sat_ecef, const double* pseudoranges,
                                const double* weights_sat, const int32_t* sys, int32_t n_sat,
                                const float* cn0, const float* los_flag, int32_t learned_mode);

/* --- test hook (no reference counterpart) ----------------------------------- */
/* Evaluates a deterministic device building block for n inputs so tests can
 * compare it bit for bit with the oracle: op 0 exp64, 1 exp32, 2 Philox words
 * (out holds 4n values for op 2). */
int gg_debug_math(int32_t op, const double* in, double* out, int32_t n);

#ifdef __cplusplus
}
#endif

#endif /* GNSS_GPU_AMD_H */
