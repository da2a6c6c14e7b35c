/* phx — MI355X-native batched fixed-step RKF78 orbit propagator with the
 * ParaHPOP full force model. C-ABI of libphx.so (parahpop_amd/libphx.so).
 *
 * All buffers are host-owned and structure-of-arrays: state6_soa holds
 * x[n], y[n], z[n], vx[n], vy[n], vz[n] (km, km/s, ICRF-aligned, Earth-centred).
 * Every call returns 0 on success and a negative code on failure; the message
 * is available from phx_last_error(). No C++ exception crosses the ABI.
 * A context owns one device, one HIP stream and its device memory; it is not
 * re-entrant (use one context per host thread or serialize calls), except the
 * serving calls phx_submit/phx_poll, which are thread-safe; phx_last_error
 * returns a copy owned by the calling thread (valid until its next call).
 */
#ifndef PHX_H
#define PHX_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* Error codes */
#define PHX_OK 0
#define PHX_EINVAL -1      /* invalid configuration, target or argument */
#define PHX_EDEVICE -2     /* HIP runtime or device failure */
#define PHX_ENOTDONE -3    /* a target did not reach its end of simulation */
#define PHX_ENOMEM -4      /* device memory budget exceeded */
#define PHX_EAGAIN -5      /* serving: result not ready yet */

typedef struct phx_ctx phx_ctx;
typedef struct phx_batch phx_batch;

typedef struct phx_stats {
    double object_steps;       /* completed RKF78 steps summed over objects */
    double kernel_seconds;     /* device time of the step kernels (HIP events) */
    double epoch_seconds;      /* device time of the epoch-record kernel */
    double wall_seconds;       /* host wall time of the call */
    double epoch_records;      /* distinct stage epochs evaluated */
    double launches;           /* kernel launches of the step path */
    int waves;                 /* waves per 64-object tile used */
    int steps_per_launch;      /* steps per launch of the fused step kernel; 1 on the split-stage path */
    int kernels_per_step;      /* kernels per step on the split-stage path; 0 on the fused path */
    int sh_groups_min;         /* split-stage path: fewest and most SH workgroups per tile over */
    int sh_groups_max;         /* the call's steps (1: the one-workgroup kernel); 0 on the fused path */
    double graph_replays;      /* split-stage path: graphs of graph_steps steps replayed in the call */
    double graph_captures;     /* split-stage path: graphs captured in the call (one per active-tile set) */
} phx_stats;

/* Creates a context from a ParaHPOP CONFIG.json: parses and validates the
 * "model" and "execution" sections (interface: src/common/driver.cpp:72-86,132-135,
 * interface/config/Model.h:5-28), loads the environment data it names
 * (gravity field, space weather, EOP, constants) and uploads all tables.
 * Replaces the reference's config::Model construction + environment().load(). */
int phx_create(const char* config_json_path, phx_ctx** out);

/* Propagates n objects for duration_s with the configured fixed step,
 * shortening each object's final step (propagators/FixedStep.h:7-47), and
 * writes final epochs (MJD2000 TDB) and states. Inputs as in TARGETS.json:
 * areas in m^2; the km^2 conversion and cr_eff = cr*srp/drag happen inside
 * (driver.cpp:99-123). Replaces paraHPOP::run(cfg, samples, duration, step,
 * maxSteps, threads, gpu=true) (engine/paraHPOP/run.h:5-13). Synchronous. */
int phx_propagate(phx_ctx* ctx, size_t n, const int64_t* id, const double* epoch_mjd2000_tdb,
                  const double* state6_soa, const double* mass_kg, const double* drag_area_m2,
                  const double* srp_area_m2, const double* cd, const double* cr, double duration_s,
                  double* out_epoch, double* out_state6_soa, phx_stats* stats);

/* Device-resident batches (the same work as phx_propagate split into
 * upload / run / download so a caller can time or interleave the device
 * part; replaces the reference's Model::upload + graph replay loop +
 * download in propagators/device/Propagator.h:33-66). */
int phx_batch_create(phx_ctx* ctx, size_t n, const int64_t* id, const double* epoch_mjd2000_tdb,
                     const double* state6_soa, const double* mass_kg, const double* drag_area_m2,
                     const double* srp_area_m2, const double* cd, const double* cr, double duration_s,
                     phx_batch** out);
/* As phx_batch_create with one duration per object (duration_s[n]): the
 * uneven batch of epochs, durations and shortened final steps runs as one
 * device-resident batch (epoch records on a uniform grid, tiles ordered
 * longest work first). Extension: the reference takes one duration per run
 * (engine/paraHPOP/run.h:5-13). */
int phx_batch_create_uneven(phx_ctx* ctx, size_t n, const int64_t* id, const double* epoch_mjd2000_tdb,
                            const double* state6_soa, const double* mass_kg, const double* drag_area_m2,
                            const double* srp_area_m2, const double* cd, const double* cr, const double* duration_s,
                            phx_batch** out);
/* phx_propagate on caller-owned device buffers, allocating no device memory
 * (no counterpart in the reference; for callers that cannot allocate: code
 * under HIP graph capture, tight serving loops, harnesses that forbid it).
 * d_epoch (n, MJD2000 TDB) and d_state6_soa (6n, SoA) are read and overwritten
 * with the final epochs and states; d_mass, d_drag_area (m^2), d_srp_area (m^2),
 * d_cd, d_cr (n each) are read. All are device pointers, written by the caller
 * before the call (synchronize the caller's stream first). Every working array
 * lives in one fixed device array of the library (21.1 MB per GPU, held
 * until the process ends): at most 16,384 objects and an epoch span of 2,043
 * grid steps (34 h at the 60 s grid). Runs the per-lane epoch path: the
 * results equal phx_propagate's with the option force_grid = 1. Synchronous on
 * the context's stream. Calls that overlap on one GPU (any context) race on
 * the array. Errors as phx_propagate (PHX_EINVAL beyond the limits). */
int phx_propagate_noalloc(phx_ctx* ctx, size_t n, double* d_epoch, double* d_state6_soa, const double* d_mass,
                          const double* d_drag_area, const double* d_srp_area, const double* d_cd, const double* d_cr,
                          double duration_s, phx_stats* stats);
/* Autotuner (no reference counterpart; the reference reads fixed
 * PARAHPOP_PERF_* overrides, AccumulateBodies.cu:20-48): times trial_steps
 * steps of this batch for each candidate shape (fused kernel: waves per
 * 64-object tile and the SH share of wave 0; split stages for shared-epoch
 * batches with drag: SH waves and species per workgroup), keeps the fastest in
 * the context's options, resets the batch and reports the winner's trial
 * statistics. */
int phx_batch_autotune(phx_batch* b, int trial_steps, phx_stats* best);
/* Runs up to max_steps further integration steps (<= 0: to completion). */
int phx_batch_run(phx_batch* b, int max_steps, phx_stats* stats);
/* 1 when every object reached its end of simulation, else 0. */
int phx_batch_done(const phx_batch* b);
int phx_batch_download(phx_batch* b, double* out_epoch, double* out_state6_soa);
/* Current epochs (MJD2000 TDB) and states of every object, finished or not
 * (diagnostic: checks part of a run; no reference counterpart). */
int phx_batch_state(phx_batch* b, double* out_epoch, double* out_state6_soa);
/* Copies the batch's initial state back in place so it can be run again. */
int phx_batch_reset(phx_batch* b);
void phx_batch_destroy(phx_batch* b);

/* Tuning options (no reference counterpart; the reference reads
 * PARAHPOP_PERF_* environment overrides, AccumulateBodies.cu:20-48):
 *   "split"            1 (default): batches and serving with drag run as split
 *                      stages (SH kernel and drag kernel per stage, k-stages in
 *                      device memory; per-lane epochs need split_species 8);
 *                      0: the fused multi-step kernel
 *   "split_waves"      4, 8 or 16 (default) waves per tile of the split SH kernel
 *   "split_groups"     1 .. 4 (default 1): with 8 or 16 waves, a batch with fewer
 *                      active tiles than compute units splits each tile's SH orders
 *                      over up to this many workgroups (min(groups, CUs / tiles));
 *                      the last workgroup to finish sums the parts in a fixed
 *                      order. Faster tails, but the SH sum is then partitioned by
 *                      the number of active tiles, so an object's last bits
 *                      (about 1e-15 relative per step) depend on the rest of its
 *                      batch and on the CU count; with 1 they do not
 *   "graph_steps"      0 .. 256 (default 8): the split stages of this many steps
 *                      are captured once as a graph and replayed while the active
 *                      tiles stay the same for all of them (0: every kernel is
 *                      launched); the results do not change
 *   "split_species"    8 (default): one eight-wave drag workgroup per tile (the
 *                      species and the drag); 1, 2 or 4: species per workgroup and
 *                      a finish kernel; 0: one four-wave drag workgroup per tile
 *   "split_cost_pg", "split_cost_corr", "split_cost_srp", "split_cost_lane",
 *   "split_cost_order", "split_cost_order0"  balancing of the split SH kernel's
 *                      order blocks (8 and 16 waves), in FP64 terms of one wave:
 *                      the cost of point gravity, tides and relativity, SRP and
 *                      the NRLMSISE-00 lane inputs on the waves that carry them,
 *                      the fixed cost of an order and the per-term factor of
 *                      order 0 (defaults 105, 246, 83, 104, 2, 1.8)
 *   "waves"            fused kernel: 1, 2 or 4 waves per 64-object tile (8 or 16
 *                      without drag)
 *   "steps_per_launch" fused kernel: integration steps per kernel launch
 *   "share0", "share_corr"  fused kernel: relative SH order-block shares of wave 0
 *                      and of the wave that also evaluates tides, relativity, SRP
 *   "sort_altitude"    1: group objects below the drag cutoff into common tiles
 *   "exact_groups_max", "grid_step", "force_grid"  epoch-record mode (exact
 *                      per-stage records up to this many (start, end) groups,
 *                      else a uniform grid of grid_step seconds)
 *   "serve_slots", "serve_record_mib", "serve_steps_per_tick"  serving pool
 * The serving loop reads the options that shape its steps (the split and
 * fused kernel shapes, the balancing costs, serve_steps_per_tick) once, at the
 * first phx_submit; later changes apply to batches only. */
int phx_set_option(phx_ctx* ctx, const char* key, double value);
double phx_get_option(const phx_ctx* ctx, const char* key);

/* Serving (continuous batching). A request is admitted when its device
 * footprint fits; its objects join the running device-resident batch at the
 * next step boundary and leave it when they finish. */
typedef struct phx_request {
    size_t n;
    const int64_t* id;
    const double* epoch_mjd2000_tdb;
    const double* state6_soa;
    const double* mass_kg;
    const double* drag_area_m2;
    const double* srp_area_m2;
    const double* cd;
    const double* cr;
    double duration_s;
} phx_request;

typedef struct phx_result {
    size_t n;
    double* out_epoch;          /* caller-owned, n entries */
    double* out_state6_soa;     /* caller-owned, 6n entries */
    double submit_time_s;       /* monotonic clock at submit */
    double finish_time_s;       /* monotonic clock when the last object finished */
} phx_result;

int phx_submit(phx_ctx* ctx, const phx_request* req, uint64_t* ticket);
int phx_poll(phx_ctx* ctx, uint64_t ticket, phx_result* result);

const char* phx_last_error(const phx_ctx* ctx);
void phx_destroy(phx_ctx* ctx);

/* Diagnostics for tests (no reference counterpart). */
/* Force model at n states (pos/vel SoA [3][n]) and one epoch (ET seconds):
 * which = -1 all, 0 central, 1 SH, 2 third bodies, 3 SRP, 4 drag, 5 corrections.
 * area_km2_over_mass, cr_eff, cd per object; acc out [3][n] km/s^2. */
int phx_force_probe(phx_ctx* ctx, size_t n, const double* pos, const double* vel, const double* mass,
                    const double* area_km2, const double* cr_eff, const double* cd, double et, int which,
                    double* acc);
/* Epoch record (352 doubles, record_len >= 352) for one ET epoch. */
int phx_epoch_record(phx_ctx* ctx, double et, double* record, size_t record_len);
/* NRLMSISE-00 gtd7d on the device for test-driver style inputs (rows of 16:
 * doy, sec, alt, lat, lon, lst, f107a, f107, ap, ap[7]); out rows of 11. */
int phx_nrlmsise_probe(phx_ctx* ctx, size_t n, const double* in, int use_ap_array, double* out);
const char* phx_version(void);

#ifdef __cplusplus
}
#endif
#endif
