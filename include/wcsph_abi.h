/*
 * wcsph_abi.h -- C-ABI of the MI355X (gfx950) hydro force path for a weakly
 * compressible SPH solver.
 *
 * Two groups of entry points:
 *
 *  1. The FDPS multiwalk drop-in. These replace the reference's CUDA backend
 *     GPU_WC::HydrDispatchKernel / GPU_WC::HydrRetrieveKernel
 *     (src/prototype.h:7-10, defined at src/kernel.cu:289-381 and :383-409).
 *     Records are passed in the reference's own binary layouts
 *     (src/class.h:63-67 RESULT::Hydro, :167-176 EPI::Hydro,
 *     :212-223 EPJ::Hydro) so FDPS's buffers are handed over without copies.
 *     include/wcsph_fdps.hpp wraps them in the exact reference signatures.
 *
 *  2. A device-resident simulation (wcsph_sim_*): cell list by counting sort,
 *     neighbour pass, force pass, FP64 kick-drift-kick and the global time
 *     step, all kept in HBM. These replace the whole per-step loop of
 *     src/main.cpp:89-133 (FDPS tree + dispatch + retrieve + the OpenMP
 *     integrator loops of src/class.h:146-163 and src/integral.h:2-12).
 *
 * Conventions: every function returns 0 on success and a nonzero WCSPH_E*
 * code on failure; wcsph_last_error() returns a message for the calling
 * thread's last failure. Nothing throws across this boundary and nothing
 * reads stdin (the reference asks for a device id on stdin,
 * src/kernel.cu:295-297; here it is wcsph_params.device).
 */
#ifndef WCSPH_ABI_H
#define WCSPH_ABI_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define WCSPH_OK 0
#define WCSPH_EINVAL 1      /* bad argument */
#define WCSPH_EHIP 2        /* HIP runtime failure */
#define WCSPH_ESTATE 3      /* call out of order (e.g. retrieve without dispatch) */
#define WCSPH_ENOMEM 4      /* allocation failed or over a stated budget */
#define WCSPH_EOVERFLOW 5   /* neighbour list or cell grid capacity exceeded */
#define WCSPH_EIO 6         /* a file could not be opened, read, written or parsed */
#define WCSPH_EDIVERGED 7   /* a simulation's state has diverged: a non-finite position,
                               or more particles in one cell than the cell list ranks */

/* ---------------- FDPS boundary records (64-bit FDPS build) ----------------
 * 3D sizes: EPI 96 B, EPJ 112 B, RESULT 40 B; 2D: 80 / 96 / 32 B. */
typedef struct { double pos[3]; double vel[3]; double smth, dens, pres, snds, visc; int64_t tag; } wcsph_epi_hydro_3d;
typedef struct { double pos[2]; double vel[2]; double smth, dens, pres, snds, visc; int64_t tag; } wcsph_epi_hydro_2d;
typedef struct { double pos[3]; double vel[3]; double dens, mass, smth, pres, snds, visc; int64_t tag; int32_t type; int32_t pad_; } wcsph_epj_hydro_3d;
typedef struct { double pos[2]; double vel[2]; double dens, mass, smth, pres, snds, visc; int64_t tag; int32_t type; int32_t pad_; } wcsph_epj_hydro_2d;
typedef struct { double acc[3]; double dt, div_v; } wcsph_result_hydro_3d;
typedef struct { double acc[2]; double dt, div_v; } wcsph_result_hydro_2d;

/* Physics and numerics parameters (one POD across the ABI). */
typedef struct {
    int32_t dim;        /* 2 or 3 */
    int32_t vsig_mode;  /* 0: v_sig = max(v_sig, mu) (src/force.h:47, gives v_sig = 0);
                           1: v_sig = max(v_sig, -mu) (src/kernel.cu:263) */
    double alpha;       /* linear artificial viscosity (src/force.h:32: 0.01) */
    double beta;        /* quadratic artificial viscosity (src/force.h:46: none) */
    double eps;         /* regulariser in mu and the viscous term (src/force.h:45,53: 0.01) */
    double c_cfl;       /* Courant factor (src/param.h:5: 0.3) */
    double rho0;        /* Tait reference density (src/init/Dam2.h:39) */
    double c0;          /* Tait reference sound speed (src/init/Dam2.h:42) */
    double gamma;       /* Tait exponent (src/init/Dam2.h:38) */
    double g;           /* gravity along the last axis (src/init/Dam2.h:106) */
    double nu;          /* kinematic viscosity for the simulation (src/EoS.h:38) */
    double rho_floor;   /* density floor after each step (src/init/Dam2.h:111; 0 = off) */
    int32_t device;     /* HIP device ordinal */
    int32_t shepard_interval; /* device-resident step: Shepard density filter (src/force.h:10-30)
                                 at init and on steps with step % interval == 0
                                 (src/main.cpp:75,105; src/param.h:12 uses 30); 0 = off */
    int32_t grad_mode;  /* 0: kernel gradient at the mean smoothing length (src/force.h:49);
                           1: the mean of the gradients at h_i and h_j, as the reference GPU
                           kernel takes it (src/kernel.cu:265), a pair then interacting within
                           2 max(h_i, h_j). The two agree when h is uniform. The FDPS drop-in
                           takes either; the device-resident paths and wcsph_force_pass take
                           1 only with one smoothing length for all particles. */
    int32_t wall_mode;  /* device-resident step only (the FDPS drop-in and wcsph_force_pass take 0):
                           0: every pair is evaluated, as the reference's functor does;
                           1: frozen walls. A wall particle never moves and is left out of the
                           global time step (src/class.h:146-163, src/integral.h:6), so its
                           only result the step uses is div v, and a wall-wall pair adds
                           exactly zero to it (dv = 0). A wave whose 64 particles are all
                           walls then takes fluid neighbours only, and wall accelerations
                           are written as zero. Every fluid result and every dt is the same;
                           a wall's div v is the same sum (the FP32 path: the same bits; the
                           mixed path: its terms may be added in another order). Needs
                           every wall at rest (checked on upload and restore). */
} wcsph_params;

/* Fill p with the reference CPU functor's constants (src/force.h:32,45-47,
 * src/param.h:5, src/init/Dam2.h:38-43,106, src/EoS.h:38). */
int wcsph_params_reference(int32_t dim, wcsph_params* p);
/* Message of the calling thread's last failure. No counterpart in the
 * reference, which asserts instead (src/kernel.cu:316,367). */
const char* wcsph_last_error(void);
/* Library version string, for logs. No counterpart in the reference. */
const char* wcsph_version(void);
/* HBM bandwidth probe: `reps` streaming copies of `bytes` on `device`;
 * *gbps = (read + write bytes) / time. Used by the bench for the measured
 * HBM peak and to calibrate hardware byte counters (kernel k_copy). No
 * counterpart in the reference. */
int wcsph_copy_probe(int32_t device, int64_t bytes, int32_t reps, double* gbps);
/* VALU rate probe: independent FMA chains in every lane of a full grid on
 * `device`, best of `reps` launches; *tflops = 2 x FMAs / time. kind 0: FP32
 * FMA (v_fma_f32) reading only VGPRs and constants, kind 1: packed FP16 FMA
 * (v_pk_fma_f16, two FMAs per instruction), kind 2: FP32 FMA with one
 * wave-uniform SGPR operand (on gfx950 it issues at half the rate of kind
 * 0). Used by the bench for the measured VALU peaks (kernel k_valu). Neither
 * probe has a counterpart in the reference. */
int wcsph_valu_probe(int32_t device, int32_t kind, int32_t reps, double* tflops);

/* ---------------- 1. FDPS multiwalk drop-in ---------------- */
/* Replaces the lazy first-call set-up of src/kernel.cu:290-305. */
int wcsph_init(const wcsph_params* p);
/* Releases the drop-in's buffers and streams, which the reference keeps as
 * static singletons for the life of the process (src/kernel.cu:42-131,239). */
int wcsph_finalize(void);

/* GPU_WC::HydrDispatchKernel (src/kernel.cu:289-381). Copies the walks into
 * pinned staging (FDPS reuses its buffers after the call), starts the
 * upload and the force kernel on the tag's stream, and returns with the work
 * in flight. tag in [0, 8). At most one outstanding batch per tag. */
int wcsph_hydro_dispatch_3d(int32_t tag, int32_t n_walk, const wcsph_epi_hydro_3d* const* epi,
                            const int32_t* n_epi, const wcsph_epj_hydro_3d* const* epj,
                            const int32_t* n_epj);
/* The same for a 2D FDPS build (F64vec of two doubles, src/class.h:167-176,
 * 212-223 with PARTICLE_SIMULATOR_TWO_DIMENSION). */
int wcsph_hydro_dispatch_2d(int32_t tag, int32_t n_walk, const wcsph_epi_hydro_2d* const* epi,
                            const int32_t* n_epi, const wcsph_epj_hydro_2d* const* epj,
                            const int32_t* n_epj);
/* GPU_WC::HydrRetrieveKernel (src/kernel.cu:383-409). Waits for the tag's
 * batch and assigns acc, dt and div_v of every i of every walk. */
int wcsph_hydro_retrieve_3d(int32_t tag, int32_t n_walk, const int32_t* ni, wcsph_result_hydro_3d** force);
/* The same for a 2D FDPS build (RESULT::Hydro of 32 B, src/class.h:63-67). */
int wcsph_hydro_retrieve_2d(int32_t tag, int32_t n_walk, const int32_t* ni, wcsph_result_hydro_2d** force);
/* Device time (ms) of the last batch's force kernels of a tag, from HIP
 * events. No counterpart in the reference (its CudaTimer is unused). */
int wcsph_hydro_last_kernel_ms(int32_t tag, float* ms);
/* Timing of the last batch of a tag: out[0] host packing (ms, wall clock,
 * all chunks), out[1] device span from the first upload to the end of the
 * result copy (ms, events), out[2] force kernels (ms, events), out[3] host
 * unpacking in retrieve (ms), out[4] chunks the batch was cut into. Waits
 * for the batch's device work. No counterpart in the reference. */
int wcsph_hydro_last_timing(int32_t tag, double* out);

/* ---------------- 2. device-resident simulation ---------------- */
#define WCSPH_PREC_FP32 0   /* FP32 pair arithmetic and accumulation */
#define WCSPH_PREC_MIXED 1  /* h-scaled FP16 pair differences on packed FP16 ops; FP32 density, pressure, accumulation */
#define WCSPH_PREC_FP64 2   /* FP64 everywhere (the GPU oracle) */

typedef struct wcsph_sim wcsph_sim;

typedef struct {
    int32_t precision;      /* WCSPH_PREC_* */
    int32_t max_neighbours; /* list capacity per particle for the FP64 path and
                               wcsph_sim_neighbours (0: 128) */
    double box_lo[3];       /* cell-grid domain; particles outside are clamped */
    double box_hi[3];       /*   into the edge cells and counted as escaped */
    int32_t cells_per_support; /* cell edge = largest support / this (1 or 2; 0: 2) */
    int32_t queue_depth;    /* fused pass: LDS queue slots per lane (24 or 32; 0: 24) */
} wcsph_sim_config;

typedef struct {
    double time, dt;
    int64_t step;
    int64_t n;
    int64_t pairs_last;     /* neighbour pairs (i != j, r < hi+hj) of the last force pass */
    int64_t pairs_total;    /* summed over all force passes */
    int64_t escaped;        /* particles clamped into the grid edge at the last binning */
    int32_t max_count;      /* largest neighbour count of the last pass */
    int32_t overflow;       /* nonzero if a neighbour list was truncated */
    int32_t nonfinite;      /* particles found with a non-finite position (sticky; the
                               step and sync calls then return WCSPH_EDIVERGED) */
    int32_t overfull;       /* nonzero once a cell held more than 256 particles (sticky) */
} wcsph_sim_stats;

/* Dam-break initial condition on a square/cubic lattice (the geometry of the
 * dam-break configurations: column col_w x depth x col_h at x in [0,col_w],
 * tank tank_l x depth x tank_h, nwall wall layers on the sides and bottom,
 * hydrostatic density, h = hfac*dx, m = rho*dx^D). 2D ignores depth. */
typedef struct {
    int32_t dim;
    int32_t nwall;
    double dx, col_w, col_h, depth, tank_l, tank_h, hfac;
} wcsph_dam_break_desc;

/* Number of particles the description generates (fluid and wall). Replaces
 * the lattice loops of src/init/Dam2.h:9-103 for the configurations' tank. */
int wcsph_dam_break_count(const wcsph_dam_break_desc* d, int64_t* n_fluid, int64_t* n_wall);
/* Host-side generation into caller SoA arrays: pos/vel [dim][n], others [n]
 * (src/init/Dam2.h:9-103; the density is hydrostatic, which the reference
 * leaves commented out at src/init/Dam2.h:73-74). */
int wcsph_dam_break_fill(const wcsph_dam_break_desc* d, const wcsph_params* p, double* pos,
                         double* vel, double* dens, double* mass, double* h, int32_t* type);
/* Device bytes a simulation of n particles needs at cfg's precision
 * (neighbour lists only for FP64), assuming a cell grid of about one cell
 * per particle; the grid an upload builds depends on the state, so this is
 * an estimate. No counterpart in the reference. */
int64_t wcsph_sim_bytes(int32_t dim, int64_t n, const wcsph_sim_config* cfg);
/* Exact device bytes of a simulation created with p and cfg and uploaded
 * with this state (pos [dim][n], h [n]): the grid is computed on the host as
 * the upload would build it. -1 on a null or inconsistent argument. The
 * serving loop admits requests by this figure. No counterpart in the
 * reference. */
int64_t wcsph_sim_bytes_state(const wcsph_params* p, const wcsph_sim_config* cfg, int64_t n, const double* pos,
                              const double* h);

/* Create / destroy the device state of one simulation: the role of FDPS's
 * ParticleSystem and tree set-up in src/main.cpp:33-67, held in HBM. */
int wcsph_sim_create(const wcsph_params* p, const wcsph_sim_config* cfg, int64_t n, wcsph_sim** out);
int wcsph_sim_destroy(wcsph_sim* s);
/* Upload a state (SoA, host): pos/vel [dim][n], dens/mass/h [n], type [n]
 * (0 fluid, 1 wall = FREEZE of src/class.h:33-36). acc, div_v and the
 * per-particle dt are then produced by wcsph_sim_init. Every value must be
 * finite (WCSPH_EINVAL otherwise). */
int wcsph_sim_upload(wcsph_sim* s, const double* pos, const double* vel, const double* dens,
                     const double* mass, const double* h, const int32_t* type);
/* Pressure, first force pass, the Shepard filter when shepard_interval > 0,
 * and gravity (src/main.cpp:69-79). Resets the clock to time 0, step 0. */
int wcsph_sim_init(wcsph_sim* s);
/* nsteps steps of src/main.cpp:89-133, queued on the simulation's stream;
 * returns without waiting. The Shepard filter runs after the drift of every
 * step whose index (0 after init) is a multiple of shepard_interval
 * (src/main.cpp:105, src/param.h:12). Returns WCSPH_EDIVERGED, queueing
 * nothing, once steps queued by an earlier call (and completed) found a
 * non-finite position or an overfull cell; such steps stop the clock and do
 * no force work, so a diverged run cannot stall the device. */
int wcsph_sim_step(wcsph_sim* s, int32_t nsteps);
/* Steps are replayed from a captured two-step hipGraph (one submission per
 * two steps) unless per-step timing is on or graphs are disabled here. No
 * counterpart in the reference. */
int wcsph_sim_set_graphs(wcsph_sim* s, int32_t enable);
/* Non-blocking: *idle = 1 when all queued work of the simulation is done
 * (then WCSPH_EDIVERGED if the state has diverged). No counterpart in the
 * reference. */
int wcsph_sim_query(wcsph_sim* s, int32_t* idle);
/* Re-bin and run one force pass on the current state (no integration):
 * acc (hydro only, no gravity), div_v and dt_i of the current positions --
 * CalcPressure and calcForceAllAndWriteBack of src/main.cpp:70-75. */
int wcsph_sim_force(wcsph_sim* s);
/* Wait for the simulation's queued work; WCSPH_EDIVERGED if the state has
 * diverged. No counterpart in the reference. */
int wcsph_sim_sync(wcsph_sim* s);
/* Clock, pair counts and bounds of the last pass; the time and step are the
 * reference's sysinfo (src/main.cpp:89-96). */
int wcsph_sim_get_stats(wcsph_sim* s, wcsph_sim_stats* st);
/* Download in the original particle order (index given at upload). Any
 * pointer may be NULL. acc is hydro + gravity after a step, hydro only after
 * wcsph_sim_force. Plays the part of reading FDPS's ParticleSystem. */
int wcsph_sim_download(wcsph_sim* s, double* pos, double* vel, double* dens, double* acc,
                       double* div_v, double* dt_i);
/* Neighbour counts (and lists of original indices, ascending, if lists is
 * not NULL) of the last force pass, in original order; lists is [n][cap].
 * The set the functor's support implies (src/force.h:49); FDPS exposes only
 * superset lists. */
int wcsph_sim_neighbours(wcsph_sim* s, int32_t* counts, int32_t* lists, int32_t cap);
/* FP64 reductions on the device: out[0] mass, out[1..3] momentum,
 * out[4] kinetic, out[5] potential, out[6] internal energy of the clamped
 * Tait law (zero below rho0),
 * out[7] dam front (max x of fluid with vertical coordinate < zfront),
 * out[8] column height (max vertical coordinate of fluid with x < xcol),
 * out[9] run-up (max vertical coordinate of fluid with x > xwall). */
/* No counterpart in the reference, which records neither conservation sums
 * nor dam-front measures. */
int wcsph_sim_diagnostics(wcsph_sim* s, double zfront, double xcol, double xwall, double* out);
/* The cell grid (for checking the neighbour predicate, which is defined
 * on cell-local coordinates): lo[3], cell edge *L, cells per axis nc[3].
 * No counterpart in the reference (FDPS builds a tree instead). */
int wcsph_sim_grid(wcsph_sim* s, double* lo, double* L, int32_t* nc);
/* Device time (ms) of the kernels of the last wcsph_sim_force call:
 * out[0] binning + reorder, out[1] separate neighbour pass (FP64 path;
 * 0 when fused), out[2] force pass. No counterpart in the reference. */
int wcsph_sim_last_force_ms(wcsph_sim* s, float* out);
/* Per-step HIP-event timing on the simulation's stream. While enabled, each
 * step records events around its stages; wcsph_sim_timing waits, returns the
 * summed device times (ms) since enabling -- out[0] kick-drift + binning +
 * reorder, out[1] neighbour pass, out[2] force pass, out[3] final stage,
 * out[4] whole steps -- with *nsteps timed, and restarts the sums. The
 * reference only prints the wall time of each step (src/main.cpp:94-96). */
int wcsph_sim_set_timing(wcsph_sim* s, int32_t enable);
int wcsph_sim_timing(wcsph_sim* s, double* out, int64_t* nsteps);

/* ---------------- 3. serving loop (continuous batching) ---------------- */
/* One dam-break request: its physics, geometry (the particle count follows
 * from geom), precision, number of steps, and arrival time in seconds after
 * the start of wcsph_serve. */
typedef struct {
    wcsph_params params;
    wcsph_dam_break_desc geom;
    int32_t precision;
    int32_t steps;
    double arrival_s;
} wcsph_request;

typedef struct {
    double arrival_s, start_s, end_s;   /* wall clock, seconds after the start */
    int64_t n, steps, pairs;            /* particles, steps run, pair interactions */
    double sim_time, front, column, kinetic;  /* state at completion */
    int32_t status;                     /* 0, or the WCSPH_E* code that ended it */
    int32_t precision_used;             /* WCSPH_PREC_*: a mixed request whose grid
                                           exceeds the mixed path's 2048 cells
                                           along x runs in FP32 */
} wcsph_request_result;

typedef struct {
    int64_t hbm_budget_bytes;  /* admission limit on the summed device footprint */
    int32_t max_slots;         /* requests running at once */
    int32_t chunk_steps;       /* steps queued per request per scheduling pass */
} wcsph_serve_config;

/* Run the whole request stream: admit arrived requests (FIFO) while a slot
 * is free and the budget allows (by each request's exact footprint,
 * wcsph_sim_bytes_state), retire finished ones and record per-request
 * timing (latency = end_s - arrival_s) in out[] (same order as reqs).
 * Device memory is one arena of hbm_budget_bytes allocated at entry;
 * requests take ranges of it, so nothing is allocated or freed on the
 * device while requests run. Mixed-precision requests without the Shepard
 * filter share grouped step kernels (one launch per stage of the step for
 * all such requests of one dimension and pair-term switches, replayed from
 * a two-step graph re-captured when the group changes); others step on a
 * stream of their own. max_slots is at most 16. Returns when every request
 * is done. No counterpart in the reference, which runs one simulation per
 * process. */
int wcsph_serve(const wcsph_serve_config* cfg, const wcsph_request* reqs, int32_t n, wcsph_request_result* out);

/* ---------------- 4. snapshots and restart ---------------- */
/* The reference's ASCII snapshot, as FDPS's writeParticleAscii writes it with
 * FileHeader::writeAscii (src/class.h:45-58) and RealPtcl::writeAscii
 * (src/class.h:129-135), called from OutputFileWithTimeInterval
 * (src/io.h:2-37): a line "%e" with the time, a line "%d" with N, then one
 * line per particle of 12 tab-separated columns, "%ld" for id and tag and
 * "%.16lf" for mass, x, y, z, vx, vy, vz, dens, snds, pres. In 2D z and vz
 * are written as 0 and snds is the sound speed; in 3D snds is written as 0.
 * pos/vel are [dim][n]. NULL id writes 0..n-1; NULL tag, snds or pres write 0. */
int wcsph_snapshot_write_ascii(const char* path, int32_t dim, int64_t n, double time, const int64_t* id,
                               const int64_t* tag, const double* mass, const double* pos, const double* vel,
                               const double* dens, const double* snds, const double* pres);
/* Reads that format back. With every array NULL it only returns *n and
 * *time; otherwise each non-NULL array must hold n entries ([dim][n] for
 * pos/vel). The reference's own reader (RealPtcl::readAscii,
 * src/class.h:136-142, used by src/io.h:39-46) hands "%.16lf" to fscanf,
 * which is not a scanf conversion; this reader parses the columns as
 * written. */
int wcsph_snapshot_read_ascii(const char* path, int32_t dim, int64_t* n, double* time, int64_t* id, int64_t* tag,
                              double* mass, double* pos, double* vel, double* dens, double* snds, double* pres);
/* The simulation's current state in that format, in the original particle
 * order: id = upload index, tag = type (0 fluid, 1 wall), pressure and sound
 * speed from the Tait law of the current density (CalcPressure,
 * src/force.h:3-9), time from the device clock. */
int wcsph_sim_write_ascii(wcsph_sim* s, const char* path);
/* Restart file: the full device state after wcsph_sim_init or a step (FP64
 * SoA in device order with the original indices, acc, div_v, dt_i, the cell
 * grid, the clock, the Shepard schedule, the next dt, params and config),
 * native byte order, with a checksum. Replaces the raw RealPtcl dump of
 * src/io.h:19-30 and InputBinary (src/io.h:60-84), which store EoS
 * pointers and so cannot be read by another process. */
int wcsph_sim_save(wcsph_sim* s, const char* path);
/* Create a simulation on `device` from a restart file. Stepping it continues
 * bit for bit as the saved simulation would have continued. Replaces
 * InputBinary (src/io.h:60-84). */
int wcsph_sim_restore(const char* path, int32_t device, wcsph_sim** out);
/* Set the clock after wcsph_sim_init (for a state read back from a
 * snapshot: the reference resumes at the file's time and step + 1,
 * src/io.h:60-84, and the Shepard schedule follows the step index). */
int wcsph_sim_set_clock(wcsph_sim* s, double time, int64_t step);
/* The parameters, configuration and particle count a simulation was created
 * with (for one made by wcsph_sim_restore); any pointer may be NULL. No
 * counterpart in the reference. */
int wcsph_sim_params(wcsph_sim* s, wcsph_params* p, wcsph_sim_config* cfg, int64_t* n);

/* ---------------- 5. one force pass on device-resident particles ---------------- */
/* The particles of a pass already live in device memory, owned by the
 * caller, in the caller's order. */
typedef struct {
    int64_t n;              /* particles */
    double box_lo[3];       /* cell-grid domain: particles outside it are clamped into */
    double box_hi[3];       /*   the edge cells (their neighbours are still all found) */
    double h_max;           /* the largest smoothing length in h[]: cells are h_max
                               wide (two per support 2 h_max); a larger h would
                               miss neighbours */
} wcsph_pass_desc;

/* One hydro force pass of CalcHydroForce (src/force.h:31-60) over all n
 * particles, as calcForceAllAndWriteBack runs it once per step
 * (src/main.cpp:75,118), with FDPS's tree and the dispatch/retrieve copies
 * replaced by a cell list built on the device: pressure and sound speed from
 * the Tait law of dens (CalcPressure, src/force.h:3-9), the general pair body
 * (per-pair mean smoothing length, so h may vary), the neighbour set
 * 0 < r^2 < (h_i + h_j)^2 of the device-resident path. Inputs pos/vel
 * [dim][n], dens, mass, h [n] and outputs acc [dim][n] (hydro only), div_v,
 * dt_i [n] are device pointers (FP64) in the caller's order; pairs, if not
 * NULL, is a device int64 that receives the pair count, or -1 when a
 * position was not finite or a cell held more than 256 particles (then the
 * outputs are not written). Queued on `stream` (a hipStream_t; NULL for
 * the default stream); returns without waiting. Workspace (about
 * 76 bytes per particle and 8 per cell) is owned by the library, one per
 * device, grown on demand; calls that overlap on one device race on it. */
int wcsph_force_pass(const wcsph_params* p, const wcsph_pass_desc* d, const double* pos, const double* vel,
                     const double* dens, const double* mass, const double* h, double* acc, double* div_v,
                     double* dt_i, int64_t* pairs, void* stream);

/* wcsph_force_pass without any allocation. No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under hipGraph
 * capture, tight serving loops, harnesses that forbid hipMalloc). Same
 * arguments, kernels and results as wcsph_force_pass, but the workspace is a
 * fixed device array of the library (about 56 MB, one copy per GPU, held
 * until the process ends), so it takes at most 524,288 particles and a grid
 * of at most 2,097,152 cells of h_max (WCSPH_EOVERFLOW beyond, before any
 * work is queued). Calls no hipMalloc, hipFree or other allocating routine.
 * All calls on one GPU share that GPU's array: calls that overlap on one GPU
 * race. */
int wcsph_force_pass_noalloc(const wcsph_params* p, const wcsph_pass_desc* d, const double* pos,
                             const double* vel, const double* dens, const double* mass, const double* h,
                             double* acc, double* div_v, double* dt_i, int64_t* pairs, void* stream);

#ifdef __cplusplus
}
#endif
#endif
