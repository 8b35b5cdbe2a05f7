/*
 * campfire_hip.h — C ABI of the MI355X gas solver (gfx950, HIP).
 *
 * This is the drop-in boundary for campfire's `FluidSim` (js/fluid.js:26-356).
 * Every entry point below replaces one method or field of that class; the
 * reference line each one stands in for is cited beside it.  The N-API addon
 * (campfire_amd/js/addon.cc) and the Python ctypes mirror (campfire_amd/fluid.py)
 * bind exactly these symbols.
 *
 * Conventions
 *  - Every `int` entry point returns CF_OK (0) or a negative CF_E* code and never
 *    throws across the ABI; cf_last_error() gives a thread-local message.
 *  - The caller owns host buffers; a cf_sim owns all of its device memory.
 *  - One HIP stream per cf_sim.  A cf_sim is not thread-safe.
 *  - Work is enqueued asynchronously on the sim's stream, like WebGPU command
 *    submission; functions that hand data back to the host synchronise.
 *  - Units are SI; grid coordinates are cell units, cell (i,j) centred at
 *    (i+0.5, j+0.5), U(i,j) at (i, j+0.5), V(i,j) at (i+0.5, j)
 *    (shaders/common.wgsl:1-2).
 */
#ifndef CAMPFIRE_HIP_H
#define CAMPFIRE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Exported constants (js/fluid.js:5-11). */
#define CF_MAX_PIECES   64
#define CF_PIECE_FLOATS 20
#define CF_ACC_GLOBAL   512
#define CF_ACC_LEN      (CF_ACC_GLOBAL + 16)
#define CF_E_SCALE      1e6
#define CF_M_SCALE      1e12
#define CF_G_SCALE      1e3
#define CF_N_FLOOR      64
#define CF_RAD_LEN      (CF_MAX_PIECES * 4 + CF_N_FLOOR)
#define CF_W_SCALE      1e8 /* faceW fixed point (shaders/geometry.wgsl:18) */
#define CF_NPARAMS      40

/* Status codes. */
#define CF_OK          0
#define CF_EINVAL     -1  /* bad argument / unknown name */
#define CF_EBADSIZE   -2  /* 'coarsest level too large' (js/fluid.js:91) */
#define CF_ENOMEM     -3  /* device allocation failed or over the HBM budget */
#define CF_EDEVICE    -4  /* a HIP runtime call failed */
#define CF_EAGAIN     -5  /* readback ring full / result not ready (js/fluid.js:317) */
#define CF_ENOTREADY  -6  /* readback slot not complete yet */
#define CF_ENOCONV    -7  /* multigrid did not reach its residual target */
#define CF_ENONFINITE -8  /* NaN/Inf detected in the fields */

/*
 * The 160-byte parameter block, field order of PARAM_FIELDS (js/fluid.js:13-24)
 * and of the WGSL `Params` uniform (shaders/common.wgsl:4-15).
 */
typedef struct cf_params {
  int32_t nx, ny; float dx, dt;
  float Tamb, rhoAmb, YOamb, grav;
  float cp, Pr, viscMult, diffMult;
  float reactB, reactTa, dHc, sO2;
  float sootFormA, sootFormTa, sootOxA, sootOxTa;
  float kappaSootC, kappaGas, exchAlpha, depth;
  float charA, charTa; int32_t nPieces; uint32_t frame;
  float heaterX0, heaterX1, heaterY1, heaterT;
  float burnerX0, burnerX1, burnerFlux, burnerT;
  float sootDensity, sigma, charO2, time;
} cf_params;

/*
 * Solver schedule.  The reference runs exactly one V-cycle per step with
 * nu = 2 red-black sweeps per level and 120 red-black sweeps on the coarsest
 * level (js/fluid.js:250,294-312; shaders/pressure.wgsl:141), 8 chemistry
 * substeps (shaders/sources.wgsl:110).  Those are the defaults.
 */
typedef struct cf_schedule {
  int32_t nu;            /* smoother sweeps (each = red + black) per level, ref 2 */
  int32_t vcycles;       /* V-cycles per step (max when tol > 0), ref 1 */
  float   tol;           /* relative residual target ||r||/||rhs||; 0 = run exactly `vcycles` */
  int32_t coarse_iters;  /* red-black sweeps on the coarsest level, ref 120;
                            0 = direct FP64 solve of the coarsest level */
  int32_t smoother_fp16; /* 1 = FP16 smoother: the multigrid levels above the
                            single-workgroup bottom solver keep p (scaled by a
                            power of two from the defect's rms) and the face
                            coefficients in FP16; sweeps, residual and
                            restriction compute in FP32.  Needs refine = 1. */
  float   cfl;           /* > 0: each step takes dt = cfl * dx / max(max|u|, umin),
                            computed on the device from the current U, V; 0 = fixed dt */
  float   umin;          /* velocity floor of the CFL rule (m/s) */
  int32_t refine;        /* 1 = defect correction at level 0: the pressure and the
                            residual b - A p in FP64, each V-cycle (FP32) solving
                            for the correction; the projection takes its gradient
                            from the FP64 pressure.  Needs tol > 0. */
  int32_t krylov;        /* 1 = flexible conjugate gradients at level 0 (FCG with
                            one stored direction), each V-cycle preconditioning
                            one iteration: the pressure, residual, direction and
                            its image under A in FP64.  Needs refine = 1. */
} cf_schedule;

typedef struct cf_sim_info {
  int32_t nx, ny;
  double  dx, dt, height, time;
  uint32_t frame;
  int32_t diff_sub;      /* diffusion substeps (js/fluid.js:190-198) */
  int32_t nlevels;       /* multigrid levels (js/fluid.js:69-91, without the 6-level cap) */
  int32_t level_nx[16], level_ny[16];
  int32_t last_vcycles;  /* V-cycles run by the last step */
  float   last_resid;    /* relative residual after the last step's projection (if measured) */
} cf_sim_info;

typedef struct cf_sim cf_sim;

/* ---- lifetime ---------------------------------------------------------- */

/* new FluidSim(device, {nx, ny, dt}) + init()  (js/fluid.js:27-47,49-160).
 * dx is explicit (the reference derives it as SCENE_WIDTH/nx, js/fluid.js:30).
 * Levels halve while both dims are even and cells > 256; an odd coarsest
 * level above 256 cells fails with CF_EBADSIZE as js/fluid.js:91 does.
 * Allocates every buffer of js/fluid.js:50-91 on the current HIP device and
 * calls cf_sim_reset(). */
int  cf_sim_create(int nx, int ny, double dx, double dt, cf_sim** out);
void cf_sim_destroy(cf_sim* sim);

/* reset()  (js/fluid.js:162-176). */
int  cf_sim_reset(cf_sim* sim);

/* ---- parameters -------------------------------------------------------- */

/* setParam(key, value)  (js/fluid.js:178).  Keys are the PARAM_FIELDS names.
 * "time" and "frame" set the sim's own counters (the app does sim.frame++,
 * js/main.js:282).  Takes effect at the next cf_sim_write_params(). */
int  cf_sim_set_param(cf_sim* sim, const char* key, double value);
int  cf_sim_get_param(const cf_sim* sim, const char* key, double* value);
/* writeParams()  (js/fluid.js:180-199): packs the 160-B block (nx, ny, dx,
 * dt, time, frame owned by the sim) and computes the diffusion substeps. */
int  cf_sim_write_params(cf_sim* sim);
/* The packed block as last written (for bit-level checks). */
int  cf_sim_get_packed_params(const cf_sim* sim, cf_params* out);
/* PARAM_FIELDS name of field `index` (0..39), or NULL. */
const char* cf_param_name(int index);

/* The solver schedule (no counterpart: the reference's is fixed).  A scene of
 * a pool takes the pool's schedule (cf_pool_set_schedule); setting another
 * one on it fails with CF_EINVAL. */
int  cf_sim_set_schedule(cf_sim* sim, const cf_schedule* s);
int  cf_sim_get_schedule(const cf_sim* sim, cf_schedule* s);

/* Tile edge of the tiled multigrid legs per level (levels 0 .. n-1 of those
 * above the single-workgroup bottom solver): 16, 32 or 64 (square tiles of
 * that edge), 48 (the wide tiles: 64 cells wide and 48 rows for the down
 * legs at nu <= 2, else 32, so that two workgroups share a CU), or 0 for the
 * default size rule.  Any choice gives bitwise the same results (a tile with
 * a halo of 2 nu + 1 cells reproduces the global half-sweeps exactly); it
 * changes only the launch shape, which the autotuner picks by timing.  No
 * counterpart in the reference (its relax dispatches are 8x8 workgroups). */
int  cf_sim_set_leg_tiles(cf_sim* sim, const int32_t* tiles, int n);
/* The tile edges in use for levels 0 .. n-1 (0 past the tiled levels, and 0
 * for a level whose legs run in a two-level pair: with FP32 storage a lone
 * scene runs the legs of levels 1-2, 3-4, ... two per launch, with tiles of
 * their own); returns the number of tiled levels. */
int  cf_sim_get_leg_tiles(const cf_sim* sim, int32_t* tiles, int n);

/* uploadPieces(Float32Array(64*20), n)  (js/fluid.js:201-204); also sets
 * nPieces.  Layout per piece (shaders/common.wgsl:17-23):
 * [x, y, angle, half, vx, vy, omega, alive, Ts[4], mflux[4], charCov[4]]. */
int  cf_sim_upload_pieces(cf_sim* sim, const float* pieces, int n);

/* ---- work -------------------------------------------------------------- */

/* clearGeometry(encoder)  (js/fluid.js:213): faceW = 0. */
int  cf_sim_clear_geometry(cf_sim* sim);
/* encodeGeometry(pass)  (js/fluid.js:215-218): cellG, fx, fy, faceW. */
int  cf_sim_geometry(cf_sim* sim);
/* nsteps x encodeStep(pass)  (js/fluid.js:225-253); time += dt each step.
 * With a residual target (schedule tol > 0) each step waits on the host for
 * the convergence flag between its V-cycles, so the call returns after the
 * last step's V-cycles have run (the projection may still be in flight). */
int  cf_sim_step(cf_sim* sim, int nsteps);
/* encodeRadiation(pass)  (js/fluid.js:220-222). */
int  cf_sim_radiation(cf_sim* sim);
/* One app frame as js/main.js:266-283 encodes it: write_params, clear
 * geometry, geometry, nsteps steps, radiation (if with_radiation). */
int  cf_sim_frame(cf_sim* sim, int nsteps, int with_radiation);

/* Run one of the seven profiled stages of encodeStepProfiled
 * (js/fluid.js:256-292): 0 advect, 1 diffuse, 2 sources, 3 forces,
 * 4 coef/rhs, 5 vcycle, 6 project.  Running 0..6 in order equals one step
 * (time is not advanced).  Stage 7 runs stages 3 and 4 as the production step
 * does, in one fused pass.  Stage 0 advects in one pass and leaves the
 * forward fields Th, Yh, Uh, Vh unwritten; stage 8 runs the reference's two
 * dispatch pairs instead and writes them. */
int  cf_sim_run_stage(cf_sim* sim, int stage);
/* encodeStepProfiled: runs nsteps steps with HIP events between the seven
 * stages and returns the summed milliseconds per stage in stage_ms[7]. */
int  cf_sim_profile_steps(cf_sim* sim, int nsteps, float* stage_ms);
/* Per-kernel timing of nsteps production steps (events after every launch):
 * up to `max` distinct kernel names (32 chars each, in `names`), summed
 * milliseconds and launch counts; *count receives the number filled. */
int  cf_sim_profile_kernels(cf_sim* sim, int nsteps, int max, char* names,
                            float* ms, int* launches, int* count);

/* Device time of the per-frame work (the measurement counterpart of
 * clearGeometry + encodeGeometry and of encodeRadiation, js/fluid.js:213-222):
 * `reps` frames' geometry and radiation (which also advances the stateful
 * smoothing of rad), the mean milliseconds of each. */
int  cf_sim_profile_frame(cf_sim* sim, int reps, float* geometry_ms, float* radiation_ms);

/* Per-step latency: nsteps steps, each bracketed by HIP events on the sim's
 * stream; ms[q] receives the device time of step q. */
int  cf_sim_step_latencies(cf_sim* sim, int nsteps, float* ms);

/* Wait for all enqueued work. */
int  cf_sim_synchronize(cf_sim* sim);

/* Status of the steps run so far; waits for them.  CF_ENONFINITE when a NaN
 * or Inf entered the sources pass (T, Y) or the CFL pass (U, V) since the
 * last call (the flag is then cleared); CF_ENOCONV when the schedule has a
 * residual target (tol > 0) and the last step ended above it; else CF_OK.
 * The reference has no counterpart: WebGPU reports only device loss and
 * validation errors (js/main.js:166-167), so a diverged field goes unnoticed
 * there. */
int  cf_sim_status(cf_sim* sim);

/* ---- readback ---------------------------------------------------------- */

/* encodeReadback(encoder)  (js/fluid.js:315-324): enqueue a copy of acc
 * (528 i32), rad (320 x vec4f) and faceW (256 i32) into the next slot of a
 * 3-deep pinned ring and clear acc.  Returns the slot (0..2) or CF_EAGAIN
 * when that slot is still held (the reference then skips the readback and
 * acc keeps accumulating). */
int  cf_sim_encode_readback(cf_sim* sim);
/* afterSubmit's completion  (js/fluid.js:327-342): block until `slot` is
 * complete (wait != 0) or test it (wait == 0, CF_ENOTREADY if pending), copy it
 * out and release the slot.  Any output pointer may be NULL. */
int  cf_sim_take_readback(cf_sim* sim, int slot, int wait,
                          int32_t* acc, float* rad, int32_t* faceW);

/* readBuffer(name)  (js/fluid.js:345-355) and the inverse (used by
 * H.blob, dev/gas_test.html:112-120).  Names: U Uh U2 V Vh V2 T Th T2 Y Yh Y2
 * SQ cellG fx fy pieces acc rad faceW and the level-0 MG arrays p rhs ax ay;
 * "p1".."p15", "rhs1".. for coarse levels, and "p64", the FP64 level-0
 * pressure of defect correction (8 B/cell).  Y* are vec4 AoS (16 B/cell), SQ
 * and cellG vec2 (8 B/cell), exactly the reference layouts.  `bytes` must
 * equal the buffer size (cf_sim_field_bytes). */
int  cf_sim_field_bytes(const cf_sim* sim, const char* name, size_t* bytes);
int  cf_sim_read_field(cf_sim* sim, const char* name, void* dst, size_t bytes);
int  cf_sim_write_field(cf_sim* sim, const char* name, const void* src, size_t bytes);

/* Fields nx, ny, dx, dt, height, time, frame, diffSub, levels of FluidSim. */
int  cf_sim_get_info(const cf_sim* sim, cf_sim_info* info);

/* Max |div(f u) - S| and the mean, after a projection (H.divErr,
 * dev/gas_test.html:101-111), computed on the device. */
int  cf_sim_div_error(cf_sim* sim, float* max_err, float* mean_err);

/* ---- scene pool: many scenes stepped together ------------------------------
 * No counterpart in the reference (one FluidSim per page).  Scenes of any
 * sizes live in one pool; each phase of a step is ONE launch over the tiles of
 * every live scene, captured once as a graph.  A scene's results are the same
 * alone and in any mix.  Scenes are admitted while their device memory fits
 * the pool's HBM budget; each is a cf_sim (borrowed through cf_pool_scene) for
 * parameters, pieces, fields and readback, on the pool's stream.  All scenes
 * use the pool's schedule.  A pool is not thread-safe. */
typedef struct cf_pool cf_pool;
typedef struct cf_scene_desc {
  int32_t nx, ny;
  double dx, dt;
} cf_scene_desc;
typedef struct cf_pool_info {
  int32_t nscenes, nslots;
  uint64_t budget, used;   /* HBM bytes */
  uint64_t cells;          /* cells of the live scenes */
  int64_t steps;           /* pool steps taken */
  int32_t launches_per_step;
  int32_t pad_;
} cf_pool_info;
/* Device bytes a scene of this grid needs (0: not a supported grid). */
uint64_t cf_scene_bytes(int nx, int ny, double dx);
int  cf_pool_create(uint64_t hbm_budget, cf_pool** out);
void cf_pool_destroy(cf_pool* pool);
/* The schedule of every scene (cf_sim_set_schedule rules).  Every scene is
 * checked before any changes: on an error no scene's schedule has changed. */
int  cf_pool_set_schedule(cf_pool* pool, const cf_schedule* sc);
/* Leg tile code (16, 32, 48 or 64, as cf_sim_set_leg_tiles) of each depth of the tiled multigrid levels,
 * for the legs of every scene at that depth; 0 (the default) takes the size
 * rule applied to the tile count of the whole depth.  Results do not depend
 * on tiles, only speed.  Takes effect at the next rebuild of the step. */
int  cf_pool_set_leg_tiles(cf_pool* pool, const int32_t* tiles, int n);
/* The tile of each depth for the live scenes (n entries, 0 beyond the
 * deepest tiled level); returns the number of tiled depths. */
int  cf_pool_get_leg_tiles(const cf_pool* pool, int32_t* tiles, int n);
/* A new scene in the first free slot; CF_EAGAIN when it does not fit the budget. */
int  cf_pool_admit(cf_pool* pool, const cf_scene_desc* desc, int* slot);
/* The scene of a slot (owned by the pool: do not cf_sim_destroy it). */
int  cf_pool_scene(cf_pool* pool, int slot, cf_sim** sim);
/* Frees a slot and its device memory. */
int  cf_pool_release(cf_pool* pool, int slot);
/* Builds the batched step of the live scenes now (tables and step graphs),
 * which cf_pool_step otherwise does on its first call after an admission or
 * a release: a serving loop calls it right after changing the scene set, so
 * the rebuild is not on the next step's latency.  Waits for the pool's
 * stream. */
int  cf_pool_prepare(cf_pool* pool);
/* Every live scene advances nsteps steps (asynchronous; with a residual
 * target the host waits between V-cycles as in cf_sim_step). */
int  cf_pool_step(cf_pool* pool, int nsteps);
/* One app frame for every scene: parameters, geometry, nsteps steps, radiation. */
int  cf_pool_frame(cf_pool* pool, int nsteps, int with_radiation);
int  cf_pool_synchronize(cf_pool* pool);
int  cf_pool_get_info(const cf_pool* pool, cf_pool_info* info);
/* cf_sim_status over every scene (each scene's flag is read and cleared):
 * the first CF_ENONFINITE scene, else the first CF_ENOCONV scene; its slot
 * in *slot (may be NULL; -1 when all are CF_OK). */
int  cf_pool_status(cf_pool* pool, int* slot);
/* Device time of each of nsteps pool steps (ms), events around every step. */
int  cf_pool_step_latencies(cf_pool* pool, int nsteps, float* ms);

/* ---- 1D premixed flame (tests/flame1d.mjs:13-71) ---------------------------
 * flameSpeed({B, Ta, dx, F, phi, L, tEnd}) on the GPU, one workgroup per case,
 * all cases in one launch.  The strip has n = Math.round(L/dx) cells (at most
 * CF_FLAME_MAXN), the first Math.round(0.01/dx) ignited at Tign. */
#define CF_FLAME_MAXN 1536
typedef struct cf_flame_case {
  double B, Ta, dx, F, phi, L, tEnd;  /* the script's flameSpeed arguments (:13) */
  double Tamb, Tign;                  /* 293.15 K and 2000 K in the script (:6, :21) */
  int32_t max_steps;                  /* cap on the device time loop; 0 = 10 million */
  int32_t pad_;
} cf_flame_case;
typedef struct cf_flame_result {
  double SL, Tb;      /* { SL, Tb } of flameSpeed (:68-70) */
  int32_t steps, n;   /* time steps taken, cells */
} cf_flame_result;
/* fp64 = 1 runs the script's double arithmetic, 0 the single-precision form.
 * Synchronous; returns CF_EBADSIZE when a case needs more than CF_FLAME_MAXN
 * or fewer than 2 cells. */
int  cf_flame1d_run(const cf_flame_case* cases, int ncases, int fp64, cf_flame_result* out);

/* cf_flame1d_run without any allocation.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (serving loops and
 * harnesses that forbid hipMalloc).  The device copies of the cases and the
 * results live in a static device array of the library (about 26 KB per GPU,
 * held until the process ends) instead of per-call hipMalloc/hipFree, and the
 * work runs on the caller's stream (a hipStream_t; NULL: the null stream),
 * waited for before returning.  At most CF_FLAME_STATIC_MAX cases per call
 * (more: CF_EINVAL).  Calls that overlap on one GPU share that array and race. */
#define CF_FLAME_STATIC_MAX 256
int  cf_flame1d_run_static(const cf_flame_case* cases, int ncases, int fp64, cf_flame_result* out,
                           void* stream);

/* Divergence compatibility, summed on the device in FP64: the net source of
 * the open cells (sum S dx^2, m^2/s) and the net outflow through the open
 * domain faces (dx sum f u.n).  Equal after a converged projection.  No
 * counterpart in the reference. */
int  cf_sim_compat_sums(cf_sim* sim, double* source, double* outflow);

/* Copies every device buffer and parameter of src into dst (same grid): a
 * snapshot, or a restore of one.  No counterpart in the reference. */
int  cf_sim_copy_state(cf_sim* dst, const cf_sim* src);

/* The bandwidth reference of the measurements: a plain streaming copy kernel
 * of this library moving `bytes` (read + write: 2 x bytes per launch), best of
 * `reps` launches, as GB/s (*gbps) and the mean launch time (*ms, may be NULL).
 * Allocates its two buffers per call.  No counterpart in the reference. */
int  cf_copy_bandwidth(uint64_t bytes, int reps, double* gbps, double* ms);

/* Thread-local text of the last error. */
const char* cf_last_error(void);
/* Library version string. */
const char* cf_version(void);

#ifdef __cplusplus
}
#endif
#endif /* CAMPFIRE_HIP_H */
