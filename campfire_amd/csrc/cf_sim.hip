// cf_sim.hip — host side of the C ABI (include/campfire_hip.h).
//
// Owns one scene: every device buffer in one allocation, one HIP stream,
// pinned upload rings for the parameter block and the pieces, the 3-deep
// pinned readback ring, and the step as a captured hipGraph that is replayed
// once per step.  The schedule of encodeStep (js/fluid.js:225-253) and of the
// V-cycle (js/fluid.js:294-312) is kept; consecutive reference dispatches
// that are pointwise compositions of each other run as one kernel.
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <string>
#include <array>
#include <algorithm>
#include <vector>

#include "cf_kernels.h"

using cf::Fields;
using cf::Level;

#define MAXLEV 16
#define UP_SLOTS 4

static thread_local char g_err[512] = "";

static int set_err(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

#define HIPCHK(call)                                                                       \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess)                                                                  \
      return set_err(CF_EDEVICE, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), __FILE__, __LINE__); \
  } while (0)

// PARAM_FIELDS (js/fluid.js:13-24)
static const char* PNAME[CF_NPARAMS] = {
    "nx", "ny", "dx", "dt", "Tamb", "rhoAmb", "YOamb", "grav",
    "cp", "Pr", "viscMult", "diffMult", "reactB", "reactTa", "dHc", "sO2",
    "sootFormA", "sootFormTa", "sootOxA", "sootOxTa",
    "kappaSootC", "kappaGas", "exchAlpha", "depth",
    "charA", "charTa", "nPieces", "frame",
    "heaterX0", "heaterX1", "heaterY1", "heaterT",
    "burnerX0", "burnerX1", "burnerFlux", "burnerT",
    "sootDensity", "sigma", "charO2", "time"};
static const char PKIND[CF_NPARAMS] = {
    'i', 'i', 'f', 'f', 'f', 'f', 'f', 'f', 'f', 'f', 'f', 'f', 'f', 'f', 'f', 'f',
    'f', 'f', 'f', 'f', 'f', 'f', 'f', 'f', 'f', 'f', 'i', 'u',
    'f', 'f', 'f', 'f', 'f', 'f', 'f', 'f', 'f', 'f', 'f', 'f'};
static_assert(sizeof(cf_params) == 160, "cf_params must be the 160-B uniform");

// Physical constants used by the host (js/config.js:12-47)
static const double PHYS_Tamb = 293.15;
static double phys_rhoAmb() { return 101325.0 * 0.0289 / (8.314462 * PHYS_Tamb); }

struct ProfTag {
  const char* name;
  double ms;
  int launches;
};

// one parameter upload: the 160-B block and the first two fields of Dyn
struct ParamUpload {
  cf_params P;
  float dsub, dt;
};

// graphs of a residual-controlled step in segments (see run_segments)
static const int SEG_HEAD = 4, SEG_CHUNK = 2;
struct Segments {
  hipGraphExec_t head = nullptr, chunk = nullptr, rem = nullptr, tail = nullptr;
  int h = 0, r = 0;  // cycles in the head and in the remainder chunk
  void drop() {
    for (hipGraphExec_t* g : {&head, &chunk, &rem, &tail})
      if (*g) { hipGraphExecDestroy(*g); *g = nullptr; }
  }
};

struct cf_sim {
  int dev = 0;
  hipStream_t st = nullptr;
  bool own_stream = true;  // false for the scenes of a pool (the pool's stream)
  bool in_pool = false;
  int nx = 0, ny = 0, nC = 0, nU = 0, nV = 0;
  double dx = 0, dt = 0, time = 0;
  uint32_t frame = 0;
  double pv[CF_NPARAMS];
  cf_params P;
  float dsub = 0;
  int diffSub = 1;
  int nlev = 0, lnx[MAXLEV], lny[MAXLEV];
  float lh2[MAXLEV];
  cf_schedule sch;
  void* block = nullptr;
  size_t block_bytes = 0;
  Fields F;
  Level L[MAXLEV];
  cf::LevelH LH[MAXLEV] = {};  // FP16 smoother: p, ax, ay of the tiled levels (below mg_bottom)
  cf_params* dP = nullptr;
  cf::Dyn* dDyn = nullptr;   // control values, right after the parameter block
  double* dNorm = nullptr;   // residual-norm partials
  double* dP64 = nullptr;    // level-0 pressure in FP64 (defect correction)
  float* dB0 = nullptr;      // level-0 right-hand side kept across correction cycles
  float* dD32 = nullptr;     // Krylov: the stored direction
  double* dBand = nullptr;   // direct coarsest solve: Cholesky band
  // upload rings
  ParamUpload* hP = nullptr;
  float* hPieces = nullptr;
  hipEvent_t evP[UP_SLOTS], evPc[UP_SLOTS];
  bool evP_used[UP_SLOTS], evPc_used[UP_SLOTS];
  int upP = 0, upPc = 0;
  // readback ring (js/fluid.js:157-159)
  uint8_t* ring = nullptr;
  hipEvent_t ringEv[3];
  bool busy[3];
  int stageIdx = 0;
  // diagnostics
  float* dPart = nullptr;
  std::vector<float> hPart;
  int* dAccPart = nullptr;  // per-workgroup partial sums of the sources kernel
  // graph of one step
  hipGraphExec_t gexec = nullptr;   // one step
  hipGraphExec_t gexec8 = nullptr;  // eight steps: one replay per 8 steps amortises the launch
  Segments seg;                     // residual-controlled steps: the step in segments
  int* hflag = nullptr;             // pinned host copy of the convergence flag
  int gkey[12] = {-1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1, -1};
  uint32_t tile_gen = 0;  // bumped when a level's leg tile changes (the graphs bake the tiles in)
  bool use_graph = true;
  bool mg_simple = false;  // one launch per reference dispatch pair (debug comparisons)
  int mg_bottom = 0;       // first level solved by the single-workgroup bottom kernel
  // two-level legs (FP32 storage, not in pools): the levels from 1 run as
  // pairs, the bottom kernel from mg_bottom_pairs on; CF_LEG_PAIRS=0 turns
  // them off, CF_PAIR_TILE=8|16 fixes their tile
  bool leg_pairs = true;
  int pair_tile = 0;
  // the projection inside the last V-cycle's level-0 up leg (CF_UP_PROJECT=0: a launch of its own)
  bool up_project = true;
  int mg_bottom_pairs = 0;
  // profiling
  bool profiling = false;
  std::vector<cf::LaunchEvents> plaunch;  // one per launch of a profiling pass
  std::vector<int> ptag;                  // its tag
  std::vector<ProfTag> tags;
  int last_vcycles = 0;
  float last_resid = -1;
};

static const size_t RING_BYTES = CF_ACC_LEN * 4 + CF_RAD_LEN * 16 + CF_MAX_PIECES * 4 * 4;

// ---- profiling of launch sites -------------------------------------------------
namespace cf {
thread_local std::vector<LaunchEvents>* prof_sink = nullptr;
}

// names of per-level launch sites ("mg_down.0" ...), static for the tag table
static const char* level_tag(const char* base, int l) {
  static char names[6][16][24];
  static const char* bases[6] = {"mg_down", "mg_up", "coef_restrict", "relax", "mg_down2", "mg_up2"};
  for (int b = 0; b < 6; b++)
    if (!strcmp(base, bases[b]) && l >= 0 && l < 16) {
      if (!names[b][l][0]) snprintf(names[b][l], sizeof(names[b][l]), "%s.%d", base, l);
      return names[b][l];
    }
  return base;
}

// attributes every launch since the previous mark to `tag`
static void mark(cf_sim* s, const char* tag) {
  if (!s->profiling) return;
  int id = -1;
  for (size_t q = 0; q < s->tags.size(); q++)
    if (!strcmp(s->tags[q].name, tag)) { id = (int)q; break; }
  if (id < 0) { s->tags.push_back({tag, 0.0, 0}); id = (int)s->tags.size() - 1; }
  while (s->ptag.size() < s->plaunch.size()) s->ptag.push_back(id);
}

// ---- step schedule ---------------------------------------------------------
// one V-cycle (js/fluid.js:294-312); coarse_coef: the bottom kernel also
// restricts the coarsest level's coefficients; skip: see launch_mg_down
// storage of the levels a tiled leg at level l touches: FP32, or with the
// FP16 smoother FP16 p and coefficients at every tiled level (the bottom
// levels stay FP32)
static cf::MgStore mg_store(const cf_sim* s, int l) {
  cf::MgStore m{};
  if (!s->sch.smoother_fp16) return m;
  const bool ch = l + 1 < s->mg_bottom;
  m.mode = 1 | (ch ? 2 : 0);
  m.f = s->LH[l];
  if (ch) m.c = s->LH[l + 1];
  m.dyn = s->dDyn;
  return m;
}

// whether a lone scene's V-cycle runs the two-level legs: FP32 storage only
static bool use_pairs(const cf_sim* s) { return s->leg_pairs && !s->mg_simple && !s->sch.smoother_fp16; }
// the first level of the bottom kernel in a lone scene's V-cycle
static int vcycle_bottom(const cf_sim* s) { return use_pairs(s) ? s->mg_bottom_pairs : s->mg_bottom; }
// whether level l's legs run in a pair with level l + 1 (first of the pair)
// or l - 1 (second): level 0 alone, then pairs while two tiled levels remain
static bool pair_first(const cf_sim* s, int l) {
  if (!use_pairs(s) || l < 1) return false;
  return (l - 1) % 2 == 0 && l + 1 < vcycle_bottom(s);
}
static bool paired(const cf_sim* s, int l) { return pair_first(s, l) || (l >= 2 && pair_first(s, l - 1)); }

// whether a Krylov iteration runs kry_update and dc_defect as one pass
// (kry_upd_def: the new direction and pressure go to the other half of d32
// and p64, and kry_settle brings them back after the step's last iteration):
// the fused path with a tiled level 0, whose down leg then starts from zero
// itself, since the fused pass cannot clear the correction that it reads
static bool kry_fused(const cf_sim* s) {
  return s->sch.tol > 0 && s->sch.refine && s->sch.krylov && !s->mg_simple && s->mg_bottom >= 1;
}

// whether the step's projection runs inside the last V-cycle's level-0 up
// leg: a fixed cycle count (no residual target, so p, not p64, is projected),
// FP32 storage, the fused path
static bool fuse_project(const cf_sim* s) {
  return !s->mg_simple && s->up_project && s->sch.tol <= 0 && !s->sch.smoother_fp16 && s->mg_bottom >= 1;
}

static void enqueue_vcycle(cf_sim* s, int nu, const int* skip, bool coarse_coef, bool project = false) {
  const int L = s->nlev;
  if (!s->mg_simple) {
    // fused: tiled down legs (one or two levels per launch), one LDS-resident
    // workgroup for the bottom levels, tiled up legs
    const int b = vcycle_bottom(s);
    for (int l = 0; l < b;) {
      if (pair_first(s, l)) {
        cf::launch_mg_down2(s->L[l], s->L[l + 1], s->L[l + 2], nu, s->pair_tile, skip, s->st);
        mark(s, level_tag("mg_down2", l));
        l += 2;
      } else {
        // a level above 0 holds the zero that the restriction into it wrote, and
        // the fused Krylov pass leaves level 0's correction to start from zero
        cf::launch_mg_down(s->L[l], s->L[l + 1], mg_store(s, l), nu, skip, s->st, l >= 1 || kry_fused(s) ? 1 : 0);
        mark(s, level_tag("mg_down", l));
        l += 1;
      }
    }
    cf::launch_mg_bottom(s->L, L, b, coarse_coef, nu, s->sch.coarse_iters, skip, s->dBand, s->st); mark(s, "mg_bottom");
    for (int l = b - 1; l >= 0;) {
      if (l >= 2 && pair_first(s, l - 1)) {
        cf::launch_mg_up2(s->L[l - 1], s->L[l], s->L[l + 1], nu, s->pair_tile, skip, s->st);
        mark(s, level_tag("mg_up2", l - 1));
        l -= 2;
      } else if (l == 0 && project) {
        if (cf::launch_mg_up_project(s->L[0], s->L[1], s->F, nu, skip, s->st)) {
          mark(s, "mg_up_project.0");
        } else {  // no such instantiation: the leg, then the projection
          cf::launch_mg_up(s->L[0], s->L[1], mg_store(s, 0), nu, skip, s->st); mark(s, level_tag("mg_up", 0));
          cf::launch_project(s->F, nullptr, s->st); mark(s, "project");
        }
        l -= 1;
      } else {
        cf::launch_mg_up(s->L[l], s->L[l + 1], mg_store(s, l), nu, skip, s->st); mark(s, level_tag("mg_up", l));
        l -= 1;
      }
    }
    return;
  }
  for (int l = 0; l < L - 1; l++) {
    for (int q = 0; q < nu; q++) {
      cf::launch_relax(s->L[l], 0, s->st); mark(s, "relax");
      cf::launch_relax(s->L[l], 1, s->st); mark(s, "relax");
    }
    cf::launch_restrict_res(s->L[l], s->L[l + 1], s->st); mark(s, "restrict_res");
  }
  cf::launch_coarsest(s->L[L - 1], s->sch.coarse_iters, s->st); mark(s, "coarsest");
  for (int l = L - 2; l >= 0; l--) {
    cf::launch_prolong(s->L[l], s->L[l + 1], s->st); mark(s, "prolong");
    for (int q = 0; q < nu; q++) {
      cf::launch_relax(s->L[l], 1, s->st); mark(s, "relax");
      cf::launch_relax(s->L[l], 0, s->st); mark(s, "relax");
    }
  }
}

// The V-cycles of one step (js/fluid.js:250).  With a residual target
// (tol > 0) up to `vcycles` cycles are enqueued, each followed by the
// relative residual norm; once the target is met the remaining launches
// return at once, so the sequence is fixed and a graph replays it.  (The
// per-dispatch path of CF_MG_SIMPLE has no skip and runs every cycle.)
// enqueue_mg_pre: what precedes the cycles (the FP16 coefficient copies,
// the control reset, the first defect); enqueue_mg_cycle: cycle c and the
// update and residual evaluation that follow it.
static __half* mg_ph0(cf_sim* s) { return s->sch.smoother_fp16 ? s->LH[0].p : nullptr; }
static void enqueue_mg_pre(cf_sim* s) {
  const bool tol = s->sch.tol > 0;
  const bool dc = tol && s->sch.refine;  // defect correction with the FP64 pressure
  // the FP16 smoother (defect correction only) reads FP16 copies of the coefficients
  __half* ph0 = mg_ph0(s);
  if (ph0) for (int l = 0; l < s->mg_bottom; l++) { cf::launch_coef_half(s->L[l], s->LH[l], s->st); mark(s, "coef_half"); }
  if (tol) { cf::launch_fill_u32(&s->dDyn->converged, 0u, 2, s->st); mark(s, "mg_control"); }  // converged, cycles
  if (dc && s->sch.krylov) {
    const cf::KryBufs kb{s->dP64, s->dD32, s->dB0};
    cf::launch_kry_start(s->L[0], ph0, s->dDyn, kb, s->dNorm, s->sch.tol, s->st);
    mark(s, "kry_start");
  } else if (dc) {
    cf::launch_dc_defect(s->L[0], ph0, s->dP64, s->dB0, s->dDyn, s->dNorm, s->sch.tol, true, s->st);
    mark(s, "dc_defect");
  }
}
static void enqueue_mg_cycle(cf_sim* s, int c, bool coarse_coef, bool project = false) {
  const bool tol = s->sch.tol > 0;
  const bool dc = tol && s->sch.refine;
  const bool kry = dc && s->sch.krylov;
  const int* skip = tol ? &s->dDyn->converged : nullptr;
  __half* ph0 = mg_ph0(s);
  const cf::KryBufs kb{s->dP64, s->dD32, s->dB0};
  enqueue_vcycle(s, s->sch.nu, skip, coarse_coef && c == 0, project);
  if (kry) {
    cf::launch_kry_step(s->L[0], ph0, s->dDyn, kb, s->dNorm, s->sch.tol, kry_fused(s), s->st); mark(s, "kry_step");
  } else if (dc) {
    cf::launch_dc_accum(s->L[0], ph0, s->dDyn, s->dP64, skip, s->st); mark(s, "dc_accum");
    cf::launch_dc_defect(s->L[0], ph0, s->dP64, s->dB0, s->dDyn, s->dNorm, s->sch.tol, false, s->st);
    mark(s, "dc_defect");
  } else if (tol) {
    cf::launch_resnorm(s->L[0], s->dDyn, s->dNorm, s->sch.tol, s->st); mark(s, "resnorm");
  }
}
// after a step's last V-cycle: the Krylov direction and pressure back into
// the first halves of d32 and p64 (kry_fused)
static void enqueue_mg_post(cf_sim* s) {
  if (!kry_fused(s)) return;
  cf::launch_kry_settle(s->L[0], s->dDyn, cf::KryBufs{s->dP64, s->dD32, s->dB0}, s->st);
  mark(s, "kry_settle");
}
static void enqueue_mg(cf_sim* s, bool coarse_coef) {
  enqueue_mg_pre(s);
  for (int c = 0; c < s->sch.vcycles; c++) enqueue_mg_cycle(s, c, coarse_coef);
  enqueue_mg_post(s);
}

static void enqueue_coef(cf_sim* s) {
  cf::launch_coef_rhs(s->F, s->st); mark(s, "coef_rhs");
  for (int l = 0; l < s->nlev - 1; l++) {
    cf::launch_coef_restrict(s->L[l], s->L[l + 1], s->st); mark(s, "coef_restrict");
  }
}

// stage granularity of encodeStepProfiled (js/fluid.js:256-292), unfused
static int enqueue_stage(cf_sim* s, int stage) {
  Fields& F = s->F;
  switch (stage) {
    case 0:
      if (s->mg_simple) return enqueue_stage(s, 8);
      cf::launch_advect(F, s->st); mark(s, "advect");
      break;
    case 8:
      cf::launch_advect_fwd(F, s->st); mark(s, "advect_fwd");
      cf::launch_advect_corr(F, s->st); mark(s, "advect_corr");
      break;
    case 1:
      cf::launch_diffuse(F, F.T2, F.Y2, F.T, F.Y, s->st); mark(s, "diffuse");
      for (int q = 1; q < s->diffSub; q += 2) {
        cf::launch_diffuse(F, F.T, F.Y, F.Th, F.Yh, s->st); mark(s, "diffuse");
        cf::launch_diffuse(F, F.Th, F.Yh, F.T, F.Y, s->st); mark(s, "diffuse");
      }
      break;
    case 2: cf::launch_sources(F, false, nullptr, nullptr, s->dAccPart, true, s->st); mark(s, "sources"); break;
    case 3: cf::launch_forces(F, s->st); mark(s, "forces"); break;
    case 4: enqueue_coef(s); break;
    case 5:
      enqueue_mg(s, false);
      break;
    case 6:
      cf::launch_project(F, s->sch.tol > 0 && s->sch.refine ? s->dP64 : nullptr, s->st); mark(s, "project");
      break;
    case 7:
      cf::launch_forces_coef(F, s->L, s->nlev, nullptr, 0, s->st); mark(s, "forces_coef");
      for (int l = 1 + cf::forces_coef_cascade(s->nlev); l < s->nlev; l++) {
        cf::launch_coef_restrict(s->L[l - 1], s->L[l], s->st); mark(s, level_tag("coef_restrict", l));
      }
      break;
    default: return set_err(CF_EINVAL, "stage %d out of range 0..8", stage);
  }
  return CF_OK;
}

// one step (js/fluid.js:225-253); the last diffusion substep runs inside
// sources, forces through the coefficient cascade run as one pass which also
// folds the accumulator partials, and the coarsest level's coefficients are
// restricted inside the bottom solver.  Split at the V-cycles:
// enqueue_step_pre runs up to them (returning the first cycle's coarse_coef),
// enqueue_mg_cycle each cycle, enqueue_step_post the projection.
static bool enqueue_step_pre(cf_sim* s) {
  Fields& F = s->F;
  const bool fused = !s->mg_simple;
  // CFL time step from the current velocities, on the device
  if (s->sch.cfl > 0) { cf::launch_cfl(F, s->sch.cfl, s->sch.umin, s->diffSub, s->st); mark(s, "cfl"); }
  enqueue_stage(s, 0);
  const float* Tin = F.T2;
  const float4* Yin = F.Y2;
  if (s->diffSub > 1) {
    cf::launch_diffuse(F, F.T2, F.Y2, F.T, F.Y, s->st); mark(s, "diffuse");
    for (int q = 1; q < s->diffSub; q += 2) {
      cf::launch_diffuse(F, F.T, F.Y, F.Th, F.Yh, s->st); mark(s, "diffuse");
      if (q + 2 < s->diffSub) { cf::launch_diffuse(F, F.Th, F.Yh, F.T, F.Y, s->st); mark(s, "diffuse"); }
    }
    Tin = F.Th;
    Yin = F.Yh;
  }
  cf::launch_sources(F, true, Tin, Yin, s->dAccPart, !fused, s->st); mark(s, "diffuse_sources");
  bool coarse_coef = false;
  if (!fused) {
    enqueue_stage(s, 3);
    enqueue_stage(s, 4);
  } else {
    cf::launch_forces_coef(F, s->L, s->nlev, s->dAccPart, (int)cf::sources_blocks(F), s->st); mark(s, "forces_coef");
    const int last_coef = s->nlev - 1 - (s->mg_bottom < s->nlev - 1 ? 1 : 0);
    for (int l = 1 + cf::forces_coef_cascade(s->nlev); l <= last_coef; l++) {
      cf::launch_coef_restrict(s->L[l - 1], s->L[l], s->st); mark(s, level_tag("coef_restrict", l));
    }
    // the coarsest coefficients are restricted once, in the first V-cycle's bottom solve
    coarse_coef = last_coef < s->nlev - 1 && s->nlev - 1 > cf::forces_coef_cascade(s->nlev);
  }
  enqueue_mg_pre(s);
  return coarse_coef;
}
static void enqueue_step_post(cf_sim* s) { enqueue_stage(s, 6); }
static void enqueue_step(cf_sim* s) {
  const bool coarse_coef = enqueue_step_pre(s);
  const bool fp = fuse_project(s);
  for (int c = 0; c < s->sch.vcycles; c++) enqueue_mg_cycle(s, c, coarse_coef, fp && c == s->sch.vcycles - 1);
  enqueue_mg_post(s);
  if (!fp) enqueue_step_post(s);
}

// ---- early exit of the residual-controlled cycles ------------------------------
// A graph is a fixed sequence, so a captured step with a residual target
// launches all `vcycles` cycles, and the ones after convergence return at
// once but still cost a launch each (a few microseconds per kernel, which
// adds up to more than the converged work on small scenes).  With a target
// the step is captured in segments instead: the head (up to and including
// the first SEG_HEAD cycles), a chunk of SEG_CHUNK cycles that is replayed
// while any scene has not converged and cycles remain, a remainder chunk for
// a cap that is not a whole number of chunks, and the tail (projection).
// The host reads the convergence flag after the head and after each chunk.
static void seg_plan(int vcycles, int* h, int* r) {
  *h = vcycles < SEG_HEAD ? vcycles : SEG_HEAD;
  *r = (vcycles - *h) % SEG_CHUNK;
}

// A graph exec must not be destroyed while a launch of it may still run: the
// runtime (and a profiler's dispatch records) still refer to its kernel nodes.
static void drop_graph(cf_sim* s) {
  if ((s->gexec || s->gexec8 || s->seg.head) && s->st) hipStreamSynchronize(s->st);
  if (s->gexec) hipGraphExecDestroy(s->gexec);
  if (s->gexec8) hipGraphExecDestroy(s->gexec8);
  s->gexec = s->gexec8 = nullptr;
  s->seg.drop();
  for (int q = 0; q < 12; q++) s->gkey[q] = -1;
}

// captures what fn enqueues on stream st
template <class Fn>
static int capture_fn(hipStream_t st, hipGraphExec_t* out, Fn fn) {
  hipGraph_t g = nullptr;
  HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  fn();
  HIPCHK(hipStreamEndCapture(st, &g));
  hipError_t e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
  hipGraphDestroy(g);
  if (e != hipSuccess) { *out = nullptr; return set_err(CF_EDEVICE, "hipGraphInstantiate: %s", hipGetErrorString(e)); }
  return CF_OK;
}

static int capture(cf_sim* s, int nsteps, hipGraphExec_t* out) {
  return capture_fn(s->st, out, [&] { for (int q = 0; q < nsteps; q++) enqueue_step(s); });
}

// residual-controlled steps on the fused path run in segments
static bool seg_mode(const cf_sim* s) { return s->sch.tol > 0 && !s->mg_simple; }

static int capture_segments(cf_sim* s) {
  Segments& g = s->seg;
  seg_plan(s->sch.vcycles, &g.h, &g.r);
  bool cc = false;
  int rc = capture_fn(s->st, &g.head, [&] {
    cc = enqueue_step_pre(s);
    for (int c = 0; c < g.h; c++) enqueue_mg_cycle(s, c, cc);
  });
  if (!rc && s->sch.vcycles > g.h)
    rc = capture_fn(s->st, &g.chunk, [&] { for (int c = 1; c <= SEG_CHUNK; c++) enqueue_mg_cycle(s, c, false); });
  if (!rc && g.r > 0) rc = capture_fn(s->st, &g.rem, [&] { for (int c = 1; c <= g.r; c++) enqueue_mg_cycle(s, c, false); });
  if (!rc) rc = capture_fn(s->st, &g.tail, [&] {
    enqueue_mg_post(s);
    enqueue_step_post(s);
  });
  return rc;
}

// one step from the segments: the host waits for the convergence flag after
// the head and after each chunk
static int run_segments(cf_sim* s) {
  Segments& g = s->seg;
  HIPCHK(hipGraphLaunch(g.head, s->st));
  int n = g.h;
  while (n < s->sch.vcycles) {
    HIPCHK(hipMemcpyAsync(s->hflag, &s->dDyn->converged, sizeof(int), hipMemcpyDeviceToHost, s->st));
    HIPCHK(hipStreamSynchronize(s->st));
    if (*s->hflag) break;
    const bool whole = s->sch.vcycles - n >= SEG_CHUNK;
    HIPCHK(hipGraphLaunch(whole ? g.chunk : g.rem, s->st));
    n += whole ? SEG_CHUNK : g.r;
  }
  HIPCHK(hipGraphLaunch(g.tail, s->st));
  return CF_OK;
}

static int ensure_graph(cf_sim* s) {
  // everything a captured step bakes in as launch arguments
  int key[12] = {s->diffSub, s->sch.nu, s->sch.vcycles, s->sch.coarse_iters, s->sch.smoother_fp16, s->nlev, 0, 0, 0,
                 s->sch.refine, s->sch.krylov, (int)s->tile_gen};
  memcpy(&key[6], &s->sch.tol, 4);
  memcpy(&key[7], &s->sch.cfl, 4);
  memcpy(&key[8], &s->sch.umin, 4);
  if ((s->gexec || s->seg.head) && !memcmp(key, s->gkey, sizeof(key))) return CF_OK;
  drop_graph(s);
  int rc;
  if (seg_mode(s)) {
    rc = capture_segments(s);
  } else {
    rc = capture(s, 1, &s->gexec);
    if (!rc) rc = capture(s, 8, &s->gexec8);
  }
  if (rc) { drop_graph(s); return rc; }
  memcpy(s->gkey, key, sizeof(key));
  return CF_OK;
}

// ---- helpers -----------------------------------------------------------------
static int param_index(const char* key) {
  for (int q = 0; q < CF_NPARAMS; q++) if (!strcmp(PNAME[q], key)) return q;
  return -1;
}

static int upload_params(cf_sim* s) {
  const int slot = s->upP;
  s->upP = (slot + 1) % UP_SLOTS;
  if (s->evP_used[slot]) HIPCHK(hipEventSynchronize(s->evP[slot]));
  s->hP[slot].P = s->P;
  s->hP[slot].dsub = s->dsub;
  s->hP[slot].dt = (float)s->dt;
  HIPCHK(hipMemcpyAsync(s->dP, &s->hP[slot], sizeof(ParamUpload), hipMemcpyHostToDevice, s->st));
  HIPCHK(hipEventRecord(s->evP[slot], s->st));
  s->evP_used[slot] = true;
  return CF_OK;
}

struct FieldRef {
  void* ptr;
  size_t bytes;
};

static bool lookup(cf_sim* s, const char* name, FieldRef& r) {
  const size_t nC = s->nC, nU = s->nU, nV = s->nV;
  Fields& F = s->F;
  struct { const char* n; void* p; size_t b; } t[] = {
      {"U", F.U, nU * 4}, {"Uh", F.Uh, nU * 4}, {"U2", F.U2, nU * 4},
      {"V", F.V, nV * 4}, {"Vh", F.Vh, nV * 4}, {"V2", F.V2, nV * 4},
      {"T", F.T, nC * 4}, {"Th", F.Th, nC * 4}, {"T2", F.T2, nC * 4},
      {"Y", F.Y, nC * 16}, {"Yh", F.Yh, nC * 16}, {"Y2", F.Y2, nC * 16},
      {"SQ", F.SQ, nC * 8}, {"cellG", F.cellG, nC * 8}, {"fx", F.fx, nU * 4}, {"fy", F.fy, nV * 4},
      {"pieces", F.pieces, CF_MAX_PIECES * CF_PIECE_FLOATS * 4}, {"acc", F.acc, CF_ACC_LEN * 4},
      {"rad", F.rad, CF_RAD_LEN * 16}, {"faceW", F.faceW, CF_MAX_PIECES * 4 * 4},
      {"p64", s->dP64, nC * 8}, {"d32", s->dD32, nC * 4}};
  for (auto& e : t)
    if (!strcmp(e.n, name)) { r.ptr = e.p; r.bytes = e.b; return true; }
  const char* pre[4] = {"rhs", "ax", "ay", "p"};
  for (int a = 0; a < 4; a++) {
    size_t n = strlen(pre[a]);
    if (strncmp(name, pre[a], n)) continue;
    const char* rest = name + n;
    int l = 0;
    if (*rest) {
      char* e;
      l = (int)strtol(rest, &e, 10);
      if (*e) continue;
    }
    if (l < 0 || l >= s->nlev) return false;
    const size_t lx = s->lnx[l], ly = s->lny[l];
    Level& Lv = s->L[l];
    if (a == 0) { r.ptr = Lv.rhs; r.bytes = lx * ly * 4; }
    else if (a == 1) { r.ptr = Lv.ax; r.bytes = (lx + 1) * ly * 4; }
    else if (a == 2) { r.ptr = Lv.ay; r.bytes = lx * (ly + 1) * 4; }
    else { r.ptr = Lv.p; r.bytes = lx * ly * 4; }
    return true;
  }
  return false;
}

// ---- C ABI -------------------------------------------------------------------
struct cf_pool {
  hipStream_t st = nullptr;
  uint64_t budget = 0, used = 0;
  std::vector<cf_sim*> slots;  // null: free slot
  cf_schedule sch;
  std::vector<cf::BatchPhase> step, frame_pre, frame_post;
  bool dirty = true;
  uint64_t key = 0, fkey = 0;
  hipGraphExec_t g1 = nullptr, g8 = nullptr;
  Segments seg;               // residual-controlled steps: the step in segments
  std::vector<size_t> cyc;    // first phase of each V-cycle of `step`; the last entry: the tail after them
  cf::Dyn** dDyns = nullptr;  // the live scenes' control blocks, for the convergence check (in `arena`)
  cf::TableArena arena;       // device tables of `step`
  cf::TableArena farena;      // device tables of frame_pre and frame_post
  int ndyn = 0;
  int* dflag = nullptr;       // all scenes converged (device), and its pinned host copy
  int* hflag = nullptr;
  bool use_graph = true;
  int nlaunch = 0;
  int64_t steps = 0;
  // bumped by every admit and release: a new scene may reuse a released one's
  // host address, so the table keys include it besides the slot pointers
  uint64_t gen = 0;
  int tile_over[MAXLEV] = {};  // cf_pool_set_leg_tiles: the leg tile of each depth (0: chosen below)
};

// The leg tile of each depth d of the tiled levels of the scenes `sc`:
// the pool's own (cf_pool_set_leg_tiles) where one is set, else the size
// rule applied to the tile counts of all those levels together (the launch
// the pool makes of them), whose many small levels fill the device with
// large tiles.  Returns the number of depths.
static int pool_depth_tiles(const cf_pool* pl, const std::vector<cf_sim*>& sc, int (&tile)[MAXLEV]) {
  int maxb = 0;
  for (auto* s : sc) maxb = s->mg_bottom > maxb ? s->mg_bottom : maxb;
  for (int d = 0; d < MAXLEV; d++) {
    tile[d] = 0;
    if (d >= maxb) continue;
    if (pl->tile_over[d]) { tile[d] = pl->tile_over[d]; continue; }
    int n64 = 0, n32 = 0;
    for (auto* s : sc)
      if (d < s->mg_bottom) {
        n64 += ((s->L[d].nx + 63) / 64) * ((s->L[d].ny + 63) / 64);
        n32 += ((s->L[d].nx + 31) / 32) * ((s->L[d].ny + 31) / 32);
      }
    tile[d] = cf::mg_tile_for_count(n64, n32);
  }
  return maxb;
}

// ---- scene pool: many scenes stepped together ---------------------------------
// Every phase of a step is one launch over the tiles of all live scenes
// (k_batch), so scenes of any mix of sizes advance together and the small
// scenes' latency-bound launches overlap with the large ones'.  A scene's own
// work (tiles, reductions, V-cycle levels) is exactly what it does alone.

// the order of a scene's extra diffusion substeps before sources, as
// enqueue_step issues them: (in, out) pairs and the fields sources reads
static void diffusion_chain(cf_sim* s, std::vector<std::array<void*, 4>>& chain, const float** Tin, const float4** Yin) {
  Fields& F = s->F;
  *Tin = F.T2;
  *Yin = F.Y2;
  if (s->diffSub > 1) {
    chain.push_back({(void*)F.T2, (void*)F.Y2, (void*)F.T, (void*)F.Y});
    for (int q = 1; q < s->diffSub; q += 2) {
      chain.push_back({(void*)F.T, (void*)F.Y, (void*)F.Th, (void*)F.Yh});
      if (q + 2 < s->diffSub) chain.push_back({(void*)F.Th, (void*)F.Yh, (void*)F.T, (void*)F.Y});
    }
    *Tin = F.Th;
    *Yin = F.Yh;
  }
}

static void pool_release_phases(std::vector<cf::BatchPhase>& v) {
  for (auto& p : v) p.release();
  v.clear();
}

static void pool_drop_graph(cf_pool* pl) {
  if ((pl->g1 || pl->g8 || pl->seg.head) && pl->st) hipStreamSynchronize(pl->st);  // see drop_graph
  if (pl->g1) hipGraphExecDestroy(pl->g1);
  if (pl->g8) hipGraphExecDestroy(pl->g8);
  pl->g1 = pl->g8 = nullptr;
  pl->seg.drop();
}

// the batched step of the live scenes, in the order of enqueue_step
static int pool_build_step(cf_pool* pl) {
  if (pl->st) hipStreamSynchronize(pl->st);  // the old graphs and tables may still be in use
  pool_drop_graph(pl);
  pool_release_phases(pl->step);
  std::vector<cf_sim*> sc;
  for (auto* s : pl->slots) if (s) sc.push_back(s);
  std::vector<cf::BatchPhase>& ph = pl->step;
  ph.reserve(64 + 16 * (size_t)pl->sch.vcycles);
  auto P = [&]() -> cf::BatchPhase& { ph.emplace_back(); return ph.back(); };
  const cf_schedule& k = pl->sch;
  if (k.cfl > 0) {
    cf::BatchPhase& a = P();
    for (auto* s : sc) cf::phase_cfl(a, s->F, k.cfl, k.umin, s->diffSub);
  }
  { cf::BatchPhase& a = P(); for (auto* s : sc) cf::phase_advect(a, s->F); }
  // extra diffusion substeps: phase d holds every scene's d-th
  std::vector<std::vector<std::array<void*, 4>>> chains(sc.size());
  std::vector<const float*> tin(sc.size());
  std::vector<const float4*> yin(sc.size());
  size_t maxc = 0;
  for (size_t q = 0; q < sc.size(); q++) {
    diffusion_chain(sc[q], chains[q], &tin[q], &yin[q]);
    maxc = chains[q].size() > maxc ? chains[q].size() : maxc;
  }
  for (size_t d = 0; d < maxc; d++) {
    cf::BatchPhase& a = P();
    for (size_t q = 0; q < sc.size(); q++)
      if (d < chains[q].size())
        cf::phase_diffuse(a, sc[q]->F, (const float*)chains[q][d][0], (const float4*)chains[q][d][1],
                          (float*)chains[q][d][2], (float4*)chains[q][d][3]);
  }
  {
    cf::BatchPhase& a = P();
    for (size_t q = 0; q < sc.size(); q++) cf::phase_sources(a, sc[q]->F, tin[q], yin[q], sc[q]->dAccPart);
  }
  {
    cf::BatchPhase& a = P();
    for (auto* s : sc) cf::phase_forces_coef(a, s->F, s->L, s->nlev, s->dAccPart, (int)cf::sources_blocks(s->F));
  }
  // coefficient restriction beyond the forces pass's cascade, by level
  std::vector<char> coarse_coef(sc.size());
  for (size_t q = 0; q < sc.size(); q++) {
    cf_sim* s = sc[q];
    const int last_coef = s->nlev - 1 - (s->mg_bottom < s->nlev - 1 ? 1 : 0);
    coarse_coef[q] = last_coef < s->nlev - 1 && s->nlev - 1 > cf::forces_coef_cascade(s->nlev);
  }
  for (int l = 1; l < MAXLEV; l++) {
    cf::BatchPhase a;
    for (auto* s : sc) {
      const int last_coef = s->nlev - 1 - (s->mg_bottom < s->nlev - 1 ? 1 : 0);
      if (l >= 1 + cf::forces_coef_cascade(s->nlev) && l <= last_coef) cf::phase_coef_restrict(a, s->L[l - 1], s->L[l]);
    }
    if (!a.empty()) ph.push_back(a);
  }
  // V-cycles; down leg d and up leg d hold level d of every scene that has one
  const bool tol = k.tol > 0, dc = tol && k.refine, kry = dc && k.krylov;
  auto kb = [](cf_sim* s) { return cf::KryBufs{s->dP64, s->dD32, s->dB0}; };
  auto ph0 = [&](cf_sim* s) { return k.smoother_fp16 ? s->LH[0].p : (__half*)nullptr; };
  if (k.smoother_fp16) {
    for (int l = 0; l < MAXLEV; l++) {
      cf::BatchPhase a;
      for (auto* s : sc)
        if (l < s->mg_bottom) cf::phase_coef_half(a, s->L[l], s->LH[l]);
      if (!a.empty()) ph.push_back(a);
    }
  }
  if (tol) { cf::BatchPhase& a = P(); for (auto* s : sc) cf::phase_fill_u32(a, &s->dDyn->converged, 0u, 2); }
  // a kernel and the fold that follows it: two phases, built as locals (a
  // reference from P() would not survive the second one's growth of ph)
  auto pair = [&](cf::BatchPhase& a, cf::BatchPhase& f) {
    ph.push_back(std::move(a));
    ph.push_back(std::move(f));
  };
  // the fused Krylov pass when every scene can take it (one kernel per phase)
  bool fused = kry;
  for (auto* s : sc) fused = fused && kry_fused(s);
  if (kry) {
    cf::BatchPhase a[5];
    for (auto* s : sc) cf::phase_kry_start(a, s->L[0], ph0(s), s->dDyn, kb(s), s->dNorm, k.tol);
    for (auto& x : a) ph.push_back(std::move(x));
  } else if (dc) {
    cf::BatchPhase a, f;
    for (auto* s : sc) cf::phase_dc_defect(a, f, s->L[0], ph0(s), s->dP64, s->dB0, s->dDyn, s->dNorm, k.tol, true);
    pair(a, f);
  }
  int maxb = 0;
  for (auto* s : sc) maxb = s->mg_bottom > maxb ? s->mg_bottom : maxb;
  // a leg's kernel depends on its tile size (chosen below per depth) and on
  // the storage of its two levels (FP16 smoother: whether the coarse level is
  // tiled too), so each level's legs form one phase per (tile, storage)
  const int modes[3] = {0, 1, 3};
  // the tile of each scene's tiled level d: the depth's (pool_depth_tiles)
  int dtile[MAXLEV];
  pool_depth_tiles(pl, sc, dtile);
  std::vector<std::array<int, MAXLEV>> ltile(sc.size());
  for (size_t q = 0; q < sc.size(); q++)
    for (int d = 0; d < MAXLEV; d++) ltile[q][d] = d < sc[q]->mg_bottom ? dtile[d] : 0;
  std::vector<int> tiles;  // every tile size a scene's level uses
  for (size_t q = 0; q < sc.size(); q++)
    for (int d = 0; d < sc[q]->mg_bottom; d++)
      if (std::find(tiles.begin(), tiles.end(), ltile[q][d]) == tiles.end()) tiles.push_back(ltile[q][d]);
  auto leg_ok = [&](size_t q, int d, int t, int m) {
    return d < sc[q]->mg_bottom && ltile[q][d] == t && mg_store(sc[q], d).mode == m;
  };
  // with a fixed cycle count and FP32 legs the projection runs inside the last
  // V-cycle's level-0 up leg of each scene whose level 0 is tiled, as the
  // scene alone runs it (the same faces, the same p: bitwise the same)
  const bool proj_in_leg = !tol && !k.smoother_fp16;
  std::vector<char> projected(sc.size(), 0);
  pl->cyc.clear();
  for (int c = 0; c < k.vcycles; c++) {
    pl->cyc.push_back(ph.size());
    for (int d = 0; d < maxb; d++) {
      for (int t : tiles)
        for (int m : modes) {
          cf::BatchPhase a;
          for (size_t q = 0; q < sc.size(); q++)
            if (leg_ok(q, d, t, m)) {
              cf_sim* s = sc[q];
              cf::phase_mg_down(a, s->L[d], s->L[d + 1], mg_store(s, d), k.nu, tol ? &s->dDyn->converged : nullptr, t,
                                d >= 1 || fused ? 1 : 0);
            }
          if (!a.empty()) ph.push_back(a);
        }
    }
    {
      cf::BatchPhase& a = P();
      for (size_t q = 0; q < sc.size(); q++) {
        cf_sim* s = sc[q];
        if (!cf::phase_mg_bottom(a, s->L, s->nlev, s->mg_bottom, coarse_coef[q] && c == 0, k.nu, k.coarse_iters,
                                 tol ? &s->dDyn->converged : nullptr, s->dBand))
          return set_err(CF_EINVAL, "pool: bottom solver does not fit a %dx%d scene", s->nx, s->ny);
      }
    }
    for (int d = maxb - 1; d >= 0; d--) {
      for (int t : tiles)
        for (int m : modes) {
          cf::BatchPhase a, b;  // b: the level-0 legs that also project
          for (size_t q = 0; q < sc.size(); q++)
            if (leg_ok(q, d, t, m)) {
              cf_sim* s = sc[q];
              if (d == 0 && c == k.vcycles - 1 && proj_in_leg && m == 0 && s->up_project && !s->mg_simple &&
                  cf::phase_mg_up_project(b, s->L[0], s->L[1], s->F, k.nu, nullptr, t)) {
                projected[q] = 1;
                continue;
              }
              cf::phase_mg_up(a, s->L[d], s->L[d + 1], mg_store(s, d), k.nu, tol ? &s->dDyn->converged : nullptr, t);
            }
          if (!a.empty()) ph.push_back(a);
          if (!b.empty()) ph.push_back(b);
        }
    }
    if (kry) {
      cf::BatchPhase a[5];
      for (auto* s : sc) cf::phase_kry_step(a, s->L[0], ph0(s), s->dDyn, kb(s), s->dNorm, k.tol, fused);
      for (auto& x : a) ph.push_back(std::move(x));
    } else if (dc) {
      {
        cf::BatchPhase& a = P();
        for (auto* s : sc) cf::phase_dc_accum(a, s->L[0], ph0(s), s->dDyn, s->dP64, &s->dDyn->converged);
      }
      {
        cf::BatchPhase a, f;
        for (auto* s : sc)
          cf::phase_dc_defect(a, f, s->L[0], ph0(s), s->dP64, s->dB0, s->dDyn, s->dNorm, k.tol, false);
        pair(a, f);
      }
    } else if (tol) {
      cf::BatchPhase a, f;
      for (auto* s : sc) cf::phase_resnorm(a, f, s->L[0], s->dDyn, s->dNorm, k.tol);
      pair(a, f);
    }
  }
  pl->cyc.push_back(ph.size());
  if (fused) { cf::BatchPhase& a = P(); for (auto* s : sc) cf::phase_kry_settle(a, s->L[0], s->dDyn, kb(s)); }
  {
    cf::BatchPhase a;
    for (size_t q = 0; q < sc.size(); q++)
      if (!projected[q]) cf::phase_project(a, sc[q]->F, dc ? sc[q]->dP64 : nullptr);
    if (!a.empty()) ph.push_back(std::move(a));
  }
  for (auto& a : ph)
    if (a.mixed) return set_err(CF_EINVAL, "pool: phase '%s' would run one kernel over scenes of another", a.tag);
  {  // the tables, and the control blocks the convergence check reads, in the step's arena
    std::vector<cf::Dyn*> dyns;
    for (auto* s : sc) dyns.push_back(s->dDyn);
    pl->ndyn = (int)dyns.size();
    std::vector<cf::BatchPhase*> all;
    for (auto& a : ph) all.push_back(&a);
    hipError_t e = pl->arena.place(all, pl->st, dyns.data(), dyns.size() * sizeof(cf::Dyn*), (void**)&pl->dDyns);
    if (e == hipSuccess && !pl->dflag) e = hipMalloc((void**)&pl->dflag, sizeof(int));
    if (e == hipSuccess && !pl->hflag) e = hipHostMalloc((void**)&pl->hflag, sizeof(int), hipHostMallocDefault);
    if (e != hipSuccess) return set_err(CF_ENOMEM, "pool tables: %s", hipGetErrorString(e));
  }
  pl->nlaunch = 0;
  for (auto& a : ph) pl->nlaunch += a.empty() ? 0 : 1;
  pl->dirty = false;
  return CF_OK;
}

static void pool_enqueue_step(cf_pool* pl) {
  for (auto& a : pl->step) a.run(pl->st);
}

static int pool_capture(cf_pool* pl, int n, hipGraphExec_t* out) {
  return capture_fn(pl->st, out, [&] { for (int q = 0; q < n; q++) pool_enqueue_step(pl); });
}

// phases [a, b) of the step
static void pool_enqueue_range(cf_pool* pl, size_t a, size_t b) {
  for (size_t q = a; q < b; q++) pl->step[q].run(pl->st);
}

// the segments of a residual-controlled step (as capture_segments for one
// scene): every cycle after the first has the same phases, so the chunks
// replay cycles 1, 2, ...
static int pool_capture_segments(cf_pool* pl) {
  Segments& g = pl->seg;
  const int V = pl->sch.vcycles;
  seg_plan(V, &g.h, &g.r);
  const std::vector<size_t>& cy = pl->cyc;
  int rc = capture_fn(pl->st, &g.head, [&] { pool_enqueue_range(pl, 0, cy[g.h]); });
  if (!rc && V > g.h) rc = capture_fn(pl->st, &g.chunk, [&] { pool_enqueue_range(pl, cy[1], cy[1 + SEG_CHUNK]); });
  if (!rc && g.r > 0) rc = capture_fn(pl->st, &g.rem, [&] { pool_enqueue_range(pl, cy[1], cy[1 + g.r]); });
  if (!rc) rc = capture_fn(pl->st, &g.tail, [&] { pool_enqueue_range(pl, cy[V], pl->step.size()); });
  return rc;
}

static int pool_run_segments(cf_pool* pl) {
  Segments& g = pl->seg;
  HIPCHK(hipGraphLaunch(g.head, pl->st));
  int n = g.h;
  while (n < pl->sch.vcycles) {
    cf::launch_all_converged(pl->dDyns, pl->ndyn, pl->dflag, pl->st);
    HIPCHK(hipMemcpyAsync(pl->hflag, pl->dflag, sizeof(int), hipMemcpyDeviceToHost, pl->st));
    HIPCHK(hipStreamSynchronize(pl->st));
    if (*pl->hflag) break;
    const bool whole = pl->sch.vcycles - n >= SEG_CHUNK;
    HIPCHK(hipGraphLaunch(whole ? g.chunk : g.rem, pl->st));
    n += whole ? SEG_CHUNK : g.r;
  }
  HIPCHK(hipGraphLaunch(g.tail, pl->st));
  return CF_OK;
}

// the batched step for the current scenes, rebuilt when scenes come or go
// or a scene's substep count changes
static int pool_ready(cf_pool* pl) {
  uint64_t key = (1469598103934665603ull ^ pl->gen) * 1099511628211ull;
  for (auto* s : pl->slots) {
    key = (key ^ (uint64_t)(uintptr_t)s) * 1099511628211ull;
    if (s) key = (key ^ (uint64_t)s->diffSub) * 1099511628211ull;
    if (s) key = (key ^ (uint64_t)s->tile_gen) * 1099511628211ull;
  }
  if (key != pl->key) { pl->dirty = true; pl->key = key; }
  if (pl->dirty) {
    int rc = pool_build_step(pl);
    if (rc) return rc;
  }
  if (pl->use_graph && !pl->g1 && !pl->seg.head) {
    int rc;
    if (pl->sch.tol > 0) {
      rc = pool_capture_segments(pl);
    } else {
      rc = pool_capture(pl, 1, &pl->g1);
      if (!rc) rc = pool_capture(pl, 8, &pl->g8);
    }
    if (rc) { pool_drop_graph(pl); return rc; }
  }
  return CF_OK;
}

namespace cf {
int set_error(int code, const char* msg) { return set_err(code, "%s", msg); }
int check_flame_case(const cf_flame_case* k) {
  if (!(k->dx > 0) || !(k->L > 0) || !(k->tEnd >= 0) || !(k->F > 0))
    return set_err(CF_EINVAL, "flame case: bad dx, L, tEnd or F");
  const double n = floor(k->L / k->dx + 0.5);  // Math.round (tests/flame1d.mjs:15)
  if (n < 2 || n > CF_FLAME_MAXN) return set_err(CF_EBADSIZE, "flame case: %g cells (2..%d)", n, CF_FLAME_MAXN);
  return CF_OK;
}
}  // namespace cf

extern "C" {

const char* cf_last_error(void) { return g_err; }
const char* cf_version(void) { return "campfire-hip 0.1 (gfx950)"; }
const char* cf_param_name(int q) { return q >= 0 && q < CF_NPARAMS ? PNAME[q] : nullptr; }

void cf_sim_destroy(cf_sim* s) {
  if (!s) return;
  if (s->st) hipStreamSynchronize(s->st);
  drop_graph(s);
  for (int q = 0; q < UP_SLOTS; q++) { hipEventDestroy(s->evP[q]); hipEventDestroy(s->evPc[q]); }
  for (int q = 0; q < 3; q++) hipEventDestroy(s->ringEv[q]);
  if (s->hP) hipHostFree(s->hP);
  if (s->hPieces) hipHostFree(s->hPieces);
  if (s->hflag) hipHostFree(s->hflag);
  if (s->ring) hipHostFree(s->ring);
  if (s->block) hipFree(s->block);
  if (s->st && s->own_stream) hipStreamDestroy(s->st);
  delete s;
}

// grid sizes and the multigrid levels of a new scene; levels halve while both
// dims are even and cells > 256 (js/fluid.js:71-91; the reference's 6-level
// cap is lifted so 1024^2 reaches a 16x16 coarsest level)
static int sim_levels(cf_sim* s, int nx, int ny, double dx, double dt) {
  if (nx < 2 || ny < 2 || nx > 16384 || ny > 16384 || !(dx > 0) || !(dt > 0))
    return set_err(CF_EINVAL, "bad grid %dx%d dx=%g dt=%g", nx, ny, dx, dt);
  s->nx = nx; s->ny = ny; s->dx = dx; s->dt = dt;
  s->nC = nx * ny; s->nU = (nx + 1) * ny; s->nV = nx * (ny + 1);
  int lx = nx, ly = ny;
  double h = dx;
  for (int l = 0; l < MAXLEV; l++) {
    s->lnx[l] = lx; s->lny[l] = ly; s->lh2[l] = (float)(h * h);
    s->nlev = l + 1;
    if (lx % 2 || ly % 2 || lx * ly <= 256) break;
    lx /= 2; ly /= 2; h *= 2;
  }
  if (s->lnx[s->nlev - 1] * s->lny[s->nlev - 1] > 256)
    return set_err(CF_EBADSIZE, "coarsest level too large (%dx%d)", s->lnx[s->nlev - 1], s->lny[s->nlev - 1]);
  for (int l = 0; l < s->nlev; l++) { s->L[l].nx = s->lnx[l]; s->L[l].ny = s->lny[l]; s->L[l].h2 = s->lh2[l]; }
  // bottom solver: the last 1-3 levels (<= 4096, 1024, 256 cells) in one workgroup
  s->mg_bottom = cf::mg_bottom_first(s->L, s->nlev);
  // with two-level legs the levels from 1 go in pairs; when they would leave
  // one over and the bottom kernel holds three levels, its first joins them
  s->mg_bottom_pairs = s->mg_bottom;
  if ((s->mg_bottom - 1) % 2 == 1 && s->nlev - s->mg_bottom == 3) s->mg_bottom_pairs = s->mg_bottom + 1;
  return CF_OK;
}

// the scene's device buffers: one allocation of 256-B aligned sub-buffers
static void sim_plan(cf_sim* s, std::vector<std::pair<void**, size_t>>& plan) {
  Fields& F = s->F;
  auto add = [&](void** p, size_t bytes) { plan.push_back({p, bytes}); };
  const size_t nC = s->nC, nU = s->nU, nV = s->nV;
  static_assert(sizeof(cf_params) % 8 == 0 && offsetof(ParamUpload, dsub) == sizeof(cf_params), "upload layout");
  add((void**)&s->dP, sizeof(cf_params) + sizeof(cf::Dyn));
  add((void**)&F.U, nU * 4); add((void**)&F.Uh, nU * 4); add((void**)&F.U2, nU * 4);
  add((void**)&F.V, nV * 4); add((void**)&F.Vh, nV * 4); add((void**)&F.V2, nV * 4);
  add((void**)&F.T, nC * 4); add((void**)&F.Th, nC * 4); add((void**)&F.T2, nC * 4);
  add((void**)&F.Y, nC * 16); add((void**)&F.Yh, nC * 16); add((void**)&F.Y2, nC * 16);
  add((void**)&F.SQ, nC * 8); add((void**)&F.cellG, nC * 8); add((void**)&F.G, nC * 8);
  add((void**)&F.fx, nU * 4); add((void**)&F.fy, nV * 4);
  add((void**)&F.pieces, CF_MAX_PIECES * CF_PIECE_FLOATS * 4);
  add((void**)&F.acc, CF_ACC_LEN * 4);
  add((void**)&F.rad, CF_RAD_LEN * 16);
  add((void**)&F.faceW, CF_MAX_PIECES * 4 * 4);
  for (int l = 0; l < s->nlev; l++) {
    const size_t lx = s->lnx[l], ly = s->lny[l];
    add((void**)&s->L[l].p, lx * ly * 4); add((void**)&s->L[l].rhs, lx * ly * 4);
    add((void**)&s->L[l].ax, (lx + 1) * ly * 4); add((void**)&s->L[l].ay, lx * (ly + 1) * 4);
  }
  for (int l = 0; l < s->mg_bottom; l++) {  // leg buffers and FP16 smoother storage of the tiled levels
    const size_t lx = s->lnx[l], ly = s->lny[l];
    add((void**)&s->L[l].q, lx * ly * 4);
    add((void**)&s->LH[l].p, lx * ly * 2); add((void**)&s->LH[l].q, lx * ly * 2);
    add((void**)&s->LH[l].ax, (lx + 1) * ly * 2); add((void**)&s->LH[l].ay, lx * (ly + 1) * 2);
  }
  for (int l = s->mg_bottom; l < s->mg_bottom_pairs; l++)  // the level the two-level legs take from the bottom
    add((void**)&s->L[l].q, (size_t)s->lnx[l] * s->lny[l] * 4);
  const unsigned nparts = (unsigned)(((s->nx + 31) / 32) * ((s->ny + 7) / 8));
  add((void**)&s->dPart, (size_t)nparts * 8);
  add((void**)&s->dAccPart, (size_t)nparts * 16);
  add((void**)&s->dNorm, 4 * sizeof(double) * 1024);
  add((void**)&s->dP64, 2 * nC * sizeof(double));  // two halves: see kry_fused
  add((void**)&s->dD32, 2 * nC * sizeof(float));
  add((void**)&s->dB0, nC * sizeof(float));
  add((void**)&s->dBand, (256 * 33 + 256) * sizeof(double));
}

static size_t plan_total(const std::vector<std::pair<void**, size_t>>& plan) {
  size_t total = 0;
  for (auto& q : plan) total += (q.second + 255) & ~(size_t)255;
  return total;
}

static int sim_create(int nx, int ny, double dx, double dt, hipStream_t shared, cf_sim** out);

int cf_sim_create(int nx, int ny, double dx, double dt, cf_sim** out) {
  return sim_create(nx, ny, dx, dt, nullptr, out);
}

// shared: the stream of a pool (borrowed), or null for a stream of its own
static int sim_create(int nx, int ny, double dx, double dt, hipStream_t shared, cf_sim** out) {
  if (!out) return set_err(CF_EINVAL, "out is NULL");
  *out = nullptr;
  cf_sim* s = new cf_sim();
  {
    const int rc = sim_levels(s, nx, ny, dx, dt);
    if (rc) { delete s; return rc; }
  }
  // defaults (js/fluid.js:35-45, js/config.js:12-47)
  const double def[CF_NPARAMS] = {
      (double)nx, (double)ny, dx, dt, PHYS_Tamb, phys_rhoAmb(), 0.232, 9.81,
      1300, 0.7, 1, 1, 2.0e10, 15000, 15.0e6, 15.0e6 / 13.1e6,
      6.0e6, 25000, 5.4e8, 19800,
      1862, 0.8, 0.05, 0.8,
      6.6, 5800, 0, 0,
      0, 0, 0, 0, 0, 0, 0, 600,
      1800, 5.670374e-8, 2.667, 0};
  memcpy(s->pv, def, sizeof(def));
  s->sch.nu = 2; s->sch.vcycles = 1; s->sch.tol = 0; s->sch.coarse_iters = 120; s->sch.smoother_fp16 = 0;
  s->sch.cfl = 0; s->sch.umin = 1.0f; s->sch.refine = 0; s->sch.krylov = 0;
  const char* ng = getenv("CF_NO_GRAPH");
  s->use_graph = !(ng && *ng && *ng != '0');
  const char* ms = getenv("CF_MG_SIMPLE");
  s->mg_simple = ms && *ms && *ms != '0';
  const char* lp = getenv("CF_LEG_PAIRS");
  s->leg_pairs = !(lp && *lp == '0');
  const char* up = getenv("CF_UP_PROJECT");
  s->up_project = !(up && *up == '0');
  const char* pt = getenv("CF_PAIR_TILE");
  s->pair_tile = pt && (atoi(pt) == 8 || atoi(pt) == 16) ? atoi(pt) : 0;

  hipError_t e = hipGetDevice(&s->dev);
  if (e != hipSuccess) { delete s; return set_err(CF_EDEVICE, "hipGetDevice: %s", hipGetErrorString(e)); }
  if (shared) {
    s->st = shared;
    s->own_stream = false;
  } else {
    e = hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking);
    if (e != hipSuccess) { s->st = nullptr; delete s; return set_err(CF_EDEVICE, "hipStreamCreate: %s", hipGetErrorString(e)); }
  }

  std::vector<std::pair<void**, size_t>> plan;
  sim_plan(s, plan);
  Fields& F = s->F;
  const size_t total = plan_total(plan);
  const unsigned nparts = (unsigned)(((nx + 31) / 32) * ((ny + 7) / 8));
  e = hipMalloc(&s->block, total);
  if (e != hipSuccess) { s->block = nullptr; cf_sim_destroy(s); return set_err(CF_ENOMEM, "hipMalloc(%zu): %s", total, hipGetErrorString(e)); }
  s->block_bytes = total;
  {
    size_t off = 0;
    for (auto& q : plan) { *q.first = (char*)s->block + off; off += (q.second + 255) & ~(size_t)255; }
  }
  s->hPart.resize((size_t)nparts * 2);
  F.P = s->dP; F.nx = nx; F.ny = ny;
  s->dDyn = (cf::Dyn*)(s->dP + 1);
  F.dyn = s->dDyn;
  F.p0 = s->L[0].p; F.rhs0 = s->L[0].rhs; F.ax0 = s->L[0].ax; F.ay0 = s->L[0].ay;
  // device buffers start zeroed, as WebGPU buffers do
  cf::launch_fill_u32(s->block, 0u, total / 4, s->st);

  for (int q = 0; q < UP_SLOTS; q++) {
    hipEventCreateWithFlags(&s->evP[q], hipEventDisableTiming);
    hipEventCreateWithFlags(&s->evPc[q], hipEventDisableTiming);
    s->evP_used[q] = s->evPc_used[q] = false;
  }
  for (int q = 0; q < 3; q++) { hipEventCreateWithFlags(&s->ringEv[q], hipEventDisableTiming); s->busy[q] = false; }
  if (hipHostMalloc((void**)&s->hP, sizeof(ParamUpload) * UP_SLOTS, hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void**)&s->hPieces, sizeof(float) * CF_MAX_PIECES * CF_PIECE_FLOATS * UP_SLOTS, hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void**)&s->ring, RING_BYTES * 3, hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void**)&s->hflag, sizeof(int), hipHostMallocDefault) != hipSuccess) {
    cf_sim_destroy(s);
    return set_err(CF_ENOMEM, "pinned host allocation failed");
  }
  int rc = cf_sim_write_params(s);
  if (rc) { cf_sim_destroy(s); return rc; }
  rc = cf_sim_reset(s);
  if (rc) { cf_sim_destroy(s); return rc; }
  e = hipStreamSynchronize(s->st);
  if (e != hipSuccess) { cf_sim_destroy(s); return set_err(CF_EDEVICE, "init: %s", hipGetErrorString(e)); }
  *out = s;
  return CF_OK;
}

int cf_sim_reset(cf_sim* s) {  // js/fluid.js:162-176
  if (!s) return set_err(CF_EINVAL, "sim is NULL");
  Fields& F = s->F;
  const float Ta = (float)PHYS_Tamb;
  uint32_t tb;
  memcpy(&tb, &Ta, 4);
  cf::launch_fill_u32(F.T, tb, s->nC, s->st);
  cf::launch_fill_u32(F.T2, tb, s->nC, s->st);
  cf::launch_fill_u32(F.Th, tb, s->nC, s->st);
  // Y = (0, YOamb, 0, 0): fill zero then set the O component through a strided view
  std::vector<float> y((size_t)s->nC * 4, 0.f);
  const float yo = (float)s->pv[6];
  for (int k = 0; k < s->nC; k++) y[(size_t)k * 4 + 1] = yo;
  HIPCHK(hipMemcpyAsync(F.Y, y.data(), y.size() * 4, hipMemcpyHostToDevice, s->st));
  HIPCHK(hipMemcpyAsync(F.Y2, y.data(), y.size() * 4, hipMemcpyHostToDevice, s->st));
  HIPCHK(hipMemcpyAsync(F.Yh, y.data(), y.size() * 4, hipMemcpyHostToDevice, s->st));
  cf::launch_fill_u32(F.U, 0, s->nU, s->st); cf::launch_fill_u32(F.Uh, 0, s->nU, s->st); cf::launch_fill_u32(F.U2, 0, s->nU, s->st);
  cf::launch_fill_u32(F.V, 0, s->nV, s->st); cf::launch_fill_u32(F.Vh, 0, s->nV, s->st); cf::launch_fill_u32(F.V2, 0, s->nV, s->st);
  cf::launch_fill_u32(F.SQ, 0, (size_t)s->nC * 2, s->st);
  cf::launch_fill_u32(F.acc, 0, CF_ACC_LEN, s->st);
  cf::launch_fill_u32(F.rad, 0, CF_RAD_LEN * 4, s->st);
  for (int l = 0; l < s->nlev; l++) cf::launch_fill_u32(s->L[l].p, 0, (size_t)s->lnx[l] * s->lny[l], s->st);
  cf::launch_fill_u32(s->dP64, 0, (size_t)s->nC * 2, s->st);
  s->time = 0;
  HIPCHK(hipMemsetAsync(&s->dDyn->time, 0, sizeof(double), s->st));
  cf::launch_fill_u32(&s->dDyn->nonfinite, 0u, 1, s->st);
  HIPCHK(hipStreamSynchronize(s->st));  // the staging vector goes out of scope
  return CF_OK;
}

int cf_sim_set_param(cf_sim* s, const char* key, double v) {  // js/fluid.js:178
  if (!s || !key) return set_err(CF_EINVAL, "NULL argument");
  if (!strcmp(key, "time")) { s->time = v; return CF_OK; }
  if (!strcmp(key, "frame")) { s->frame = (uint32_t)(int64_t)v; return CF_OK; }
  const int q = param_index(key);
  if (q < 0) return set_err(CF_EINVAL, "unknown parameter '%s'", key);
  if (q <= 3) return set_err(CF_EINVAL, "'%s' is owned by the sim", key);
  if (q == 26 && !(v >= 0 && v <= CF_MAX_PIECES)) return set_err(CF_EINVAL, "nPieces %g outside 0..%d", v, CF_MAX_PIECES);
  s->pv[q] = v;
  return CF_OK;
}

int cf_sim_get_param(const cf_sim* s, const char* key, double* v) {
  if (!s || !key || !v) return set_err(CF_EINVAL, "NULL argument");
  if (!strcmp(key, "time")) { *v = s->time; return CF_OK; }
  if (!strcmp(key, "frame")) { *v = s->frame; return CF_OK; }
  const int q = param_index(key);
  if (q < 0) return set_err(CF_EINVAL, "unknown parameter '%s'", key);
  *v = s->pv[q];
  return CF_OK;
}

int cf_sim_write_params(cf_sim* s) {  // js/fluid.js:180-199
  if (!s) return set_err(CF_EINVAL, "sim is NULL");
  s->pv[0] = s->nx; s->pv[1] = s->ny; s->pv[2] = s->dx; s->pv[3] = s->dt;
  s->pv[39] = s->time; s->pv[27] = (double)s->frame;
  uint32_t* raw = (uint32_t*)&s->P;
  for (int q = 0; q < CF_NPARAMS; q++) {
    if (PKIND[q] == 'f') { float f = (float)s->pv[q]; memcpy(&raw[q], &f, 4); }
    else if (PKIND[q] == 'i') { int32_t iv = (int32_t)s->pv[q]; memcpy(&raw[q], &iv, 4); }
    else raw[q] = (uint32_t)(int64_t)s->pv[q];
  }
  // explicit diffusion stability: dt_sub <= dx^2 / (4.5 alpha(2200 K)), odd count <= 5
  const double Tmax = 2200, mu = 1.716e-5 * pow(Tmax / 273.15, 1.5) * 383.55 / (Tmax + 110.4);
  const double alpha = mu / 0.7 * s->pv[11] / (phys_rhoAmb() * PHYS_Tamb / Tmax);
  // with a CFL time step the count is fixed by the largest dt the rule allows
  const double dtmax = s->sch.cfl > 0 ? (double)s->sch.cfl * s->dx / (double)s->sch.umin : s->dt;
  int n = (int)ceil(dtmax / (s->dx * s->dx / (4.5 * alpha)));
  n = n % 2 ? n : n + 1;
  if (n > 5) n = 5;
  s->diffSub = n;
  s->dsub = (float)(s->dt / n);
  return upload_params(s);
}

int cf_sim_get_packed_params(const cf_sim* s, cf_params* out) {
  if (!s || !out) return set_err(CF_EINVAL, "NULL argument");
  *out = s->P;
  return CF_OK;
}

// the rules a schedule must meet on scene s (CF_OK or CF_EINVAL with a message)
static int sched_check(const cf_sim* s, const cf_schedule* sc) {
  if (sc->nu < 1 || sc->nu > 3 || sc->vcycles < 1 || sc->coarse_iters < 0 || !(sc->tol >= 0))
    return set_err(CF_EINVAL, "bad schedule (nu 1..3, vcycles >= 1, coarse_iters >= 0)");
  if (sc->coarse_iters == 0 && (s->lnx[s->nlev - 1] > 32 || s->mg_simple))
    return set_err(CF_EINVAL, "direct coarsest solve needs a coarsest level <= 32 cells wide (fused path)");
  if (!(sc->cfl >= 0) || (sc->cfl > 0 && !(sc->umin > 0)))
    return set_err(CF_EINVAL, "bad schedule (cfl >= 0; umin > 0 with a CFL step)");
  if (sc->refine && !(sc->tol > 0)) return set_err(CF_EINVAL, "bad schedule (refine needs tol > 0)");
  if (sc->krylov && !sc->refine) return set_err(CF_EINVAL, "bad schedule (krylov needs refine = 1)");
  if (sc->smoother_fp16 && (!sc->refine || s->mg_simple))
    return set_err(CF_EINVAL, "FP16 smoother needs defect correction (refine = 1, tol > 0) and the fused path");
  return CF_OK;
}

// applies a schedule that sched_check accepted
static int sched_apply(cf_sim* s, const cf_schedule* sc) {
  // switching defect correction on: the FP64 pressure starts from the FP32 one
  if (sc->refine && !s->sch.refine) cf::launch_p32_to_64(s->L[0].p, s->dP64, s->nC, s->st);
  const bool cfl_changed = sc->cfl != s->sch.cfl || sc->umin != s->sch.umin;
  s->sch = *sc;
  // the diffusion substep count follows the largest dt of a CFL rule
  if (cfl_changed) return cf_sim_write_params(s);
  return CF_OK;
}

int cf_sim_set_schedule(cf_sim* s, const cf_schedule* sc) {
  if (!s || !sc) return set_err(CF_EINVAL, "NULL argument");
  // a pool's batched step is built from the pool's schedule: a scene of a
  // pool takes it from cf_pool_set_schedule only
  if (s->in_pool && memcmp(sc, &s->sch, sizeof(*sc)))
    return set_err(CF_EINVAL, "scene belongs to a pool: set the schedule with cf_pool_set_schedule");
  int rc = sched_check(s, sc);
  if (rc) return rc;
  return sched_apply(s, sc);
}

int cf_sim_set_leg_tiles(cf_sim* s, const int32_t* tiles, int n) {
  if (!s || (n > 0 && !tiles) || n < 0) return set_err(CF_EINVAL, "bad argument");
  if (n > s->mg_bottom) return set_err(CF_EINVAL, "%d tile entries for %d tiled levels", n, s->mg_bottom);
  for (int l = 0; l < n; l++)
    if (tiles[l] != 0 && tiles[l] != 16 && tiles[l] != 32 && tiles[l] != 48 && tiles[l] != 64)
      return set_err(CF_EINVAL, "leg tile %d of level %d is not 0, 16, 32, 48 or 64", tiles[l], l);
  bool changed = false;
  for (int l = 0; l < s->mg_bottom; l++) {
    const int t = l < n ? tiles[l] : 0;
    if (s->L[l].tile != t) { s->L[l].tile = t; changed = true; }
  }
  if (changed) s->tile_gen++;
  return CF_OK;
}

int cf_sim_get_leg_tiles(const cf_sim* s, int32_t* tiles, int n) {
  if (!s || (n > 0 && !tiles)) return set_err(CF_EINVAL, "bad argument");
  for (int l = 0; l < n; l++) tiles[l] = l < s->mg_bottom && !paired(s, l) ? cf::mg_tile(s->L[l]) : 0;
  return s->mg_bottom;
}

int cf_sim_get_schedule(const cf_sim* s, cf_schedule* sc) {
  if (!s || !sc) return set_err(CF_EINVAL, "NULL argument");
  *sc = s->sch;
  return CF_OK;
}

int cf_sim_upload_pieces(cf_sim* s, const float* data, int n) {  // js/fluid.js:201-204
  if (!s || !data) return set_err(CF_EINVAL, "NULL argument");
  if (n < 0 || n > CF_MAX_PIECES) return set_err(CF_EINVAL, "nPieces %d outside 0..%d", n, CF_MAX_PIECES);
  s->pv[26] = n;
  const int slot = s->upPc;
  s->upPc = (slot + 1) % UP_SLOTS;
  if (s->evPc_used[slot]) HIPCHK(hipEventSynchronize(s->evPc[slot]));
  float* h = s->hPieces + (size_t)slot * CF_MAX_PIECES * CF_PIECE_FLOATS;
  memcpy(h, data, sizeof(float) * CF_MAX_PIECES * CF_PIECE_FLOATS);
  HIPCHK(hipMemcpyAsync(s->F.pieces, h, sizeof(float) * CF_MAX_PIECES * CF_PIECE_FLOATS, hipMemcpyHostToDevice, s->st));
  HIPCHK(hipEventRecord(s->evPc[slot], s->st));
  s->evPc_used[slot] = true;
  return CF_OK;
}

int cf_sim_clear_geometry(cf_sim* s) {  // js/fluid.js:213
  if (!s) return set_err(CF_EINVAL, "sim is NULL");
  cf::launch_fill_u32(s->F.faceW, 0, CF_MAX_PIECES * 4, s->st);
  HIPCHK(hipGetLastError());
  return CF_OK;
}

int cf_sim_geometry(cf_sim* s) {  // js/fluid.js:215-218
  if (!s) return set_err(CF_EINVAL, "sim is NULL");
  cf::launch_geometry(s->F, s->st);
  HIPCHK(hipGetLastError());
  return CF_OK;
}

int cf_sim_radiation(cf_sim* s) {  // js/fluid.js:220-222
  if (!s) return set_err(CF_EINVAL, "sim is NULL");
  const int np = s->P.nPieces < CF_MAX_PIECES ? s->P.nPieces : CF_MAX_PIECES;
  cf::launch_radiation(s->F, np * 4 + CF_N_FLOOR, s->st);
  HIPCHK(hipGetLastError());
  return CF_OK;
}


int cf_sim_step(cf_sim* s, int nsteps) {  // js/fluid.js:225-253
  if (!s) return set_err(CF_EINVAL, "sim is NULL");
  if (nsteps < 0) return set_err(CF_EINVAL, "nsteps < 0");
  if (s->use_graph) {
    int rc = ensure_graph(s);
    if (rc) return rc;
    int q = 0;
    if (s->seg.head) {
      for (; q < nsteps; q++)
        if ((rc = run_segments(s))) return rc;
    }
    for (; q + 8 <= nsteps; q += 8) HIPCHK(hipGraphLaunch(s->gexec8, s->st));
    for (; q < nsteps; q++) HIPCHK(hipGraphLaunch(s->gexec, s->st));
  } else {
    for (int q = 0; q < nsteps; q++) enqueue_step(s);
    HIPCHK(hipGetLastError());
  }
  // with a CFL step the simulated time lives on the device (cf_sim_get_info)
  if (s->sch.cfl <= 0) s->time += s->dt * nsteps;
  s->last_vcycles = s->sch.vcycles;
  return CF_OK;
}

int cf_sim_frame(cf_sim* s, int nsteps, int with_radiation) {  // js/main.js:266-283
  int rc;
  if ((rc = cf_sim_write_params(s))) return rc;
  if ((rc = cf_sim_clear_geometry(s))) return rc;
  if ((rc = cf_sim_geometry(s))) return rc;
  if ((rc = cf_sim_step(s, nsteps))) return rc;
  if (with_radiation && (rc = cf_sim_radiation(s))) return rc;
  return CF_OK;
}

int cf_sim_run_stage(cf_sim* s, int stage) {
  if (!s) return set_err(CF_EINVAL, "sim is NULL");
  int rc = enqueue_stage(s, stage);
  if (rc) return rc;
  HIPCHK(hipGetLastError());
  return CF_OK;
}


int cf_sim_profile_steps(cf_sim* s, int nsteps, float* stage_ms) {
  if (!s || !stage_ms) return set_err(CF_EINVAL, "NULL argument");
  std::vector<hipEvent_t> ev((size_t)nsteps * 8);
  for (auto& e : ev) HIPCHK(hipEventCreate(&e));
  for (int q = 0; q < 7; q++) stage_ms[q] = 0;
  for (int n = 0; n < nsteps; n++) {
    hipEventRecord(ev[n * 8], s->st);
    for (int st = 0; st < 7; st++) {
      enqueue_stage(s, st);
      hipEventRecord(ev[n * 8 + st + 1], s->st);
    }
    s->time += s->dt;
  }
  HIPCHK(hipStreamSynchronize(s->st));
  for (int n = 0; n < nsteps; n++)
    for (int st = 0; st < 7; st++) {
      float ms = 0;
      hipEventElapsedTime(&ms, ev[n * 8 + st], ev[n * 8 + st + 1]);
      stage_ms[st] += ms;
    }
  for (auto& e : ev) hipEventDestroy(e);
  return CF_OK;
}

// Per-kernel timing of the production step: an event after every launch site;
// returns, per kernel name, the summed milliseconds and the launch count.
int cf_sim_profile_kernels(cf_sim* s, int nsteps, int max, char* names, float* ms, int* launches, int* count) {
  if (!s || !names || !ms || !launches || !count) return set_err(CF_EINVAL, "NULL argument");
  s->plaunch.clear(); s->ptag.clear(); s->tags.clear();
  s->profiling = true;
  cf::prof_sink = &s->plaunch;
  for (int n = 0; n < nsteps; n++) { enqueue_step(s); s->time += s->dt; }
  cf::prof_sink = nullptr;
  s->profiling = false;
  mark(s, "other");  // launches after the last mark (none in a step)
  hipError_t e = hipStreamSynchronize(s->st);
  for (size_t q = 0; q < s->plaunch.size(); q++) {
    float t = 0;
    if (e == hipSuccess && hipEventElapsedTime(&t, s->plaunch[q].start, s->plaunch[q].stop) == hipSuccess) {
      s->tags[s->ptag[q]].ms += t;
      s->tags[s->ptag[q]].launches++;
    }
  }
  for (auto& ev : s->plaunch) { hipEventDestroy(ev.start); hipEventDestroy(ev.stop); }
  s->plaunch.clear(); s->ptag.clear();
  if (e != hipSuccess) return set_err(CF_EDEVICE, "profile_kernels: %s", hipGetErrorString(e));
  int n = 0;
  for (size_t q = 0; q < s->tags.size() && n < max; q++) {
    if (!s->tags[q].launches) continue;
    snprintf(names + 32 * n, 32, "%s", s->tags[q].name);
    ms[n] = (float)s->tags[q].ms;
    launches[n] = s->tags[q].launches;
    n++;
  }
  *count = n;
  return CF_OK;
}

int cf_sim_profile_frame(cf_sim* s, int reps, float* geometry_ms, float* radiation_ms) {
  if (!s || !geometry_ms || !radiation_ms || reps < 1) return set_err(CF_EINVAL, "bad argument");
  hipEvent_t ev[3];
  for (auto& e : ev) HIPCHK(hipEventCreate(&e));
  const int np = s->P.nPieces < CF_MAX_PIECES ? s->P.nPieces : CF_MAX_PIECES;
  float g = 0, r = 0;
  for (int q = 0; q < reps; q++) {
    HIPCHK(hipEventRecord(ev[0], s->st));
    cf::launch_fill_u32(s->F.faceW, 0, CF_MAX_PIECES * 4, s->st);
    cf::launch_geometry(s->F, s->st);
    HIPCHK(hipEventRecord(ev[1], s->st));
    cf::launch_radiation(s->F, np * 4 + CF_N_FLOOR, s->st);
    HIPCHK(hipEventRecord(ev[2], s->st));
    HIPCHK(hipEventSynchronize(ev[2]));
    float a = 0, b = 0;
    HIPCHK(hipEventElapsedTime(&a, ev[0], ev[1]));
    HIPCHK(hipEventElapsedTime(&b, ev[1], ev[2]));
    g += a;
    r += b;
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  *geometry_ms = g / reps;
  *radiation_ms = r / reps;
  return CF_OK;
}

int cf_sim_step_latencies(cf_sim* s, int nsteps, float* ms) {
  if (!s || !ms || nsteps < 0) return set_err(CF_EINVAL, "bad argument");
  std::vector<hipEvent_t> ev((size_t)nsteps + 1);
  for (auto& e : ev) HIPCHK(hipEventCreate(&e));
  if (s->use_graph) {
    int rc = ensure_graph(s);
    if (rc) return rc;
  }
  HIPCHK(hipEventRecord(ev[0], s->st));
  for (int q = 0; q < nsteps; q++) {
    if (s->use_graph && s->seg.head) {
      int rc = run_segments(s);
      if (rc) return rc;
    } else if (s->use_graph) {
      HIPCHK(hipGraphLaunch(s->gexec, s->st));
    } else {
      enqueue_step(s);
    }
    HIPCHK(hipEventRecord(ev[q + 1], s->st));
    if (s->sch.cfl <= 0) s->time += s->dt;
  }
  s->last_vcycles = s->sch.vcycles;
  HIPCHK(hipStreamSynchronize(s->st));
  for (int q = 0; q < nsteps; q++) hipEventElapsedTime(&ms[q], ev[q], ev[q + 1]);
  for (auto& e : ev) hipEventDestroy(e);
  return CF_OK;
}

int cf_sim_synchronize(cf_sim* s) {
  if (!s) return set_err(CF_EINVAL, "sim is NULL");
  HIPCHK(hipStreamSynchronize(s->st));
  return CF_OK;
}

int cf_sim_encode_readback(cf_sim* s) {  // js/fluid.js:315-324
  if (!s) return set_err(CF_EINVAL, "sim is NULL");
  const int i = s->stageIdx;
  if (s->busy[i]) return set_err(CF_EAGAIN, "readback ring slot %d busy", i);
  uint8_t* h = s->ring + (size_t)i * RING_BYTES;
  HIPCHK(hipMemcpyAsync(h, s->F.acc, CF_ACC_LEN * 4, hipMemcpyDeviceToHost, s->st));
  HIPCHK(hipMemcpyAsync(h + CF_ACC_LEN * 4, s->F.rad, CF_RAD_LEN * 16, hipMemcpyDeviceToHost, s->st));
  HIPCHK(hipMemcpyAsync(h + CF_ACC_LEN * 4 + CF_RAD_LEN * 16, s->F.faceW, CF_MAX_PIECES * 16, hipMemcpyDeviceToHost, s->st));
  cf::launch_fill_u32(s->F.acc, 0, CF_ACC_LEN, s->st);
  HIPCHK(hipEventRecord(s->ringEv[i], s->st));
  s->busy[i] = true;
  s->stageIdx = (i + 1) % 3;
  return i;
}

int cf_sim_take_readback(cf_sim* s, int slot, int wait, int32_t* acc, float* rad, int32_t* faceW) {  // js/fluid.js:327-342
  if (!s) return set_err(CF_EINVAL, "sim is NULL");
  if (slot < 0 || slot > 2 || !s->busy[slot]) return set_err(CF_EINVAL, "readback slot %d not pending", slot);
  if (wait) {
    HIPCHK(hipEventSynchronize(s->ringEv[slot]));
  } else {
    hipError_t e = hipEventQuery(s->ringEv[slot]);
    if (e == hipErrorNotReady) return set_err(CF_ENOTREADY, "readback slot %d not complete", slot);
    if (e != hipSuccess) { s->busy[slot] = false; return set_err(CF_EDEVICE, "readback: %s", hipGetErrorString(e)); }
  }
  const uint8_t* h = s->ring + (size_t)slot * RING_BYTES;
  if (acc) memcpy(acc, h, CF_ACC_LEN * 4);
  if (rad) memcpy(rad, h + CF_ACC_LEN * 4, CF_RAD_LEN * 16);
  if (faceW) memcpy(faceW, h + CF_ACC_LEN * 4 + CF_RAD_LEN * 16, CF_MAX_PIECES * 16);
  s->busy[slot] = false;
  return CF_OK;
}

int cf_sim_field_bytes(const cf_sim* s, const char* name, size_t* bytes) {
  if (!s || !name || !bytes) return set_err(CF_EINVAL, "NULL argument");
  FieldRef r;
  if (!lookup(const_cast<cf_sim*>(s), name, r)) return set_err(CF_EINVAL, "no buffer '%s'", name);
  *bytes = r.bytes;
  return CF_OK;
}

int cf_sim_read_field(cf_sim* s, const char* name, void* dst, size_t bytes) {  // js/fluid.js:345-355
  if (!s || !name || !dst) return set_err(CF_EINVAL, "NULL argument");
  FieldRef r;
  if (!lookup(s, name, r)) return set_err(CF_EINVAL, "no buffer '%s'", name);
  if (bytes != r.bytes) return set_err(CF_EINVAL, "'%s' is %zu bytes, got %zu", name, r.bytes, bytes);
  HIPCHK(hipMemcpyAsync(dst, r.ptr, bytes, hipMemcpyDeviceToHost, s->st));
  HIPCHK(hipStreamSynchronize(s->st));
  return CF_OK;
}

int cf_sim_write_field(cf_sim* s, const char* name, const void* src, size_t bytes) {
  if (!s || !name || !src) return set_err(CF_EINVAL, "NULL argument");
  FieldRef r;
  if (!lookup(s, name, r)) return set_err(CF_EINVAL, "no buffer '%s'", name);
  if (bytes != r.bytes) return set_err(CF_EINVAL, "'%s' is %zu bytes, got %zu", name, r.bytes, bytes);
  HIPCHK(hipMemcpyAsync(r.ptr, src, bytes, hipMemcpyHostToDevice, s->st));
  HIPCHK(hipStreamSynchronize(s->st));
  return CF_OK;
}

int cf_sim_get_info(const cf_sim* s, cf_sim_info* info) {
  if (!s || !info) return set_err(CF_EINVAL, "NULL argument");
  memset(info, 0, sizeof(*info));
  info->nx = s->nx; info->ny = s->ny; info->dx = s->dx; info->dt = s->dt;
  info->height = s->dx * s->ny; info->time = s->time; info->frame = s->frame;
  info->diff_sub = s->diffSub; info->nlevels = s->nlev;
  for (int l = 0; l < s->nlev; l++) { info->level_nx[l] = s->lnx[l]; info->level_ny[l] = s->lny[l]; }
  info->last_vcycles = s->last_vcycles; info->last_resid = s->last_resid;
  if (s->sch.cfl > 0 || s->sch.tol > 0) {
    // device-side control values of the last step
    cf::Dyn d;
    HIPCHK(hipStreamSynchronize(s->st));
    HIPCHK(hipMemcpy(&d, s->dDyn, sizeof(d), hipMemcpyDeviceToHost));
    if (s->sch.cfl > 0) { info->time = d.time; info->dt = d.dt; }
    if (s->sch.tol > 0) { info->last_vcycles = d.cycles; info->last_resid = d.resid; }
  }
  return CF_OK;
}

// the status of the steps run so far; reads and clears the non-finite flag
static int sim_status(cf_sim* s) {
  cf::Dyn d;
  HIPCHK(hipStreamSynchronize(s->st));
  HIPCHK(hipMemcpy(&d, s->dDyn, sizeof(d), hipMemcpyDeviceToHost));
  if (d.nonfinite) {
    cf::launch_fill_u32(&s->dDyn->nonfinite, 0u, 1, s->st);
    HIPCHK(hipStreamSynchronize(s->st));
    return set_err(CF_ENONFINITE, "non-finite values in the fields of the %dx%d scene (%s%s%s)", s->nx, s->ny,
                   d.nonfinite & 1 ? "T or Y" : "", d.nonfinite == 3 ? ", " : "", d.nonfinite & 2 ? "U or V" : "");
  }
  if (s->sch.tol > 0 && !d.converged)
    return set_err(CF_ENOCONV, "residual %.3g above the target %.3g after %d V-cycles (%dx%d scene)", d.resid,
                   s->sch.tol, d.cycles, s->nx, s->ny);
  return CF_OK;
}

int cf_sim_status(cf_sim* s) {
  if (!s) return set_err(CF_EINVAL, "sim is NULL");
  return sim_status(s);
}

int cf_sim_div_error(cf_sim* s, float* max_err, float* mean_err) {
  if (!s || !max_err || !mean_err) return set_err(CF_EINVAL, "NULL argument");
  const unsigned nb = cf::div_error_blocks(s->F);
  cf::launch_div_error(s->F, s->dPart, s->st);
  HIPCHK(hipMemcpyAsync(s->hPart.data(), s->dPart, (size_t)nb * 8, hipMemcpyDeviceToHost, s->st));
  HIPCHK(hipStreamSynchronize(s->st));
  double mx = 0, sum = 0;
  for (unsigned b = 0; b < nb; b++) { mx = fmax(mx, s->hPart[2 * b]); sum += s->hPart[2 * b + 1]; }
  *max_err = (float)mx;
  *mean_err = (float)(sum / s->nC);
  return CF_OK;
}

int cf_sim_copy_state(cf_sim* dst, const cf_sim* src) {
  if (!dst || !src) return set_err(CF_EINVAL, "NULL argument");
  if (dst->nx != src->nx || dst->ny != src->ny || dst->block_bytes != src->block_bytes)
    return set_err(CF_EINVAL, "copy_state: grids differ (%dx%d, %dx%d)", src->nx, src->ny, dst->nx, dst->ny);
  HIPCHK(hipStreamSynchronize(src->st));
  HIPCHK(hipMemcpyAsync(dst->block, src->block, src->block_bytes, hipMemcpyDeviceToDevice, dst->st));
  // the host mirror of what the copied parameter block holds
  memcpy(dst->pv, src->pv, sizeof(dst->pv));
  dst->P = src->P;
  dst->dsub = src->dsub;
  dst->diffSub = src->diffSub;
  dst->dt = src->dt;
  dst->time = src->time;
  dst->frame = src->frame;
  HIPCHK(hipStreamSynchronize(dst->st));
  return CF_OK;
}

int cf_sim_compat_sums(cf_sim* s, double* source, double* outflow) {
  if (!s || !source || !outflow) return set_err(CF_EINVAL, "NULL argument");
  const unsigned nb = cf::compat_blocks(s->F);
  std::vector<double> h((size_t)nb * 2);
  cf::launch_compat(s->F, s->dNorm, s->st);
  HIPCHK(hipMemcpyAsync(h.data(), s->dNorm, h.size() * sizeof(double), hipMemcpyDeviceToHost, s->st));
  HIPCHK(hipStreamSynchronize(s->st));
  double a = 0, b = 0;
  for (unsigned q = 0; q < nb; q++) { a += h[2 * q]; b += h[2 * q + 1]; }
  *source = a;
  *outflow = b;
  return CF_OK;
}

int cf_flame1d_run(const cf_flame_case* cases, int ncases, int fp64, cf_flame_result* out) {
  if (!cases || !out || ncases < 1) return set_err(CF_EINVAL, "bad argument");
  std::vector<cf_flame_case> c(cases, cases + ncases);
  for (auto& k : c) {
    const int rc = cf::check_flame_case(&k);
    if (rc) return rc;
    if (k.max_steps <= 0) k.max_steps = 10000000;
  }
  cf_flame_case* dc = nullptr;
  cf_flame_result* dr = nullptr;
  hipStream_t st = nullptr;
  int rc = CF_OK;
  hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc((void**)&dc, sizeof(cf_flame_case) * ncases);
  if (e == hipSuccess) e = hipMalloc((void**)&dr, sizeof(cf_flame_result) * ncases);
  if (e == hipSuccess) e = hipMemcpyAsync(dc, c.data(), sizeof(cf_flame_case) * ncases, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) {
    cf::launch_flame1d(dc, ncases, fp64 != 0, dr, st);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(out, dr, sizeof(cf_flame_result) * ncases, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) rc = set_err(CF_EDEVICE, "flame1d: %s", hipGetErrorString(e));
  if (dc) hipFree(dc);
  if (dr) hipFree(dr);
  if (st) hipStreamDestroy(st);
  return rc;
}

int cf_copy_bandwidth(uint64_t bytes, int reps, double* gbps, double* ms) {
  if (!gbps || reps < 1 || bytes < 4096 || bytes % 16) return set_err(CF_EINVAL, "bad argument");
  void *a = nullptr, *b = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = CF_OK;
  hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc(&a, bytes);
  if (e == hipSuccess) e = hipMalloc(&b, bytes);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  float best = 1e30f, total = 0.f;
  if (e == hipSuccess) {
    cf::launch_fill_u32(a, 0x3f800000u, bytes / 4, st);
    cf::launch_bw_copy(a, b, bytes, st);  // warm-up
    for (int q = 0; q < reps && e == hipSuccess; q++) {
      e = hipEventRecord(e0, st);
      cf::launch_bw_copy(a, b, bytes, st);
      if (e == hipSuccess) e = hipEventRecord(e1, st);
      if (e == hipSuccess) e = hipEventSynchronize(e1);
      float t = 0.f;
      if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
      best = t < best ? t : best;
      total += t;
    }
  }
  if (e != hipSuccess) rc = set_err(CF_EDEVICE, "copy bandwidth: %s", hipGetErrorString(e));
  else {
    *gbps = 2.0 * (double)bytes / (best * 1e-3) / 1e9;
    if (ms) *ms = total / reps;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  if (st) (void)hipStreamDestroy(st);
  return rc;
}

uint64_t cf_scene_bytes(int nx, int ny, double dx) {
  cf_sim* s = new cf_sim();
  uint64_t total = 0;
  if (sim_levels(s, nx, ny, dx, 1.0) == CF_OK) {
    std::vector<std::pair<void**, size_t>> plan;
    sim_plan(s, plan);
    total = plan_total(plan);
  }
  delete s;
  return total;
}

int cf_pool_create(uint64_t hbm_budget, cf_pool** out) {
  if (!out) return set_err(CF_EINVAL, "out is NULL");
  *out = nullptr;
  cf_pool* pl = new cf_pool();
  pl->budget = hbm_budget;
  pl->sch.nu = 2; pl->sch.vcycles = 1; pl->sch.tol = 0; pl->sch.coarse_iters = 120; pl->sch.smoother_fp16 = 0;
  pl->sch.cfl = 0; pl->sch.umin = 1.0f; pl->sch.refine = 0; pl->sch.krylov = 0;
  const char* ng = getenv("CF_NO_GRAPH");
  pl->use_graph = !(ng && *ng && *ng != '0');
  hipError_t e = hipStreamCreateWithFlags(&pl->st, hipStreamNonBlocking);
  if (e != hipSuccess) { delete pl; return set_err(CF_EDEVICE, "hipStreamCreate: %s", hipGetErrorString(e)); }
  *out = pl;
  return CF_OK;
}

void cf_pool_destroy(cf_pool* pl) {
  if (!pl) return;
  hipStreamSynchronize(pl->st);
  pool_drop_graph(pl);
  pool_release_phases(pl->step);
  pool_release_phases(pl->frame_pre);
  pool_release_phases(pl->frame_post);
  pl->arena.free();
  pl->farena.free();
  if (pl->dflag) (void)hipFree(pl->dflag);
  if (pl->hflag) hipHostFree(pl->hflag);
  for (auto* s : pl->slots) if (s) cf_sim_destroy(s);
  hipStreamDestroy(pl->st);
  delete pl;
}

int cf_pool_set_schedule(cf_pool* pl, const cf_schedule* sc) {
  if (!pl || !sc) return set_err(CF_EINVAL, "NULL argument");
  // validate with a scene's rules even when the pool is empty
  if (sc->nu < 1 || sc->nu > 3 || sc->vcycles < 1 || sc->coarse_iters < 0 || !(sc->tol >= 0) || !(sc->cfl >= 0) ||
      (sc->refine && !(sc->tol > 0)) || (sc->smoother_fp16 && !sc->refine) || (sc->krylov && !sc->refine))
    return set_err(CF_EINVAL, "bad schedule");
  // every scene must accept it before any of them changes
  for (auto* s : pl->slots)
    if (s) { int rc = sched_check(s, sc); if (rc) return rc; }
  for (auto* s : pl->slots)
    if (s) { int rc = sched_apply(s, sc); if (rc) return rc; }
  pl->sch = *sc;
  pl->dirty = true;
  return CF_OK;
}

int cf_pool_set_leg_tiles(cf_pool* pl, const int32_t* tiles, int n) {
  if (!pl || (n > 0 && !tiles) || n < 0) return set_err(CF_EINVAL, "bad argument");
  if (n > MAXLEV) return set_err(CF_EINVAL, "%d tile entries for at most %d depths", n, MAXLEV);
  for (int d = 0; d < n; d++)
    if (tiles[d] != 0 && tiles[d] != 16 && tiles[d] != 32 && tiles[d] != 48 && tiles[d] != 64)
      return set_err(CF_EINVAL, "leg tile %d of depth %d is not 0, 16, 32, 48 or 64", tiles[d], d);
  for (int d = 0; d < MAXLEV; d++) {
    const int t = d < n ? tiles[d] : 0;
    if (pl->tile_over[d] != t) { pl->tile_over[d] = t; pl->dirty = true; }
  }
  return CF_OK;
}

int cf_pool_get_leg_tiles(const cf_pool* pl, int32_t* tiles, int n) {
  if (!pl || (n > 0 && !tiles) || n < 0) return set_err(CF_EINVAL, "bad argument");
  std::vector<cf_sim*> sc;
  for (auto* s : pl->slots) if (s) sc.push_back(s);
  int t[MAXLEV];
  const int maxb = pool_depth_tiles(pl, sc, t);
  for (int d = 0; d < n; d++) tiles[d] = d < MAXLEV ? t[d] : 0;
  return maxb;
}

int cf_pool_admit(cf_pool* pl, const cf_scene_desc* d, int* slot) {
  if (!pl || !d || !slot) return set_err(CF_EINVAL, "NULL argument");
  const uint64_t need = cf_scene_bytes(d->nx, d->ny, d->dx);
  if (!need) return set_err(CF_EBADSIZE, "pool: unsupported grid %dx%d", d->nx, d->ny);
  if (pl->used + need > pl->budget)
    return set_err(CF_EAGAIN, "pool: %llu bytes needed, %llu of %llu in use", (unsigned long long)need,
                   (unsigned long long)pl->used, (unsigned long long)pl->budget);
  cf_sim* s = nullptr;
  int rc = sim_create(d->nx, d->ny, d->dx, d->dt, pl->st, &s);
  if (rc) return rc;
  rc = sched_check(s, &pl->sch);
  if (!rc) rc = sched_apply(s, &pl->sch);
  if (rc) { cf_sim_destroy(s); return rc; }
  s->in_pool = true;
  int q = 0;
  while (q < (int)pl->slots.size() && pl->slots[q]) q++;
  if (q == (int)pl->slots.size()) pl->slots.push_back(nullptr);
  pl->slots[q] = s;
  pl->used += s->block_bytes;
  pl->dirty = true;
  pl->gen++;
  *slot = q;
  return CF_OK;
}

int cf_pool_scene(cf_pool* pl, int slot, cf_sim** sim) {
  if (!pl || !sim) return set_err(CF_EINVAL, "NULL argument");
  if (slot < 0 || slot >= (int)pl->slots.size() || !pl->slots[slot]) return set_err(CF_EINVAL, "pool: no scene in slot %d", slot);
  *sim = pl->slots[slot];
  return CF_OK;
}

int cf_pool_release(cf_pool* pl, int slot) {
  if (!pl) return set_err(CF_EINVAL, "pool is NULL");
  if (slot < 0 || slot >= (int)pl->slots.size() || !pl->slots[slot]) return set_err(CF_EINVAL, "pool: no scene in slot %d", slot);
  HIPCHK(hipStreamSynchronize(pl->st));
  pl->used -= pl->slots[slot]->block_bytes;
  cf_sim_destroy(pl->slots[slot]);
  pl->slots[slot] = nullptr;
  pl->dirty = true;
  pl->gen++;
  // the per-frame tables hold the released scene's device pointers
  pool_release_phases(pl->frame_pre);
  pool_release_phases(pl->frame_post);
  return CF_OK;
}

int cf_pool_prepare(cf_pool* pl) {
  if (!pl) return set_err(CF_EINVAL, "pool is NULL");
  for (auto* s : pl->slots)
    if (s) return pool_ready(pl);
  return CF_OK;
}

int cf_pool_step(cf_pool* pl, int nsteps) {
  if (!pl) return set_err(CF_EINVAL, "pool is NULL");
  if (nsteps < 0) return set_err(CF_EINVAL, "nsteps < 0");
  int live = 0;
  for (auto* s : pl->slots) live += s ? 1 : 0;
  if (!live || !nsteps) return CF_OK;
  int rc = pool_ready(pl);
  if (rc) return rc;
  int q = 0;
  if (pl->use_graph && pl->seg.head) {
    for (; q < nsteps; q++)
      if ((rc = pool_run_segments(pl))) return rc;
  } else if (pl->use_graph) {
    for (; q + 8 <= nsteps; q += 8) HIPCHK(hipGraphLaunch(pl->g8, pl->st));
    for (; q < nsteps; q++) HIPCHK(hipGraphLaunch(pl->g1, pl->st));
  } else {
    for (; q < nsteps; q++) pool_enqueue_step(pl);
    HIPCHK(hipGetLastError());
  }
  for (auto* s : pl->slots)
    if (s && s->sch.cfl <= 0) s->time += s->dt * nsteps;
  pl->steps += nsteps;
  return CF_OK;
}

// one app frame for every scene (js/main.js:266-283): parameters, geometry,
// nsteps steps, radiation; geometry and radiation are batched too.  As with
// cf_sim_frame, the caller advances each scene's frame counter.
int cf_pool_frame(cf_pool* pl, int nsteps, int with_radiation) {
  if (!pl) return set_err(CF_EINVAL, "pool is NULL");
  std::vector<cf_sim*> sc;
  for (auto* s : pl->slots) if (s) sc.push_back(s);
  if (sc.empty()) return CF_OK;
  for (auto* s : sc) { int rc = cf_sim_write_params(s); if (rc) return rc; }
  // per-frame tables: rebuilt when the piece counts or the scenes change
  uint64_t key = (1469598103934665603ull ^ pl->gen) * 1099511628211ull;
  for (auto* s : pl->slots) {
    key = (key ^ (uint64_t)(uintptr_t)s) * 1099511628211ull;
    if (s) key = (key ^ (uint64_t)s->P.nPieces) * 1099511628211ull;
  }
  if (key != pl->fkey || pl->frame_pre.empty()) {
    HIPCHK(hipStreamSynchronize(pl->st));
    pool_release_phases(pl->frame_pre);
    pool_release_phases(pl->frame_post);
    pl->frame_pre.resize(2);
    pl->frame_post.resize(2);
    for (auto* s : sc) {
      cf::phase_fill_u32(pl->frame_pre[0], s->F.faceW, 0u, CF_MAX_PIECES * 4);
      cf::phase_geometry(pl->frame_pre[1], s->F);
      const int np = s->P.nPieces < CF_MAX_PIECES ? s->P.nPieces : CF_MAX_PIECES;
      cf::phase_gas_terms(pl->frame_post[0], s->F);
      cf::phase_radiation(pl->frame_post[1], s->F, np * 4 + CF_N_FLOOR);
    }
    std::vector<cf::BatchPhase*> all;
    for (auto& a : pl->frame_pre) all.push_back(&a);
    for (auto& a : pl->frame_post) all.push_back(&a);
    hipError_t e = pl->farena.place(all, pl->st);
    if (e != hipSuccess) return set_err(CF_ENOMEM, "pool frame tables: %s", hipGetErrorString(e));
    pl->fkey = key;
  }
  for (auto& a : pl->frame_pre) a.run(pl->st);
  int rc = cf_pool_step(pl, nsteps);
  if (rc) return rc;
  if (with_radiation) for (auto& a : pl->frame_post) a.run(pl->st);
  HIPCHK(hipGetLastError());
  return CF_OK;
}

int cf_pool_synchronize(cf_pool* pl) {
  if (!pl) return set_err(CF_EINVAL, "pool is NULL");
  HIPCHK(hipStreamSynchronize(pl->st));
  return CF_OK;
}

int cf_pool_status(cf_pool* pl, int* slot) {
  if (!pl) return set_err(CF_EINVAL, "pool is NULL");
  if (slot) *slot = -1;
  HIPCHK(hipStreamSynchronize(pl->st));
  int first = CF_OK;
  char msg[512] = "";
  for (size_t q = 0; q < pl->slots.size(); q++) {
    if (!pl->slots[q]) continue;
    const int rc = sim_status(pl->slots[q]);  // every scene's flag is read and cleared
    if (rc && rc != CF_ENOCONV && rc != CF_ENONFINITE) return rc;
    // a non-finite scene outranks one that only missed its residual target
    if (rc && (first == CF_OK || (rc == CF_ENONFINITE && first == CF_ENOCONV))) {
      first = rc;
      snprintf(msg, sizeof(msg), "slot %d: %s", (int)q, g_err);
      if (slot) *slot = (int)q;
    }
  }
  if (first) return set_err(first, "%s", msg);
  return CF_OK;
}

int cf_pool_get_info(const cf_pool* pl, cf_pool_info* info) {
  if (!pl || !info) return set_err(CF_EINVAL, "NULL argument");
  memset(info, 0, sizeof(*info));
  for (auto* s : pl->slots) {
    if (!s) continue;
    info->nscenes++;
    info->cells += (uint64_t)s->nC;
  }
  info->nslots = (int32_t)pl->slots.size();
  info->budget = pl->budget;
  info->used = pl->used;
  info->steps = pl->steps;
  info->launches_per_step = pl->nlaunch;
  return CF_OK;
}

int cf_pool_step_latencies(cf_pool* pl, int nsteps, float* ms) {
  if (!pl || !ms || nsteps < 0) return set_err(CF_EINVAL, "bad argument");
  int rc = pool_ready(pl);
  if (rc) return rc;
  std::vector<hipEvent_t> ev((size_t)nsteps + 1);
  for (auto& e : ev) HIPCHK(hipEventCreate(&e));
  HIPCHK(hipEventRecord(ev[0], pl->st));
  for (int q = 0; q < nsteps; q++) {
    if (pl->use_graph && pl->seg.head) {
      rc = pool_run_segments(pl);
      if (rc) return rc;
    } else if (pl->use_graph) {
      HIPCHK(hipGraphLaunch(pl->g1, pl->st));
    } else {
      pool_enqueue_step(pl);
    }
    HIPCHK(hipEventRecord(ev[q + 1], pl->st));
  }
  HIPCHK(hipStreamSynchronize(pl->st));
  for (int q = 0; q < nsteps; q++) hipEventElapsedTime(&ms[q], ev[q], ev[q + 1]);
  for (auto& e : ev) hipEventDestroy(e);
  for (auto* s : pl->slots)
    if (s && s->sch.cfl <= 0) s->time += s->dt * nsteps;
  pl->steps += nsteps;
  return CF_OK;
}

}  // extern "C"
