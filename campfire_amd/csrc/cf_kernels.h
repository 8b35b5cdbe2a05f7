// cf_kernels.h — device buffers of one scene and the kernel launchers.
#pragma once
#include <hip/hip_ext.h>
#include <hip/hip_fp16.h>
#include <string.h>
#include <type_traits>
#include <vector>
#include "cf_device.h"

namespace cf {

// Start/stop events of one launch, recorded by the runtime around the
// dispatch packet itself, so their interval is the kernel's execution alone.
struct LaunchEvents {
  hipEvent_t start, stop;
};
// While set (a profiling pass of the host library, on its calling thread),
// every launch appends its events here.
extern thread_local std::vector<LaunchEvents>* prof_sink;

// ---- batched launches ------------------------------------------------------------
// A kernel's work is written once as Op::run(args, tile): what one workgroup
// does for tile `tile` of one scene.  The single-scene kernels call it with
// their (XCD-remapped) block index; k_batch runs many scenes of any sizes in
// one launch from a table of work items, each naming an argument block (one
// per scene) and a tile of it.  A scene's tiles and their arithmetic are the
// same either way, so its results do not depend on what it is batched with.
struct Item {
  int arg, tile;
};

// the waves per SIMD a batched kernel must fit (Op::kMinWaves, else 1): its
// register budget, so the batched form keeps the single-scene kernel's
// occupancy (the argument-block indirection costs the batched form registers)
template <class Op, class = void>
struct MinWaves {
  static constexpr int v = 1;
};
template <class Op>
struct MinWaves<Op, std::void_t<decltype(Op::kMinWaves)>> {
  static constexpr int v = Op::kMinWaves;
};

template <class Op>
__global__ __launch_bounds__(Op::kThreads, MinWaves<Op>::v) void k_batch(const typename Op::Args* __restrict__ args,
                                                                         const Item* __restrict__ items, int n) {
  const Item it = items[xcd_tile(blockIdx.x, n)];
  Op::run(args[it.arg], it.tile);
}

template <typename... KArgs, typename... Args>
inline void launch(void (*k)(KArgs...), dim3 grid, dim3 block, unsigned shmem, hipStream_t s, Args... args);

// One batched launch: every scene's argument block and its work items,
// collected on the host (add) and copied to the device once (upload); run()
// then replays the launch, also under graph capture.
struct BatchPhase {
  void (*fn)(const BatchPhase&, hipStream_t) = nullptr;
  dim3 block;
  size_t asize = 0;
  int nargs = 0;
  std::vector<unsigned char> hargs;
  std::vector<Item> hitems;
  void* dargs = nullptr;
  Item* ditems = nullptr;
  const char* tag = "";
  // set when a scene of another kernel was offered: one launch runs one
  // kernel over every item, so such a phase must not run
  bool mixed = false;

  template <class A>
  void add(const A& a, int ntiles) {
    asize = sizeof(A);
    const size_t off = hargs.size();
    hargs.resize(off + sizeof(A));
    memcpy(&hargs[off], &a, sizeof(A));
    for (int t = 0; t < ntiles; t++) hitems.push_back(Item{nargs, t});
    nargs++;
  }
  bool empty() const { return hitems.empty(); }
  bool owned = false;  // dargs/ditems allocated by upload() (else they live in a TableArena)
  hipError_t upload(hipStream_t s) {
    if (empty()) return hipSuccess;
    hipError_t e = hipMalloc(&dargs, hargs.size());
    if (e == hipSuccess) e = hipMalloc((void**)&ditems, hitems.size() * sizeof(Item));
    owned = true;
    if (e == hipSuccess) e = hipMemcpyAsync(dargs, hargs.data(), hargs.size(), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(ditems, hitems.data(), hitems.size() * sizeof(Item), hipMemcpyHostToDevice, s);
    return e;
  }
  void release() {
    if (owned) {
      if (dargs) (void)hipFree(dargs);
      if (ditems) (void)hipFree(ditems);
    }
    dargs = nullptr;
    ditems = nullptr;
    owned = false;
  }
  void run(hipStream_t s) const {
    if (fn && !empty() && !mixed) fn(*this, s);
  }
};

// Device storage for the tables of a set of phases: one buffer that only
// grows, filled by one host-to-device copy per rebuild, so rebuilding a pool's
// step allocates nothing once the buffer is large enough.  The caller
// guarantees that no launch still reads the old tables (a stream sync).
struct TableArena {
  void* dev = nullptr;
  size_t cap = 0;
  std::vector<unsigned char> host;  // kept until the next place(): the copy's source
  static size_t al(size_t b) { return (b + 255) & ~(size_t)255; }
  // lays out every phase's argument blocks and items, plus `nextra` bytes of
  // `extra` (returned in *extra_dev), and points the phases at them
  hipError_t place(const std::vector<BatchPhase*>& phases, hipStream_t s, const void* extra = nullptr,
                   size_t nextra = 0, void** extra_dev = nullptr) {
    size_t total = al(nextra);
    for (BatchPhase* p : phases)
      if (!p->empty()) total += al(p->hargs.size()) + al(p->hitems.size() * sizeof(Item));
    if (total == 0) return hipSuccess;
    if (total > cap) {
      if (dev) (void)hipFree(dev);
      dev = nullptr;
      const size_t want = total + total / 2;
      hipError_t e = hipMalloc(&dev, want);
      if (e != hipSuccess) { cap = 0; return e; }
      cap = want;
    }
    host.assign(total, 0);
    size_t off = 0;
    unsigned char* d = (unsigned char*)dev;
    if (nextra) {
      memcpy(&host[0], extra, nextra);
      if (extra_dev) *extra_dev = d;
      off = al(nextra);
    }
    for (BatchPhase* p : phases) {
      p->release();
      if (p->empty()) continue;
      memcpy(&host[off], p->hargs.data(), p->hargs.size());
      p->dargs = d + off;
      off += al(p->hargs.size());
      memcpy(&host[off], p->hitems.data(), p->hitems.size() * sizeof(Item));
      p->ditems = (Item*)(d + off);
      off += al(p->hitems.size() * sizeof(Item));
    }
    return hipMemcpyAsync(dev, host.data(), total, hipMemcpyHostToDevice, s);
  }
  void free() {
    if (dev) (void)hipFree(dev);
    dev = nullptr;
    cap = 0;
  }
};

template <class Op>
void run_batch(const BatchPhase& ph, hipStream_t s) {
  launch(k_batch<Op>, dim3((unsigned)ph.hitems.size()), ph.block, 0, s, (const typename Op::Args*)ph.dargs, ph.ditems,
         (int)ph.hitems.size());
}

// start a phase of Op (or check that it is one) and add one scene
template <class Op>
void phase_add(BatchPhase& ph, const typename Op::Args& a, int ntiles, dim3 block, const char* tag) {
  void (*fn)(const BatchPhase&, hipStream_t) = &run_batch<Op>;
  if (ph.fn && ph.fn != fn) {
    ph.mixed = true;
    return;
  }
  ph.fn = fn;
  ph.block = block;
  ph.tag = tag;
  ph.add(a, ntiles);
}

template <typename... KArgs, typename... Args>
inline void launch(void (*k)(KArgs...), dim3 grid, dim3 block, unsigned shmem, hipStream_t s, Args... args) {
  if (prof_sink) {
    LaunchEvents e{nullptr, nullptr};
    if (hipEventCreate(&e.start) == hipSuccess && hipEventCreate(&e.stop) == hipSuccess) {
      hipExtLaunchKernelGGL(k, grid, block, shmem, s, e.start, e.stop, 0, args...);
      prof_sink->push_back(e);
      return;
    }
  }
  hipLaunchKernelGGL(k, grid, block, shmem, s, args...);
}

// Per-scene control values in device memory, right after the parameter block.
// dsub and dt arrive with every parameter upload (fixed dt) or are written by
// the CFL kernel at the start of a step; the rest is scratch of the control
// kernels.
struct Dyn {
  float dsub;          // diffusion substep dt (DiffParams.dt, js/fluid.js:196-198)
  float dt;            // this step's dt (mirrors the parameter block's dt)
  uint32_t umax_bits;  // CFL: max |u| over the faces, as float bits (atomicMax)
  uint32_t cfl_done;   // CFL: workgroups finished
  int32_t converged;   // V-cycle loop: relative residual below tol this step
  int32_t cycles;      // V-cycles run this step
  float resid;         // relative residual after the last V-cycle
  float sig, isig;     // FP16 smoother: p is stored as p * sig (a power of two), read as h * isig
  double time;         // simulated time accumulated by the CFL kernel
  // Krylov (schedule krylov = 1): ||b||^2; (d, B d) of the stored direction
  // (0: none yet); this iteration's step and direction coefficients
  double kbb, kdq, kalpha, kbeta;
  double kbest;        // Krylov: lowest relative residual of this solve
  int32_t kfall;       // Krylov: fallen back to defect correction for this solve
  // non-finite values seen since the host last read this (bit 0: T or Y
  // entering the sources pass, bit 1: U or V in the CFL pass)
  int32_t nonfinite;
};

// Every device buffer of one scene (the `buf` dictionary of js/fluid.js:52-68
// plus the level-0 multigrid arrays).  Passed by value as a kernel argument.
struct Fields {
  const cf_params* P;  // the 160-B parameter block in device memory
  Dyn* dyn;            // control values right after it
  int nx, ny;
  float *U, *Uh, *U2, *V, *Vh, *V2, *T, *Th, *T2;
  float4 *Y, *Yh, *Y2;
  float2 *SQ, *cellG;
  float *fx, *fy, *pieces;
  int *acc, *faceW;
  float4* rad;
  float *p0, *rhs0, *ax0, *ay0;
  float2* G;  // radiation: each cell's gas terms of the frame (emission, visible emission)
};

// The parameter block, read through the constant address space: it does not
// change while a kernel runs, so the compiler reads it with scalar loads into
// SGPRs also where the pointer itself comes from memory (a batched kernel's
// argument block), where it would otherwise load it into some 40 VGPRs.
__device__ __forceinline__ cf_params load_params(const cf_params* p) {
  static_assert(sizeof(cf_params) % 4 == 0, "whole words");
  const __attribute__((address_space(4))) int* q = (const __attribute__((address_space(4))) int*)(uintptr_t)p;
  cf_params r;
  int* w = reinterpret_cast<int*>(&r);
#pragma unroll
  for (int k = 0; k < (int)(sizeof(cf_params) / 4); k++) w[k] = q[k];
  return r;
}

// One multigrid level (the `Level` uniform of shaders/pressure.wgsl:7 and its arrays).
struct Level {
  int nx, ny;
  float h2;
  float *p, *rhs, *ax, *ay;
  // levels smoothed by tiled legs: p between the down and the up leg.  A
  // tiled leg reads p over its tile plus a halo that neighbouring tiles own,
  // so it must not write p in place: a neighbour that starts late would read
  // smoothed values in its halo.  The down leg writes q, the up leg reads q
  // and writes p.
  float* q;
  int tile;  // edge of the tiled legs' tiles at this level (16, 32, 64); 0: the size rule of mg_tile
};

// FP16 storage of a level for the FP16 smoother (schedule smoother_fp16 = 1):
// p as p * dyn->sig, the face coefficients as they are.  The level's FP32
// arrays stay: rhs is always FP32, and the FP32 coefficients feed the
// coefficient restriction and the FP64 defect.
struct LevelH {
  __half *p, *ax, *ay;
  __half* q;  // as Level::q
};

// storage of the two levels a tiled multigrid leg touches
struct MgStore {
  int mode;       // 0: both FP32; 1: fine FP16, coarse FP32; 3: both FP16
  LevelH f, c;    // FP16 arrays of the fine and the coarse level (mode bits 0 and 1)
  const Dyn* dyn; // sig, isig (mode != 0)
};

// advection in one pass (cf_advect.hip); the two launches below are the
// reference's dispatch pairs and also write the forward fields Th, Yh, Uh, Vh
void launch_advect(const Fields& F, hipStream_t s);
void launch_advect_fwd(const Fields& F, hipStream_t s);
void launch_advect_corr(const Fields& F, hipStream_t s);
// one diffusion substep of length dyn->dsub
void launch_diffuse(const Fields& F, const float* Tin, const float4* Yin, float* Tout, float4* Yout, hipStream_t s);
// sources (+ fused last diffusion substep) and the fold of its per-workgroup
// accumulator partials (part: 4 ints per workgroup, sources_blocks() of them)
unsigned sources_blocks(const Fields& F);
void launch_sources(const Fields& F, bool fuse, const float* Tin, const float4* Yin, int* part, bool finish,
                    hipStream_t s);
void launch_forces(const Fields& F, hipStream_t s);
void launch_coef_rhs(const Fields& F, hipStream_t s);
// forces + coef + rhs0 + coefficient restriction of the first
// forces_coef_cascade(nlev) coarse levels, in one pass; its first workgroup
// also folds nparts sources partials into acc (nparts = 0: none pending)
int forces_coef_cascade(int nlev);
void launch_forces_coef(const Fields& F, const Level* L, int nlev, const int* part, int nparts, hipStream_t s);
// fused multigrid (cf_mg.hip); skip (may be null): the launch does nothing
// once *skip is non-zero (the residual target of the step has been met)
// zero_p: the leg starts from p = 0 without reading the fine p (a level above
// 0, which holds the zero the restriction into it wrote, or a Krylov
// iteration's level-0 correction, which kry_upd_def does not clear)
bool launch_mg_down(const Level& f, const Level& c, const MgStore& st, int nu, const int* skip, hipStream_t s,
                    int zero_p = 0);
bool launch_mg_up(const Level& f, const Level& c, const MgStore& st, int nu, const int* skip, hipStream_t s);
// the level-0 up leg of a step's last V-cycle with the projection of the
// velocities fused in (FP32 storage; replaces the k_project launch)
bool launch_mg_up_project(const Level& f, const Level& c, const Fields& F, int nu, const int* skip, hipStream_t s);
// the legs of two levels f, c = f + 1 in one launch (FP32 storage), n = f + 2;
// tile: the edge of a workgroup's tile in cells of c (16 or 8; 0: mg_pair_tile)
bool launch_mg_down2(const Level& f, const Level& c, const Level& n, int nu, int tile, const int* skip,
                     hipStream_t s);
bool launch_mg_up2(const Level& f, const Level& c, const Level& n, int nu, int tile, const int* skip, hipStream_t s);
int mg_pair_tile(const Level& c);
// FP16 copies of a level's face coefficients (once per step, FP16 smoother)
void launch_coef_half(const Level& L, const LevelH& H, hipStream_t s);
// tile code of a level's legs (64, 32, 16: square tiles of that edge; 48: the
// wide tiles, 64 cells wide): Level::tile, or by default a rule on the level's
// size (on the tile count of a pool's depth); a batched phase holds one tile code
int mg_tile(const Level& f);
int mg_tile_rule(const Level& f);
int mg_tile_for_count(int n64, int n32);
// first level of the single-workgroup bottom solver (the last 1-3 levels)
int mg_bottom_first(const Level* L, int nlev);
// coarse_coef: restrict the coarsest level's coefficients inside the kernel
// coarse_iters = 0: direct FP64 solve of the coarsest level (<= 32 cells wide)
// in `band` (scratch of (256 * 33 + 256) doubles)
bool launch_mg_bottom(const Level* L, int nlev, int first, bool coarse_coef, int nu, int coarse_iters, const int* skip,
                      double* band, hipStream_t s);
// step control (cf_control.hip)
// dt = cfl * dx / max(max|u|, umin) into the parameter block and dyn, dsub =
// dt / nsub, time += dt
void launch_cfl(const Fields& F, float cfl, float umin, int nsub, hipStream_t s);
// ||r|| / ||rhs|| of level 0 after a V-cycle into dyn->resid (FP64 sums);
// sets dyn->converged when below tol.  part: 2 doubles per workgroup
unsigned resnorm_blocks(const Level& L0);
// defect correction at level 0: r = b - A p64 (FP64) into L0.rhs, L0.p = 0
// (ph0 non-null: its FP16 p instead), ||r||/||b|| and the FP16 scale into
// dyn; then, after a V-cycle on the correction, p64 += L0.p (or ph0 * isig)
// Krylov iteration (schedule krylov = 1): flexible CG at level 0 with the
// V-cycle as preconditioner.  After a V-cycle left z in p (or ph0 * isig),
// kry_step runs kry_dots (the dot products with B z), kry_coef (alpha,
// beta), kry_update (the direction and the FP64 pressure), then dc_defect
// and resid_fold (the residual from the pressure, the next V-cycle's
// right-hand side, the convergence flag).  part holds 3 doubles per
// workgroup.  A residual above twice the lowest of the solve so far switches
// the rest of the solve to defect correction (alpha = 1, beta = 0), which
// stalls where CG diverges: at a right-hand side outside the range of a
// singular system.
//
// With `fused` (the V-cycle's level-0 down leg starting from zero itself),
// kry_update and dc_defect run as one pass, kry_upd_def: it writes the new
// direction and pressure into the other half of p64 and d32 (2n entries
// each), the current half given by the parity of the iterations done
// (Dyn.cycles), and kry_settle copies them back into the first half after
// the last iteration of a step, where every other reader finds them.
struct KryBufs {
  double* p64;  // the pressure (2 n: the halves kry_upd_def alternates)
  float* d32;   // the stored direction (FP32: the residual is recomputed from p64; 2 n)
  float* b0;    // the saved right-hand side
};
void launch_kry_step(const Level& L0, __half* ph0, Dyn* dyn, const KryBufs& k, double* part, float tol, bool fused,
                     hipStream_t s);
void launch_kry_settle(const Level& L0, Dyn* dyn, const KryBufs& k, hipStream_t s);
void phase_kry_settle(BatchPhase& ph, const Level& L0, Dyn* dyn, const KryBufs& k);
// *out = 1 when the converged flag of each of the n scenes is set, else 0
void launch_all_converged(Dyn* const* d, int n, int* out, hipStream_t s);
// the kernels of launch_kry_start / launch_kry_step, one phase each
void phase_kry_step(BatchPhase (&ph)[5], const Level& L0, __half* ph0, Dyn* dyn, const KryBufs& k, double* part,
                    float tol, bool fused);
void launch_dc_defect(const Level& L0, __half* ph0, double* p64, float* b0, Dyn* dyn, double* part, float tol,
                      bool first, hipStream_t s);
// the Krylov start: a line search along the previous pressure (in p64) from
// zero, which stays as the stored direction
void launch_kry_start(const Level& L0, __half* ph0, Dyn* dyn, const KryBufs& k, double* part, float tol,
                      hipStream_t s);
void launch_dc_accum(const Level& L0, const __half* ph0, const Dyn* dyn, double* p64, const int* skip, hipStream_t s);
void launch_p32_to_64(const float* p, double* p64, int n, hipStream_t s);
void launch_resnorm(const Level& L0, Dyn* dyn, double* part, float tol, hipStream_t s);
void launch_coef_restrict(const Level& f, const Level& c, hipStream_t s);
void launch_relax(const Level& L, int color, hipStream_t s);
void launch_restrict_res(const Level& f, const Level& c, hipStream_t s);
void launch_prolong(const Level& f, const Level& c, hipStream_t s);
void launch_coarsest(const Level& L, int iters, hipStream_t s);
// p64 (may be null): gradients from the FP64 pressure of defect correction
void launch_project(const Fields& F, const double* p64, hipStream_t s);
void launch_geometry(const Fields& F, hipStream_t s);
// radiation: the gas terms of every cell into F.G, then the rays of ntargets targets
void launch_radiation(const Fields& F, int ntargets, hipStream_t s);
void launch_fill_u32(void* a, uint32_t v, size_t n, hipStream_t s);
// a streaming copy of `bytes` (a multiple of 16): the bandwidth reference
void launch_bw_copy(const void* a, void* b, size_t bytes, hipStream_t s);
// tests/flame1d.mjs on the GPU (cf_flame1d.hip): one workgroup per case
void launch_flame1d(const cf_flame_case* cases, int ncases, bool fp64, cf_flame_result* out, hipStream_t s);
// the checks of one flame case (CF_OK, or a code with the message set)
int check_flame_case(const cf_flame_case* k);
// sets cf_last_error()'s text and returns code
int set_error(int code, const char* msg);
unsigned div_error_blocks(const Fields& F);
// FP64 net source and open-boundary outflow: 2 doubles per workgroup
unsigned compat_blocks(const Fields& F);
void launch_compat(const Fields& F, double* part, hipStream_t s);
void launch_div_error(const Fields& F, float* part, hipStream_t s);

// ---- batched phases: add one scene's launch of each kernel to a phase ----
// (the arguments mean what they mean for the single-scene launchers above)
void phase_advect(BatchPhase& ph, const Fields& F);
void phase_diffuse(BatchPhase& ph, const Fields& F, const float* Tin, const float4* Yin, float* Tout, float4* Yout);
void phase_sources(BatchPhase& ph, const Fields& F, const float* Tin, const float4* Yin, int* part);
void phase_forces_coef(BatchPhase& ph, const Fields& F, const Level* L, int nlev, const int* part, int nparts);
void phase_coef_restrict(BatchPhase& ph, const Level& f, const Level& c);
void phase_cfl(BatchPhase& ph, const Fields& F, float cfl, float umin, int nsub);
void phase_resnorm(BatchPhase& ph, BatchPhase& fold, const Level& L0, Dyn* dyn, double* part, float tol);
void phase_dc_defect(BatchPhase& ph, BatchPhase& fold, const Level& L0, __half* ph0, double* p64, float* b0,
                     Dyn* dyn, double* part, float tol, bool first);
void phase_kry_start(BatchPhase (&ph)[5], const Level& L0, __half* ph0, Dyn* dyn, const KryBufs& k, double* part,
                     float tol);
void phase_dc_accum(BatchPhase& ph, const Level& L0, const __half* ph0, const Dyn* dyn, double* p64, const int* skip);
void phase_coef_half(BatchPhase& ph, const Level& L, const LevelH& H);
void phase_project(BatchPhase& ph, const Fields& F, const double* p64);
void phase_geometry(BatchPhase& ph, const Fields& F);
void phase_gas_terms(BatchPhase& ph, const Fields& F);
void phase_radiation(BatchPhase& ph, const Fields& F, int ntargets);
void phase_fill_u32(BatchPhase& ph, void* a, uint32_t v, size_t n);
// tile: the legs' tile edge (16, 32, 64), as the pool chose it for this level
bool phase_mg_down(BatchPhase& ph, const Level& f, const Level& c, const MgStore& st, int nu, const int* skip,
                   int tile, int zero_p = 0);
bool phase_mg_up(BatchPhase& ph, const Level& f, const Level& c, const MgStore& st, int nu, const int* skip,
                 int tile);
// the last V-cycle's level-0 up leg with the projection (as launch_mg_up_project);
// false: no instantiation for that tile and nu
bool phase_mg_up_project(BatchPhase& ph, const Level& f, const Level& c, const Fields& F, int nu, const int* skip,
                         int t);
bool phase_mg_bottom(BatchPhase& ph, const Level* L, int nlev, int first, bool coarse_coef, int nu, int coarse_iters,
                     const int* skip, double* band);

}  // namespace cf
