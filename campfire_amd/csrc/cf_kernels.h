// cf_kernels.h — device buffers of one scene and the kernel launchers.
#pragma once
#include <hip/hip_ext.h>
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

// Every device buffer of one scene (the `buf` dictionary of js/fluid.js:52-68
// plus the level-0 multigrid arrays).  Passed by value as a kernel argument.
struct Fields {
  const cf_params* P;  // the 160-B parameter block in device memory
  int nx, ny;
  float *U, *Uh, *U2, *V, *Vh, *V2, *T, *Th, *T2;
  float4 *Y, *Yh, *Y2;
  float2 *SQ, *cellG;
  float *fx, *fy, *pieces;
  int *acc, *faceW;
  float4* rad;
  float *p0, *rhs0, *ax0, *ay0;
};

// One multigrid level (the `Level` uniform of shaders/pressure.wgsl:7 and its arrays).
struct Level {
  int nx, ny;
  float h2;
  float *p, *rhs, *ax, *ay;
};

// advection in one pass (cf_advect.hip); the two launches below are the
// reference's dispatch pairs and also write the forward fields Th, Yh, Uh, Vh
void launch_advect(const Fields& F, hipStream_t s);
void launch_advect_fwd(const Fields& F, hipStream_t s);
void launch_advect_corr(const Fields& F, hipStream_t s);
void launch_diffuse(const Fields& F, const float* Tin, const float4* Yin, float* Tout, float4* Yout, float dts, hipStream_t s);
// sources (+ fused last diffusion substep) and the fold of its per-workgroup
// accumulator partials (part: 4 ints per workgroup, sources_blocks() of them)
unsigned sources_blocks(const Fields& F);
void launch_sources(const Fields& F, bool fuse, const float* Tin, const float4* Yin, float dts, int* part, bool finish,
                    hipStream_t s);
void launch_forces(const Fields& F, hipStream_t s);
void launch_coef_rhs(const Fields& F, hipStream_t s);
// forces + coef + rhs0 + coefficient restriction of the first
// forces_coef_cascade(nlev) coarse levels, in one pass; its first workgroup
// also folds nparts sources partials into acc (nparts = 0: none pending)
int forces_coef_cascade(int nlev);
void launch_forces_coef(const Fields& F, const Level* L, int nlev, const int* part, int nparts, hipStream_t s);
// fused multigrid (cf_mg.hip)
bool launch_mg_down(const Level& f, const Level& c, int nu, hipStream_t s);
bool launch_mg_up(const Level& f, const Level& c, int nu, hipStream_t s);
// first level of the single-workgroup bottom solver (the last 1-3 levels)
int mg_bottom_first(const Level* L, int nlev);
// coarse_coef: restrict the coarsest level's coefficients inside the kernel
bool launch_mg_bottom(const Level* L, int nlev, int first, bool coarse_coef, int nu, int coarse_iters, hipStream_t s);
void launch_coef_restrict(const Level& f, const Level& c, hipStream_t s);
void launch_relax(const Level& L, int color, hipStream_t s);
void launch_restrict_res(const Level& f, const Level& c, hipStream_t s);
void launch_prolong(const Level& f, const Level& c, hipStream_t s);
void launch_coarsest(const Level& L, int iters, hipStream_t s);
void launch_project(const Fields& F, hipStream_t s);
void launch_geometry(const Fields& F, hipStream_t s);
void launch_radiation(const Fields& F, int ntargets, hipStream_t s);
void launch_fill_u32(void* a, uint32_t v, size_t n, hipStream_t s);
// tests/flame1d.mjs on the GPU (cf_flame1d.hip): one workgroup per case
void launch_flame1d(const cf_flame_case* cases, int ncases, bool fp64, cf_flame_result* out, hipStream_t s);
unsigned div_error_blocks(const Fields& F);
void launch_div_error(const Fields& F, float* part, hipStream_t s);

}  // namespace cf
