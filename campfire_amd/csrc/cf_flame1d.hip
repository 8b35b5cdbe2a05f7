// cf_flame1d.hip — the 1D premixed flame of tests/flame1d.mjs:13-71 on the GPU.
//
// One workgroup of 1024 threads runs one flame: the strip (<= 1536 cells, two
// per thread) stays in LDS for the whole run, and each time step is
//   1. diffusion with Sutherland rho*D and one-step Arrhenius reaction (:32-46)
//   2. continuity u(x) from the dilatation S: a workgroup scan (:49)
//   3. first-order upwind advection (:51-57)
//   4. front position (furthest cell above 1200 K) and peak T: workgroup max
// with the script's front tracking and speed estimate (:61-70).  Many flames
// (the script's table rows) run concurrently, one per workgroup.  Templated
// on the arithmetic type: double for the script's numbers, float for the
// single-precision variant.
#include <hip/hip_runtime.h>

#include "cf_kernels.h"

namespace cf {

namespace {

constexpr int FL_T = 1024;

__device__ __forceinline__ double m_pow(double a, double b) { return pow(a, b); }
__device__ __forceinline__ float m_pow(float a, float b) { return powf(a, b); }
__device__ __forceinline__ double m_exp(double a) { return exp(a); }
__device__ __forceinline__ float m_exp(float a) { return expf(a); }
__device__ __forceinline__ double m_log(double a) { return log(a); }
__device__ __forceinline__ float m_log(float a) { return logf(a); }
__device__ __forceinline__ double m_abs(double a) { return fabs(a); }
__device__ __forceinline__ float m_abs(float a) { return fabsf(a); }
__device__ __forceinline__ double m_min(double a, double b) { return fmin(a, b); }
__device__ __forceinline__ float m_min(float a, float b) { return fminf(a, b); }
__device__ __forceinline__ double m_max(double a, double b) { return fmax(a, b); }
__device__ __forceinline__ float m_max(float a, float b) { return fmaxf(a, b); }

// tests/flame1d.mjs:9 — Sutherland viscosity
template <typename R>
__device__ __forceinline__ R mu1(R T) {
  return R(1.716e-5) * m_pow(T / R(273.15), R(1.5)) * R(383.55) / (T + R(110.4));
}

}  // namespace

template <typename R>
__global__ __launch_bounds__(FL_T) void k_flame1d(const cf_flame_case* __restrict__ cases,
                                                  cf_flame_result* __restrict__ out) {
  __shared__ R sT[CF_FLAME_MAXN], sF[CF_FLAME_MAXN], sO[CF_FLAME_MAXN];
  __shared__ R su[CF_FLAME_MAXN + 1];
  __shared__ R wscan[FL_T / 64];
  __shared__ int wlast[FL_T / 64];
  __shared__ R wTmax[FL_T / 64];
  const cf_flame_case C = cases[blockIdx.x];
  const R B = R(C.B), Ta = R(C.Ta), dx = R(C.dx), F = R(C.F), phi = R(C.phi), L = R(C.L), tEnd = R(C.tEnd);
  // tests/flame1d.mjs:6-8
  const R Tamb = R(C.Tamb), rhoAmb = R(1.2), cp = R(1300), Pr = R(0.7), dHc = R(15e6), sO2 = R(15e6 / 13.1e6),
          YO = R(0.232);
  // Math.round(L / dx) cells, Math.round(0.01 / dx) of them ignited (:15, :21)
  const int n = (int)floor(C.L / C.dx + 0.5), nign = (int)floor(0.01 / C.dx + 0.5);
  const R YFst = YO / sO2;  // :16-18
  const R YF0 = phi * YFst / (R(1) + phi * YFst) * R(1), YO0 = YO / (R(1) + phi * YFst);
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int i0 = 2 * tid;
  for (int q = 0; q < 2; q++) {  // :19-21
    const int i = i0 + q;
    if (i < n) {
      const bool ign = i < nign;
      sT[i] = ign ? R(C.Tign) : Tamb;
      sF[i] = ign ? R(0) : YF0;
      sO[i] = ign ? m_max(R(0), YO0 - sO2 * YF0) : YO0;
    }
  }
  if (tid == 0) su[0] = R(0);
  const R alphaMax = mu1(R(2300)) / Pr * F / (rhoAmb * Tamb / R(2300));  // :22-23
  const R dt = m_min(R(0.2) * dx * dx / alphaMax, R(0.2) * dx / R(6));
  R t = R(0), Tb = R(0), at = R(0), ax = R(0), bt = R(0), bx = R(0), lt = R(0), lx = R(0);
  bool have_a = false, have_b = false;
  int steps = 0;
  __syncthreads();
  while (t < tEnd && steps < C.max_steps) {
    // 1. diffusion and reaction (:32-46)
    R Tn[2], Fn[2], On[2], S[2];
#pragma unroll
    for (int q = 0; q < 2; q++) {
      const int i = i0 + q;
      Tn[q] = Fn[q] = On[q] = S[q] = R(0);
      if (i < n) {
        const R Ti = sT[i], Fi = sF[i], Oi = sO[i];
        const R rho = rhoAmb * Tamb / Ti;
        R dT = R(0), dF = R(0), dO = R(0);
        if (i - 1 >= 0) {
          const R Tj = sT[i - 1];
          const R a = mu1(R(0.5) * (Ti + Tj)) / Pr * F;
          dT += a * (Tj - Ti); dF += a * (sF[i - 1] - Fi); dO += a * (sO[i - 1] - Oi);
        }
        if (i + 1 < n) {
          const R Tj = sT[i + 1];
          const R a = mu1(R(0.5) * (Ti + Tj)) / Pr * F;
          dT += a * (Tj - Ti); dF += a * (sF[i + 1] - Fi); dO += a * (sO[i + 1] - Oi);
        }
        const R c = dt / (rho * dx * dx);
        R tn = Ti + c * dT, fn = Fi + c * dF, on = Oi + c * dO;
        const R wr = B * rho * fn * on * m_exp(-Ta / tn) / F;
        const R d = m_min(m_min(wr * dt, fn), on / sO2);
        fn -= d; on -= sO2 * d; tn += d * dHc / cp;
        S[q] = m_log(tn / Ti) / dt;
        Tn[q] = tn; Fn[q] = fn; On[q] = on;
      }
    }
    // 2. continuity du/dx = S (:49) as a workgroup scan of S*dx
    const R s0 = i0 < n ? S[0] * dx : R(0), s1 = i0 + 1 < n ? S[1] * dx : R(0);
    const R pair = s0 + s1;
    R inc = pair;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const R v = __shfl_up(inc, o, 64);
      if (lane >= o) inc += v;
    }
    if (lane == 63) wscan[w] = inc;
    __syncthreads();  // everyone has read the old state
#pragma unroll
    for (int q = 0; q < 2; q++) {
      const int i = i0 + q;
      if (i < n) { sT[i] = Tn[q]; sF[i] = Fn[q]; sO[i] = On[q]; }
    }
    if (tid == 0) {
      R acc = R(0);
      for (int k = 0; k < FL_T / 64; k++) { const R v = wscan[k]; wscan[k] = acc; acc += v; }
    }
    __syncthreads();
    const R off = wscan[w] + inc - pair;
    if (i0 < n) su[i0 + 1] = off + s0;
    if (i0 + 1 < n) su[i0 + 2] = off + s0 + s1;
    __syncthreads();
    // 3. upwind advection (:51-57)
#pragma unroll
    for (int q = 0; q < 2; q++) {
      const int i = i0 + q;
      if (i < n) {
        const R uc = R(0.5) * (su[i] + su[i + 1]);
        const int j = uc > R(0) ? i - 1 : i + 1;
        const int k = j < 0 ? 0 : (j > n - 1 ? n - 1 : j);
        const R cfl = m_abs(uc) * dt / dx;
        Tn[q] = sT[i] + cfl * (sT[k] - sT[i]);
        Fn[q] = sF[i] + cfl * (sF[k] - sF[i]);
        On[q] = sO[i] + cfl * (sO[k] - sO[i]);
      }
    }
    __syncthreads();
    int last = -1;
    R tmax = R(0);
#pragma unroll
    for (int q = 0; q < 2; q++) {
      const int i = i0 + q;
      if (i < n) {
        sT[i] = Tn[q]; sF[i] = Fn[q]; sO[i] = On[q];
        if (Tn[q] > R(1200)) last = i;
        tmax = m_max(tmax, Tn[q]);
      }
    }
    t += dt;
    steps++;
    // 4. front and peak temperature (:61-64)
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      last = max(last, __shfl_xor(last, o, 64));
      tmax = m_max(tmax, __shfl_xor(tmax, o, 64));
    }
    if (lane == 0) { wlast[w] = last; wTmax[w] = tmax; }
    __syncthreads();
    int lastAll = -1;
    R tmaxAll = R(0);
    for (int k = 0; k < FL_T / 64; k++) { lastAll = max(lastAll, wlast[k]); tmaxAll = m_max(tmaxAll, wTmax[k]); }
    const R xf = lastAll >= 0 ? (R(lastAll) + R(0.5)) * dx : R(0);
    Tb = m_max(Tb, tmaxAll);
    if (!have_a && xf > R(0.3) * L) { have_a = true; at = t; ax = xf; }
    if (!have_b && xf > R(0.75) * L) { have_b = true; bt = t; bx = xf; }
    lt = t; lx = xf;
    if (xf > L * R(0.8)) break;
  }
  if (tid == 0) {  // :68-70
    if (!have_b) { bt = lt; bx = lx; }
    cf_flame_result r;
    r.Tb = (double)Tb;
    r.steps = steps;
    r.n = n;
    r.SL = (!have_a || steps == 0 || bt <= at) ? 0.0 : (double)((bx - ax) / (bt - at) * Tamb / Tb);
    out[blockIdx.x] = r;
  }
}

void launch_flame1d(const cf_flame_case* cases, int ncases, bool fp64, cf_flame_result* out, hipStream_t s) {
  if (fp64) hipLaunchKernelGGL(k_flame1d<double>, dim3(ncases), dim3(FL_T), 0, s, cases, out);
  else hipLaunchKernelGGL(k_flame1d<float>, dim3(ncases), dim3(FL_T), 0, s, cases, out);
}

}  // namespace cf

// ---- the allocation-free variant of cf_flame1d_run -------------------------------
// The device staging of the cases and the results lives in one file-scope
// array, `scratch`, instead of per-call allocations: the cases at offset 0, the
// results after CF_FLAME_STATIC_MAX cases.  Each device holds its own copy,
// until the process ends.  Host code must not pass `scratch` itself to a
// kernel or a copy (in host code the name is a stand-in, not a device
// address): the device address comes from hipGetSymbolAddress, once per
// device, and is kept in g_scratch.
#include <mutex>
#include <string.h>

namespace {
constexpr size_t SCR_CASES = 0;
constexpr size_t SCR_RESULTS = ((CF_FLAME_STATIC_MAX * sizeof(cf_flame_case) + 255) / 256) * 256;
constexpr size_t SCR_BYTES = SCR_RESULTS + CF_FLAME_STATIC_MAX * sizeof(cf_flame_result);
constexpr int SCR_MAX_DEVICES = 64;
__device__ __align__(256) unsigned char scratch[SCR_BYTES];

std::mutex g_scratch_mu;
void* g_scratch[SCR_MAX_DEVICES] = {};  // device address of `scratch` on each device, once looked up

// the device address of this device's `scratch` (looked up on the first call
// on the device, with it current); null with a message on failure
void* scratch_ptr(int* code) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess || dev < 0 || dev >= SCR_MAX_DEVICES) {
    *code = cf::set_error(CF_EDEVICE, "hipGetDevice failed or device out of range");
    return nullptr;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      *code = cf::set_error(CF_EDEVICE, "hipGetSymbolAddress(scratch) failed");
      return nullptr;
    }
    g_scratch[dev] = p;
  }
  return g_scratch[dev];
}
}  // namespace

extern "C" int cf_flame1d_run_static(const cf_flame_case* cases, int ncases, int fp64, cf_flame_result* out,
                                     void* stream) {
  if (!cases || !out || ncases < 1) return cf::set_error(CF_EINVAL, "bad argument (cases, out non-null, ncases >= 1)");
  if (ncases > CF_FLAME_STATIC_MAX)
    return cf::set_error(CF_EINVAL, "more cases than the static staging holds (CF_FLAME_STATIC_MAX)");
  for (int q = 0; q < ncases; q++) {
    const int rc = cf::check_flame_case(&cases[q]);
    if (rc) return rc;
  }
  int code = CF_OK;
  unsigned char* scr = (unsigned char*)scratch_ptr(&code);
  if (!scr) return code;
  hipStream_t st = (hipStream_t)stream;
  cf_flame_case* dc = (cf_flame_case*)(scr + SCR_CASES);
  cf_flame_result* dr = (cf_flame_result*)(scr + SCR_RESULTS);
  // the cases as cf_flame1d_run passes them: a zero step cap means 10 million
  hipError_t e = hipMemcpyAsync(dc, cases, sizeof(cf_flame_case) * ncases, hipMemcpyHostToDevice, st);
  for (int q = 0; q < ncases && e == hipSuccess; q++)
    if (cases[q].max_steps <= 0) {
      const int32_t cap = 10000000;
      e = hipMemcpyAsync((unsigned char*)&dc[q] + offsetof(cf_flame_case, max_steps), &cap, sizeof cap,
                         hipMemcpyHostToDevice, st);
    }
  if (e == hipSuccess) {
    cf::launch_flame1d(dc, ncases, fp64 != 0, dr, st);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(out, dr, sizeof(cf_flame_result) * ncases, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return cf::set_error(CF_EDEVICE, hipGetErrorString(e));
  return CF_OK;
}
