// cf_device.h — device-side helpers shared by the gas-step kernels (gfx950).
//
// Arithmetic keeps the reference's operation order (shaders/*.wgsl) and is
// compiled without FMA contraction, so FP32 results match the FP32 CPU oracle
// to the ULP differences of the transcendental functions.  Constants are
// written K(decimal), a decimal literal rounded to float, which is how WGSL
// abstract-float literals reach f32.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "../../include/campfire_hip.h"

#define K(x) ((float)(x))
#define CF_SOLID_NONE 0xffffffffu

namespace cf {

// A workgroup barrier that orders LDS only.  The threads' global loads and
// stores stay in flight across it and are waited for where their values are
// used (or at the kernel's end); __syncthreads' global fence would wait for
// every one of them at each barrier.  For barriers that exchange data through
// LDS alone.
__device__ __forceinline__ void lds_sync() { __barrier(__CLK_LOCAL_MEM_FENCE); }

__device__ __forceinline__ float mixf(float a, float b, float t) { return a * (K(1) - t) + b * t; }
__device__ __forceinline__ float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }
__device__ __forceinline__ int iclamp(int x, int lo, int hi) { return min(max(x, lo), hi); }
// i32(f) truncates; floor values far outside the grid saturate
__device__ __forceinline__ int toint(float x) {
  return (int)fminf(fmaxf(x, -1073741824.0f), 1073741824.0f);
}
__device__ __forceinline__ int iround(float x) { return (int)rintf(x); }

// x / d correctly rounded, for d > 0 given y = 1/d correctly rounded:
// Markstein's correction step (P. Markstein, IBM J. Res. Dev. 34(1), 1990;
// Muller et al., Handbook of Floating-Point Arithmetic, the theorem on
// division with a correctly rounded reciprocal).  q = x*y is within one ulp
// of x/d, the FMA residual x - q*d is exact, and one more FMA rounds to the
// nearest quotient; the quotient takes the sign of x, as x/d does for zero.
// The theorem assumes no underflow, so a nonzero dividend below 2^-100 (a
// mass fraction that has burnt out, say) takes the IEEE division.  A few
// instructions where a division takes eleven, for divisors reused many times.
__device__ __forceinline__ float div_rcp(float x, float d, float y) {
  const float q = x * y;
  const float m = copysignf(fmaf(fmaf(-q, d, x), y, q), x);
  const uint32_t ax = __float_as_uint(x) & 0x7fffffffu;
  if (__builtin_expect(ax - 1u < 0x0d7fffffu, 0)) {  // 0 < |x| < 2^-100 (2^-100 = 0x0d800000)
    // an empty volatile statement keeps the compiler from computing this
    // rare division on every call and selecting
    asm volatile("");
    return x / d;
  }
  return m;
}

// div_rcp without the small-dividend branch, for loops where the branch
// costs more than the division saves (the coarsest sweeps, the advection
// traces).  Equal to x / d for x = 0 and |x| >= 2^-100; a nonzero dividend
// below that (never a pressure update or a velocity of this solver, which
// are zero or many orders above it) may round its quotient one ulp apart.
__device__ __forceinline__ float div_rcp_fast(float x, float d, float y) {
  const float q = x * y;
  return copysignf(fmaf(fmaf(-q, d, x), y, q), x);
}

// x^1.5 as x*sqrt(x): within ~1 ulp of the correctly rounded power (WGSL's pow
// promises no better), a tenth of the cost of powf, which dominated the
// stencils that evaluate Sutherland's law five times per face
__device__ __forceinline__ float pow15(float x) { return x * sqrtf(x); }

// gas closures (shaders/common.wgsl:33-42).  The divisions of a temperature
// by the constant 273.15 and of a diffusivity by Pr are correctly rounded
// quotients through div_rcp_fast (the reciprocal of 273.15 folds at compile
// time; 1/Pr is one division per thread)
__device__ __forceinline__ float rhoOf(const cf_params& P, float T) { return P.rhoAmb * P.Tamb / T; }
__device__ __forceinline__ float muOf(const cf_params& P, float T) {
  return K(1.716e-5) * pow15(div_rcp_fast(T, K(273.15), K(1) / K(273.15))) * K(273.15 + 110.4) / (T + K(110.4)) *
         P.viscMult;
}
// iPr = 1/P.Pr, computed once by the caller
__device__ __forceinline__ float rhoDOf(const cf_params& P, float T, float iPr) {
  return div_rcp_fast(
             K(1.716e-5) * pow15(div_rcp_fast(T, K(273.15), K(1) / K(273.15))) * K(273.15 + 110.4) / (T + K(110.4)),
             P.Pr, iPr) *
         P.diffMult;
}
__device__ __forceinline__ float rhoDOf(const cf_params& P, float T) { return rhoDOf(P, T, K(1) / P.Pr); }
__device__ __forceinline__ float invRho(const cf_params& P, float T) { return T / (P.rhoAmb * P.Tamb); }

// Hardware places workgroup b on XCD (b mod 8).  Give each XCD a contiguous
// run of tiles so that stencil halos are shared inside one L2.
__device__ __forceinline__ int xcd_tile(int b, int nb) {
  const int full = (nb / 8) * 8;
  if (b >= full) return b;
  return (b % 8) * (nb / 8) + b / 8;
}

// Sum of 64 wave lanes in the order of the reference's workgroup tree
// (for s = 32..1: red[li] += red[li+s]); lane 0 holds the result.
__device__ __forceinline__ float wave_tree_sum(float v) {
#pragma unroll
  for (int s = 32; s > 0; s >>= 1) v += __shfl_down(v, s, 64);
  return v;
}
__device__ __forceinline__ float wave_tree_max(float v) {
#pragma unroll
  for (int s = 32; s > 0; s >>= 1) v = fmaxf(v, __shfl_down(v, s, 64));
  return v;
}

// ---- boundary-aware bilinear samplers in grid coordinates (js/gpu.js:16-65) ----
// cell fields: outside sides/top -> ambient; below the floor -> row 0
__device__ __forceinline__ float fC(const float* __restrict__ X, int nx, int ny, float amb, int i, int j) {
  if (i < 0 || i >= nx || j >= ny) return amb;
  return X[max(j, 0) * nx + i];
}
__device__ __forceinline__ float4 fC4(const float4* __restrict__ X, int nx, int ny, float4 amb, int i, int j) {
  if (i < 0 || i >= nx || j >= ny) return amb;
  return X[max(j, 0) * nx + i];
}
__device__ __forceinline__ float sC(const float* __restrict__ X, int nx, int ny, float amb, float gx, float gy) {
  float fx = gx - K(0.5), fy = gy - K(0.5);
  float bx = floorf(fx), by = floorf(fy);
  float tx = fx - bx, ty = fy - by;
  int i = toint(bx), j = toint(by);
  return mixf(mixf(fC(X, nx, ny, amb, i, j), fC(X, nx, ny, amb, i + 1, j), tx),
              mixf(fC(X, nx, ny, amb, i, j + 1), fC(X, nx, ny, amb, i + 1, j + 1), tx), ty);
}
__device__ __forceinline__ float4 mix4(float4 a, float4 b, float t) {
  return make_float4(mixf(a.x, b.x, t), mixf(a.y, b.y, t), mixf(a.z, b.z, t), mixf(a.w, b.w, t));
}
__device__ __forceinline__ float4 sC4(const float4* __restrict__ X, int nx, int ny, float4 amb, float gx, float gy) {
  float fx = gx - K(0.5), fy = gy - K(0.5);
  float bx = floorf(fx), by = floorf(fy);
  float tx = fx - bx, ty = fy - by;
  int i = toint(bx), j = toint(by);
  return mix4(mix4(fC4(X, nx, ny, amb, i, j), fC4(X, nx, ny, amb, i + 1, j), tx),
              mix4(fC4(X, nx, ny, amb, i, j + 1), fC4(X, nx, ny, amb, i + 1, j + 1), tx), ty);
}
__device__ __forceinline__ float2 mmC(const float* __restrict__ X, int nx, int ny, float amb, float gx, float gy) {
  int i = toint(floorf(gx - K(0.5))), j = toint(floorf(gy - K(0.5)));
  float a = fC(X, nx, ny, amb, i, j), c = fC(X, nx, ny, amb, i + 1, j);
  float d = fC(X, nx, ny, amb, i, j + 1), e = fC(X, nx, ny, amb, i + 1, j + 1);
  return make_float2(fminf(fminf(a, c), fminf(d, e)), fmaxf(fmaxf(a, c), fmaxf(d, e)));
}
__device__ __forceinline__ void mmC4(const float4* __restrict__ X, int nx, int ny, float4 amb, float gx, float gy,
                                     float4& lo, float4& hi) {
  int i = toint(floorf(gx - K(0.5))), j = toint(floorf(gy - K(0.5)));
  float4 a = fC4(X, nx, ny, amb, i, j), c = fC4(X, nx, ny, amb, i + 1, j);
  float4 d = fC4(X, nx, ny, amb, i, j + 1), e = fC4(X, nx, ny, amb, i + 1, j + 1);
  lo = make_float4(fminf(fminf(a.x, c.x), fminf(d.x, e.x)), fminf(fminf(a.y, c.y), fminf(d.y, e.y)),
                   fminf(fminf(a.z, c.z), fminf(d.z, e.z)), fminf(fminf(a.w, c.w), fminf(d.w, e.w)));
  hi = make_float4(fmaxf(fmaxf(a.x, c.x), fmaxf(d.x, e.x)), fmaxf(fmaxf(a.y, c.y), fmaxf(d.y, e.y)),
                   fmaxf(fmaxf(a.z, c.z), fmaxf(d.z, e.z)), fmaxf(fmaxf(a.w, c.w), fmaxf(d.w, e.w)));
}
// U faces: outside -> 0 (still air), below the floor the no-slip ghost -U(i,0)
__device__ __forceinline__ float fU(const float* __restrict__ X, int nx, int ny, int i, int j) {
  if (i < 0 || i > nx || j >= ny) return K(0);
  if (j < 0) return -X[i];
  return X[j * (nx + 1) + i];
}
// V faces: outside -> 0, below the floor clamps to row 0
__device__ __forceinline__ float fV(const float* __restrict__ X, int nx, int ny, int i, int j) {
  if (i < 0 || i >= nx || j > ny) return K(0);
  return X[max(j, 0) * nx + i];
}
__device__ __forceinline__ float sU(const float* __restrict__ X, int nx, int ny, float gx, float gy) {
  float fx = gx - K(0), fy = gy - K(0.5);
  float bx = floorf(fx), by = floorf(fy);
  float tx = fx - bx, ty = fy - by;
  int i = toint(bx), j = toint(by);
  return mixf(mixf(fU(X, nx, ny, i, j), fU(X, nx, ny, i + 1, j), tx),
              mixf(fU(X, nx, ny, i, j + 1), fU(X, nx, ny, i + 1, j + 1), tx), ty);
}
__device__ __forceinline__ float sV(const float* __restrict__ X, int nx, int ny, float gx, float gy) {
  float fx = gx - K(0.5), fy = gy - K(0);
  float bx = floorf(fx), by = floorf(fy);
  float tx = fx - bx, ty = fy - by;
  int i = toint(bx), j = toint(by);
  return mixf(mixf(fV(X, nx, ny, i, j), fV(X, nx, ny, i + 1, j), tx),
              mixf(fV(X, nx, ny, i, j + 1), fV(X, nx, ny, i + 1, j + 1), tx), ty);
}
__device__ __forceinline__ float2 mmU(const float* __restrict__ X, int nx, int ny, float gx, float gy) {
  int i = toint(floorf(gx - K(0))), j = toint(floorf(gy - K(0.5)));
  float a = fU(X, nx, ny, i, j), c = fU(X, nx, ny, i + 1, j), d = fU(X, nx, ny, i, j + 1), e = fU(X, nx, ny, i + 1, j + 1);
  return make_float2(fminf(fminf(a, c), fminf(d, e)), fmaxf(fmaxf(a, c), fmaxf(d, e)));
}
__device__ __forceinline__ float2 mmV(const float* __restrict__ X, int nx, int ny, float gx, float gy) {
  int i = toint(floorf(gx - K(0.5))), j = toint(floorf(gy - K(0)));
  float a = fV(X, nx, ny, i, j), c = fV(X, nx, ny, i + 1, j), d = fV(X, nx, ny, i, j + 1), e = fV(X, nx, ny, i + 1, j + 1);
  return make_float2(fminf(fminf(a, c), fminf(d, e)), fmaxf(fmaxf(a, c), fmaxf(d, e)));
}

// RK2 midpoint departure point of grid position g over dt (shaders/advect.wgsl:28-34)
__device__ __forceinline__ float2 trace(const float* __restrict__ U, const float* __restrict__ V, int nx, int ny,
                                        float dx, float gx, float gy, float dt) {
  float vx = sU(U, nx, ny, gx, gy) / dx, vy = sV(V, nx, ny, gx, gy) / dx;
  float hdt = K(0.5) * dt;
  float mx_ = gx - hdt * vx, my_ = gy - hdt * vy;
  float wx = sU(U, nx, ny, mx_, my_) / dx, wy = sV(V, nx, ny, mx_, my_) / dx;
  return make_float2(gx - dt * wx, gy - dt * wy);
}

}  // namespace cf
