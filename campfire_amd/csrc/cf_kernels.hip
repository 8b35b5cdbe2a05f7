// cf_kernels.hip — gfx950 kernels of the campfire gas step.
//
// One thread per cell or face, workgroups of 32x8 threads (four wave64s, each
// wave two 32-cell rows so every row access is a contiguous 128-B segment),
// a 1-D launch remapped so each XCD owns a contiguous band of tiles.  The
// integer accumulators are reduced over 8x8 cell tiles exactly as the
// reference's 8x8 workgroups do (shaders/sources.wgsl:160-172), so `acc` gets
// the same set of integer contributions.
#include <climits>

#include "cf_kernels.h"

namespace cf {

// -------------------------------------------------------------------------
// tile decode: TX x TY threads per workgroup over an (ex x ey) index domain
// -------------------------------------------------------------------------
// (i, j) of this thread in tile t of a TX x TY tiling of an ex-wide domain
template <int TX, int TY>
__device__ __forceinline__ bool tile_ij(int t, int ex, int ey, int& i, int& j) {
  const int gx = (ex + TX - 1) / TX;
  i = (t % gx) * TX + threadIdx.x;
  j = (t / gx) * TY + threadIdx.y;
  return true;
}
template <int TX, int TY>
__device__ __forceinline__ bool tile_ij(int ex, int ey, int& i, int& j) {
  const int gx = (ex + TX - 1) / TX, gy = (ey + TY - 1) / TY;
  const int t = xcd_tile(blockIdx.x, gx * gy);
  i = (t % gx) * TX + threadIdx.x;
  j = (t / gx) * TY + threadIdx.y;
  return true;
}

// -------------------------------------------------------------------------
// advection, forward pass: scalForward + velForward (shaders/advect.wgsl:36-47,68-76)
// -------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_advect_fwd(Fields F) {
  const cf_params P = load_params(F.P);
  const int nx = P.nx, ny = P.ny;
  int i, j;
  tile_ij<32, 8>(nx + 1, ny + 1, i, j);
  const float dt = P.dt, dx = P.dx;
  if (i < nx && j < ny) {
    const int k = j * nx + i;
    float2 d = trace(F.U, F.V, nx, ny, dx, (float)i + K(0.5), (float)j + K(0.5), dt);
    F.Th[k] = sC(F.T, nx, ny, P.Tamb, d.x, d.y);
    F.Yh[k] = sC4(F.Y, nx, ny, make_float4(0.f, P.YOamb, 0.f, 0.f), d.x, d.y);
  }
  if (i <= nx && j < ny) {
    float2 d = trace(F.U, F.V, nx, ny, dx, (float)i, (float)j + K(0.5), dt);
    F.Uh[j * (nx + 1) + i] = sU(F.U, nx, ny, d.x, d.y);
  }
  if (i < nx && j <= ny) {
    float2 d = trace(F.U, F.V, nx, ny, dx, (float)i + K(0.5), (float)j, dt);
    F.Vh[j * nx + i] = sV(F.V, nx, ny, d.x, d.y);
  }
}

// -------------------------------------------------------------------------
// advection, MacCormack correction: scalCorrect + velCorrect (:49-66,78-103)
// -------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_advect_corr(Fields F) {
  const cf_params P = load_params(F.P);
  const int nx = P.nx, ny = P.ny;
  int i, j;
  tile_ij<32, 8>(nx + 1, ny + 1, i, j);
  const float dt = P.dt, dx = P.dx;
  if (i < nx && j < ny) {
    const int k = j * nx + i;
    const float4 aY = make_float4(0.f, P.YOamb, 0.f, 0.f);
    const float gx = (float)i + K(0.5), gy = (float)j + K(0.5);
    float2 d = trace(F.U, F.V, nx, ny, dx, gx, gy, dt);
    float2 f = trace(F.U, F.V, nx, ny, dx, gx, gy, -dt);
    float2 mT = mmC(F.T, nx, ny, P.Tamb, d.x, d.y);
    float4 lo, hi;
    mmC4(F.Y, nx, ny, aY, d.x, d.y, lo, hi);
    const float Thk = F.Th[k];
    const float4 Yhk = F.Yh[k], Yk = F.Y[k];
    const float Tc = Thk + K(0.5) * (F.T[k] - sC(F.Th, nx, ny, P.Tamb, f.x, f.y));
    const float4 sb = sC4(F.Yh, nx, ny, aY, f.x, f.y);
    const float4 Yc = make_float4(Yhk.x + K(0.5) * (Yk.x - sb.x), Yhk.y + K(0.5) * (Yk.y - sb.y),
                                  Yhk.z + K(0.5) * (Yk.z - sb.z), Yhk.w + K(0.5) * (Yk.w - sb.w));
    const bool inT = Tc >= mT.x && Tc <= mT.y;
    const bool inY = Yc.x >= lo.x && Yc.y >= lo.y && Yc.z >= lo.z && Yc.w >= lo.w &&
                     Yc.x <= hi.x && Yc.y <= hi.y && Yc.z <= hi.z && Yc.w <= hi.w;
    const float mixC = fminf(Yc.x, Yc.y / P.sO2);
    const float mixL = fminf(Yhk.x, Yhk.y / P.sO2);
    if (inT && inY && mixC <= mixL + K(1e-5)) {
      F.T2[k] = Tc; F.Y2[k] = Yc;
    } else {
      F.T2[k] = Thk; F.Y2[k] = Yhk;
    }
  }
  if (i <= nx && j < ny) {
    const int k = j * (nx + 1) + i;
    const float gx = (float)i, gy = (float)j + K(0.5);
    float2 b = trace(F.U, F.V, nx, ny, dx, gx, gy, -dt);
    float back = sU(F.Uh, nx, ny, b.x, b.y);
    float2 d = trace(F.U, F.V, nx, ny, dx, gx, gy, dt);
    float2 m = mmU(F.U, nx, ny, d.x, d.y);
    F.U2[k] = clampf(F.Uh[k] + K(0.5) * (F.U[k] - back), m.x, m.y);
  }
  if (i < nx && j <= ny) {
    const int k = j * nx + i;
    const float gx = (float)i + K(0.5), gy = (float)j;
    float2 b = trace(F.U, F.V, nx, ny, dx, gx, gy, -dt);
    float back = sV(F.Vh, nx, ny, b.x, b.y);
    float2 d = trace(F.U, F.V, nx, ny, dx, gx, gy, dt);
    float2 m = mmV(F.V, nx, ny, d.x, d.y);
    F.V2[k] = clampf(F.Vh[k] + K(0.5) * (F.V[k] - back), m.x, m.y);
  }
}

// -------------------------------------------------------------------------
// explicit diffusion of one cell (shaders/diffuse.wgsl:18-50)
// -------------------------------------------------------------------------
// The weight a = f * rhoD(0.5 (T0 + T1)) of a face is the same for both of its
// cells (the face temperature is symmetric in them), so a 32 x 8 workgroup
// evaluates each face once into LDS and both cells read the one value: every
// thread forms its cell's east and north faces, and the first 40 threads (one
// wave) the tile's 8 west and 32 south edge faces.  ax[r][c]: the x-face west
// of the tile's column c in row r (c = 32: the east edge); ay[r][c]: the
// y-face south of row r (r = 8: the north edge).
struct DiffFaces {
  float ax[8][33];
  float ay[9][32];
};

struct NoHook {
  __device__ void operator()() const {}
};

// Cell (i, j) of a 32 x 8 tile at thread (threadIdx.x, threadIdx.y); every
// thread of the workgroup calls it (a workgroup barrier inside), `in` false
// for a thread whose cell lies outside the domain (To, Yo are then unused).
// `before_barrier` runs once the cell's loads are issued, before the barrier:
// what it stores to LDS is visible to the whole workgroup after the call.
template <class Hook = NoHook>
__device__ __forceinline__ void diffuse_cell(const cf_params& P, int nx, int ny, const float* __restrict__ Tin,
                                             const float4* __restrict__ Yin, const float* __restrict__ fx,
                                             const float* __restrict__ fy, float dts, int i, int j, bool in,
                                             DiffFaces& sh, float& To, float4& Yo, Hook before_barrier = Hook{}) {
  const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * 32 + tx;
  const int x0 = i - tx, y0 = j - ty;
  const int k = in ? j * nx + i : 0;
  // Every load is issued first, branch-free (a missing neighbour reads the
  // cell itself, and its term is not added), so the cell's loads are in
  // flight together; each term is then added only where the reference adds
  // it, in its order (west, east, south, north)
  const bool has[4] = {in && i > 0, in && i < nx - 1, in && j > 0, in && j < ny - 1};
  const int kn[4] = {has[0] ? k - 1 : k, has[1] ? k + 1 : k, has[2] ? k - nx : k, has[3] ? k + nx : k};
  float f[4], Tn[4];
  float4 Yn[4];
  const int kx = in ? j * (nx + 1) + i : 0, ky = in ? j * nx + i : 0;
  f[0] = fx[kx];
  f[1] = fx[kx + 1];
  f[2] = fy[ky];
  f[3] = fy[ky + (in ? nx : 0)];
  const float T0 = Tin[k];
  const float4 Y0 = Yin[k];
#pragma unroll
  for (int d = 0; d < 4; d++) {
    Tn[d] = Tin[kn[d]];
    Yn[d] = Yin[kn[d]];
  }
  // an edge face of the first 40 threads: the inner cell (ci, cj) of the
  // tile, the outer one to its west (tid < 8, row tid) or south (column tid - 8)
  const bool west = tid < 8;
  const int ci = west ? x0 : x0 + tid - 8, cj = west ? y0 + tid : y0;
  const bool eface = tid < 40 && ci < nx && cj < ny && (west ? ci > 0 : cj > 0);
  const int ek = eface ? cj * nx + ci : 0, eo = eface ? (west ? ek - 1 : ek - nx) : 0;
  const float eT0 = Tin[ek], eT1 = Tin[eo];
  const float ef = eface ? (west ? fx[cj * (nx + 1) + ci] : fy[ek]) : K(0);
  before_barrier();
  const float iPr = K(1) / P.Pr;
  auto weight = [&](bool ex, float fd, float ta, float tb) {
    return ex && fd > K(0) ? fd * rhoDOf(P, K(0.5) * (ta + tb), iPr) : K(0);
  };
  sh.ax[ty][tx + 1] = weight(has[1], f[1], T0, Tn[1]);
  sh.ay[ty + 1][tx] = weight(has[3], f[3], T0, Tn[3]);
  if (tid < 40) {
    const float w = weight(eface, ef, eT0, eT1);
    if (west) sh.ax[tid][0] = w;
    else sh.ay[0][tid - 8] = w;
  }
  __syncthreads();
  const float a[4] = {sh.ax[ty][tx], sh.ax[ty][tx + 1], sh.ay[ty][tx], sh.ay[ty + 1][tx]};
  float sT = K(0);
  float4 sY = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
  for (int d = 0; d < 4; d++) {
    if (has[d] && f[d] > K(0)) {
      sT += a[d] * (Tn[d] - T0);
      sY.x += a[d] * (Yn[d].x - Y0.x); sY.y += a[d] * (Yn[d].y - Y0.y);
      sY.z += a[d] * (Yn[d].z - Y0.z); sY.w += a[d] * (Yn[d].w - Y0.w);
    }
  }
  const float c = dts / (rhoOf(P, T0) * P.dx * P.dx);
  To = T0 + c * sT;
  Yo = make_float4(fmaxf(Y0.x + c * sY.x * K(1), K(0)), fmaxf(Y0.y + c * sY.y * K(1), K(0)),
                   fmaxf(Y0.z + c * sY.z * K(1), K(0)), fmaxf(Y0.w + c * sY.w * K(0), K(0)));
}

struct DiffuseOp {
  struct Args {
    Fields F;
    const float* Tin;
    const float4* Yin;
    float* Tout;
    float4* Yout;
    int vg;  // tiles of this scene's launch
  };
  static constexpr int kThreads = 256;
  static __device__ __forceinline__ void run(const Args& a_, int vb) {
    const Fields& F = a_.F;
    const float* __restrict__ Tin = a_.Tin;
    const float4* __restrict__ Yin = a_.Yin;
    float* __restrict__ Tout = a_.Tout;
    float4* __restrict__ Yout = a_.Yout;

  __shared__ DiffFaces sfaces;
  const cf_params P = load_params(F.P);
  const float dts = F.dyn->dsub;
  const int nx = P.nx, ny = P.ny;
  int i, j;
  tile_ij<32, 8>(vb, nx, ny, i, j);
  const bool in = i < nx && j < ny;
  float To;
  float4 Yo;
  diffuse_cell(P, nx, ny, Tin, Yin, F.fx, F.fy, dts, i, j, in, sfaces, To, Yo);
  if (!in) return;
  Tout[j * nx + i] = To;
  Yout[j * nx + i] = Yo;

  }
};
__global__ __launch_bounds__(256) void k_diffuse(const DiffuseOp::Args a_) {
  DiffuseOp::run(a_, xcd_tile(blockIdx.x, gridDim.x));
}

// -------------------------------------------------------------------------
// pointwise sources of one cell (shaders/sources.wgsl:41-157)
// -------------------------------------------------------------------------
// tamb4 = powf(P.Tamb, 4), evaluated once per workgroup by the caller
// What a cell's sources read besides T and Y, loaded by the caller with its
// other loads: cellG (sdf, solid id), the velocities of its four faces (for
// the entrainment) and T2 (the temperature after advection, for S)
struct SourceIn {
  float2 g;
  float uw, ue, vs, vn;
  float t2;
};
__device__ __forceinline__ SourceIn source_in(const Fields& F, int nx, int i, int j) {
  SourceIn r;
  r.g = F.cellG[j * nx + i];
  r.uw = F.U[j * (nx + 1) + i];
  r.ue = F.U[j * (nx + 1) + i + 1];
  r.vs = F.V[j * nx + i];
  r.vn = F.V[(j + 1) * nx + i];
  r.t2 = F.T2[j * nx + i];
  return r;
}

__device__ __forceinline__ void sources_cell(const cf_params& P, const Fields& F, int i, int j, float Tc, float4 Yc,
                                             const SourceIn& in, float& Tout, float4& Yout, float2& SQout,
                                             float& qrel, float& qrad, float& tpeak, int* sacc, float tamb4) {
  const float dx = P.dx, dt = P.dt, cp = P.cp, sO2 = P.sO2, dHc = P.dHc;
  // reciprocals for the correctly rounded quotients of div_rcp
  const float isO2 = K(1) / sO2, icp = K(1) / cp;
  const float2 g = in.g;
  const float sdf = g.x;
  // ids come from the geometry pass (piece*4 + face) or are SOLID_NONE;
  // anything else is treated as no solid rather than indexed
  uint32_t sid = __float_as_uint(g.y);
  if (sid >= (uint32_t)(CF_MAX_PIECES * 4)) sid = CF_SOLID_NONE;
  float massSrc = K(0);
  const float cellA = dx * dx;
  qrel = K(0); qrad = K(0); tpeak = K(0);
  if (sdf < K(0) && sid != CF_SOLID_NONE) {
    // inside wood: hold the face temperature (:50-55)
    Tout = F.pieces[(sid >> 2) * CF_PIECE_FLOATS + 8 + (sid & 3u)];
    Yout = Yc;
    SQout = make_float2(0.f, 0.f);
    return;
  }
  if (sid != CF_SOLID_NONE && sdf < K(2.0) * dx) {
    // wood coupling over the interface band (:57-94)
    const uint32_t pi = sid >> 2, fi = sid & 3u;
    const float* pc = F.pieces + pi * CF_PIECE_FLOATS;
    const float Ts = pc[8 + fi];
    const float band = (float)F.faceW[sid] / K(1.0e8);
    const float w = fmaxf(K(0), K(1) - fabsf(sdf - dx) / dx) / dx * (K(2.0) * pc[3]) / fmaxf(band, K(1e-6));
    const float rho = rhoOf(P, Tc);
    const float kg = rhoDOf(P, K(0.5) * (Tc + Ts)) * cp;
    const float dist = fmaxf(sdf, K(0.5) * dx);
    const float rate = kg * w / (dist * rho * cp);
    const float dTc = (Ts - Tc) * (K(1) - expf(-rate * dt));
    Tc += dTc;
    const float toWood = -rho * cp * dTc * cellA;
    atomicAdd(&sacc[(pi * 4u + fi) * 2u], iround(toWood * K(1e6)));
    const float mf = pc[12 + fi] * w;
    if (mf > K(0)) {
      const float madd = mf * dt;
      const float r = rhoOf(P, Tc);
      Yc = make_float4((r * Yc.x + madd) / (r + madd), (r * Yc.y + K(0)) / (r + madd),
                       (r * Yc.z + K(0)) / (r + madd), (r * Yc.w + K(0)) / (r + madd));
      Tc = (r * Tc + madd * Ts) / (r + madd);
      massSrc += mf;
    }
    const float cov = pc[16 + fi];
    if (cov > K(0) && Yc.y > K(0)) {
      const float r = rhoOf(P, Tc);
      const float kc = P.charA * expf(-P.charTa / Ts) * cov;
      const float mO2 = fminf(kc * r * Yc.y * w * dt, K(0.9) * r * Yc.y);
      const float mC = mO2 / P.charO2;
      Yc = make_float4((r * Yc.x + K(0)) / (r + mC), (r * Yc.y + -mO2) / (r + mC),
                       (r * Yc.z + (mC + mO2)) / (r + mC), (r * Yc.w + K(0)) / (r + mC));
      Tc = (r * Tc + mC * Ts) / (r + mC);
      massSrc += mC / dt;
      atomicAdd(&sacc[(pi * 4u + fi) * 2u + 1u], iround(mC * cellA * K(1e12)));
    }
  }
  // test sources (:96-107)
  const float x = ((float)i + K(0.5)) * dx, y = ((float)j + K(0.5)) * dx;
  if (P.heaterT > K(0) && x > P.heaterX0 && x < P.heaterX1 && y < P.heaterY1) Tc = P.heaterT;
  if (P.burnerFlux > K(0) && j == 0 && x > P.burnerX0 && x < P.burnerX1) {
    const float mf = P.burnerFlux / dx;
    const float madd = mf * dt;
    const float r = rhoOf(P, Tc);
    Yc = make_float4((r * Yc.x + madd) / (r + madd), (r * Yc.y + K(0)) / (r + madd),
                     (r * Yc.z + K(0)) / (r + madd), (r * Yc.w + K(0)) / (r + madd));
    Tc = (r * Tc + madd * P.burnerT) / (r + madd);
    massSrc += mf;
  }
  // one-step Arrhenius, 8 substeps with rho frozen (:109-120).  The rate's
  // last factor, exp(-Ta/T), is positive and finite, so where the fuel
  // fraction is zero the product is that zero (of the same sign) with any
  // positive factor: the exponential (and its quotient) is evaluated only
  // where the fuel fraction is not zero, a branch that waves of fuel-free
  // cells skip.  Where the fraction is +0, every dF is a zero (fminf returns
  // the zero even beside a NaN rate), so each substep adds zeros to Y, T and
  // burned: x + (+-0) = x for any x but a zero (denormals are kept), +0 +
  // (+-0) = +0, and burned starts at +0.  Such a cell leaves the substeps
  // bitwise unchanged unless one of the values it adds to is -0 (-0 + +0 =
  // +0), so those cells skip the whole loop.
  const float rho0 = rhoOf(P, Tc);
  const float h = dt / (float)8;
  float burned = K(0);
  constexpr uint32_t kNegZero = 0x80000000u;
  const bool inert = __float_as_uint(Yc.x) == 0u && __float_as_uint(Yc.y) != kNegZero &&
                     __float_as_uint(Yc.z) != kNegZero && __float_as_uint(Yc.w) != kNegZero &&
                     __float_as_uint(Tc) != kNegZero;
  if (!inert) {
#pragma unroll 1
    for (int s = 0; s < 8; s++) {
      float ea = K(1);
      if (Yc.x != K(0)) ea = expf(-P.reactTa / Tc);
      const float wr = P.reactB * rho0 * Yc.x * Yc.y * ea;
      const float dF = fminf(wr * h, fminf(Yc.x, div_rcp(Yc.y, sO2, isO2)));
      Yc.x += -dF; Yc.y += -sO2 * dF; Yc.z += (K(1) + sO2) * dF; Yc.w += K(0);
      Tc += div_rcp(dF * dHc, cp, icp);
      burned += dF;
    }
  }
  // soot formation and oxidation (:122-127)
  // (the same argument for the soot rates: a zero factor decides the product.
  // In a cell that skipped the substeps, formation adds zeros (dSf is a zero)
  // to values none of which is -0, and so does oxidation where the soot
  // fraction is +0 as well: those cells skip them too)
  float ef = K(1), eo = K(1);
  if (!inert) {
    if (Yc.x != K(0)) ef = expf(-P.sootFormTa / Tc);
    const float dSf = fminf(P.sootFormA * ef * Yc.x * dt, Yc.x);
    Yc.x += -dSf; Yc.y += K(0); Yc.z += K(0); Yc.w += dSf;
  }
  if (!inert || __float_as_uint(Yc.w) != 0u) {
    if (Yc.w != K(0) && Yc.y != K(0)) eo = expf(-P.sootOxTa / Tc);
    const float dSo = fminf(P.sootOxA * Yc.w * Yc.y * eo * dt, fminf(Yc.w, div_rcp(Yc.y, sO2, isO2)));
    Yc.x += K(0); Yc.y += -sO2 * dSo; Yc.z += (K(1) + sO2) * dSo; Yc.w += -dSo;
    Tc += div_rcp(dSo * dHc, cp, icp);
    burned += dSo;
  }
  qrel = rho0 * burned * dHc / dt;
  // optically thin loss, linearised implicit (:130-136)
  const float r1 = rhoOf(P, Tc);
  const float fv = div_rcp(r1 * Yc.w, P.sootDensity, K(1) / P.sootDensity);
  const float kap = P.kappaSootC * fv * Tc + P.kappaGas * Yc.z;
  const float T3 = Tc * Tc * Tc;
  qrad = K(4.0) * kap * P.sigma * (T3 * Tc - tamb4);
  const float rc = r1 * cp, irc = K(1) / rc;  // two quotients by r1 * cp
  const float dTr = div_rcp(qrad * dt, rc, irc) / (K(1) + div_rcp(K(16.0) * kap * P.sigma * T3 * dt, rc, irc));
  Tc -= dTr;
  // entrainment through the front/back faces (:140-148)
  const float ucx = K(0.5) * (in.uw + in.ue);
  const float ucy = K(0.5) * (in.vs + in.vn);
  const float me = K(2.0) * P.exchAlpha * sqrtf(ucx * ucx + ucy * ucy) * P.rhoAmb / P.depth;
  if (me > K(0)) {
    const float madd = me * dt;
    const float r = rhoOf(P, Tc);
    const float den = r + madd, iden = K(1) / den;  // five quotients by one denominator
    Yc = make_float4(div_rcp(r * Yc.x + madd * K(0), den, iden), div_rcp(r * Yc.y + madd * P.YOamb, den, iden),
                     div_rcp(r * Yc.z + madd * K(0), den, iden), div_rcp(r * Yc.w + madd * K(0), den, iden));
    Tc = div_rcp(r * Tc + madd * P.Tamb, den, iden);
    massSrc += me;
  }
  // write (:150-156)
  Tc = fmaxf(Tc, K(0.5) * P.Tamb);
  Yc = make_float4(fmaxf(Yc.x, K(0)), fmaxf(Yc.y, K(0)), fmaxf(Yc.z, K(0)), fmaxf(Yc.w, K(0)));
  const float S = massSrc / rhoOf(P, Tc) + logf(Tc / in.t2) / dt;
  Tout = Tc;
  Yout = Yc;
  tpeak = Tc;
  SQout = make_float2(S, qrel);
}

// sources, optionally fused with the last diffusion substep (which then reads
// Tin/Yin): a pointwise composition, identical to running the two in turn.
//
// The fixed-point accumulators receive the reference's integer contributions
// (per band cell and per 8x8 tile, shaders/sources.wgsl:72,92,169-171).
// Integer sums are order independent, so they are gathered per workgroup:
// face entries in LDS and flushed with one global atomic each, the three
// global sums as per-workgroup partials folded by k_acc_finish.  Same-address
// atomics from every 8x8 tile would otherwise serialise on one L2 line.
template <bool FUSE>
struct SourcesOp {
  struct Args {
    Fields F;
    const float* Tin;
    const float4* Yin;
    int* part;
    int vg;  // tiles of this scene's launch
  };
  static constexpr int kThreads = 256;
  static constexpr int kMinWaves = 5;
  static __device__ __forceinline__ void run(const Args& a_, int vb) {
    const Fields& F = a_.F;
    const float* __restrict__ Tin = a_.Tin;
    const float4* __restrict__ Yin = a_.Yin;
    int* __restrict__ part = a_.part;

  __shared__ float red[3][8][33];
  __shared__ DiffFaces sfaces;  // FUSE: the diffusion substep's face weights
  __shared__ int sacc[CF_ACC_GLOBAL];
  __shared__ int sglob[4][3];
  __shared__ float stamb4;
  __shared__ int sbad;
  const int tid = threadIdx.y * 32 + threadIdx.x;
  const cf_params P = load_params(F.P);
  // the workgroup's accumulators and flag, and Tamb^4 (one wave evaluates the
  // power), ready after the next workgroup barrier: in the fused form the
  // diffusion's, reached once every load is issued, so that the power holds
  // back no wave's loads
  auto init = [&]() {
    for (int q = tid; q < CF_ACC_GLOBAL; q += 256) sacc[q] = 0;
    if (tid == 0) { stamb4 = powf(P.Tamb, K(4.0)); sbad = 0; }
  };
  if constexpr (!FUSE) {
    init();
    __syncthreads();
  }
  const int nx = P.nx, ny = P.ny;
  int i, j;
  tile_ij<32, 8>(vb, nx, ny, i, j);
  float qrel = K(0), qrad = K(0), tpeak = K(0);
  const bool in = i < nx && j < ny;
  const int k = in ? j * nx + i : 0;
  SourceIn sin_{};
  if (in) sin_ = source_in(F, nx, i, j);  // issued with the diffusion's loads
  float Tc = K(0);
  float4 Yc = make_float4(0.f, 0.f, 0.f, 0.f);
  if constexpr (FUSE) diffuse_cell(P, nx, ny, Tin, Yin, F.fx, F.fy, F.dyn->dsub, i, j, in, sfaces, Tc, Yc, init);
  else if (in) { Tc = F.T[k]; Yc = F.Y[k]; }
  const float tamb4 = stamb4;
  if (in) {
    // the clamps below would hide a NaN (max(NaN, x) = x), so the inputs are checked
    if (!(isfinite(Tc) && isfinite(Yc.x) && isfinite(Yc.y) && isfinite(Yc.z) && isfinite(Yc.w))) sbad = 1;
    float To;
    float4 Yo;
    float2 sq;
    sources_cell(P, F, i, j, Tc, Yc, sin_, To, Yo, sq, qrel, qrad, tpeak, sacc, tamb4);
    F.T[k] = To;
    F.Y[k] = Yo;
    F.SQ[k] = sq;
  }
  // 8x8-tile reduction in the reference's tree order (:160-172)
  const float dx = P.dx, dt = P.dt;
  red[0][threadIdx.y][threadIdx.x] = qrel * dx * dx * dt;
  red[1][threadIdx.y][threadIdx.x] = qrad * dx * dx * dt;
  red[2][threadIdx.y][threadIdx.x] = tpeak;
  __syncthreads();
  const int w = tid >> 6, lane = tid & 63;
  const int lx = lane & 7, ly = lane >> 3;
  float vq = red[0][ly][w * 8 + lx], vr = red[1][ly][w * 8 + lx], vt = red[2][ly][w * 8 + lx];
  vq = wave_tree_sum(vq);
  vr = wave_tree_sum(vr);
  vt = wave_tree_max(vt);
  if (lane == 0) {
    sglob[w][0] = iround(vq * K(1e3));
    sglob[w][1] = iround(vr * K(1e3));
    sglob[w][2] = (int)vt;
  }
  __syncthreads();
  for (int q = tid; q < CF_ACC_GLOBAL; q += 256)
    if (sacc[q] != 0) atomicAdd(&F.acc[q], sacc[q]);
  if (tid == 0) {
    // two's-complement wrap-around, as repeated i32 atomicAdd gives
    const unsigned sq_ = (unsigned)sglob[0][0] + (unsigned)sglob[1][0] + (unsigned)sglob[2][0] + (unsigned)sglob[3][0];
    const unsigned sr_ = (unsigned)sglob[0][1] + (unsigned)sglob[1][1] + (unsigned)sglob[2][1] + (unsigned)sglob[3][1];
    const int mt = max(max(sglob[0][2], sglob[1][2]), max(sglob[2][2], sglob[3][2]));
    part[4 * vb + 0] = (int)sq_;
    part[4 * vb + 1] = (int)sr_;
    part[4 * vb + 2] = mt;
    part[4 * vb + 3] = sbad;
  }

  }
};
template <bool FUSE>
__global__ __launch_bounds__(256) void k_sources(const typename SourcesOp<FUSE>::Args a_) {
  SourcesOp<FUSE>::run(a_, xcd_tile(blockIdx.x, gridDim.x));
}

// Folds the per-workgroup partials of k_sources into acc[512..514]: integer
// sums with two's-complement wrap-around and a max, so the order is irrelevant.
// The partials' non-finite flags go into dyn->nonfinite (bit 0).
// Called by 256 threads; `scratch` holds 3 x 256 ints of LDS.
__device__ __forceinline__ void acc_fold(const int* __restrict__ part, int nparts, int* __restrict__ acc, Dyn* dyn,
                                         int* scratch) {
  const int t = threadIdx.y * blockDim.x + threadIdx.x;
  unsigned* sq = (unsigned*)scratch;
  unsigned* sr = sq + 256;
  int* st = (int*)(sr + 256);
  unsigned a = 0, b = 0;
  int m = INT_MIN, bad = 0;
  for (int q = t; q < nparts; q += 256) {
    a += (unsigned)part[4 * q];
    b += (unsigned)part[4 * q + 1];
    m = max(m, part[4 * q + 2]);
    bad |= part[4 * q + 3];
  }
  sq[t] = a; sr[t] = b; st[t] = m;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) { sq[t] += sq[t + s]; sr[t] += sr[t + s]; st[t] = max(st[t], st[t + s]); }
    __syncthreads();
  }
  if (t == 0) {
    atomicAdd(&acc[CF_ACC_GLOBAL], (int)sq[0]);
    atomicAdd(&acc[CF_ACC_GLOBAL + 1], (int)sr[0]);
    atomicMax(&acc[CF_ACC_GLOBAL + 2], st[0]);
  }
  if (bad) atomicOr(&dyn->nonfinite, 1);
  __syncthreads();
}

__global__ __launch_bounds__(256) void k_acc_finish(const int* __restrict__ part, int nparts, int* __restrict__ acc,
                                                    Dyn* dyn) {
  __shared__ int scratch[3 * 256];
  acc_fold(part, nparts, acc, dyn, scratch);
}

// -------------------------------------------------------------------------
// momentum: viscosity, buoyancy, entrainment drag (shaders/forces.wgsl:34-101)
// -------------------------------------------------------------------------
__device__ __forceinline__ float Tcl(const float* __restrict__ T, int nx, int ny, int i, int j) {
  return T[iclamp(j, 0, ny - 1) * nx + iclamp(i, 0, nx - 1)];
}
// reciprocals of the per-launch constants the momentum and coefficient
// updates divide by: their quotients are correctly rounded through div_rcp
struct MomRecip {
  float iTamb, iDepth, iRT;  // 1/Tamb, 1/depth, 1/(rhoAmb Tamb)
};
__device__ __forceinline__ MomRecip mom_recip(const cf_params& P) {
  return MomRecip{K(1) / P.Tamb, K(1) / P.depth, K(1) / (P.rhoAmb * P.Tamb)};
}
__device__ __forceinline__ float exchangef(const cf_params& P, float spd, float Tf, const MomRecip& r) {
  return expf(div_rcp(K(-2.0) * P.exchAlpha * spd * div_rcp(Tf, P.Tamb, r.iTamb), P.depth, r.iDepth) * P.dt);
}
// invRho of the closures with the reciprocal of rhoAmb Tamb
__device__ __forceinline__ float invRhoR(const cf_params& P, float T, const MomRecip& r) {
  return div_rcp(T, P.rhoAmb * P.Tamb, r.iRT);
}
// Momentum updates, written against an accessor so the same arithmetic runs
// from global memory (k_forces) or from an LDS tile (k_forces_coef).
//   T(i, j)   cell temperature at clamped indices (forces.wgsl:17-19)
//   U2/V2     advected face velocities, raw
//   fx/fy     open fractions
struct GlobalAcc {
  const float *T, *U2, *V2, *fx, *fy;
  int nx, ny;
  __device__ float Tc(int i, int j) const { return T[iclamp(j, 0, ny - 1) * nx + iclamp(i, 0, nx - 1)]; }
  // viscosity of the cell's temperature
  __device__ float mu(const cf_params& P, int i, int j) const { return muOf(P, Tc(i, j)); }
  __device__ float U(int i, int j) const { return U2[j * (nx + 1) + i]; }
  __device__ float V(int i, int j) const { return V2[j * nx + i]; }
  __device__ float FX(int i, int j) const { return fx[j * (nx + 1) + i]; }
  __device__ float FY(int i, int j) const { return fy[j * nx + i]; }
};

// new U at face (i, j), 0 <= i <= nx, 0 <= j < ny (shaders/forces.wgsl:44-71)
template <class A>
__device__ __forceinline__ float force_u(const cf_params& P, const A& a, int i, int j, const MomRecip& rc) {
  const int nx = P.nx, ny = P.ny;
  const float dt = P.dt;
  const float idx2 = K(1) / (P.dx * P.dx);
  auto u2 = [&](int x, int y) { return a.FX(x, y) > K(0) ? a.U(x, y) : K(0); };
  if (!(a.FX(i, j) > K(0))) return K(0);
  if (i == 0) return fminf(a.U(1, j), K(0));
  if (i == nx) return fmaxf(a.U(nx - 1, j), K(0));
  const float u = a.U(i, j);
  const float TL = a.Tc(i - 1, j), TR = a.Tc(i, j);
  const float rhoF = rhoOf(P, K(0.5) * (TL + TR));
  const float muE = a.mu(P, i, j), muW = a.mu(P, i - 1, j);
  // the north and south corner viscosities: the mean of four cells, or at
  // the top (floor) row the face's own two; the temperature is chosen first
  // and muOf evaluated once (at the edge the neighbour face is read at the
  // face itself and not used; Tc clamps)
  const bool hn = j < ny - 1, hs = j > 0;
  const float uN = hn ? u2(i, hn ? j + 1 : j) : u;
  const float muN = muOf(P, hn ? K(0.25) * (TL + TR + a.Tc(i - 1, j + 1) + a.Tc(i, j + 1)) : K(0.5) * (TL + TR));
  const float uS = hs ? u2(i, hs ? j - 1 : j) : -u;
  const float muS = muOf(P, hs ? K(0.25) * (TL + TR + a.Tc(i - 1, j - 1) + a.Tc(i, j - 1)) : K(0.5) * (TL + TR));
  const float visc = (muE * (u2(i + 1, j) - u) + muW * (u2(i - 1, j) - u) + muN * (uN - u) + muS * (uS - u)) * idx2 / rhoF;
  const float vAvg = K(0.25) * (a.V(i - 1, j) + a.V(i, j) + a.V(i - 1, j + 1) + a.V(i, j + 1));
  return (u + dt * visc) * exchangef(P, sqrtf(u * u + vAvg * vAvg), K(0.5) * (TL + TR), rc);
}

// new V at face (i, j), 0 <= i < nx, 0 <= j <= ny (shaders/forces.wgsl:72-100)
template <class A>
__device__ __forceinline__ float force_v(const cf_params& P, const A& a, int i, int j, const MomRecip& rc) {
  const int nx = P.nx, ny = P.ny;
  const float dt = P.dt;
  const float idx2 = K(1) / (P.dx * P.dx);
  auto v2 = [&](int x, int y) { return a.FY(x, y) > K(0) ? a.V(x, y) : K(0); };
  if (j == 0 || !(a.FY(i, j) > K(0))) return K(0);
  if (j == ny) return fmaxf(a.V(i, ny - 1), K(0));
  const float v = a.V(i, j);
  const float TB = a.Tc(i, j - 1), TT = a.Tc(i, j);
  const float Tf = K(0.5) * (TB + TT);
  const float rhoF = rhoOf(P, Tf);
  const float muN = a.mu(P, i, j), muS = a.mu(P, i, j - 1);
  // the east and west corner viscosities, as the corners of force_u
  const bool he = i < nx - 1, hw = i > 0;
  const float vE = he ? v2(he ? i + 1 : i, j) : v;
  const float muE = muOf(P, he ? K(0.25) * (TB + TT + a.Tc(i + 1, j - 1) + a.Tc(i + 1, j)) : Tf);
  const float vW = hw ? v2(hw ? i - 1 : i, j) : v;
  const float muW = muOf(P, hw ? K(0.25) * (TB + TT + a.Tc(i - 1, j - 1) + a.Tc(i - 1, j)) : Tf);
  const float visc = (muN * (v2(i, j + 1) - v) + muS * (v2(i, j - 1) - v) + muE * (vE - v) + muW * (vW - v)) * idx2 / rhoF;
  const float buoy = P.grav * (div_rcp(Tf, P.Tamb, rc.iTamb) - K(1));
  const float uAvg = K(0.25) * (a.U(i, j - 1) + a.U(i + 1, j - 1) + a.U(i, j) + a.U(i + 1, j));
  return (v + dt * (visc + buoy)) * exchangef(P, sqrtf(uAvg * uAvg + v * v), Tf, rc);
}

// face coefficients a = f / rho (shaders/pressure.wgsl:30-43)
template <class A>
__device__ __forceinline__ float coef_u(const cf_params& P, const A& a, int i, int j, const MomRecip& rc) {
  const int nx = P.nx;
  const float Tf = K(0.5) * (a.Tc(iclamp(i - 1, 0, nx - 1), j) + a.Tc(min(i, nx - 1), j));
  return a.FX(i, j) * invRhoR(P, Tf, rc);
}
template <class A>
__device__ __forceinline__ float coef_v(const cf_params& P, const A& a, int i, int j, const MomRecip& rc) {
  const int ny = P.ny;
  const float Tf = K(0.5) * (a.Tc(i, iclamp(j - 1, 0, ny - 1)) + a.Tc(i, min(j, ny - 1)));
  return a.FY(i, j) * invRhoR(P, Tf, rc);
}

__global__ __launch_bounds__(256) void k_forces(Fields F) {
  const cf_params P = load_params(F.P);
  const int nx = P.nx, ny = P.ny;
  int i, j;
  tile_ij<32, 8>(nx + 1, ny + 1, i, j);
  const GlobalAcc a{F.T, F.U2, F.V2, F.fx, F.fy, nx, ny};
  const MomRecip rc = mom_recip(P);
  if (i <= nx && j < ny) F.U[j * (nx + 1) + i] = force_u(P, a, i, j, rc);
  if (i < nx && j <= ny) F.V[j * nx + i] = force_v(P, a, i, j, rc);
}

// Fused momentum update + projection set-up on 32x32-cell tiles:
// forces (shaders/forces.wgsl:34-101), coef + rhs0 (shaders/pressure.wgsl:30-54)
// and the coefficient restriction (:57-67) cascaded through the coarse levels
// the tile halves into (5 levels: 16^2 .. 1^2 coarse cells per tile).
// T, U2, V2, fx and fy of the tile and a one-cell halo are staged in LDS with
// coalesced loads; every face update then reads LDS.  Each tile also computes
// the face column and row just past its edge (owned by the next tile) for its
// own rhs0 and coarse coefficients, instead of reading them back.
struct CoefLevels {
  Level L[16];
  int ncascade;  // coarse levels produced here (<= 5)
};

constexpr int FC_T = 32;                 // tile edge in cells
constexpr int FC_R = FC_T + 2;           // T region: [x0-1, x0+32] x [y0-1, y0+32]
constexpr int FC_UX = FC_T + 3;          // U faces [x0-1, x0+33] x rows [y0-1, y0+32]
constexpr int FC_VY = FC_T + 3;          // V faces cols [x0-1, x0+32] x rows [y0-1, y0+33]

struct TileAcc {
  const float *T, *MU, *U2, *V2, *fx, *fy;  // LDS, region-relative; MU: muOf of T
  int x0, y0;
  __device__ float Tc(int i, int j) const { return T[(j - y0 + 1) * FC_R + (i - x0 + 1)]; }
  __device__ float mu(const cf_params&, int i, int j) const { return MU[(j - y0 + 1) * FC_R + (i - x0 + 1)]; }
  __device__ float U(int i, int j) const { return U2[(j - y0 + 1) * FC_UX + (i - x0 + 1)]; }
  __device__ float FX(int i, int j) const { return fx[(j - y0 + 1) * FC_UX + (i - x0 + 1)]; }
  __device__ float V(int i, int j) const { return V2[(j - y0 + 1) * FC_R + (i - x0 + 1)]; }
  __device__ float FY(int i, int j) const { return fy[(j - y0 + 1) * FC_R + (i - x0 + 1)]; }
};

// Coarse level LV's face coefficients over a forces tile (TX = FC_T) from
// the tile's fine faces in LDS: sax (FC_T rows of FC_T + 1 U faces) and say
// (FC_T + 1 rows of FC_T V faces).  The face (I, J) is the restriction
// cascade's binary tree over 2^LV fine faces (ax: column 2^LV I, rows
// 2^LV J + k; ay: row 2^LV J, columns 2^LV I + k), each node 0.5 (a + b) of
// its two children as the level-by-level restriction forms it
// (shaders/pressure.wgsl:57-67), so the values are bitwise the cascade's.
template <int LV>
__device__ __forceinline__ void cascade_level(const Level& C, const float* sax, const float* say, int x0, int y0,
                                              int tid) {
  constexpr int TX = FC_T, cw = TX >> LV, nax = (cw + 1) * cw, N = 1 << LV;  // square tile: nay = nax
  const int xc = x0 >> LV, yc = y0 >> LV;
  for (int kk = tid; kk < 2 * nax; kk += 256) {
    const bool isx = kk < nax;
    const int k2 = isx ? kk : kk - nax;
    const int I = isx ? k2 % (cw + 1) : k2 % cw, J = isx ? k2 / (cw + 1) : k2 / cw;
    // fine face k of the tree: ax at (row N J + k, column N I), ay at (row N J, column N I + k)
    const int base = isx ? (J * N) * (TX + 1) + I * N : (J * N) * TX + I * N;
    const int step = isx ? TX + 1 : 1;
    const float* src = isx ? sax : say;
    float v[N];
#pragma unroll
    for (int k = 0; k < N; k++) v[k] = src[base + k * step];
#pragma unroll
    for (int w = N / 2; w >= 1; w /= 2)
#pragma unroll
      for (int k = 0; k < w; k++) v[k] = K(0.5) * (v[2 * k] + v[2 * k + 1]);
    const int gI = xc + I, gJ = yc + J;
    if (isx) {
      if (gJ < C.ny && ((I < cw && gI < C.nx) || gI == C.nx)) C.ax[gJ * (C.nx + 1) + gI] = v[0];
    } else {
      if (gI < C.nx && ((J < cw && gJ < C.ny) || gJ == C.ny)) C.ay[gJ * C.nx + gI] = v[0];
    }
  }
}

struct ForcesCoefOp {
  struct Args {
    Fields F;
    CoefLevels CL;
    const int* part;
    int nparts;
    int vg;  // tiles of this scene's launch; workgroup vg (when nparts > 0) folds the partials
  };
  static constexpr int kThreads = 256;
  static __device__ __forceinline__ void run(const Args& a_, int vb) {
    const Fields& F = a_.F;
    const CoefLevels& CL = a_.CL;
    const int* __restrict__ part = a_.part;
    int nparts = a_.nparts;

  constexpr int TX = FC_T, TY = FC_T;
  // One LDS pool: the staged T, its viscosity (muOf once per cell rather
  // than once per adjacent face), U2 and V2, all dead once the faces are
  // computed, then fx and fy.  The faces' forces and coefficients are kept
  // in registers and stored over the dead region, the forces for rhs0 and
  // then the coefficients for the cascade, so the tile needs 28 KB and five
  // workgroups share a CU.
  constexpr int OFF_MU = FC_R * FC_R, OFF_U = 2 * FC_R * FC_R, OFF_V = OFF_U + FC_R * FC_UX,
                OFF_FX = OFF_V + FC_VY * FC_R, OFF_FY = OFF_FX + FC_R * FC_UX, POOL = OFF_FY + FC_VY * FC_R;
  constexpr int NFU = (TX + 1) * TY, NFV = TX * (TY + 1), RF = (NFU + 255) / 256;
  static_assert(NFU == NFV && 2 * NFU <= OFF_FX && 3 * 256 <= POOL, "face arrays fit over the staged inputs");
  __shared__ float pool[POOL];
  float* sT = pool;
  float* sMU = pool + OFF_MU;
  float* sU = pool + OFF_U;
  float* sV = pool + OFF_V;
  float* sfx = pool + OFF_FX;
  float* sfy = pool + OFF_FY;
  float (*su)[TX + 1] = reinterpret_cast<float (*)[TX + 1]>(pool);
  float (*sv)[TX] = reinterpret_cast<float (*)[TX]>(pool + NFU);
  float (*sax)[TX + 1] = reinterpret_cast<float (*)[TX + 1]>(pool);
  float (*say)[TX] = reinterpret_cast<float (*)[TX]>(pool + NFU);
  const cf_params P = load_params(F.P);
  const MomRecip rc = mom_recip(P);
  const int tid = threadIdx.y * blockDim.x + threadIdx.x;  // 32 x 8 threads
  // the workgroup after the tiles folds the sources partials into acc, so
  // that no tile's work waits for it
  if (vb == a_.vg) {
    if (nparts > 0) acc_fold(part, nparts, F.acc, F.dyn, (int*)pool);
    return;
  }
  const int nx = P.nx, ny = P.ny;
  const int gx = (nx + TX - 1) / TX;
  const int t = vb;
  const int x0 = (t % gx) * TX, y0 = (t / gx) * TY;
  // ---- stage inputs (cells at clamped indices; faces outside the domain read
  // 0, never used).  Every load is issued before the first store, branch-free
  // (a face outside the domain loads face 0 and stores 0), so a thread's loads
  // are in flight together; the source term of the tile's cells for rhs0 is
  // loaded here too
  constexpr int NTC = FC_R * FC_R, NUC = FC_R * FC_UX, NVC = FC_VY * FC_R;
  constexpr int QT = (NTC + 255) / 256, QU = (NUC + 255) / 256, QV = (NVC + 255) / 256, QS = TX * TY / 256;
  static_assert(TX * TY % 256 == 0, "whole rounds of the tile's cells");
  float sqv[QS];
  {
    float tv[QT], uv[QU], fxv[QU], vv[QV], fyv[QV];
#pragma unroll
    for (int q = 0; q < QT; q++) {
      const int k = min(tid + 256 * q, NTC - 1);
      const int a = k % FC_R, b = k / FC_R;
      tv[q] = F.T[iclamp(y0 - 1 + b, 0, ny - 1) * nx + iclamp(x0 - 1 + a, 0, nx - 1)];
    }
#pragma unroll
    for (int q = 0; q < QU; q++) {
      const int k = min(tid + 256 * q, NUC - 1);
      const int i = x0 - 1 + k % FC_UX, j = y0 - 1 + k / FC_UX;
      const int idx = i >= 0 && i <= nx && j >= 0 && j < ny ? j * (nx + 1) + i : 0;
      uv[q] = F.U2[idx];
      fxv[q] = F.fx[idx];
    }
#pragma unroll
    for (int q = 0; q < QV; q++) {
      const int k = min(tid + 256 * q, NVC - 1);
      const int i = x0 - 1 + k % FC_R, j = y0 - 1 + k / FC_R;
      const int idx = i >= 0 && i < nx && j >= 0 && j <= ny ? j * nx + i : 0;
      vv[q] = F.V2[idx];
      fyv[q] = F.fy[idx];
    }
#pragma unroll
    for (int q = 0; q < QS; q++) {
      const int k = tid + 256 * q;
      const int i = x0 + k % TX, j = y0 + k / TX;
      sqv[q] = F.SQ[i < nx && j < ny ? j * nx + i : 0].x;
    }
#pragma unroll
    for (int q = 0; q < QT; q++) {
      const int k = tid + 256 * q;
      if (k < NTC) {
        sT[k] = tv[q];
        sMU[k] = muOf(P, tv[q]);
      }
    }
#pragma unroll
    for (int q = 0; q < QU; q++) {
      const int k = tid + 256 * q;
      if (k < NUC) {
        const int i = x0 - 1 + k % FC_UX, j = y0 - 1 + k / FC_UX;
        const bool in = i >= 0 && i <= nx && j >= 0 && j < ny;
        sU[k] = in ? uv[q] : 0.f;
        sfx[k] = in ? fxv[q] : 0.f;
      }
    }
#pragma unroll
    for (int q = 0; q < QV; q++) {
      const int k = tid + 256 * q;
      if (k < NVC) {
        const int i = x0 - 1 + k % FC_R, j = y0 - 1 + k / FC_R;
        const bool in = i >= 0 && i < nx && j >= 0 && j <= ny;
        sV[k] = in ? vv[q] : 0.f;
        sfy[k] = in ? fyv[q] : 0.f;
      }
    }
  }
  __syncthreads();
  const TileAcc acc{sT, sMU, sU, sV, sfx, sfy, x0, y0};
  // ---- U faces [x0, x0+32] x [y0, y0+32) and V faces [x0, x0+32) x
  // [y0, y0+32]: forces and coefficients (face r of this thread: q = tid + 256 r)
  float fu[RF], fa[RF], fv[RF], fb[RF];
#pragma unroll
  for (int r = 0; r < RF; r++) {
    const int q = tid + 256 * r;
    fu[r] = fa[r] = fv[r] = fb[r] = 0.f;
    if (q < NFU) {
      const int li = q % (TX + 1), lj = q / (TX + 1);
      const int i = x0 + li, j = y0 + lj;
      if (i <= nx && j < ny) {
        const float u = force_u(P, acc, i, j, rc), a = coef_u(P, acc, i, j, rc);
        fu[r] = u; fa[r] = a;
        if (li < TX || i == nx) { F.U[j * (nx + 1) + i] = u; F.ax0[j * (nx + 1) + i] = a; }
      }
    }
    if (q < NFV) {
      const int li = q % TX, lj = q / TX;
      const int i = x0 + li, j = y0 + lj;
      if (i < nx && j <= ny) {
        const float v = force_v(P, acc, i, j, rc), b = coef_v(P, acc, i, j, rc);
        fv[r] = v; fb[r] = b;
        if (lj < TY || j == ny) { F.V[j * nx + i] = v; F.ay0[j * nx + i] = b; }
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < RF; r++) {
    const int q = tid + 256 * r;
    if (q < NFU) su[q / (TX + 1)][q % (TX + 1)] = fu[r];
    if (q < NFV) sv[q / TX][q % TX] = fv[r];
  }
  __syncthreads();
  // ---- rhs0 (:46-54)
#pragma unroll
  for (int r = 0; r < QS; r++) {
    const int q = tid + 256 * r;
    const int li = q % TX, lj = q / TX;
    const int i = x0 + li, j = y0 + lj;
    if (i < nx && j < ny) {
      const float fe = acc.FX(i + 1, j), fw = acc.FX(i, j);
      const float fn = acc.FY(i, j + 1), fs = acc.FY(i, j);
      const float div = (fe * su[lj][li + 1] - fw * su[lj][li] + fn * sv[lj + 1][li] - fs * sv[lj][li]) / P.dx;
      const float open = fe + fw + fn + fs;
      F.rhs0[j * nx + i] = open > K(0) ? (div - sqv[r]) / P.dt : K(0);
    }
  }
  // ---- cascade of coefficient restrictions (:57-67), one pass: the face
  // (I, J) of coarse level l over the tile is the cascade's binary tree of
  // 2^l fine faces (ax: column 2^l I, rows 2^l J + k; ay: row 2^l J,
  // columns 2^l I + k), each node 0.5 (a + b) of its two children as the
  // level-by-level restriction forms it, so every level comes from the fine
  // faces in LDS after one barrier
  if (CL.ncascade > 0) {
    __syncthreads();
#pragma unroll
    for (int r = 0; r < RF; r++) {
      const int q = tid + 256 * r;
      if (q < NFU) sax[q / (TX + 1)][q % (TX + 1)] = fa[r];
      if (q < NFV) say[q / TX][q % TX] = fb[r];
    }
    __syncthreads();
  }
  if (CL.ncascade >= 1) cascade_level<1>(CL.L[1], &sax[0][0], &say[0][0], x0, y0, tid);
  if (CL.ncascade >= 2) cascade_level<2>(CL.L[2], &sax[0][0], &say[0][0], x0, y0, tid);
  if (CL.ncascade >= 3) cascade_level<3>(CL.L[3], &sax[0][0], &say[0][0], x0, y0, tid);
  if (CL.ncascade >= 4) cascade_level<4>(CL.L[4], &sax[0][0], &say[0][0], x0, y0, tid);
  if (CL.ncascade >= 5) cascade_level<5>(CL.L[5], &sax[0][0], &say[0][0], x0, y0, tid);

  }
};
__global__ __launch_bounds__(256) void k_forces_coef(const ForcesCoefOp::Args a_) {
  ForcesCoefOp::run(a_, xcd_tile(blockIdx.x, gridDim.x));
}

__global__ __launch_bounds__(256) void k_coef_rhs(Fields F) {
  const cf_params P = load_params(F.P);
  const int nx = P.nx, ny = P.ny;
  int i, j;
  tile_ij<32, 8>(nx + 1, ny + 1, i, j);
  const float* __restrict__ T = F.T;
  if (i <= nx && j < ny) {
    const int k = j * (nx + 1) + i;
    const float Tf = K(0.5) * (T[j * nx + iclamp(i - 1, 0, nx - 1)] + T[j * nx + min(i, nx - 1)]);
    F.ax0[k] = F.fx[k] * invRho(P, Tf);
  }
  if (i < nx && j <= ny) {
    const int k = j * nx + i;
    const float Tf = K(0.5) * (T[iclamp(j - 1, 0, ny - 1) * nx + i] + T[min(j, ny - 1) * nx + i]);
    F.ay0[k] = F.fy[k] * invRho(P, Tf);
  }
  if (i < nx && j < ny) {
    const float fe = F.fx[j * (nx + 1) + i + 1], fw = F.fx[j * (nx + 1) + i];
    const float fn = F.fy[(j + 1) * nx + i], fs = F.fy[j * nx + i];
    const float div = (fe * F.U[j * (nx + 1) + i + 1] - fw * F.U[j * (nx + 1) + i] + fn * F.V[(j + 1) * nx + i] -
                       fs * F.V[j * nx + i]) / P.dx;
    const float open = fe + fw + fn + fs;
    F.rhs0[j * nx + i] = open > K(0) ? (div - F.SQ[j * nx + i].x) / P.dt : K(0);
  }
}

// coefficient restriction to the next coarser level (:57-67)
struct CoefRestrictOp {
  struct Args {
    Level f;
    Level c;
    int vg;  // tiles of this scene's launch
  };
  static constexpr int kThreads = 256;
  static __device__ __forceinline__ void run(const Args& a_, int vb) {
    const Level& f = a_.f;
    const Level& c = a_.c;

  int I, J;
  tile_ij<32, 8>(vb, c.nx + 1, c.ny + 1, I, J);
  const int nx = f.nx, nxc = c.nx, nyc = c.ny;
  if (I <= nxc && J < nyc) c.ax[J * (nxc + 1) + I] = K(0.5) * (f.ax[(2 * J) * (nx + 1) + 2 * I] + f.ax[(2 * J + 1) * (nx + 1) + 2 * I]);
  if (I < nxc && J <= nyc) c.ay[J * nxc + I] = K(0.5) * (f.ay[(2 * J) * nx + 2 * I] + f.ay[(2 * J) * nx + 2 * I + 1]);

  }
};
__global__ __launch_bounds__(256) void k_coef_restrict(const CoefRestrictOp::Args a_) {
  CoefRestrictOp::run(a_, xcd_tile(blockIdx.x, gridDim.x));
}

// FP16 copies of one level's face coefficients for the FP16 smoother
// (round to nearest even; a face stays closed, 0, exactly)
struct CoefHalfOp {
  struct Args {
    Level L;
    LevelH H;
    int vg;  // tiles of this scene's launch
  };
  static constexpr int kThreads = 256;
  static __device__ __forceinline__ void run(const Args& a_, int vb) {
    const Level& L = a_.L;
    int i, j;
    tile_ij<32, 8>(vb, L.nx + 1, L.ny + 1, i, j);
    if (i <= L.nx && j < L.ny) a_.H.ax[j * (L.nx + 1) + i] = __float2half(L.ax[j * (L.nx + 1) + i]);
    if (i < L.nx && j <= L.ny) a_.H.ay[j * L.nx + i] = __float2half(L.ay[j * L.nx + i]);
  }
};
__global__ __launch_bounds__(256) void k_coef_half(const CoefHalfOp::Args a_) {
  CoefHalfOp::run(a_, xcd_tile(blockIdx.x, gridDim.x));
}

// 5-point stencil: sum a*p_nb and the diagonal (:77-85)
__device__ __forceinline__ void stencil(const Level& L, const float* __restrict__ p, int i, int j, float& off, float& diag) {
  const int nx = L.nx, ny = L.ny;
  const float w = L.ax[j * (nx + 1) + i], e = L.ax[j * (nx + 1) + i + 1];
  const float s = L.ay[j * nx + i], n = L.ay[(j + 1) * nx + i];
  off = K(0); diag = K(0);
  if (i > 0) { off += w * p[j * nx + i - 1]; diag += w; } else { diag += K(2.0) * w; }
  if (i < nx - 1) { off += e * p[j * nx + i + 1]; diag += e; } else { diag += K(2.0) * e; }
  if (j > 0) { off += s * p[(j - 1) * nx + i]; diag += s; }
  if (j < ny - 1) { off += n * p[(j + 1) * nx + i]; diag += n; } else { diag += K(2.0) * n; }
}

// one red or black half-sweep (:87-94); thread per cell of that colour
__global__ __launch_bounds__(256) void k_relax(Level L, int color) {
  int ii, j;
  tile_ij<32, 8>((L.nx + 1) / 2, L.ny, ii, j);
  const int i = 2 * ii + ((j + color) & 1);
  if (i >= L.nx || j >= L.ny) return;
  float off, diag;
  stencil(L, L.p, i, j, off, diag);
  const int k = j * L.nx + i;
  L.p[k] = diag > K(0) ? (off - L.rhs[k] * L.h2) / diag : K(0);
}

__device__ __forceinline__ float residual(const Level& L, int i, int j) {
  float off, diag;
  stencil(L, L.p, i, j, off, diag);
  if (diag <= K(0)) return K(0);
  return L.rhs[j * L.nx + i] - (off - diag * L.p[j * L.nx + i]) / L.h2;
}

// residual restricted to the coarse rhs; coarse p = 0 (:104-113)
__global__ __launch_bounds__(256) void k_restrict_res(Level f, Level c) {
  int I, J;
  tile_ij<32, 8>(c.nx, c.ny, I, J);
  if (I >= c.nx || J >= c.ny) return;
  const float r = residual(f, 2 * I, 2 * J) + residual(f, 2 * I + 1, 2 * J) + residual(f, 2 * I, 2 * J + 1) +
                  residual(f, 2 * I + 1, 2 * J + 1);
  c.rhs[J * c.nx + I] = K(0.25) * r;
  c.p[J * c.nx + I] = K(0);
}

// ---- step control ------------------------------------------------------------
// The last workgroup to finish (a ticket from an atomic counter) completes
// the reduction, so each control value costs one launch and the step stays a
// fixed sequence of launches that a graph can replay.

// CFL time step: dt = cfl * dx / max(max |u|, umin) over every U and V face.
// Max is order independent: per-workgroup maxima meet in one atomicMax on
// the float bits (non-negative floats order like their bit patterns).
struct CflOp {
  struct Args {
    Fields F;
    float cfl;
    float umin;
    int nsub;
    int vg;  // tiles of this scene's launch
  };
  static constexpr int kThreads = 256;
  static __device__ __forceinline__ void run(const Args& a_, int vb) {
    const Fields& F = a_.F;
    float cfl = a_.cfl;
    float umin = a_.umin;
    int nsub = a_.nsub;

  __shared__ unsigned wm[4];
  __shared__ int last;
  const cf_params P = load_params(F.P);
  const int nx = P.nx, ny = P.ny;
  const int nU = (nx + 1) * ny, nV = nx * (ny + 1), n = nU > nV ? nU : nV;
  const int tid = threadIdx.x;
  // max |u| as the max of the bit patterns of |u|: the same order for finite
  // values, and Inf and NaN (above every finite pattern) are not dropped as
  // fmaxf would drop a NaN
  unsigned m = 0u;
  for (int k = vb * 256 + tid; k < n; k += a_.vg * 256) {
    if (k < nU) m = max(m, __float_as_uint(fabsf(F.U[k])));
    if (k < nV) m = max(m, __float_as_uint(fabsf(F.V[k])));
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) m = max(m, (unsigned)__shfl_xor((int)m, o, 64));
  if ((tid & 63) == 0) wm[tid >> 6] = m;
  __syncthreads();
  if (tid == 0) {
    const unsigned b = max(max(wm[0], wm[1]), max(wm[2], wm[3]));
    atomicMax(&F.dyn->umax_bits, b);
    __threadfence();
    last = atomicAdd(&F.dyn->cfl_done, 1u) == a_.vg - 1;
  }
  __syncthreads();
  if (last && tid == 0) {
    __threadfence();
    Dyn* d = F.dyn;
    const unsigned ub = atomicOr(&d->umax_bits, 0u);
    if (ub >= 0x7f800000u) atomicOr(&d->nonfinite, 2);  // Inf or NaN
    const float umax = __uint_as_float(ub);
    const float dt = cfl * P.dx / fmaxf(umax, umin);
    ((cf_params*)F.P)->dt = dt;
    d->dt = dt;
    d->dsub = (float)((double)dt / (double)nsub);
    d->time += (double)dt;
    atomicExch(&d->umax_bits, 0u);
    atomicExch(&d->cfl_done, 0u);
  }

  }
};
__global__ __launch_bounds__(256) void k_cfl(const CflOp::Args a_) {
  CflOp::run(a_, xcd_tile(blockIdx.x, gridDim.x));
}

// Fixed-order sums of NV per-thread values over a launch of vg workgroups
// of 256 threads, in two kernels: each wave folds its lanes as a tree (lane
// i += lane i + o, o = 32 ... 1) and each workgroup stores ((w0 + w1) + w2) +
// w3 of its waves in part[NV * vb + q] (store_partials); after the kernel
// boundary one workgroup folds the partials the same way, thread t adding
// partials t, t + 256, ... (fold_partials, totals in thread 0).  A last-
// workgroup ticket would save the second launch but needs a device-scope
// fence in every workgroup, which on this GPU writes back and invalidates
// the XCD's L2 each time: in kernels that stream FP64 vectors that costs
// more than the data movement.
template <int NV>
__device__ __forceinline__ void wave_block_sum(double (&v)[NV]) {
  __shared__ double w[NV][4];
  const int tid = threadIdx.x;
#pragma unroll
  for (int q = 0; q < NV; q++) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v[q] += __shfl_down(v[q], o, 64);
    if ((tid & 63) == 0) w[q][tid >> 6] = v[q];
  }
  __syncthreads();
  if (tid == 0) {
#pragma unroll
    for (int q = 0; q < NV; q++) v[q] = ((w[q][0] + w[q][1]) + w[q][2]) + w[q][3];
  }
}
template <int NV>
__device__ __forceinline__ void store_partials(double (&v)[NV], double* __restrict__ part, int vb) {
  wave_block_sum<NV>(v);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NV; q++) part[NV * vb + q] = v[q];
  }
}
template <int NV>
__device__ __forceinline__ void fold_partials(double (&v)[NV], const double* __restrict__ part, int vg) {
#pragma unroll
  for (int q = 0; q < NV; q++) {
    double a = 0;
    for (int b = threadIdx.x; b < vg; b += 256) a += part[NV * b + q];
    v[q] = a;
  }
  wave_block_sum<NV>(v);
}

// ||r|| / ||rhs|| over level 0 (sums of squares in FP64), after a V-cycle:
// per-workgroup partials here, then resid_fold adds them in a fixed order, so
// the value does not depend on scheduling.  Below tol it sets
// dyn->converged, and the remaining V-cycles of the step are skipped.
struct ResnormOp {
  struct Args {
    Level L;
    Dyn* d;
    double* part;
    int vg;  // tiles of this scene's launch
  };
  static constexpr int kThreads = 256;
  // residual(L, i, j) without branches: the neighbours are loaded from
  // clamped indices first, then the same FP32 terms in the same order
  static __device__ __forceinline__ float residual_bf(const Level& L, int i, int j) {
    const int nx = L.nx, ny = L.ny, k = j * nx + i;
    const bool hw = i > 0, he = i < nx - 1, hs = j > 0, hn = j < ny - 1;
    const float w = L.ax[j * (nx + 1) + i], e = L.ax[j * (nx + 1) + i + 1];
    const float s = L.ay[k], n = L.ay[k + nx];
    const float pw = L.p[hw ? k - 1 : k], pe = L.p[he ? k + 1 : k];
    const float ps = L.p[hs ? k - nx : k], pn = L.p[hn ? k + nx : k], pc = L.p[k];
    float off = K(0), diag = K(0);
    off = hw ? off + w * pw : off;
    diag = hw ? diag + w : diag + K(2.0) * w;
    off = he ? off + e * pe : off;
    diag = he ? diag + e : diag + K(2.0) * e;
    off = hs ? off + s * ps : off;
    diag = hs ? diag + s : diag;
    off = hn ? off + n * pn : off;
    diag = hn ? diag + n : diag + K(2.0) * n;
    const float r = L.rhs[k] - (off - diag * pc) / L.h2;
    return diag <= K(0) ? K(0) : r;
  }
  static __device__ __forceinline__ void run(const Args& a_, int vb) {
    const Level& L = a_.L;
    if (a_.d->converged) return;
    const int n = L.nx * L.ny;
    double v[2] = {0, 0};  // ||r||^2, ||rhs||^2
    for (int k = vb * 256 + (int)threadIdx.x; k < n; k += a_.vg * 256) {
      const int j = k / L.nx, i = k - j * L.nx;
      const float r = residual_bf(L, i, j);
      const double b = (double)L.rhs[k];
      v[0] += (double)r * (double)r;
      v[1] += b * b;
    }
    store_partials<2>(v, a_.part, vb);
  }
};
__global__ __launch_bounds__(256) void k_resnorm(const ResnormOp::Args a_) {
  ResnormOp::run(a_, xcd_tile(blockIdx.x, gridDim.x));
}

// ---- defect correction at level 0 (schedule refine = 1) ----------------------
// The FP32 pressure cannot hold a solution whose residual is below ~1e-5 of
// ||rhs|| on a burning scene (rounding of p alone leaves that much), so the
// outer iteration keeps p in FP64 and solves A e = b - A p with FP32 V-cycles.

// A v at cell (i, j) of level 0 in FP64: bv = (sum a v_nb - diag v) / h2
// with the stencil of shaders/pressure.wgsl:77-94 (twice the face coefficient
// at the open sides and the top, none at the floor), the coefficients
// promoted; returns whether the cell is open (diag > 0).  Branch-free: the
// four neighbours are loaded from clamped indices before any arithmetic, so a
// thread's loads are in flight together, and the terms are then added in the
// fixed order w, e, s, n.  val(k) gives v at cell k.
template <class V>
__device__ __forceinline__ bool apply64(const Level& L, const V& val, int i, int j, double vc, double& bv) {
  const int nx = L.nx, ny = L.ny, k = j * nx + i;
  const bool hw = i > 0, he = i < nx - 1, hs = j > 0, hn = j < ny - 1;
  const double w = L.ax[j * (nx + 1) + i], e = L.ax[j * (nx + 1) + i + 1];
  const double s = L.ay[k], n = L.ay[k + nx];
  const double vw = val(hw ? k - 1 : k), ve = val(he ? k + 1 : k);
  const double vs = val(hs ? k - nx : k), vn = val(hn ? k + nx : k);
  double off = 0, diag = 0;
  off = hw ? off + w * vw : off;
  diag = hw ? diag + w : diag + 2.0 * w;
  off = he ? off + e * ve : off;
  diag = he ? diag + e : diag + 2.0 * e;
  off = hs ? off + s * vs : off;
  diag = hs ? diag + s : diag;
  off = hn ? off + n * vn : off;
  diag = hn ? diag + n : diag + 2.0 * n;
  bv = (off - diag * vc) / (double)L.h2;
  return diag > 0;
}

// the residual b - A p of one cell with p in FP64, 0 at a closed cell
__device__ __forceinline__ double residual64(const Level& L, const double* __restrict__ p, int i, int j, double b) {
  double bv;
  const int k = j * L.nx + i;
  const bool open = apply64(L, [&](int q) { return p[q]; }, i, j, p[k], bv);
  return open ? b - bv : 0.0;
}

// end of a residual evaluation, in the fold's thread 0: ||r|| / ||b|| into
// dyn, converged when below tol; the FP16 smoother's scale for the next
// correction: p is stored scaled by the power of two that brings the rms of
// this residual into [0.5, 1), so the correction's values sit in the normal
// range of FP16 (the scaling itself is exact)
__device__ __forceinline__ double resid_done(Dyn* d, double rr, double bb, int ncell, float tol, bool count) {
  const double r = bb > 0 ? sqrt(rr / bb) : sqrt(rr);
  int ex = 0;
  const double rms = sqrt(rr / (double)ncell);
  if (rms > 0) {
    frexp(rms, &ex);
    ex = ex < -60 ? -60 : (ex > 60 ? 60 : ex);
  }
  d->sig = (float)ldexp(1.0, -ex);
  d->isig = (float)ldexp(1.0, ex);
  d->resid = (float)r;
  if (count) d->cycles += 1;
  if (r < (double)tol) d->converged = 1;
  return r;
}

// r = b - A p64 into the level's rhs (the correction equation) with e = 0 as
// its starting guess, and per-workgroup partials of ||r||^2 and ||b||^2
// (FP64) for resid_fold.  first: b is the level's rhs as the forces pass
// wrote it, saved to b0 for the later iterations.
struct DcDefectOp {
  struct Args {
    Level L;
    __half* ph0;  // FP16 smoother: the correction's start e = 0 goes here instead of L.p
    double* p64;
    float* b0;
    Dyn* d;
    double* part;
    int first;
    int vg;  // tiles of this scene's launch
  };
  static constexpr int kThreads = 256;
  // FIRST: b from the level's rhs (saved to b0), else from b0; H: the
  // correction's start e = 0 goes to ph0 (FP16 smoother) instead of L.p
  template <bool FIRST, bool H>
  static __device__ __forceinline__ void sweep(const Args& a_, int vb, double (&v)[2]) {
    const Level& L = a_.L;
    const double* __restrict__ p64 = a_.p64;
    float* __restrict__ b0 = a_.b0;
    const int n = L.nx * L.ny;
    for (int k = vb * 256 + (int)threadIdx.x; k < n; k += a_.vg * 256) {
      const float bf = FIRST ? L.rhs[k] : b0[k];
      if (FIRST) b0[k] = bf;
      const double b = (double)bf;
      const int j = k / L.nx, i = k - j * L.nx;
      const double r = residual64(L, p64, i, j, b);
      L.rhs[k] = (float)r;
      if (H) a_.ph0[k] = __float2half(0.f);
      else L.p[k] = K(0);
      v[0] += r * r;
      v[1] += b * b;
    }
  }
  static __device__ __forceinline__ void run(const Args& a_, int vb) {
    if (a_.d->converged) return;
    double v[2] = {0, 0};  // ||r||^2, ||b||^2
    if (a_.first) {
      if (a_.ph0) sweep<true, true>(a_, vb, v);
      else sweep<true, false>(a_, vb, v);
    } else {
      if (a_.ph0) sweep<false, true>(a_, vb, v);
      else sweep<false, false>(a_, vb, v);
    }
    store_partials<2>(v, a_.part, vb);
  }
};
__global__ __launch_bounds__(256) void k_dc_defect(const DcDefectOp::Args a_) {
  DcDefectOp::run(a_, xcd_tile(blockIdx.x, gridDim.x));
}

// One workgroup per scene after dc_defect, resnorm or the Krylov start and
// step kernels: folds their partials (||r||^2, ||b||^2) and ends the residual
// evaluation: ||r|| / ||b||, the convergence flag, the FP16 smoother's scale
// and the cycle count.  Krylov (krylov = 1) divides by the ||b||^2 saved at
// the start; the start's residual is the lowest so far, and after a step one
// above twice the lowest of the solve switches the rest of the solve to
// defect correction.
struct ResidFoldOp {
  struct Args {
    Dyn* d;
    const double* part;
    int vg;     // workgroups of the partials' launch
    int ncell;  // level-0 cells
    float tol;
    int first;
    int krylov;
  };
  static constexpr int kThreads = 256;
  static __device__ __forceinline__ void run(const Args& a_, int) {
    Dyn* d = a_.d;
    if (d->converged) return;
    double v[2];
    fold_partials<2>(v, a_.part, a_.vg);
    if (threadIdx.x != 0) return;
    if (a_.krylov && a_.first) {
      d->kbest = resid_done(d, v[0], d->kbb, a_.ncell, a_.tol, false);
      d->kfall = 0;
      return;
    }
    if (a_.krylov && !d->kfall) {
      const double r = resid_done(d, v[0], d->kbb, a_.ncell, a_.tol, true);
      if (r > 2.0 * d->kbest) d->kfall = 1;
      if (r < d->kbest) d->kbest = r;
      return;
    }
    resid_done(d, v[0], v[1], a_.ncell, a_.tol, !a_.first);
  }
};
__global__ __launch_bounds__(256) void k_resid_fold(const ResidFoldOp::Args a_) { ResidFoldOp::run(a_, 0); }

// p64 += e after a correction V-cycle
struct DcAccumOp {
  struct Args {
    Level L;
    const __half* ph0;  // FP16 smoother: e is ph0 * isig
    const Dyn* dyn;
    double* p64;
    const int* skip;
    int vg;  // tiles of this scene's launch
  };
  static constexpr int kThreads = 256;
  static __device__ __forceinline__ void run(const Args& a_, int vb) {
    const Level& L = a_.L;
    double* __restrict__ p64 = a_.p64;
    const int* __restrict__ skip = a_.skip;

  if (*skip) return;
  const int n = L.nx * L.ny;
  if (a_.ph0) {
    const float is = a_.dyn->isig;
    for (int k = vb * 256 + threadIdx.x; k < n; k += a_.vg * 256) p64[k] += (double)(__half2float(a_.ph0[k]) * is);
  } else {
    for (int k = vb * 256 + threadIdx.x; k < n; k += a_.vg * 256) p64[k] += (double)L.p[k];
  }

  }
};
__global__ __launch_bounds__(256) void k_dc_accum(const DcAccumOp::Args a_) {
  DcAccumOp::run(a_, xcd_tile(blockIdx.x, gridDim.x));
}

// z of the last V-cycle at cell k: p (FP32 smoother) or, H, ph0 * isig
template <bool H>
__device__ __forceinline__ double kry_z(const Level& L, const __half* __restrict__ ph0, float is, int k) {
  return H ? (double)(__half2float(ph0[k]) * is) : (double)L.p[k];
}

// Flexible conjugate gradients with one stored direction (Notay's FCG(1)),
// the V-cycle's z = M r as the preconditioned residual, the residual
// recomputed from the pressure each iteration (dc_defect) rather than
// carried:
//   beta = (B z, d) / (d, B d)   (= (z, B d): B is symmetric),
//   d <- z - beta d,  alpha = (r, z) / (d, B d),  p64 += alpha d,
//   r = b - B p64,
// where (d, B d) of the new direction expands to (z, B z) - beta (B z, d_old),
// so one pass over z (kry_dots) gives the three dot products; r enters them
// as the FP32 right-hand side the V-cycle solved for.  kry_coef folds them
// into alpha and beta; after a fallback to defect correction it sets
// alpha = 1, beta = 0 (p64 += z) without them.
struct KryDotsOp {
  struct Args {
    Level L;
    const __half* ph0;
    Dyn* d;
    KryBufs k;
    double* part;
    int vg;
    int pp;  // kry_upd_def follows: the current direction is in the half of the iterations' parity
  };
  static constexpr int kThreads = 256;
  template <bool H>
  static __device__ __forceinline__ void sweep(const Args& a_, int vb, float is, double (&v)[3]) {
    const Level& L = a_.L;
    const float* __restrict__ d32 = a_.k.d32 + (a_.pp ? (size_t)(a_.d->cycles & 1) * (size_t)(L.nx * L.ny) : 0);
    const auto zat = [&](int q) { return kry_z<H>(L, a_.ph0, is, q); };
    const int n = L.nx * L.ny;
    for (int k = vb * 256 + (int)threadIdx.x; k < n; k += a_.vg * 256) {
      const int j = k / L.nx, i = k - j * L.nx;
      const double z = zat(k), dk = (double)d32[k], r = (double)L.rhs[k];
      double w;
      if (!apply64(L, zat, i, j, z, w)) w = 0.0;
      v[0] += w * dk;
      v[1] += r * z;
      v[2] += z * w;
    }
  }
  static __device__ __forceinline__ void run(const Args& a_, int vb) {
    Dyn* d = a_.d;
    if (d->converged || d->kfall) return;
    double v[3] = {0, 0, 0};  // (B z, d), (r, z), (z, B z)
    if (a_.ph0) sweep<true>(a_, vb, d->isig, v);
    else sweep<false>(a_, vb, 1.f, v);
    store_partials<3>(v, a_.part, vb);
  }
};
__global__ __launch_bounds__(256) void k_kry_dots(const KryDotsOp::Args a_) {
  KryDotsOp::run(a_, xcd_tile(blockIdx.x, gridDim.x));
}

// one workgroup per scene: alpha, beta and (d, B d) from kry_dots' partials
struct KryCoefOp {
  struct Args {
    Dyn* d;
    const double* part;
    int vg;
  };
  static constexpr int kThreads = 256;
  static __device__ __forceinline__ void run(const Args& a_, int) {
    Dyn* d = a_.d;
    if (d->converged) return;
    if (d->kfall) {
      if (threadIdx.x == 0) { d->kalpha = 1.0; d->kbeta = 0.0; }
      return;
    }
    double v[3];
    fold_partials<3>(v, a_.part, a_.vg);
    if (threadIdx.x != 0) return;
    const double beta = d->kdq != 0 ? v[0] / d->kdq : 0.0;
    const double dq = v[2] - beta * v[0];
    d->kalpha = dq != 0 ? v[1] / dq : 0.0;
    d->kbeta = beta;
    d->kdq = dq;
  }
};
__global__ __launch_bounds__(256) void k_kry_coef(const KryCoefOp::Args a_) { KryCoefOp::run(a_, 0); }

// d <- z - beta d (stored in FP32), p64 += alpha d along the stored d (own
// cell only; dc_defect follows)
struct KryUpdateOp {
  struct Args {
    Level L;
    const __half* ph0;
    Dyn* d;
    KryBufs k;
    int vg;
  };
  static constexpr int kThreads = 256;
  template <bool H>
  static __device__ __forceinline__ void sweep(const Args& a_, int vb, float is, double alpha, double beta) {
    const Level& L = a_.L;
    double* __restrict__ p64 = a_.k.p64;
    float* __restrict__ d32 = a_.k.d32;
    const int n = L.nx * L.ny;
    for (int k = vb * 256 + (int)threadIdx.x; k < n; k += a_.vg * 256) {
      const float dn = (float)(kry_z<H>(L, a_.ph0, is, k) - beta * (double)d32[k]);
      const double x = p64[k] + alpha * (double)dn;
      d32[k] = dn;
      p64[k] = x;
    }
  }
  static __device__ __forceinline__ void run(const Args& a_, int vb) {
    Dyn* d = a_.d;
    if (d->converged) return;
    if (a_.ph0) sweep<true>(a_, vb, d->isig, d->kalpha, d->kbeta);
    else sweep<false>(a_, vb, 1.f, d->kalpha, d->kbeta);
  }
};
__global__ __launch_bounds__(256) void k_kry_update(const KryUpdateOp::Args a_) {
  KryUpdateOp::run(a_, xcd_tile(blockIdx.x, gridDim.x));
}

// kry_update and dc_defect (after the start) in one pass: the new direction
// d <- z - beta d and pressure p64 + alpha d of a cell and of its four
// neighbours are formed from z, the stored d and p64 (kry_update's
// arithmetic), so r = b - B p64 (dc_defect's) needs no second pass over
// p64.  Half P = cycles & 1 of p64 and d32 holds the current values; the new
// ones go to the other half, so no cell's store meets a neighbour's read.
// The correction's start e = 0 is not stored: the next V-cycle's level-0
// down leg starts from zero (zero_p).  Same workgroups, loop and partials as
// dc_defect, so resid_fold sums the same values in the same order.
struct KryUpdDefOp {
  struct Args {
    Level L;
    const __half* ph0;
    Dyn* d;
    KryBufs k;
    double* part;
    int vg;
  };
  static constexpr int kThreads = 256;
  template <bool H>
  static __device__ __forceinline__ void sweep(const Args& a_, int vb, float is, double alpha, double beta, int P,
                                               double (&v)[2]) {
    const Level& L = a_.L;
    const int n = L.nx * L.ny;
    const double* __restrict__ pin = a_.k.p64 + (size_t)P * n;
    double* __restrict__ pout = a_.k.p64 + (size_t)(1 - P) * n;
    const float* __restrict__ din = a_.k.d32 + (size_t)P * n;
    float* __restrict__ dout = a_.k.d32 + (size_t)(1 - P) * n;
    const float* __restrict__ b0 = a_.k.b0;
    const auto dnew = [&](int q) { return (float)(kry_z<H>(L, a_.ph0, is, q) - beta * (double)din[q]); };
    const auto xnew = [&](int q) { return pin[q] + alpha * (double)dnew(q); };
    for (int k = vb * 256 + (int)threadIdx.x; k < n; k += a_.vg * 256) {
      const float dk = dnew(k);
      const double xk = pin[k] + alpha * (double)dk;
      dout[k] = dk;
      pout[k] = xk;
      const double b = (double)b0[k];
      const int j = k / L.nx, i = k - j * L.nx;
      double bv;
      const double r = apply64(L, xnew, i, j, xk, bv) ? b - bv : 0.0;
      L.rhs[k] = (float)r;
      v[0] += r * r;
      v[1] += b * b;
    }
  }
  static __device__ __forceinline__ void run(const Args& a_, int vb) {
    Dyn* d = a_.d;
    if (d->converged) return;
    double v[2] = {0, 0};  // ||r||^2, ||b||^2
    const int P = d->cycles & 1;
    if (a_.ph0) sweep<true>(a_, vb, d->isig, d->kalpha, d->kbeta, P, v);
    else sweep<false>(a_, vb, 1.f, d->kalpha, d->kbeta, P, v);
    store_partials<2>(v, a_.part, vb);
  }
};
__global__ __launch_bounds__(256) void k_kry_upd_def(const KryUpdDefOp::Args a_) {
  KryUpdDefOp::run(a_, xcd_tile(blockIdx.x, gridDim.x));
}

// after a step's last iteration: the direction and pressure back into the
// first half of d32 and p64 when an odd number of iterations left them in
// the second (run whether or not the step converged)
struct KrySettleOp {
  struct Args {
    Level L;
    const Dyn* d;
    KryBufs k;
    int vg;
  };
  static constexpr int kThreads = 256;
  static __device__ __forceinline__ void run(const Args& a_, int vb) {
    if (!(a_.d->cycles & 1)) return;
    const int n = a_.L.nx * a_.L.ny;
    double* __restrict__ p64 = a_.k.p64;
    float* __restrict__ d32 = a_.k.d32;
    for (int k = vb * 256 + (int)threadIdx.x; k < n; k += a_.vg * 256) {
      p64[k] = p64[n + k];
      d32[k] = d32[n + k];
    }
  }
};
__global__ __launch_bounds__(256) void k_kry_settle(const KrySettleOp::Args a_) {
  KrySettleOp::run(a_, xcd_tile(blockIdx.x, gridDim.x));
}

// Krylov start from the previous step's pressure p (in p64): a line search
// along p (rounded to FP32, the direction's storage) from zero.  kry_start
// saves b and stores d = p with the partials of (p, b), (p, B p), (b, b);
// kry_start_coef folds them into
// a = (p, b) / (p, B p) ((p, B p) = 0: a = 0 and no stored direction);
// kry_scale sets x = a p; dc_defect then forms r = b - B x.
struct KryStartOp {
  struct Args {
    Level L;
    Dyn* d;
    KryBufs k;
    double* part;
    int vg;
  };
  static constexpr int kThreads = 256;
  static __device__ __forceinline__ void run(const Args& a_, int vb) {
    const Level& L = a_.L;
    const double* __restrict__ p64 = a_.k.p64;
    float* __restrict__ d32 = a_.k.d32;
    float* __restrict__ b0 = a_.k.b0;
    const auto pat = [&](int q) { return (double)(float)p64[q]; };
    const int n = L.nx * L.ny;
    double v[3] = {0, 0, 0};  // (p, b), (p, B p), (b, b)
    for (int k = vb * 256 + (int)threadIdx.x; k < n; k += a_.vg * 256) {
      const int j = k / L.nx, i = k - j * L.nx;
      const float bf = L.rhs[k];
      const double b = (double)bf, p = pat(k);
      double bp;
      if (!apply64(L, pat, i, j, p, bp)) bp = 0.0;
      b0[k] = bf;
      d32[k] = (float)p;
      v[0] += p * b;
      v[1] += p * bp;
      v[2] += b * b;
    }
    store_partials<3>(v, a_.part, vb);
  }
};
__global__ __launch_bounds__(256) void k_kry_start(const KryStartOp::Args a_) {
  KryStartOp::run(a_, xcd_tile(blockIdx.x, gridDim.x));
}

struct KryStartCoefOp {
  struct Args {
    Dyn* d;
    const double* part;
    int vg;
  };
  static constexpr int kThreads = 256;
  static __device__ __forceinline__ void run(const Args& a_, int) {
    Dyn* d = a_.d;
    double v[3];
    fold_partials<3>(v, a_.part, a_.vg);
    if (threadIdx.x != 0) return;
    d->kalpha = v[1] != 0 ? v[0] / v[1] : 0.0;
    d->kdq = v[1];
    d->kbb = v[2];
  }
};
__global__ __launch_bounds__(256) void k_kry_start_coef(const KryStartCoefOp::Args a_) { KryStartCoefOp::run(a_, 0); }

struct KryScaleOp {
  struct Args {
    Level L;
    Dyn* d;
    KryBufs k;
    int vg;
  };
  static constexpr int kThreads = 256;
  static __device__ __forceinline__ void run(const Args& a_, int vb) {
    double* __restrict__ p64 = a_.k.p64;
    const float* __restrict__ d32 = a_.k.d32;
    const double a = a_.d->kalpha;
    const int n = a_.L.nx * a_.L.ny;
    for (int k = vb * 256 + (int)threadIdx.x; k < n; k += a_.vg * 256) p64[k] = a * (double)d32[k];
  }
};
__global__ __launch_bounds__(256) void k_kry_scale(const KryScaleOp::Args a_) {
  KryScaleOp::run(a_, xcd_tile(blockIdx.x, gridDim.x));
}

__global__ __launch_bounds__(256) void k_p32_to_64(const float* __restrict__ p, double* __restrict__ p64, int n) {
  for (int k = blockIdx.x * 256 + threadIdx.x; k < n; k += gridDim.x * 256) p64[k] = (double)p[k];
}

// bilinear coarse correction (:116-129)
__global__ __launch_bounds__(256) void k_prolong(Level f, Level c) {
  int i, j;
  tile_ij<32, 8>(f.nx, f.ny, i, j);
  if (i >= f.nx || j >= f.ny) return;
  const int nxc = c.nx, nyc = c.ny;
  const int I = i / 2, J = j / 2;
  const int I2 = iclamp(I + ((i & 1) == 1 ? 1 : -1), 0, nxc - 1);
  const int J2 = iclamp(J + ((j & 1) == 1 ? 1 : -1), 0, nyc - 1);
  const float* __restrict__ pc = c.p;
  const float cc = K(0.5625) * pc[J * nxc + I] + K(0.1875) * (pc[J * nxc + I2] + pc[J2 * nxc + I]) + K(0.0625) * pc[J2 * nxc + I2];
  const int nx = f.nx;
  // same diagonal as stencil(): w-term, e-term, then s, n (order kept)
  float d2 = K(0);
  {
    const float w = f.ax[j * (nx + 1) + i], e = f.ax[j * (nx + 1) + i + 1];
    const float s = f.ay[j * nx + i], n = f.ay[(j + 1) * nx + i];
    if (i > 0) d2 += w; else d2 += K(2.0) * w;
    if (i < nx - 1) d2 += e; else d2 += K(2.0) * e;
    if (j > 0) d2 += s;
    if (j < f.ny - 1) d2 += n; else d2 += K(2.0) * n;
  }
  f.p[j * nx + i] += d2 > K(0) ? cc : K(0);
}

// coarsest level: `iters` red-black sweeps from zero in one workgroup (:131-156)
__global__ __launch_bounds__(256) void k_coarsest(Level L, int iters) {
  __shared__ float wp[256];
  const int n = L.nx * L.ny;
  const int k = threadIdx.x;
  const int nx = L.nx, ny = L.ny;
  const int i = k % nx, j = k / nx;
  float w = 0, e = 0, s = 0, nn = 0, r = 0;
  if (k < n) {
    wp[k] = K(0);
    w = L.ax[j * (nx + 1) + i]; e = L.ax[j * (nx + 1) + i + 1];
    s = L.ay[j * nx + i]; nn = L.ay[(j + 1) * nx + i];
    r = L.rhs[k] * L.h2;
  }
  __syncthreads();
  for (int it = 0; it < iters; it++) {
    for (int c = 0; c < 2; c++) {
      if (k < n && ((i + j) & 1) == c) {
        float off = K(0), diag = K(0);
        if (i > 0) { off += w * wp[k - 1]; diag += w; } else { diag += K(2.0) * w; }
        if (i < nx - 1) { off += e * wp[k + 1]; diag += e; } else { diag += K(2.0) * e; }
        if (j > 0) { off += s * wp[k - nx]; diag += s; }
        if (j < ny - 1) { off += nn * wp[k + nx]; diag += nn; } else { diag += K(2.0) * nn; }
        wp[k] = diag > K(0) ? (off - r) / diag : K(0);
      }
      __syncthreads();
    }
  }
  if (k < n) L.p[k] = wp[k];
}

// velocity correction u -= dt (1/rho) grad p (:159-180)
// P64: the pressure differences come from the FP64 pressure of defect
// correction (rounded to FP32 after the subtraction), and the cells' FP32 p
// is refreshed from it
template <bool P64>
struct ProjectOp {
  struct Args {
    Fields F;
    const double* p64;
    int vg;  // tiles of this scene's launch
  };
  static constexpr int kThreads = 256;
  static __device__ __forceinline__ void run(const Args& a_, int vb) {
    const Fields& F = a_.F;
    const double* __restrict__ p64 = a_.p64;

  const cf_params P = load_params(F.P);
  const int nx = P.nx, ny = P.ny;
  int i, j;
  tile_ij<32, 8>(vb, nx + 1, ny + 1, i, j);
  const float dt = P.dt, dx = P.dx;
  const float* __restrict__ p = F.p0;
  const float* __restrict__ T = F.T;
  if (i <= nx && j < ny) {
    const int k = j * (nx + 1) + i;
    if (F.fx[k] > K(0)) {
      float dp;
      if (P64) {
        const double pl = i > 0 ? p64[j * nx + max(i - 1, 0)] : -p64[j * nx];
        const double pr = i < nx ? p64[j * nx + min(i, nx - 1)] : -p64[j * nx + nx - 1];
        dp = (float)(pr - pl);
      } else {
        const float pl = i > 0 ? p[j * nx + max(i - 1, 0)] : -p[j * nx];
        const float pr = i < nx ? p[j * nx + min(i, nx - 1)] : -p[j * nx + nx - 1];
        dp = pr - pl;
      }
      const float Tf = K(0.5) * (T[j * nx + iclamp(i - 1, 0, nx - 1)] + T[j * nx + min(i, nx - 1)]);
      F.U[k] -= dt * invRho(P, Tf) * dp / dx;
    } else F.U[k] = K(0);
  }
  if (i < nx && j <= ny) {
    const int k = j * nx + i;
    if (j > 0 && F.fy[k] > K(0)) {
      float dp;
      if (P64) {
        const double pb = p64[(j - 1) * nx + i];
        const double pt = j < ny ? p64[min(j, ny - 1) * nx + i] : -p64[(ny - 1) * nx + i];
        dp = (float)(pt - pb);
      } else {
        const float pb = p[(j - 1) * nx + i];
        const float pt = j < ny ? p[min(j, ny - 1) * nx + i] : -p[(ny - 1) * nx + i];
        dp = pt - pb;
      }
      const float Tf = K(0.5) * (T[(j - 1) * nx + i] + T[min(j, ny - 1) * nx + i]);
      F.V[k] -= dt * invRho(P, Tf) * dp / dx;
    } else F.V[k] = K(0);
  }
  if (P64 && i < nx && j < ny) F.p0[j * nx + i] = (float)p64[j * nx + i];

  }
};
template <bool P64>
__global__ __launch_bounds__(256) void k_project(const typename ProjectOp<P64>::Args a_) {
  ProjectOp<P64>::run(a_, xcd_tile(blockIdx.x, gridDim.x));
}

// -------------------------------------------------------------------------
// geometry: rasterise the pieces (shaders/geometry.wgsl:20-52, common.wgsl:47-62)
// -------------------------------------------------------------------------
// a piece's pose for the SDF: centre, cos and sin of the angle, half size,
// and whether the piece is active (pieces[7] > 0)
struct PiecePose {
  float x, y, c, s, h;
  int on;
};
// signed distance to the nearest active piece and its face id
// (shaders/geometry.wgsl:20-29, the box SDF of common.wgsl:47-62), from the
// poses computed once per workgroup.  `cand` lists, in increasing index
// order, the pieces that can be nearest anywhere in the workgroup's
// rectangle (every other piece is farther everywhere there), so the loop
// over them finds the same minimum and the same first piece reaching it as
// the loop over all pieces.
__device__ __forceinline__ float nearest_pose(const PiecePose* __restrict__ pp, const int* __restrict__ cand, int nc,
                                              float px, float py, uint32_t& bid) {
  float best = K(1e9);
  bid = CF_SOLID_NONE;
  for (int c = 0; c < nc; c++) {
    const int n = cand[c];
    const PiecePose q = pp[n];
    const float dxp = px - q.x, dyp = py - q.y;
    const float lx = q.c * dxp + q.s * dyp, ly = -q.s * dxp + q.c * dyp;
    const float qx = fabsf(lx) - q.h, qy = fabsf(ly) - q.h;
    const float mxq = fmaxf(qx, K(0)), myq = fmaxf(qy, K(0));
    const float d = sqrtf(mxq * mxq + myq * myq) + fminf(fmaxf(qx, qy), K(0));
    if (d < best) {
      best = d;
      uint32_t face;
      if (fabsf(lx) > fabsf(ly)) face = lx > K(0) ? 0u : 2u;
      else face = ly > K(0) ? 1u : 3u;
      bid = (uint32_t)n * 4u + face;
    }
  }
  return best;
}

struct GeometryOp {
  struct Args {
    Fields F;
    int vg;  // tiles of this scene's launch
  };
  static constexpr int kThreads = 256;
  static __device__ __forceinline__ void run(const Args& a_, int vb) {
    const Fields& F = a_.F;

  const cf_params P = load_params(F.P);
  const int nx = P.nx, ny = P.ny;
  // the pieces' poses, once per workgroup: every thread evaluates each piece
  // at three points, and the angle's cos and sin would dominate
  __shared__ PiecePose pp[CF_MAX_PIECES];
  __shared__ float lo[CF_MAX_PIECES], hi[CF_MAX_PIECES];
  __shared__ int cand[CF_MAX_PIECES];
  __shared__ int ncand;
  const int np = min(P.nPieces, CF_MAX_PIECES);
  const int tid = threadIdx.y * blockDim.x + threadIdx.x;
  const float dx = P.dx;
  // the workgroup's points (cell centres and the faces of 32 x 8 indices)
  // lie in the rectangle [rx0, rx1] x [ry0, ry1]
  const int gx = (nx + 1 + 31) / 32;
  const float rx0 = (float)((vb % gx) * 32) * dx, ry0 = (float)((vb / gx) * 8) * dx;
  const float rx1 = rx0 + K(32) * dx, ry1 = ry0 + K(8) * dx;
  if (tid < np) {
    const float* pc = F.pieces + tid * CF_PIECE_FLOATS;
    const bool on = pc[7] > K(0);
    pp[tid] = PiecePose{pc[0], pc[1], cosf(pc[2]), sinf(pc[2]), pc[3], on ? 1 : 0};
    // bounds of the piece's SDF over the rectangle: a point at distance r from
    // the centre is at least r - sqrt(2) h (the circumscribed circle) and at
    // most r - h (the inscribed circle) from the box, inside or out
    const float ex = fmaxf(fmaxf(rx0 - pc[0], pc[0] - rx1), K(0)), ey = fmaxf(fmaxf(ry0 - pc[1], pc[1] - ry1), K(0));
    const float fx_ = fmaxf(fabsf(pc[0] - rx0), fabsf(pc[0] - rx1)), fy_ = fmaxf(fabsf(pc[1] - ry0), fabsf(pc[1] - ry1));
    lo[tid] = on ? sqrtf(ex * ex + ey * ey) - K(1.41421357) * pc[3] : K(3e38);
    hi[tid] = on ? sqrtf(fx_ * fx_ + fy_ * fy_) - pc[3] : K(3e38);
  }
  __syncthreads();
  if (tid == 0) {
    // a piece can be nearest somewhere in the rectangle only if its lower
    // bound is below every piece's upper bound there; the margin covers the
    // rounding of the bounds and of the SDFs
    float mhi = K(3e38);
    for (int n = 0; n < np; n++) mhi = fminf(mhi, hi[n]);
    const float lim = mhi + K(1e-5) + K(1e-3) * dx;
    int c = 0;
    for (int n = 0; n < np; n++)
      if (pp[n].on && lo[n] <= lim) cand[c++] = n;
    ncand = c;
  }
  __syncthreads();
  const int nc = ncand;
  int i, j;
  tile_ij<32, 8>(vb, nx + 1, ny + 1, i, j);
  uint32_t id;
  if (i < nx && j < ny) {
    const float d = nearest_pose(pp, cand, nc, ((float)i + K(0.5)) * dx, ((float)j + K(0.5)) * dx, id);
    F.cellG[j * nx + i] = make_float2(d, __uint_as_float(id));
    if (id != CF_SOLID_NONE && d >= K(0) && d < K(2.0) * dx) {
      const float w = fmaxf(K(0), K(1) - fabsf(d - dx) / dx) / dx;
      atomicAdd(&F.faceW[id], iround(w * dx * dx * K(1.0e8)));
    }
  }
  if (i <= nx && j < ny) {
    const float d = nearest_pose(pp, cand, nc, (float)i * dx, ((float)j + K(0.5)) * dx, id);
    F.fx[j * (nx + 1) + i] = clampf(K(0.5) + d / dx, K(0), K(1));
  }
  if (i < nx && j <= ny) {
    const float d = nearest_pose(pp, cand, nc, ((float)i + K(0.5)) * dx, (float)j * dx, id);
    F.fy[j * nx + i] = j == 0 ? K(0) : clampf(K(0.5) + d / dx, K(0), K(1));
  }

  }
};
__global__ __launch_bounds__(256) void k_geometry(const GeometryOp::Args a_) {
  GeometryOp::run(a_, xcd_tile(blockIdx.x, gridDim.x));
}

// -------------------------------------------------------------------------
// radiation onto wood faces and floor probes (shaders/radiation.wgsl:28-158)
// one wave64 per target, one stratified ray per lane
// -------------------------------------------------------------------------
#define CF_PI K(3.14159265)
__device__ __forceinline__ float fractf_(float x) { return x - floorf(x); }
__device__ __forceinline__ float planck580(float T) {
  const float l = K(580e-9);
  const float x = K(1.438777e-2) / (l * T);
  if (x > K(80.0)) return K(0);
  return K(1.191042e-16) / (l * l * l * l * l) / (expf(x) - K(1)) * K(1e-9);
}
__device__ __forceinline__ float hitBox(float ox, float oy, float dxr, float dyr, const float* pc, float& face) {
  const float c = cosf(pc[2]), s = sinf(pc[2]);
  const float qx = ox - pc[0], qy = oy - pc[1];
  const float lox = c * qx + s * qy, loy = -s * qx + c * qy;
  const float ldx = c * dxr + s * dyr, ldy = -s * dxr + c * dyr;
  const float h = pc[3];
  const float invx = K(1) / (fabsf(ldx) < K(1e-9) ? K(1e-9) : ldx);
  const float invy = K(1) / (fabsf(ldy) < K(1e-9) ? K(1e-9) : ldy);
  const float t1x = (-h - lox) * invx, t1y = (-h - loy) * invy;
  const float t2x = (h - lox) * invx, t2y = (h - loy) * invy;
  const float tminx = fminf(t1x, t2x), tminy = fminf(t1y, t2y), tmaxx = fmaxf(t1x, t2x), tmaxy = fmaxf(t1y, t2y);
  const float tn = fmaxf(tminx, tminy), tf = fminf(tmaxx, tmaxy);
  if (tf < K(0) || tn > tf || tn < K(1e-5)) { face = K(0); return K(-1); }
  if (tminx > tminy) face = ldx < K(0) ? K(0) : K(2);
  else face = ldy < K(0) ? K(1) : K(3);
  return tn;
}
// the gas terms of one nearest-cell sample with T and Y of that cell
__device__ __forceinline__ void gasTerms(const cf_params& P, float Tg, float4 Yg, float& g0, float& g1) {
  const float fv = rhoOf(P, Tg) * Yg.w / P.sootDensity;
  const float kap = P.kappaSootC * fv * Tg + P.kappaGas * Yg.z;
  const float T2 = Tg * Tg;
  g0 = kap * P.sigma * (T2 * T2 - powf(P.Tamb, K(4.0))) / CF_PI;
  g1 = K(6.3) * fv / K(580e-9) * planck580(Tg);
}
// gasAt's two terms (:53-62) of every cell, once per frame: a ray sample
// then reads 8 bytes instead of T and Y (20 bytes) and evaluates no
// exponential.  The terms depend on the cell alone, so the integrals are
// unchanged.
struct GasTermsOp {
  struct Args {
    Fields F;
    int vg;  // tiles of this scene's launch
  };
  static constexpr int kThreads = 256;
  static __device__ __forceinline__ void run(const Args& a_, int vb) {
    const Fields& F = a_.F;
    const cf_params P = load_params(F.P);
    const int n = P.nx * P.ny;
    for (int k = vb * 256 + (int)threadIdx.x; k < n; k += a_.vg * 256) {
      float g0, g1;
      gasTerms(P, F.T[k], F.Y[k], g0, g1);
      F.G[k] = make_float2(g0, g1);
    }
  }
};
__global__ __launch_bounds__(256) void k_gas_terms(const GasTermsOp::Args a_) {
  GasTermsOp::run(a_, xcd_tile(blockIdx.x, gridDim.x));
}

// One workgroup per target, 64 rays (:64-158).  Each ray's march is split
// over RG lanes: lane g takes the samples m = g, g + RG, ... of every
// segment, and the lanes' partial integrals are added by a fixed tree
// (shuffles over the group).  The per-ray terms then meet in LDS and one wave
// adds them over the rays in the reference's workgroup tree order.  With
// RG = 16 a target is 16 waves, so a scene's 4·nPieces + 64 targets fill the
// chip's SIMDs instead of one wave each; the ray-piece tests of a segment are
// shared among the group's lanes too.
struct RadiationOp {
  struct Args {
    Fields F;
    int vg;  // tiles of this scene's launch
  };
  static constexpr int RG = 16;
  static constexpr int kThreads = 64 * RG;
  static __device__ __forceinline__ void run(const Args& a_, int vb) {
    const Fields& F = a_.F;
    __shared__ float sq[64], sv[64];
    // a vector load here: held in SGPRs the block costs this kernel VGPRs elsewhere (84 against 60)
    const cf_params P = *F.P;
    const int tgt = vb;
    const int li = (int)threadIdx.x / RG;  // the ray
    const int g = (int)threadIdx.x % RG;   // its lane in the ray's group
    const int nFaces = min(P.nPieces, CF_MAX_PIECES) * 4;
    float ox = 0, oy = 0, nrx = 0, nry = 1;
    bool valid = false;
    if (tgt < nFaces) {
      const float* pc = F.pieces + (tgt / 4) * CF_PIECE_FLOATS;
      const int f = tgt % 4;
      if (pc[7] > K(0)) {
        const float c = cosf(pc[2]), s = sinf(pc[2]);
        float lnx = 1, lny = 0;
        if (f == 1) { lnx = 0; lny = 1; } else if (f == 2) { lnx = -1; lny = 0; } else if (f == 3) { lnx = 0; lny = -1; }
        nrx = c * lnx - s * lny; nry = s * lnx + c * lny;
        const float tgx = -nry, tgy = nrx;
        const float u = (fractf_((float)li * K(0.618034) + (float)P.frame * K(0.1234)) - K(0.5)) * K(1.6) * pc[3];
        const float hh = pc[3] + K(1e-4);
        ox = pc[0] + nrx * hh + tgx * u;
        oy = pc[1] + nry * hh + tgy * u;
        valid = true;
      }
    } else {
      const int fi = tgt - nFaces;
      if (fi < CF_N_FLOOR) {
        ox = ((float)fi + K(0.5)) / (float)CF_N_FLOOR * (float)P.nx * P.dx;
        oy = K(1e-4);
        nrx = 0; nry = 1;
        valid = true;
      }
    }
    float dq = K(0), dv = K(0);
    if (valid) {  // uniform over the workgroup
      const float jit = fractf_((float)P.frame * K(0.7548777) + (float)tgt * K(0.5698403));
      const float a = (((float)li + jit) / K(64) - K(0.5)) * CF_PI;
      const float cosn = cosf(a);
      const float base = atan2f(nry, nrx);
      float dxr = cosf(base + a), dyr = sinf(base + a);
      float J = K(0), Jv = K(0);
      float Twall = P.Tamb;
      bool bounced = false;
      const float ds = P.dx;
      const float W = (float)P.nx * P.dx, H = (float)P.ny * P.dx;
      const int npc = min(P.nPieces, CF_MAX_PIECES);
      for (int step = 0; step < 4096; step++) {
        // the nearest piece along the ray: lane g of the ray's group tests
        // pieces g, g + RG, ...; the group then takes the smallest entry
        // distance, the lowest piece index among equal ones, which is the
        // piece the reference's loop in index order keeps (:108-113)
        float tHit = K(1e9);
        int hitN = -1, hitF = 0;
        for (int n = g; n < npc; n += RG) {
          const float* pc = F.pieces + n * CF_PIECE_FLOATS;
          if (!(pc[7] > K(0))) continue;
          float face;
          const float hx = hitBox(ox, oy, dxr, dyr, pc, face);
          if (hx > K(0) && hx < tHit) { tHit = hx; hitN = n; hitF = (int)face; }
        }
#pragma unroll
        for (int o = RG / 2; o > 0; o >>= 1) {
          const float t2 = __shfl_xor(tHit, o, RG);
          const int n2 = __shfl_xor(hitN, o, RG), f2 = __shfl_xor(hitF, o, RG);
          const bool take = t2 < tHit || (t2 == tHit && n2 >= 0 && (hitN < 0 || n2 < hitN));
          tHit = take ? t2 : tHit;
          hitN = take ? n2 : hitN;
          hitF = take ? f2 : hitF;
        }
        const float hitT = hitN >= 0 ? F.pieces[hitN * CF_PIECE_FLOATS + 8 + hitF] : K(0);
        float tFloor = K(1e9);
        if (dyr < K(-1e-6)) tFloor = -oy / dyr;
        const float tEnd = fminf(tHit, tFloor);
        float tx = K(1e9), ty = K(1e9);
        if (dxr > K(1e-6)) tx = (W - ox) / dxr; else if (dxr < K(-1e-6)) tx = -ox / dxr;
        if (dyr > K(1e-6)) ty = (H - oy) / dyr;
        const float tOut = fminf(fminf(tx, ty), tEnd);
        const int ns = toint(ceilf(tOut / ds));
        // gasAt (:53-62) at this lane's samples, RU at a time: every sample's
        // terms are loaded before any is used (a sample outside the grid
        // loads cell 0 and contributes 0)
        constexpr int RU = 8;
        for (int m0 = g; m0 < ns; m0 += RU * RG) {
          float2 gg[RU];
          bool in[RU];
#pragma unroll
          for (int u = 0; u < RU; u++) {
            const int m = m0 + u * RG;
            const float t = ((float)m + K(0.5)) * tOut / (float)ns;
            const float px = ox + dxr * t, py = oy + dyr * t;
            const int i = toint(px / P.dx), j = toint(py / P.dx);
            in[u] = m < ns && !(i < 0 || j < 0 || i >= P.nx || j >= P.ny);
            gg[u] = F.G[in[u] ? j * P.nx + i : 0];
          }
#pragma unroll
          for (int u = 0; u < RU; u++) {
            const float g0 = in[u] ? gg[u].x : K(0), g1 = in[u] ? gg[u].y : K(0);
            J += g0 * tOut / (float)ns;
            Jv += g1 * tOut / (float)ns;
          }
        }
        if (tHit < tFloor && tHit < K(1e8)) { Twall = hitT; break; }
        if (tFloor < K(1e8) && tFloor <= tHit && !bounced) {
          ox = ox + dxr * tFloor + K(0);
          oy = oy + dyr * tFloor + K(1e-5);
          dyr = -dyr;
          bounced = true;
          continue;
        }
        Twall = P.Tamb;
        break;
      }
      // the ray's integrals: the group's partials added by a fixed tree
#pragma unroll
      for (int o = RG / 2; o > 0; o >>= 1) {
        J += __shfl_xor(J, o, RG);
        Jv += __shfl_xor(Jv, o, RG);
      }
      const float w = CF_PI / K(64) * cosn;
      const float T2 = Twall * Twall;
      dq = w * (K(0.5) * P.sigma * T2 * T2 + K(2.0) * J);
      dv = w * (K(2.0) * Jv);
    }
    if (g == 0) { sq[li] = dq; sv[li] = dv; }
    __syncthreads();
    if (threadIdx.x < 64) {
      dq = wave_tree_sum(sq[threadIdx.x]);
      dv = wave_tree_sum(sv[threadIdx.x]);
      if (threadIdx.x == 0 && valid) {
        const float4 old = F.rad[tgt];
        const float am = old.x <= K(0) ? K(1.0) : K(0.35);
        F.rad[tgt] = make_float4(mixf(old.x, dq, am), mixf(old.y, dv, am), K(0), K(0));
      }
    }
  }
};
__global__ __launch_bounds__(RadiationOp::kThreads) void k_radiation(const RadiationOp::Args a_) {
  RadiationOp::run(a_, xcd_tile(blockIdx.x, gridDim.x));
}

// -------------------------------------------------------------------------
// utilities
// -------------------------------------------------------------------------
struct FillOp {
  struct Args {
    uint32_t* a;
    uint32_t v;
    size_t n;
    int vg;  // tiles of this scene's launch
  };
  static constexpr int kThreads = 256;
  static __device__ __forceinline__ void run(const Args& a_, int vb) {
    uint32_t* __restrict__ a = a_.a;
    uint32_t v = a_.v;
    size_t n = a_.n;

  size_t q = (size_t)vb * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)a_.vg * blockDim.x;
  for (; q < n; q += stride) a[q] = v;

  }
};
__global__ __launch_bounds__(256) void k_fill_u32(uint32_t* __restrict__ a, uint32_t v, size_t n) {
  const FillOp::Args a_{a, v, n, (int)gridDim.x};
  FillOp::run(a_, xcd_tile(blockIdx.x, gridDim.x));
}

// The bandwidth reference: a plain streaming copy, 16 B per lane per access,
// as many workgroups as fill the chip a few times over, each striding the
// array.  Its rate is the practical HBM peak the stencil kernels are held to.
__global__ __launch_bounds__(256) void k_bw_copy(const float4* __restrict__ a, float4* __restrict__ b, size_t n) {
  const size_t stride = (size_t)gridDim.x * 256;
  for (size_t k = (size_t)blockIdx.x * 256 + threadIdx.x; k < n; k += stride) b[k] = a[k];
}
void launch_bw_copy(const void* a, void* b, size_t bytes, hipStream_t s) {
  const size_t n = bytes / 16;
  unsigned g = (unsigned)((n + 255) / 256);
  if (g > 256 * 16) g = 256 * 16;
  launch(k_bw_copy, dim3(g), dim3(256), 0, s, (const float4*)a, (float4*)b, n);
}

// |div(f u) - S| after projection (the H.divErr diagnostic, dev/gas_test.html:101-111):
// per-workgroup max (as uint bits) and sum
__global__ __launch_bounds__(256) void k_div_error(Fields F, float* __restrict__ part) {
  __shared__ float smax[256], ssum[256];
  const cf_params P = load_params(F.P);
  const int nx = P.nx, ny = P.ny;
  int i, j;
  tile_ij<32, 8>(nx, ny, i, j);
  float e = K(0);
  if (i < nx && j < ny) {
    const int iu = j * (nx + 1) + i, iv = j * nx + i;
    const float div = (F.fx[iu + 1] * F.U[iu + 1] - F.fx[iu] * F.U[iu] + F.fy[iv + nx] * F.V[iv + nx] - F.fy[iv] * F.V[iv]) / P.dx;
    e = fabsf(div - F.SQ[iv].x);
  }
  const int t = threadIdx.y * 32 + threadIdx.x;
  smax[t] = e; ssum[t] = e;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) { smax[t] = fmaxf(smax[t], smax[t + s]); ssum[t] += ssum[t + s]; }
    __syncthreads();
  }
  if (t == 0) { part[2 * blockIdx.x] = smax[0]; part[2 * blockIdx.x + 1] = ssum[0]; }
}

// Divergence compatibility in FP64: the net source of the open cells
// (sum S dx^2) and the net outflow through the open domain faces
// (dx sum f u.n: -U at i = 0, +U at i = nx, -V at j = 0, +V at j = ny).  After
// a converged projection div(f u) = S in every open cell, so the two agree.
__global__ __launch_bounds__(256) void k_compat(Fields F, double* __restrict__ part) {
  __shared__ double ws[4], wo[4];
  const cf_params P = load_params(F.P);
  const int nx = P.nx, ny = P.ny;
  const double dx = (double)P.dx;
  double src = 0, out = 0;
  const int n = (nx + 1) * (ny + 1);
  for (int q = blockIdx.x * 256 + threadIdx.x; q < n; q += gridDim.x * 256) {
    const int i = q % (nx + 1), j = q / (nx + 1);
    if (i < nx && j < ny) {
      const int iu = j * (nx + 1) + i, iv = j * nx + i;
      const float open = F.fx[iu] + F.fx[iu + 1] + F.fy[iv] + F.fy[iv + nx];
      if (open > K(0)) src += (double)F.SQ[iv].x * dx * dx;
    }
    if (j < ny && (i == 0 || i == nx)) {
      const int iu = j * (nx + 1) + i;
      const double f = (double)F.fx[iu] * (double)F.U[iu] * dx;
      out += i == 0 ? -f : f;
    }
    if (i < nx && (j == 0 || j == ny)) {
      const int iv = j * nx + i;
      const double f = (double)F.fy[iv] * (double)F.V[iv] * dx;
      out += j == 0 ? -f : f;
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    src += __shfl_down(src, o, 64);
    out += __shfl_down(out, o, 64);
  }
  if ((threadIdx.x & 63) == 0) { ws[threadIdx.x >> 6] = src; wo[threadIdx.x >> 6] = out; }
  __syncthreads();
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = ((ws[0] + ws[1]) + ws[2]) + ws[3];
    part[2 * blockIdx.x + 1] = ((wo[0] + wo[1]) + wo[2]) + wo[3];
  }
}

// ---- launchers ----------------------------------------------------------
static inline unsigned nblk(int ex, int ey) { return (unsigned)(((ex + 31) / 32) * ((ey + 7) / 8)); }
static const dim3 B328(32, 8);

void launch_advect_fwd(const Fields& F, hipStream_t s) {
  launch(k_advect_fwd, dim3(nblk(F.nx + 1, F.ny + 1)), B328, 0, s, F);
}
void launch_advect_corr(const Fields& F, hipStream_t s) {
  launch(k_advect_corr, dim3(nblk(F.nx + 1, F.ny + 1)), B328, 0, s, F);
}
void launch_diffuse(const Fields& F, const float* Tin, const float4* Yin, float* Tout, float4* Yout, hipStream_t s) {
  const int nb = (int)nblk(F.nx, F.ny);
  launch(k_diffuse, dim3(nb), B328, 0, s, DiffuseOp::Args{F, Tin, Yin, Tout, Yout, nb});
}
unsigned sources_blocks(const Fields& F) { return nblk(F.nx, F.ny); }
void launch_sources(const Fields& F, bool fuse, const float* Tin, const float4* Yin, int* part, bool finish,
                    hipStream_t s) {
  const unsigned nb = nblk(F.nx, F.ny);
  if (fuse) launch(k_sources<true>, dim3(nb), B328, 0, s, SourcesOp<true>::Args{F, Tin, Yin, part, (int)nb});
  else launch(k_sources<false>, dim3(nb), B328, 0, s, SourcesOp<false>::Args{F, Tin, Yin, part, (int)nb});
  if (finish) launch(k_acc_finish, dim3(1), dim3(256), 0, s, (const int*)part, (int)nb, F.acc, F.dyn);
}
void launch_forces(const Fields& F, hipStream_t s) {
  launch(k_forces, dim3(nblk(F.nx + 1, F.ny + 1)), B328, 0, s, F);
}
int forces_coef_cascade(int nlev) {
  // a 32x32 tile halves into coarse tiles down to 1x1 cell: five levels
  return nlev - 1 < 5 ? nlev - 1 : 5;
}
void launch_forces_coef(const Fields& F, const Level* L, int nlev, const int* part, int nparts, hipStream_t s) {
  CoefLevels CL;
  for (int l = 0; l < nlev && l < 16; l++) CL.L[l] = L[l];
  CL.ncascade = forces_coef_cascade(nlev);
  const unsigned nb = (unsigned)(((F.nx + FC_T - 1) / FC_T) * ((F.ny + FC_T - 1) / FC_T));
  launch(k_forces_coef, dim3(nb + (nparts > 0 ? 1 : 0)), dim3(32, 8), 0, s,
         ForcesCoefOp::Args{F, CL, part, nparts, (int)nb});
}
void launch_coef_rhs(const Fields& F, hipStream_t s) {
  launch(k_coef_rhs, dim3(nblk(F.nx + 1, F.ny + 1)), B328, 0, s, F);
}
void launch_coef_restrict(const Level& f, const Level& c, hipStream_t s) {
  const int nb = (int)nblk(c.nx + 1, c.ny + 1);
  launch(k_coef_restrict, dim3(nb), B328, 0, s, CoefRestrictOp::Args{f, c, nb});
}
void launch_relax(const Level& L, int color, hipStream_t s) {
  launch(k_relax, dim3(nblk((L.nx + 1) / 2, L.ny)), B328, 0, s, L, color);
}
void launch_restrict_res(const Level& f, const Level& c, hipStream_t s) {
  launch(k_restrict_res, dim3(nblk(c.nx, c.ny)), B328, 0, s, f, c);
}
void launch_prolong(const Level& f, const Level& c, hipStream_t s) {
  launch(k_prolong, dim3(nblk(f.nx, f.ny)), B328, 0, s, f, c);
}
void launch_coarsest(const Level& L, int iters, hipStream_t s) {
  launch(k_coarsest, dim3(1), dim3(256), 0, s, L, iters);
}
void launch_project(const Fields& F, const double* p64, hipStream_t s) {
  const int nb = (int)nblk(F.nx + 1, F.ny + 1);
  if (p64) launch(k_project<true>, dim3(nb), B328, 0, s, ProjectOp<true>::Args{F, p64, nb});
  else launch(k_project<false>, dim3(nb), B328, 0, s, ProjectOp<false>::Args{F, p64, nb});
}
void launch_geometry(const Fields& F, hipStream_t s) {
  const int nb = (int)nblk(F.nx + 1, F.ny + 1);
  launch(k_geometry, dim3(nb), B328, 0, s, GeometryOp::Args{F, nb});
}
static int gas_terms_blocks(const Fields& F) {
  int b = (F.nx * F.ny + 255) / 256;
  return b > 2048 ? 2048 : b;
}
void launch_radiation(const Fields& F, int ntargets, hipStream_t s) {
  if (ntargets <= 0) return;
  const int gb = gas_terms_blocks(F);
  launch(k_gas_terms, dim3(gb), dim3(256), 0, s, GasTermsOp::Args{F, gb});
  launch(k_radiation, dim3(ntargets), dim3(RadiationOp::kThreads), 0, s, RadiationOp::Args{F, ntargets});
}
void launch_cfl(const Fields& F, float cfl, float umin, int nsub, hipStream_t s) {
  const int n = (F.nx + 1) * F.ny > F.nx * (F.ny + 1) ? (F.nx + 1) * F.ny : F.nx * (F.ny + 1);
  unsigned b = (unsigned)((n + 1023) / 1024);
  if (b > 1024) b = 1024;
  launch(k_cfl, dim3(b), dim3(256), 0, s, CflOp::Args{F, cfl, umin, nsub, (int)b});
}
void launch_dc_defect(const Level& L0, __half* ph0, double* p64, float* b0, Dyn* dyn, double* part, float tol,
                      bool first, hipStream_t s) {
  const int nb = (int)resnorm_blocks(L0);
  launch(k_dc_defect, dim3(nb), dim3(256), 0, s, DcDefectOp::Args{L0, ph0, p64, b0, dyn, part, first ? 1 : 0, nb});
  launch(k_resid_fold, dim3(1), dim3(256), 0, s,
         ResidFoldOp::Args{dyn, part, nb, L0.nx * L0.ny, tol, first ? 1 : 0, 0});
}
void launch_kry_start(const Level& L0, __half* ph0, Dyn* dyn, const KryBufs& k, double* part, float tol,
                      hipStream_t s) {
  const int nb = (int)resnorm_blocks(L0);
  launch(k_kry_start, dim3(nb), dim3(256), 0, s, KryStartOp::Args{L0, dyn, k, part, nb});
  launch(k_kry_start_coef, dim3(1), dim3(256), 0, s, KryStartCoefOp::Args{dyn, part, nb});
  launch(k_kry_scale, dim3(nb), dim3(256), 0, s, KryScaleOp::Args{L0, dyn, k, nb});
  launch(k_dc_defect, dim3(nb), dim3(256), 0, s, DcDefectOp::Args{L0, ph0, k.p64, k.b0, dyn, part, 0, nb});
  launch(k_resid_fold, dim3(1), dim3(256), 0, s, ResidFoldOp::Args{dyn, part, nb, L0.nx * L0.ny, tol, 1, 1});
}
void launch_kry_step(const Level& L0, __half* ph0, Dyn* dyn, const KryBufs& k, double* part, float tol, bool fused,
                     hipStream_t s) {
  const int nb = (int)resnorm_blocks(L0);
  launch(k_kry_dots, dim3(nb), dim3(256), 0, s, KryDotsOp::Args{L0, ph0, dyn, k, part, nb, fused ? 1 : 0});
  launch(k_kry_coef, dim3(1), dim3(256), 0, s, KryCoefOp::Args{dyn, part, nb});
  if (fused) {
    launch(k_kry_upd_def, dim3(nb), dim3(256), 0, s, KryUpdDefOp::Args{L0, ph0, dyn, k, part, nb});
  } else {
    launch(k_kry_update, dim3(nb), dim3(256), 0, s, KryUpdateOp::Args{L0, ph0, dyn, k, nb});
    launch(k_dc_defect, dim3(nb), dim3(256), 0, s, DcDefectOp::Args{L0, ph0, k.p64, k.b0, dyn, part, 0, nb});
  }
  launch(k_resid_fold, dim3(1), dim3(256), 0, s, ResidFoldOp::Args{dyn, part, nb, L0.nx * L0.ny, tol, 0, 1});
}
void launch_kry_settle(const Level& L0, Dyn* dyn, const KryBufs& k, hipStream_t s) {
  const int nb = (int)resnorm_blocks(L0);
  launch(k_kry_settle, dim3(nb), dim3(256), 0, s, KrySettleOp::Args{L0, dyn, k, nb});
}
__global__ __launch_bounds__(256) void k_all_converged(Dyn* const* __restrict__ d, int n, int* __restrict__ out) {
  __shared__ int open;
  if (threadIdx.x == 0) open = 0;
  __syncthreads();
  for (int q = threadIdx.x; q < n; q += 256)
    if (!d[q]->converged) atomicOr(&open, 1);
  __syncthreads();
  if (threadIdx.x == 0) *out = open ? 0 : 1;
}
void launch_all_converged(Dyn* const* d, int n, int* out, hipStream_t s) {
  launch(k_all_converged, dim3(1), dim3(256), 0, s, d, n, out);
}
void launch_coef_half(const Level& L, const LevelH& H, hipStream_t s) {
  const int nb = (int)nblk(L.nx + 1, L.ny + 1);
  launch(k_coef_half, dim3(nb), B328, 0, s, CoefHalfOp::Args{L, H, nb});
}
void launch_dc_accum(const Level& L0, const __half* ph0, const Dyn* dyn, double* p64, const int* skip,
                     hipStream_t s) {
  unsigned b = (unsigned)((L0.nx * L0.ny + 1023) / 1024);
  if (b > 2048) b = 2048;
  launch(k_dc_accum, dim3(b), dim3(256), 0, s, DcAccumOp::Args{L0, ph0, dyn, p64, skip, (int)b});
}
void launch_p32_to_64(const float* p, double* p64, int n, hipStream_t s) {
  unsigned b = (unsigned)((n + 1023) / 1024);
  launch(k_p32_to_64, dim3(b > 2048 ? 2048 : b), dim3(256), 0, s, p, p64, n);
}
unsigned resnorm_blocks(const Level& L0) {
  unsigned b = (unsigned)((L0.nx * L0.ny + 1023) / 1024);
  return b > 1024 ? 1024 : b;
}
void launch_resnorm(const Level& L0, Dyn* dyn, double* part, float tol, hipStream_t s) {
  const int nb = (int)resnorm_blocks(L0);
  launch(k_resnorm, dim3(nb), dim3(256), 0, s, ResnormOp::Args{L0, dyn, part, nb});
  launch(k_resid_fold, dim3(1), dim3(256), 0, s, ResidFoldOp::Args{dyn, part, nb, L0.nx * L0.ny, tol, 0, 0});
}
void launch_fill_u32(void* a, uint32_t v, size_t n, hipStream_t s) {
  if (n == 0) return;
  unsigned b = (unsigned)((n + 255) / 256);
  if (b > 4096) b = 4096;
  launch(k_fill_u32, dim3(b), dim3(256), 0, s, (uint32_t*)a, v, n);
}
// ---- batched phases (one scene's contribution each) ----
void phase_diffuse(BatchPhase& ph, const Fields& F, const float* Tin, const float4* Yin, float* Tout, float4* Yout) {
  const int nb = (int)nblk(F.nx, F.ny);
  phase_add<DiffuseOp>(ph, DiffuseOp::Args{F, Tin, Yin, Tout, Yout, nb}, nb, B328, "diffuse");
}
void phase_sources(BatchPhase& ph, const Fields& F, const float* Tin, const float4* Yin, int* part) {
  const int nb = (int)nblk(F.nx, F.ny);
  phase_add<SourcesOp<true>>(ph, SourcesOp<true>::Args{F, Tin, Yin, part, nb}, nb, B328, "diffuse_sources");
}
void phase_forces_coef(BatchPhase& ph, const Fields& F, const Level* L, int nlev, const int* part, int nparts) {
  CoefLevels CL;
  for (int l = 0; l < nlev && l < 16; l++) CL.L[l] = L[l];
  CL.ncascade = forces_coef_cascade(nlev);
  const int nb = ((F.nx + FC_T - 1) / FC_T) * ((F.ny + FC_T - 1) / FC_T);
  phase_add<ForcesCoefOp>(ph, ForcesCoefOp::Args{F, CL, part, nparts, nb}, nb + (nparts > 0 ? 1 : 0), dim3(32, 8),
                          "forces_coef");
}
void phase_coef_restrict(BatchPhase& ph, const Level& f, const Level& c) {
  const int nb = (int)nblk(c.nx + 1, c.ny + 1);
  phase_add<CoefRestrictOp>(ph, CoefRestrictOp::Args{f, c, nb}, nb, B328, "coef_restrict");
}
void phase_cfl(BatchPhase& ph, const Fields& F, float cfl, float umin, int nsub) {
  const int n = (F.nx + 1) * F.ny > F.nx * (F.ny + 1) ? (F.nx + 1) * F.ny : F.nx * (F.ny + 1);
  int b = (n + 1023) / 1024;
  if (b > 1024) b = 1024;
  phase_add<CflOp>(ph, CflOp::Args{F, cfl, umin, nsub, b}, b, dim3(256), "cfl");
}
void phase_resnorm(BatchPhase& ph, BatchPhase& fold, const Level& L0, Dyn* dyn, double* part, float tol) {
  const int nb = (int)resnorm_blocks(L0);
  phase_add<ResnormOp>(ph, ResnormOp::Args{L0, dyn, part, nb}, nb, dim3(256), "resnorm");
  phase_add<ResidFoldOp>(fold, ResidFoldOp::Args{dyn, part, nb, L0.nx * L0.ny, tol, 0, 0}, 1, dim3(256),
                         "resid_fold");
}
void phase_dc_defect(BatchPhase& ph, BatchPhase& fold, const Level& L0, __half* ph0, double* p64, float* b0,
                     Dyn* dyn, double* part, float tol, bool first) {
  const int nb = (int)resnorm_blocks(L0);
  phase_add<DcDefectOp>(ph, DcDefectOp::Args{L0, ph0, p64, b0, dyn, part, first ? 1 : 0, nb}, nb, dim3(256),
                        "dc_defect");
  phase_add<ResidFoldOp>(fold, ResidFoldOp::Args{dyn, part, nb, L0.nx * L0.ny, tol, first ? 1 : 0, 0}, 1, dim3(256),
                         "resid_fold");
}
void phase_kry_start(BatchPhase (&ph)[5], const Level& L0, __half* ph0, Dyn* dyn, const KryBufs& k, double* part,
                     float tol) {
  const int nb = (int)resnorm_blocks(L0);
  phase_add<KryStartOp>(ph[0], KryStartOp::Args{L0, dyn, k, part, nb}, nb, dim3(256), "kry_start");
  phase_add<KryStartCoefOp>(ph[1], KryStartCoefOp::Args{dyn, part, nb}, 1, dim3(256), "kry_start_coef");
  phase_add<KryScaleOp>(ph[2], KryScaleOp::Args{L0, dyn, k, nb}, nb, dim3(256), "kry_scale");
  phase_add<DcDefectOp>(ph[3], DcDefectOp::Args{L0, ph0, k.p64, k.b0, dyn, part, 0, nb}, nb, dim3(256), "dc_defect");
  phase_add<ResidFoldOp>(ph[4], ResidFoldOp::Args{dyn, part, nb, L0.nx * L0.ny, tol, 1, 1}, 1, dim3(256),
                         "resid_fold");
}
void phase_kry_step(BatchPhase (&ph)[5], const Level& L0, __half* ph0, Dyn* dyn, const KryBufs& k, double* part,
                    float tol, bool fused) {
  const int nb = (int)resnorm_blocks(L0);
  phase_add<KryDotsOp>(ph[0], KryDotsOp::Args{L0, ph0, dyn, k, part, nb, fused ? 1 : 0}, nb, dim3(256), "kry_dots");
  phase_add<KryCoefOp>(ph[1], KryCoefOp::Args{dyn, part, nb}, 1, dim3(256), "kry_coef");
  if (fused) {
    phase_add<KryUpdDefOp>(ph[2], KryUpdDefOp::Args{L0, ph0, dyn, k, part, nb}, nb, dim3(256), "kry_upd_def");
  } else {
    phase_add<KryUpdateOp>(ph[2], KryUpdateOp::Args{L0, ph0, dyn, k, nb}, nb, dim3(256), "kry_update");
    phase_add<DcDefectOp>(ph[3], DcDefectOp::Args{L0, ph0, k.p64, k.b0, dyn, part, 0, nb}, nb, dim3(256),
                          "dc_defect");
  }
  phase_add<ResidFoldOp>(ph[4], ResidFoldOp::Args{dyn, part, nb, L0.nx * L0.ny, tol, 0, 1}, 1, dim3(256),
                         "resid_fold");
}
void phase_kry_settle(BatchPhase& ph, const Level& L0, Dyn* dyn, const KryBufs& k) {
  const int nb = (int)resnorm_blocks(L0);
  phase_add<KrySettleOp>(ph, KrySettleOp::Args{L0, dyn, k, nb}, nb, dim3(256), "kry_settle");
}
void phase_dc_accum(BatchPhase& ph, const Level& L0, const __half* ph0, const Dyn* dyn, double* p64, const int* skip) {
  int b = (L0.nx * L0.ny + 1023) / 1024;
  if (b > 2048) b = 2048;
  phase_add<DcAccumOp>(ph, DcAccumOp::Args{L0, ph0, dyn, p64, skip, b}, b, dim3(256), "dc_accum");
}
void phase_coef_half(BatchPhase& ph, const Level& L, const LevelH& H) {
  const int nb = (int)nblk(L.nx + 1, L.ny + 1);
  phase_add<CoefHalfOp>(ph, CoefHalfOp::Args{L, H, nb}, nb, B328, "coef_half");
}
void phase_project(BatchPhase& ph, const Fields& F, const double* p64) {
  const int nb = (int)nblk(F.nx + 1, F.ny + 1);
  if (p64) phase_add<ProjectOp<true>>(ph, ProjectOp<true>::Args{F, p64, nb}, nb, B328, "project");
  else phase_add<ProjectOp<false>>(ph, ProjectOp<false>::Args{F, p64, nb}, nb, B328, "project");
}
void phase_geometry(BatchPhase& ph, const Fields& F) {
  const int nb = (int)nblk(F.nx + 1, F.ny + 1);
  phase_add<GeometryOp>(ph, GeometryOp::Args{F, nb}, nb, B328, "geometry");
}
void phase_gas_terms(BatchPhase& ph, const Fields& F) {
  const int gb = gas_terms_blocks(F);
  phase_add<GasTermsOp>(ph, GasTermsOp::Args{F, gb}, gb, dim3(256), "gas_terms");
}
void phase_radiation(BatchPhase& ph, const Fields& F, int ntargets) {
  if (ntargets > 0)
    phase_add<RadiationOp>(ph, RadiationOp::Args{F, ntargets}, ntargets, dim3(RadiationOp::kThreads), "radiation");
}
void phase_fill_u32(BatchPhase& ph, void* a, uint32_t v, size_t n) {
  int b = (int)((n + 255) / 256);
  if (b > 4096) b = 4096;
  if (n > 0) phase_add<FillOp>(ph, FillOp::Args{(uint32_t*)a, v, n, b}, b, dim3(256), "fill");
}

unsigned compat_blocks(const Fields& F) {
  const unsigned b = (unsigned)(((F.nx + 1) * (F.ny + 1) + 2047) / 2048);
  return b > 1024 ? 1024 : b;
}
void launch_compat(const Fields& F, double* part, hipStream_t s) {
  launch(k_compat, dim3(compat_blocks(F)), dim3(256), 0, s, F, part);
}
unsigned div_error_blocks(const Fields& F) { return nblk(F.nx, F.ny); }
void launch_div_error(const Fields& F, float* part, hipStream_t s) {
  launch(k_div_error, dim3(nblk(F.nx, F.ny)), B328, 0, s, F, part);
}

}  // namespace cf
