/*
 * cf_oracle.c — CPU oracle of campfire's per-timestep gas solver.
 *
 * TEST INFRASTRUCTURE ONLY: loaded by tests/, the smoke check and the
 * benchmark's CPU-baseline leg as the checker.  Never part of the product.
 *
 * Each function restates one reference kernel or host routine; the file:line
 * it follows is cited on it.  Arithmetic keeps the reference's operation
 * order (left-associative, no fused multiply-add) so that the REAL=float build
 * reproduces FP32 results up to the ULP differences of exp/log/pow/sin/cos.
 * Built twice: -DCFO_DOUBLE gives the FP64 oracle (symbols cfo64_*), without it
 * the FP32 oracle (cfo32_*).
 *
 * WGSL semantics used (WGSL spec): mix(a,b,t) = a*(1-t) + b*t; clamp(x,l,h) =
 * min(max(x,l),h); length(v) = sqrt(dot(v,v)); i32(f) truncates toward zero;
 * round() is round-half-to-even; abstract-float literals are evaluated in
 * double and converted (the K() macro below).
 *
 * Pinning (what checks each part): tests/test_oracle_*.py — closed forms of
 * the transport closures, sampler invariants, conservation of diffusion,
 * single-cell chemistry balances, the char-rate range of
 * tests/char_test.mjs:27-31, projection against a direct sparse solve, the
 * geometry partition of unity and the radiation black-enclosure limit.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "cf_oracle.h"
#include "half.h"

/* CFO_MUTANT = k > 0 builds the oracle with one deliberate mistake (mutant k,
 * marked MUTANT(k) where it is planted) so that the oracle's own checks can
 * be shown to catch it: tests/test_oracle_pins.py compiles each mutant and
 * asserts that the check aimed at it fails.  The oracle proper is CFO_MUTANT
 * 0, where every MUTANT(k) is false. */
#ifndef CFO_MUTANT
#define CFO_MUTANT 0
#endif
#define MUTANT(k) (CFO_MUTANT == (k))

#ifdef CFO_DOUBLE
typedef double REAL;
#define PFX(n) cfo64_##n
#define EXP exp
#define LOG log
#define POW pow
#define SQRT sqrt
#define FLOOR floor
#define CEIL ceil
#define FABS fabs
#define COS cos
#define SIN sin
#define ATAN2 atan2
#define FMIN fmin
#define FMAX fmax
#define RINT rint
#else
typedef float REAL;
#define PFX(n) cfo32_##n
#define EXP expf
#define LOG logf
#define POW powf
#define SQRT sqrtf
#define FLOOR floorf
#define CEIL ceilf
#define FABS fabsf
#define COS cosf
#define SIN sinf
#define ATAN2 atan2f
#define FMIN fminf
#define FMAX fmaxf
#define RINT rintf
#endif
#define K(x) ((REAL)(x))

#define SOLID_NONE 0xffffffffu
#define MAXLEV 16

typedef struct { REAL x, y; } v2;
typedef struct { REAL x, y, z, w; } v4;

/* PARAM_FIELDS (js/fluid.js:13-24): name and kind. */
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

typedef struct {
  int nx, ny, nC, nU, nV;
  double dx, dt, time;
  uint32_t frame;
  double pv[CF_NPARAMS];  /* the `params` dictionary (js/fluid.js:35-45) */
  cf_params P;            /* packed block (js/fluid.js:180-189) */
  float dsub;             /* DiffParams.dt (js/fluid.js:196-198) */
  int diffSub;
  int nlev, lnx[MAXLEV], lny[MAXLEV];
  float lh2[MAXLEV];      /* Level.h2 (js/fluid.js:82) */
  REAL *U, *Uh, *U2, *V, *Vh, *V2, *T, *Th, *T2;
  REAL *Y, *Yh, *Y2;      /* vec4 per cell, AoS */
  REAL *SQ;               /* vec2 per cell */
  REAL *sdf; uint32_t *sid;
  REAL *fx, *fy;
  float pieces[CF_MAX_PIECES * CF_PIECE_FLOATS];
  int32_t acc[CF_ACC_LEN], faceW[CF_MAX_PIECES * 4];
  REAL rad[CF_RAD_LEN * 4];
  REAL *p[MAXLEV], *rhs[MAXLEV], *ax[MAXLEV], *ay[MAXLEV];
  double* p64;            /* defect correction: level-0 pressure in FP64 */
  REAL* b0;               /* defect correction: level-0 right-hand side */
  float* d32;             /* Krylov: the stored direction (FP32) */
  double kbb, kdq;        /* Krylov: ||b||^2, (d, A d) of the stored direction (0: none) */
  double kbest;           /* Krylov: lowest ||r|| / ||b|| of this solve */
  int kfall;              /* Krylov: fallen back to defect correction for this solve */
  float sig, isig;        /* FP16 smoother: p stored as p * sig, a power of two */
  cf_schedule sch;
  int last_vcycles;
  double last_resid;
} S;

#define PP(f) ((REAL)s->P.f)

/* ---------------- shaders/common.wgsl ---------------- */
static inline int idxC(const S* s, int i, int j) { return j * s->nx + i; }       /* :29 */
static inline int idxU(const S* s, int i, int j) { return j * (s->nx + 1) + i; } /* :30 */
static inline int idxV(const S* s, int i, int j) { return j * s->nx + i; }       /* :31 */
static inline int imax(int a, int b) { return a > b ? a : b; }
static inline int imin(int a, int b) { return a < b ? a : b; }
static inline int iclamp(int x, int lo, int hi) { return imin(imax(x, lo), hi); }
static inline REAL mn(REAL a, REAL b) { return FMIN(a, b); }
static inline REAL mx(REAL a, REAL b) { return FMAX(a, b); }
static inline REAL mixr(REAL a, REAL b, REAL t) { return a * (K(1) - t) + b * t; }
static inline REAL clampr(REAL x, REAL lo, REAL hi) { return mn(mx(x, lo), hi); }
static inline int toint(REAL x) { /* i32(f): truncation; saturate far outside the grid */
  if (x > K(1073741824.0)) x = K(1073741824.0);
  if (x < K(-1073741824.0)) x = K(-1073741824.0);
  return (int)x;
}
static inline int32_t iround(REAL x) { return (int32_t)RINT(x); }
static inline v4 V4(REAL x, REAL y, REAL z, REAL w) { v4 r = {x, y, z, w}; return r; }
static inline v4 getY(const REAL* Y, int k) { return V4(Y[4 * k], Y[4 * k + 1], Y[4 * k + 2], Y[4 * k + 3]); }
static inline void putY(REAL* Y, int k, v4 v) { Y[4 * k] = v.x; Y[4 * k + 1] = v.y; Y[4 * k + 2] = v.z; Y[4 * k + 3] = v.w; }

static inline REAL rhoOf(const S* s, REAL T) { return PP(rhoAmb) * PP(Tamb) / T; } /* :33 */
/* pow(x, 1.5) written x*sqrt(x): the same power, within an ulp of the
 * correctly rounded value (WGSL's pow is looser), and what the kernels use */
static inline REAL pow15(REAL x) { return x * SQRT(x); }
static inline REAL muOf(const S* s, REAL T) {                                     /* :36-38 */
  return K(1.716e-5) * pow15(T / K(273.15)) * K(273.15 + 110.4) / (T + K(110.4)) * PP(viscMult);
}
static inline REAL rhoDOf(const S* s, REAL T) {                                   /* :40-42 */
  return K(1.716e-5) * pow15(T / K(273.15)) * K(273.15 + 110.4) / (T + K(110.4)) / PP(Pr) * PP(diffMult);
}
static inline v4 ambY(const S* s) { return V4(K(0), PP(YOamb), K(0), K(0)); }   /* :44 */

static inline const float* piece(const S* s, uint32_t n) { return s->pieces + n * CF_PIECE_FLOATS; }
/* shaders/common.wgsl:47-53 */
static REAL sdBox(v2 p, const float* pc) {
  REAL c = COS((REAL)pc[2]), sn = SIN((REAL)pc[2]);
  REAL dxp = p.x - (REAL)pc[0], dyp = p.y - (REAL)pc[1];
  REAL lx = c * dxp + sn * dyp, ly = -sn * dxp + c * dyp;
  REAL qx = FABS(lx) - (REAL)pc[3], qy = FABS(ly) - (REAL)pc[3];
  REAL mxq = mx(qx, K(0)), myq = mx(qy, K(0));
  return SQRT(mxq * mxq + myq * myq) + mn(mx(qx, qy), K(0));
}
/* shaders/common.wgsl:56-62 */
static uint32_t faceOf(v2 p, const float* pc) {
  REAL c = COS((REAL)pc[2]), sn = SIN((REAL)pc[2]);
  REAL dxp = p.x - (REAL)pc[0], dyp = p.y - (REAL)pc[1];
  REAL lx = c * dxp + sn * dyp, ly = -sn * dxp + c * dyp;
  if (FABS(lx) > FABS(ly)) return lx > K(0) ? 0u : 2u;
  return ly > K(0) ? 1u : 3u;
}

/* ---------------- generated samplers (js/gpu.js:16-65) ---------------- */
/* kind C: outside sides/top -> ambient, floor zero-gradient (js/gpu.js:31-35) */
static inline REAL fC(const S* s, const REAL* X, REAL amb, int i, int j) {
  if (i < 0 || i >= s->nx || j >= s->ny) return amb;
  return X[idxC(s, i, imax(j, 0))];
}
static REAL sC(const S* s, const REAL* X, REAL amb, v2 g) {                       /* :36-39 */
  REAL fxv = g.x - K(0.5), fyv = g.y - K(0.5);
  REAL bx = FLOOR(fxv), by = FLOOR(fyv);
  REAL tx = fxv - bx, ty = fyv - by;
  int i = toint(bx), j = toint(by);
  return mixr(mixr(fC(s, X, amb, i, j), fC(s, X, amb, i + 1, j), tx),
              mixr(fC(s, X, amb, i, j + 1), fC(s, X, amb, i + 1, j + 1), tx), ty);
}
static v2 mmC(const S* s, const REAL* X, REAL amb, v2 g) {                        /* :20-24 */
  int i = toint(FLOOR(g.x - K(0.5))), j = toint(FLOOR(g.y - K(0.5)));
  REAL a = fC(s, X, amb, i, j), c = fC(s, X, amb, i + 1, j), d = fC(s, X, amb, i, j + 1), e = fC(s, X, amb, i + 1, j + 1);
  v2 r = {mn(mn(a, c), mn(d, e)), mx(mx(a, c), mx(d, e))};
  return r;
}
static inline v4 fC4(const S* s, const REAL* X, v4 amb, int i, int j) {
  if (i < 0 || i >= s->nx || j >= s->ny) return amb;
  return getY(X, idxC(s, i, imax(j, 0)));
}
static inline v4 mix4(v4 a, v4 b, REAL t) {
  return V4(mixr(a.x, b.x, t), mixr(a.y, b.y, t), mixr(a.z, b.z, t), mixr(a.w, b.w, t));
}
static v4 sC4(const S* s, const REAL* X, v4 amb, v2 g) {
  REAL fxv = g.x - K(0.5), fyv = g.y - K(0.5);
  REAL bx = FLOOR(fxv), by = FLOOR(fyv);
  REAL tx = fxv - bx, ty = fyv - by;
  int i = toint(bx), j = toint(by);
  return mix4(mix4(fC4(s, X, amb, i, j), fC4(s, X, amb, i + 1, j), tx),
              mix4(fC4(s, X, amb, i, j + 1), fC4(s, X, amb, i + 1, j + 1), tx), ty);
}
static void mmC4(const S* s, const REAL* X, v4 amb, v2 g, v4* lo, v4* hi) {       /* :25-29 */
  int i = toint(FLOOR(g.x - K(0.5))), j = toint(FLOOR(g.y - K(0.5)));
  v4 a = fC4(s, X, amb, i, j), c = fC4(s, X, amb, i + 1, j), d = fC4(s, X, amb, i, j + 1), e = fC4(s, X, amb, i + 1, j + 1);
  *lo = V4(mn(mn(a.x, c.x), mn(d.x, e.x)), mn(mn(a.y, c.y), mn(d.y, e.y)), mn(mn(a.z, c.z), mn(d.z, e.z)), mn(mn(a.w, c.w), mn(d.w, e.w)));
  *hi = V4(mx(mx(a.x, c.x), mx(d.x, e.x)), mx(mx(a.y, c.y), mx(d.y, e.y)), mx(mx(a.z, c.z), mx(d.z, e.z)), mx(mx(a.w, c.w), mx(d.w, e.w)));
}
/* kind U: outside -> 0, below the floor the no-slip ghost -U(i,0) (js/gpu.js:43-46) */
static inline REAL fU(const S* s, const REAL* X, int i, int j) {
  if (i < 0 || i > s->nx || j >= s->ny) return K(0);
  if (j < 0) return -X[idxU(s, i, 0)];
  return X[idxU(s, i, j)];
}
/* kind V: outside -> 0, floor clamps (js/gpu.js:47-48) */
static inline REAL fV(const S* s, const REAL* X, int i, int j) {
  if (i < 0 || i >= s->nx || j > s->ny) return K(0);
  return X[idxV(s, i, imax(j, 0))];
}
/* s_U / s_V (js/gpu.js:54-57), offset (0,0.5) for U and (0.5,0) for V */
static REAL sU(const S* s, const REAL* X, v2 g) {
  REAL fxv = g.x - K(0), fyv = g.y - K(0.5);
  REAL bx = FLOOR(fxv), by = FLOOR(fyv);
  REAL tx = fxv - bx, ty = fyv - by;
  int i = toint(bx), j = toint(by);
  return mixr(mixr(fU(s, X, i, j), fU(s, X, i + 1, j), tx), mixr(fU(s, X, i, j + 1), fU(s, X, i + 1, j + 1), tx), ty);
}
static REAL sV(const S* s, const REAL* X, v2 g) {
  REAL fxv = g.x - K(0.5), fyv = g.y - K(0);
  REAL bx = FLOOR(fxv), by = FLOOR(fyv);
  REAL tx = fxv - bx, ty = fyv - by;
  int i = toint(bx), j = toint(by);
  return mixr(mixr(fV(s, X, i, j), fV(s, X, i + 1, j), tx), mixr(fV(s, X, i, j + 1), fV(s, X, i + 1, j + 1), tx), ty);
}
/* mm_U / mm_V (js/gpu.js:58-62) */
static v2 mmU(const S* s, const REAL* X, v2 g) {
  int i = toint(FLOOR(g.x - K(0))), j = toint(FLOOR(g.y - K(0.5)));
  REAL a = fU(s, X, i, j), c = fU(s, X, i + 1, j), d = fU(s, X, i, j + 1), e = fU(s, X, i + 1, j + 1);
  v2 r = {mn(mn(a, c), mn(d, e)), mx(mx(a, c), mx(d, e))};
  return r;
}
static v2 mmV(const S* s, const REAL* X, v2 g) {
  int i = toint(FLOOR(g.x - K(0.5))), j = toint(FLOOR(g.y - K(0)));
  REAL a = fV(s, X, i, j), c = fV(s, X, i + 1, j), d = fV(s, X, i, j + 1), e = fV(s, X, i + 1, j + 1);
  v2 r = {mn(mn(a, c), mn(d, e)), mx(mx(a, c), mx(d, e))};
  return r;
}

/* ---------------- shaders/advect.wgsl ---------------- */
/* vel (:28) and trace (:31-34): RK2 midpoint back-trace */
static v2 vel(const S* s, v2 g) {
  v2 r = {sU(s, s->U, g) / PP(dx), sV(s, s->V, g) / PP(dx)};
  return r;
}
static v2 trace(const S* s, v2 g, REAL dt) {
  v2 v = vel(s, g);
  REAL hdt = K(0.5) * dt;
  v2 gm = {g.x - hdt * v.x, g.y - hdt * v.y};
  v2 w = vel(s, gm);
  v2 r = {g.x - dt * w.x, g.y - dt * w.y};
  return r;
}

/* velForward (:36-47) */
static void k_velForward(S* s) {
  const int nx = s->nx, ny = s->ny;
  const REAL dt = PP(dt);
#pragma omp parallel for schedule(static)
  for (int j = 0; j <= ny; j++)
    for (int i = 0; i <= nx; i++) {
      if (i <= nx && j < ny) {
        v2 g = {(REAL)i, (REAL)j + K(0.5)};
        s->Uh[idxU(s, i, j)] = sU(s, s->U, trace(s, g, dt));
      }
      if (i < nx && j <= ny) {
        v2 g = {(REAL)i + K(0.5), (REAL)j};
        s->Vh[idxV(s, i, j)] = sV(s, s->V, trace(s, g, dt));
      }
    }
}

/* velCorrect (:49-66): MacCormack, clamped to the departure stencil */
static void k_velCorrect(S* s) {
  const int nx = s->nx, ny = s->ny;
  const REAL dt = PP(dt);
#pragma omp parallel for schedule(static)
  for (int j = 0; j <= ny; j++)
    for (int i = 0; i <= nx; i++) {
      if (i <= nx && j < ny) {
        int k = idxU(s, i, j);
        v2 g = {(REAL)i, (REAL)j + K(0.5)};
        REAL back = sU(s, s->Uh, trace(s, g, -dt));
        v2 m = mmU(s, s->U, trace(s, g, dt));
        s->U2[k] = clampr(s->Uh[k] + K(0.5) * (s->U[k] - back), m.x, m.y);
      }
      if (i < nx && j <= ny) {
        int k = idxV(s, i, j);
        v2 g = {(REAL)i + K(0.5), (REAL)j};
        REAL back = sV(s, s->Vh, trace(s, g, -dt));
        v2 m = mmV(s, s->V, trace(s, g, dt));
        s->V2[k] = clampr(s->Vh[k] + K(0.5) * (s->V[k] - back), m.x, m.y);
      }
    }
}

/* scalForward (:68-76) */
static void k_scalForward(S* s) {
  const int nx = s->nx, ny = s->ny;
  const REAL dt = PP(dt), Tamb = PP(Tamb);
  const v4 aY = ambY(s);
#pragma omp parallel for schedule(static)
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) {
      int k = idxC(s, i, j);
      v2 g = {(REAL)i + K(0.5), (REAL)j + K(0.5)};
      v2 d = trace(s, g, dt);
      s->Th[k] = sC(s, s->T, Tamb, d);
      putY(s->Yh, k, sC4(s, s->Y, aY, d));
    }
}

/* scalCorrect (:78-103): MacCormack candidate kept only inside the departure
 * bounds and without creating co-existing fuel/oxidiser beyond first order. */
static void k_scalCorrect(S* s) {
  const int nx = s->nx, ny = s->ny;
  const REAL dt = PP(dt), Tamb = PP(Tamb), sO2 = PP(sO2);
  const v4 aY = ambY(s);
#pragma omp parallel for schedule(static)
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) {
      int k = idxC(s, i, j);
      v2 g = {(REAL)i + K(0.5), (REAL)j + K(0.5)};
      v2 d = trace(s, g, dt);
      v2 f = trace(s, g, -dt);
      v2 mT = mmC(s, s->T, Tamb, d);
      v4 lo, hi;
      mmC4(s, s->Y, aY, d, &lo, &hi);
      REAL Thk = s->Th[k];
      v4 Yhk = getY(s->Yh, k), Yk = getY(s->Y, k);
      REAL Tc = Thk + K(0.5) * (s->T[k] - sC(s, s->Th, Tamb, f));
      v4 sb = sC4(s, s->Yh, aY, f);
      v4 Yc = V4(Yhk.x + K(0.5) * (Yk.x - sb.x), Yhk.y + K(0.5) * (Yk.y - sb.y),
                 Yhk.z + K(0.5) * (Yk.z - sb.z), Yhk.w + K(0.5) * (Yk.w - sb.w));
      int inT = Tc >= mT.x && Tc <= mT.y;
      int inY = Yc.x >= lo.x && Yc.y >= lo.y && Yc.z >= lo.z && Yc.w >= lo.w &&
                Yc.x <= hi.x && Yc.y <= hi.y && Yc.z <= hi.z && Yc.w <= hi.w;
      REAL mixC = mn(Yc.x, Yc.y / sO2);
      REAL mixL = mn(Yhk.x, Yhk.y / sO2);
      if (inT && inY && mixC <= mixL + K(1e-5)) {
        s->T2[k] = Tc; putY(s->Y2, k, Yc);
      } else {
        s->T2[k] = Thk; putY(s->Y2, k, Yhk);
      }
    }
}

/* ---------------- shaders/diffuse.wgsl:18-50 ---------------- */
static void k_diffuse(S* s, const REAL* Tin, const REAL* Yin, REAL* Tout, REAL* Yout) {
  const int nx = s->nx, ny = s->ny;
  const REAL dts = (REAL)s->dsub, dx = PP(dx);
#pragma omp parallel for schedule(static)
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) {
      int k = idxC(s, i, j);
      REAL T0 = Tin[k];
      v4 Y0 = getY(Yin, k);
      REAL sT = K(0);
      v4 sY = V4(0, 0, 0, 0);
      int nb[4] = {-1, -1, -1, -1};
      REAL fr[4] = {0, 0, 0, 0};
      if (i > 0) { fr[0] = s->fx[idxU(s, i, j)]; nb[0] = idxC(s, i - 1, j); }          /* west */
      if (i < nx - 1) { fr[1] = s->fx[idxU(s, i + 1, j)]; nb[1] = idxC(s, i + 1, j); } /* east */
      if (j > 0) { fr[2] = s->fy[idxV(s, i, j)]; nb[2] = idxC(s, i, j - 1); }          /* south */
      if (j < ny - 1) { fr[3] = s->fy[idxV(s, i, j + 1)]; nb[3] = idxC(s, i, j + 1); } /* north */
      for (int q = 0; q < 4; q++) {
        if (nb[q] < 0 || !(fr[q] > K(0))) continue;
        int n = nb[q];
        REAL a = fr[q] * rhoDOf(s, K(0.5) * (T0 + Tin[n]));
        v4 Yn = getY(Yin, n);
        sT += a * (Tin[n] - T0);
        sY.x += a * (Yn.x - Y0.x); sY.y += a * (Yn.y - Y0.y);
        sY.z += a * (Yn.z - Y0.z); sY.w += a * (Yn.w - Y0.w);
      }
      REAL c = dts / (rhoOf(s, T0) * dx * dx);
      Tout[k] = T0 + c * sT;
      putY(Yout, k, V4(mx(Y0.x + c * sY.x * K(1), K(0)), mx(Y0.y + c * sY.y * K(1), K(0)),
                       mx(Y0.z + c * sY.z * K(1), K(0)), mx(Y0.w + c * sY.w * K(0), K(0))));
    }
}

/* ---------------- shaders/sources.wgsl:34-173 ---------------- */
static inline void acc_add(S* s, int idx, int32_t v) {
#pragma omp atomic
  s->acc[idx] += v;
}
/* one cell of the sources kernel; returns qrel, qrad, tpeak for the reduction */
static void sources_cell(S* s, int i, int j, REAL* qrel_o, REAL* qrad_o, REAL* tpeak_o) {
  const REAL dx = PP(dx), dt = PP(dt), cp = PP(cp), Tamb = PP(Tamb), sO2 = PP(sO2), dHc = PP(dHc);
  REAL qrel = K(0), qrad = K(0), tpeak = K(0);
  int k = idxC(s, i, j);
  REAL Tc = s->T[k];
  v4 Yc = getY(s->Y, k);
  REAL sdf = s->sdf[k];
  uint32_t sid = s->sid[k];
  if (sid >= CF_MAX_PIECES * 4) sid = SOLID_NONE; /* ids are piece*4+face or none */
  REAL massSrc = K(0);
  REAL cellA = dx * dx;
  if (sdf < K(0) && sid != SOLID_NONE) {                                         /* :50-55 */
    const float* pc = piece(s, sid >> 2);
    s->T[k] = (REAL)pc[8 + (sid & 3u)];
    s->SQ[2 * k] = K(0); s->SQ[2 * k + 1] = K(0);
  } else {
    if (sid != SOLID_NONE && sdf < K(2.0) * dx) {                                /* :57-94 */
      uint32_t pi = sid >> 2, fi = sid & 3u;
      const float* pc = piece(s, pi);
      REAL Ts = (REAL)pc[8 + fi];
      REAL band = (REAL)s->faceW[sid] / K(1.0e8);
      REAL w = mx(K(0), K(1) - FABS(sdf - dx) / dx) / dx * (K(2.0) * (REAL)pc[3]) / mx(band, K(1e-6));
      REAL rho = rhoOf(s, Tc);
      REAL kg = rhoDOf(s, K(0.5) * (Tc + Ts)) * cp;
      REAL dist = mx(sdf, K(0.5) * dx);
      REAL rate = kg * w / (dist * rho * cp);
      REAL dTc = (Ts - Tc) * (K(1) - EXP(-rate * dt));
      Tc += dTc;
      REAL toWood = -rho * cp * dTc * cellA;
      acc_add(s, (int)((pi * 4u + fi) * 2u), iround(toWood * K(1e6)));
      REAL mf = (REAL)pc[12 + fi] * w;                                            /* pyrolysis */
      if (mf > K(0)) {
        REAL madd = mf * dt;
        REAL r = rhoOf(s, Tc);
        Yc = V4((r * Yc.x + madd) / (r + madd), (r * Yc.y + K(0)) / (r + madd),
                (r * Yc.z + K(0)) / (r + madd), (r * Yc.w + K(0)) / (r + madd));
        Tc = (r * Tc + madd * Ts) / (r + madd);
        massSrc += mf;
      }
      REAL cov = (REAL)pc[16 + fi];                                               /* char oxidation */
      if (cov > K(0) && Yc.y > K(0)) {
        REAL r = rhoOf(s, Tc);
        REAL kc = PP(charA) * EXP(-PP(charTa) / Ts) * cov;
        REAL mO2 = mn(kc * r * Yc.y * w * dt, K(0.9) * r * Yc.y);
        REAL mC = mO2 / PP(charO2);
        Yc = V4((r * Yc.x + K(0)) / (r + mC), (r * Yc.y + -mO2) / (r + mC),
                (r * Yc.z + (mC + mO2)) / (r + mC), (r * Yc.w + K(0)) / (r + mC));
        Tc = (r * Tc + mC * Ts) / (r + mC);
        massSrc += mC / dt;
        acc_add(s, (int)((pi * 4u + fi) * 2u + 1u), iround(mC * cellA * K(1e12)));
      }
    }
    REAL x = ((REAL)i + K(0.5)) * dx, y = ((REAL)j + K(0.5)) * dx;              /* :96-107 */
    if (PP(heaterT) > K(0) && x > PP(heaterX0) && x < PP(heaterX1) && y < PP(heaterY1)) Tc = PP(heaterT);
    if (PP(burnerFlux) > K(0) && j == 0 && x > PP(burnerX0) && x < PP(burnerX1)) {
      REAL mf = PP(burnerFlux) / dx;
      REAL madd = mf * dt;
      REAL r = rhoOf(s, Tc);
      Yc = V4((r * Yc.x + madd) / (r + madd), (r * Yc.y + K(0)) / (r + madd),
              (r * Yc.z + K(0)) / (r + madd), (r * Yc.w + K(0)) / (r + madd));
      Tc = (r * Tc + madd * PP(burnerT)) / (r + madd);
      massSrc += mf;
    }
    REAL rho0 = rhoOf(s, Tc);                                                     /* :109-120 */
    const int NS = 8;
    REAL h = dt / (REAL)NS;
    REAL burned = K(0);
    for (int q = 0; q < NS; q++) {
      REAL wr = PP(reactB) * rho0 * Yc.x * Yc.y * EXP(-PP(reactTa) / Tc);
      REAL dF = mn(wr * h, mn(Yc.x, Yc.y / sO2));
      Yc.x += -dF; Yc.y += -sO2 * dF; Yc.z += (K(1) + sO2) * dF; Yc.w += K(0);
      Tc += dF * dHc / cp;
      burned += dF;
    }
    REAL dSf = mn(PP(sootFormA) * EXP(-PP(sootFormTa) / Tc) * Yc.x * dt, Yc.x);   /* :122-127 */
    Yc.x += -dSf; Yc.y += K(0); Yc.z += K(0); Yc.w += dSf;
    REAL dSo = mn(PP(sootOxA) * Yc.w * Yc.y * EXP(-PP(sootOxTa) / Tc) * dt, mn(Yc.w, Yc.y / sO2));
    Yc.x += K(0); Yc.y += -sO2 * dSo; Yc.z += (K(1) + sO2) * dSo; Yc.w += -dSo;
    Tc += dSo * dHc / cp;
    burned += dSo;
    qrel = rho0 * burned * dHc / dt;                                              /* :128 */
    REAL r1 = rhoOf(s, Tc);                                                       /* :130-136 */
    REAL fv = r1 * Yc.w / PP(sootDensity);
    REAL kap = PP(kappaSootC) * fv * Tc + PP(kappaGas) * Yc.z;
    REAL T3 = Tc * Tc * Tc;
    qrad = K(4.0) * kap * PP(sigma) * (T3 * Tc - POW(Tamb, K(4.0)));
    REAL dTr = qrad * dt / (r1 * cp) / (K(1) + K(16.0) * kap * PP(sigma) * T3 * dt / (r1 * cp));
    Tc -= dTr;
    REAL ucx = K(0.5) * (s->U[idxU(s, i, j)] + s->U[idxU(s, i + 1, j)]);         /* :140-148 */
    REAL ucy = K(0.5) * (s->V[idxV(s, i, j)] + s->V[idxV(s, i, j + 1)]);
    REAL me = K(2.0) * PP(exchAlpha) * SQRT(ucx * ucx + ucy * ucy) * PP(rhoAmb) / PP(depth);
    if (me > K(0)) {
      REAL madd = me * dt;
      REAL r = rhoOf(s, Tc);
      v4 a = ambY(s);
      Yc = V4((r * Yc.x + madd * a.x) / (r + madd), (r * Yc.y + madd * a.y) / (r + madd),
              (r * Yc.z + madd * a.z) / (r + madd), (r * Yc.w + madd * a.w) / (r + madd));
      Tc = (r * Tc + madd * Tamb) / (r + madd);
      massSrc += me;
    }
    Tc = mx(Tc, K(0.5) * Tamb);                                                   /* :150-156 */
    Yc = V4(mx(Yc.x, K(0)), mx(Yc.y, K(0)), mx(Yc.z, K(0)), mx(Yc.w, K(0)));
    REAL S_ = massSrc / rhoOf(s, Tc) + LOG(Tc / s->T2[k]) / dt;
    s->T[k] = Tc;
    putY(s->Y, k, Yc);
    tpeak = Tc;
    s->SQ[2 * k] = S_; s->SQ[2 * k + 1] = qrel;
  }
  *qrel_o = qrel; *qrad_o = qrad; *tpeak_o = tpeak;
}

static void k_sources(S* s) {
  const int nx = s->nx, ny = s->ny;
  const int gx = (nx + 7) / 8, gy = (ny + 7) / 8;
  const REAL dx = PP(dx), dt = PP(dt);
#pragma omp parallel for schedule(static)
  for (int w = 0; w < gx * gy; w++) {
    int wx = w % gx, wy = w / gx;
    REAL rq[64], rr[64], rt[64];
    for (int li = 0; li < 64; li++) {
      int i = wx * 8 + (li & 7), j = wy * 8 + (li >> 3);
      REAL qrel = K(0), qrad = K(0), tpeak = K(0);
      if (i < nx && j < ny) sources_cell(s, i, j, &qrel, &qrad, &tpeak);
      rq[li] = qrel * dx * dx * dt;                                               /* :160-162 */
      rr[li] = qrad * dx * dx * dt;
      rt[li] = tpeak;
    }
    for (int st = 32; st > 0; st >>= 1)                                           /* :164-167 */
      for (int li = 0; li < st; li++) {
        rq[li] += rq[li + st]; rr[li] += rr[li + st]; rt[li] = mx(rt[li], rt[li + st]);
      }
    acc_add(s, CF_ACC_GLOBAL, iround(rq[0] * K(1e3)));                            /* :168-172 */
    acc_add(s, CF_ACC_GLOBAL + 1, iround(rr[0] * K(1e3)));
    int32_t tp = (int32_t)rt[0];
#pragma omp critical(cfo_accmax)
    if (tp > s->acc[CF_ACC_GLOBAL + 2]) s->acc[CF_ACC_GLOBAL + 2] = tp;
  }
}

/* ---------------- shaders/forces.wgsl:34-101 ---------------- */
static inline REAL Tcl(const S* s, int i, int j) {                                /* :17-19 */
  return s->T[idxC(s, iclamp(i, 0, s->nx - 1), iclamp(j, 0, s->ny - 1))];
}
static inline REAL u2f(const S* s, int i, int j) { int k = idxU(s, i, j); return s->fx[k] > K(0) ? s->U2[k] : K(0); }
static inline REAL v2f(const S* s, int i, int j) { int k = idxV(s, i, j); return s->fy[k] > K(0) ? s->V2[k] : K(0); }
static inline REAL exchange(const S* s, REAL spd, REAL Tf) {                     /* :30-32 */
  return EXP(K(-2.0) * PP(exchAlpha) * spd * (Tf / PP(Tamb)) / PP(depth) * PP(dt));
}
static void k_forces(S* s) {
  const int nx = s->nx, ny = s->ny;
  const REAL dx = PP(dx), dt = PP(dt);
  const REAL idx2 = (MUTANT(7) ? K(2) : K(1)) / (dx * dx);
#pragma omp parallel for schedule(static)
  for (int j = 0; j <= ny; j++)
    for (int i = 0; i <= nx; i++) {
      if (i <= nx && j < ny) {                                                    /* U faces */
        int k = idxU(s, i, j);
        if (!(s->fx[k] > K(0))) {
          s->U[k] = K(0);
        } else if (i == 0) {
          s->U[k] = mn(s->U2[idxU(s, 1, j)], K(0));
        } else if (i == nx) {
          s->U[k] = mx(s->U2[idxU(s, nx - 1, j)], K(0));
        } else {
          REAL u = s->U2[k];
          REAL TL = Tcl(s, i - 1, j), TR = Tcl(s, i, j);
          REAL rhoF = rhoOf(s, K(0.5) * (TL + TR));
          REAL muE = muOf(s, TR), muW = muOf(s, TL);
          REAL uN = u, muN = muOf(s, K(0.5) * (TL + TR));
          if (j < ny - 1) { uN = u2f(s, i, j + 1); muN = muOf(s, K(0.25) * (TL + TR + Tcl(s, i - 1, j + 1) + Tcl(s, i, j + 1))); }
          if (MUTANT(8) && j < ny - 1) muN = muOf(s, K(0.5) * (TL + TR));  /* corner mu from the face's own cells */
          REAL uS = MUTANT(9) ? u : -u, muS = muOf(s, K(0.5) * (TL + TR));
          if (j > 0) { uS = u2f(s, i, j - 1); muS = muOf(s, K(0.25) * (TL + TR + Tcl(s, i - 1, j - 1) + Tcl(s, i, j - 1))); }
          REAL visc = (muE * (u2f(s, i + 1, j) - u) + muW * (u2f(s, i - 1, j) - u) + muN * (uN - u) + muS * (uS - u)) * idx2 / rhoF;
          REAL vAvg = K(0.25) * (s->V2[idxV(s, i - 1, j)] + s->V2[idxV(s, i, j)] + s->V2[idxV(s, i - 1, j + 1)] + s->V2[idxV(s, i, j + 1)]);
          s->U[k] = (u + dt * visc) * exchange(s, SQRT(u * u + vAvg * vAvg), K(0.5) * (TL + TR));
        }
      }
      if (i < nx && j <= ny) {                                                    /* V faces */
        int k = idxV(s, i, j);
        if (j == 0 || !(s->fy[k] > K(0))) {
          s->V[k] = K(0);
        } else if (j == ny) {
          s->V[k] = mx(s->V2[idxV(s, i, ny - 1)], K(0));
        } else {
          REAL v = s->V2[k];
          REAL TB = Tcl(s, i, j - 1), TT = Tcl(s, i, j);
          REAL Tf = K(0.5) * (TB + TT);
          REAL rhoF = rhoOf(s, Tf);
          REAL muN = muOf(s, TT), muS = muOf(s, TB);
          REAL vE = v, muE = muOf(s, Tf);
          if (i < nx - 1) { vE = v2f(s, i + 1, j); muE = muOf(s, K(0.25) * (TB + TT + Tcl(s, i + 1, j - 1) + Tcl(s, i + 1, j))); }
          REAL vW = v, muW = muOf(s, Tf);
          if (i > 0) { vW = v2f(s, i - 1, j); muW = muOf(s, K(0.25) * (TB + TT + Tcl(s, i - 1, j - 1) + Tcl(s, i - 1, j))); }
          REAL visc = (muN * (v2f(s, i, j + 1) - v) + muS * (v2f(s, i, j - 1) - v) + muE * (vE - v) + muW * (vW - v)) * idx2 / rhoF;
          REAL buoy = PP(grav) * (Tf / PP(Tamb) - K(1));
          REAL uAvg = K(0.25) * (s->U2[idxU(s, i, j - 1)] + s->U2[idxU(s, i + 1, j - 1)] + s->U2[idxU(s, i, j)] + s->U2[idxU(s, i + 1, j)]);
          s->V[k] = (v + dt * (visc + buoy)) * exchange(s, SQRT(uAvg * uAvg + v * v), Tf);
        }
      }
    }
}

/* ---------------- shaders/pressure.wgsl ---------------- */
static inline REAL invRho(const S* s, REAL T) { return T / (PP(rhoAmb) * PP(Tamb)); } /* :27 */

static void k_coef(S* s) {                                                         /* :30-43 */
  const int nx = s->nx, ny = s->ny;
  REAL *ax = s->ax[0], *ay = s->ay[0];
#pragma omp parallel for schedule(static)
  for (int j = 0; j <= ny; j++)
    for (int i = 0; i <= nx; i++) {
      if (i <= nx && j < ny) {
        int k = idxU(s, i, j);
        REAL Tf = K(0.5) * (s->T[idxC(s, iclamp(i - 1, 0, nx - 1), j)] + s->T[idxC(s, imin(i, nx - 1), j)]);
        ax[k] = s->fx[k] * invRho(s, Tf);
      }
      if (i < nx && j <= ny) {
        int k = idxV(s, i, j);
        REAL Tf = K(0.5) * (s->T[idxC(s, i, iclamp(j - 1, 0, ny - 1))] + s->T[idxC(s, i, imin(j, ny - 1))]);
        ay[k] = s->fy[k] * invRho(s, Tf);
      }
    }
}

static void k_rhs0(S* s) {                                                         /* :46-54 */
  const int nx = s->nx, ny = s->ny;
  const REAL dx = PP(dx), dt = PP(dt);
#pragma omp parallel for schedule(static)
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) {
      REAL fe = s->fx[idxU(s, i + 1, j)], fw = s->fx[idxU(s, i, j)];
      REAL fn = s->fy[idxV(s, i, j + 1)], fs = s->fy[idxV(s, i, j)];
      REAL div = (fe * s->U[idxU(s, i + 1, j)] - fw * s->U[idxU(s, i, j)] + fn * s->V[idxV(s, i, j + 1)] - fs * s->V[idxV(s, i, j)]) / dx;
      REAL open = fe + fw + fn + fs;
      s->rhs[0][idxC(s, i, j)] = open > K(0) ? (div - s->SQ[2 * idxC(s, i, j)]) / dt : K(0);
    }
}

static void k_coefRestrict(S* s, int l) {                                          /* :57-67 */
  const int nx = s->lnx[l], ny = s->lny[l], nxc = nx / 2, nyc = ny / 2;
  const REAL *ax = s->ax[l], *ay = s->ay[l];
  REAL *axc = s->ax[l + 1], *ayc = s->ay[l + 1];
  for (int J = 0; J <= nyc; J++)
    for (int I = 0; I <= nxc; I++) {
      if (I <= nxc && J < nyc) axc[J * (nxc + 1) + I] = K(0.5) * (ax[(2 * J) * (nx + 1) + 2 * I] + ax[(2 * J + 1) * (nx + 1) + 2 * I]);
      if (I < nxc && J <= nyc) ayc[J * nxc + I] = K(0.5) * (ay[(2 * J) * nx + 2 * I] + ay[(2 * J) * nx + 2 * I + 1]);
    }
}

/* stencil (:77-85): sum a*p_nb and the diagonal; Dirichlet p = 0 at open sides/top */
static inline void stencil(const S* s, int l, const REAL* p, int i, int j, REAL* off_o, REAL* diag_o) {
  const int nx = s->lnx[l], ny = s->lny[l];
  const REAL *ax = s->ax[l], *ay = s->ay[l];
  REAL w = ax[j * (nx + 1) + i], e = ax[j * (nx + 1) + i + 1], so = ay[j * nx + i], n = ay[(j + 1) * nx + i];
  REAL off = K(0), diag = K(0);
  if (i > 0) { off += w * p[j * nx + i - 1]; diag += w; } else { diag += K(2.0) * w; }
  if (i < nx - 1) { off += e * p[j * nx + i + 1]; diag += e; } else { diag += K(2.0) * e; }
  if (j > 0) { off += so * p[(j - 1) * nx + i]; diag += so; }
  if (j < ny - 1) { off += n * p[(j + 1) * nx + i]; diag += n; } else { diag += K(2.0) * n; }
  *off_o = off; *diag_o = diag;
}

static void k_relax(S* s, int l, int color) {                                      /* :87-94 */
  const int nx = s->lnx[l], ny = s->lny[l];
  REAL* p = s->p[l];
  const REAL* rhs = s->rhs[l];
  const REAL h2 = (REAL)s->lh2[l];
#pragma omp parallel for schedule(static) if (nx * ny > 16384)
  for (int j = 0; j < ny; j++)
    for (int i = ((j + color) & 1); i < nx; i += 2) {
      REAL off, diag;
      stencil(s, l, p, i, j, &off, &diag);
      int k = j * nx + i;
      p[k] = diag > K(0) ? (off - rhs[k] * h2) / diag : K(0);
    }
}

static inline REAL residual(const S* s, int l, int i, int j) {                     /* :96-101 */
  REAL off, diag;
  stencil(s, l, s->p[l], i, j, &off, &diag);
  int k = j * s->lnx[l] + i;
  if (diag <= K(0)) return K(0);
  return s->rhs[l][k] - (off - diag * s->p[l][k]) / (REAL)s->lh2[l];
}

static void k_restrictRes(S* s, int l) {                                           /* :104-113 */
  const int nxc = s->lnx[l] / 2, nyc = s->lny[l] / 2;
#pragma omp parallel for schedule(static) if (nxc * nyc > 4096)
  for (int J = 0; J < nyc; J++)
    for (int I = 0; I < nxc; I++) {
      REAL r = residual(s, l, 2 * I, 2 * J) + residual(s, l, 2 * I + 1, 2 * J) +
               residual(s, l, 2 * I, 2 * J + 1) + residual(s, l, 2 * I + 1, 2 * J + 1);
      s->rhs[l + 1][J * nxc + I] = (MUTANT(5) ? K(0.5) : K(0.25)) * r;
      if (!MUTANT(6)) s->p[l + 1][J * nxc + I] = K(0);
    }
}

static void k_prolong(S* s, int l) {                                               /* :116-129 */
  const int nx = s->lnx[l], ny = s->lny[l], nxc = nx / 2, nyc = ny / 2;
  REAL* p = s->p[l];
  const REAL* pc = s->p[l + 1];
#pragma omp parallel for schedule(static) if (nx * ny > 16384)
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) {
      int I = i / 2, J = j / 2;
      int I2 = iclamp(I + ((i & 1) == 1 ? 1 : -1), 0, nxc - 1);
      int J2 = iclamp(J + ((j & 1) == 1 ? 1 : -1), 0, nyc - 1);
      if (MUTANT(4)) {  /* mirror instead of clamp at the edges */
        I2 = I + ((i & 1) == 1 ? 1 : -1); I2 = I2 < 0 ? 1 : (I2 > nxc - 1 ? nxc - 2 : I2);
        J2 = J + ((j & 1) == 1 ? 1 : -1); J2 = J2 < 0 ? 1 : (J2 > nyc - 1 ? nyc - 2 : J2);
      }
      const REAL w3 = MUTANT(3) ? K(0.0625) : K(0.1875), w1 = MUTANT(3) ? K(0.1875) : K(0.0625);
      REAL c = K(0.5625) * pc[J * nxc + I] + w3 * (pc[J * nxc + I2] + pc[J2 * nxc + I]) + w1 * pc[J2 * nxc + I2];
      REAL off, diag;
      stencil(s, l, p, i, j, &off, &diag);
      p[j * nx + i] += diag > K(0) ? c : K(0);
    }
}

/* coarsest (:131-156): `iters` red-black sweeps from zero */
static void k_coarsest(S* s, int l, int iters) {
  const int nx = s->lnx[l], ny = s->lny[l];
  REAL* p = s->p[l];
  for (int k = 0; k < nx * ny; k++) p[k] = K(0);
  for (int it = 0; it < iters; it++)
    for (int c = 0; c < 2; c++) k_relax(s, l, c);
}

/* the build's schedule option coarse_iters = 0: a direct FP64 solve of the
 * coarsest level, A p = -rhs h2 with A the relaxation's operator (diag on
 * the diagonal, minus the coefficients to existing neighbours, a closed cell
 * an identity row), by dense Cholesky */
static void coarse_direct(S* s, int l) {
  const int nx = s->lnx[l], ny = s->lny[l], N = nx * ny;
  const REAL *ax = s->ax[l], *ay = s->ay[l];
  double* A = calloc((size_t)N * N, sizeof(double));
  double* y = calloc((size_t)N, sizeof(double));
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) {
      const int k = j * nx + i;
      REAL off, diag;
      stencil(s, l, s->p[l], i, j, &off, &diag);
      if (diag > K(0)) {
        A[(size_t)k * N + k] = (double)diag;
        if (i > 0) A[(size_t)k * N + k - 1] = -(double)ax[j * (nx + 1) + i];
        if (i < nx - 1) A[(size_t)k * N + k + 1] = -(double)ax[j * (nx + 1) + i + 1];
        if (j > 0) A[(size_t)k * N + k - nx] = -(double)ay[j * nx + i];
        if (j < ny - 1) A[(size_t)k * N + k + nx] = -(double)ay[(j + 1) * nx + i];
        y[k] = -(double)s->rhs[l][k] * (double)s->lh2[l];
      } else {
        A[(size_t)k * N + k] = 1.0;
      }
    }
  for (int c = 0; c < N; c++) {            /* A = L L^T in the lower triangle */
    double d = A[(size_t)c * N + c];
    for (int k = 0; k < c; k++) d -= A[(size_t)c * N + k] * A[(size_t)c * N + k];
    d = sqrt(d);
    A[(size_t)c * N + c] = d;
    for (int r = c + 1; r < N; r++) {
      double v = A[(size_t)r * N + c];
      for (int k = 0; k < c; k++) v -= A[(size_t)r * N + k] * A[(size_t)c * N + k];
      A[(size_t)r * N + c] = v / d;
    }
  }
  for (int r = 0; r < N; r++) {            /* L z = y */
    double v = y[r];
    for (int k = 0; k < r; k++) v -= A[(size_t)r * N + k] * y[k];
    y[r] = v / A[(size_t)r * N + r];
  }
  for (int r = N - 1; r >= 0; r--) {       /* L^T x = z */
    double v = y[r];
    for (int k = r + 1; k < N; k++) v -= A[(size_t)k * N + r] * y[k];
    y[r] = v / A[(size_t)r * N + r];
  }
  for (int k = 0; k < N; k++) s->p[l][k] = (REAL)y[k];
  free(A);
  free(y);
}

/* level-0 relative residual ||r||2 / ||rhs||2 (convergence test of the
 * converged-multigrid schedule; the reference runs a fixed single V-cycle) */
static double resid_norm(const S* s) {
  double rr = 0, bb = 0;
  for (int j = 0; j < s->ny; j++)
    for (int i = 0; i < s->nx; i++) {
      double r = (double)residual(s, 0, i, j), b = (double)s->rhs[0][j * s->nx + i];
      rr += r * r; bb += b * b;
    }
  return bb > 0 ? sqrt(rr / bb) : sqrt(rr);
}

/* FP16 smoother: the power of two bringing the rms of a residual with
 * ||r||^2 = rr into [0.5, 1) scales the stored correction */
static void set_sig(S* s, double rr) {
  int ex = 0;
  const double rms = sqrt(rr / ((double)s->nx * (double)s->ny));
  if (rms > 0) {
    frexp(rms, &ex);
    ex = ex < -60 ? -60 : (ex > 60 ? 60 : ex);
  }
  s->sig = (float)ldexp(1.0, -ex);
  s->isig = (float)ldexp(1.0, ex);
}

/* defect correction (the build's schedule option refine = 1): r = b - A p64
 * in FP64 into rhs[0] with p[0] = 0 as the correction's starting guess;
 * returns ||r|| / ||b|| */
static double dc_defect(S* s, int first) {
  const int nx = s->nx, ny = s->ny;
  const REAL* ax = s->ax[0];
  const REAL* ay = s->ay[0];
  const double* p = s->p64;
  const double h2 = (double)s->lh2[0];
  double rr = 0, bb = 0;
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) {
      const int k = j * nx + i;
      if (first) s->b0[k] = s->rhs[0][k];
      const double b = (double)s->b0[k];
      const double w = ax[j * (nx + 1) + i], e = ax[j * (nx + 1) + i + 1];
      const double so = ay[j * nx + i], n = ay[(j + 1) * nx + i];
      double off = 0, diag = 0;
      if (i > 0) { off += w * p[k - 1]; diag += w; } else diag += 2.0 * w;
      if (i < nx - 1) { off += e * p[k + 1]; diag += e; } else diag += 2.0 * e;
      if (j > 0) { off += so * p[k - nx]; diag += so; }
      if (j < ny - 1) { off += n * p[k + nx]; diag += n; } else diag += 2.0 * n;
      const double r = diag <= 0 ? 0.0 : b - (off - diag * p[k]) / h2;
      s->rhs[0][k] = (REAL)r;
      s->p[0][k] = 0;
      rr += r * r; bb += b * b;
    }
  set_sig(s, rr);
  return bb > 0 ? sqrt(rr / bb) : sqrt(rr);
}

/* Sum of v[0..n) in the order of the GPU's two-pass reduction (the
 * build's convention for FP64 norms and dot products, not the reference's):
 * vg workgroups of 256 threads, thread t of workgroup b adding the cells
 * b*256 + t + m*vg*256 in turn; each 64-lane wave folding its lanes as a tree
 * (lane i += lane i+o for o = 32, 16, ..., 1); each workgroup's partial
 * ((w0 + w1) + w2) + w3 of its four waves; then the partials folded the same
 * way by one workgroup (thread t adding partials t, t+256, ...).  vg is
 * ceil(n / 1024) capped at 1024. */
static double gpu_sum(const double* v, int n) {
  int vg = (n + 1023) / 1024;
  if (vg > 1024) vg = 1024;
  if (vg < 1) vg = 1;
  double* part = malloc(sizeof(double) * (size_t)vg);
  double lane[256];
  for (int pass = 0; pass < 2; pass++) {
    const int nb = pass == 0 ? vg : 1;
    for (int b = 0; b < nb; b++) {
      for (int t = 0; t < 256; t++) {
        double a = 0;
        if (pass == 0)
          for (long k = (long)b * 256 + t; k < n; k += (long)vg * 256) a += v[k];
        else
          for (int q = t; q < vg; q += 256) a += part[q];
        lane[t] = a;
      }
      for (int w = 0; w < 4; w++) {
        double* L = lane + 64 * w;
        for (int o = 32; o > 0; o >>= 1)
          for (int i = 0; i < o; i++) L[i] = L[i] + L[i + o];
      }
      const double r = ((lane[0] + lane[64]) + lane[128]) + lane[192];
      if (pass == 0) part[b] = r;
      else { free(part); return r; }
    }
  }
  free(part);
  return 0;
}

/* B z at level 0 for the Krylov iteration: (sum a z_nb - diag z) / h2 in FP64
 * with the relaxation's boundary terms, 0 at a closed cell */
static int apply64_open(const S* s, int i, int j) {
  const int nx = s->nx, ny = s->ny;
  const REAL *ax = s->ax[0], *ay = s->ay[0];
  double diag = 0;
  diag += i > 0 ? (double)ax[j * (nx + 1) + i] : 2.0 * (double)ax[j * (nx + 1) + i];
  diag += i < nx - 1 ? (double)ax[j * (nx + 1) + i + 1] : 2.0 * (double)ax[j * (nx + 1) + i + 1];
  if (j > 0) diag += (double)ay[j * nx + i];
  diag += j < ny - 1 ? (double)ay[(j + 1) * nx + i] : 2.0 * (double)ay[(j + 1) * nx + i];
  return diag > 0;
}
static double apply64(const S* s, const double* z, int i, int j) {
  const int nx = s->nx, ny = s->ny;
  const REAL *ax = s->ax[0], *ay = s->ay[0];
  const double w = ax[j * (nx + 1) + i], e = ax[j * (nx + 1) + i + 1];
  const double so = ay[j * nx + i], n = ay[(j + 1) * nx + i];
  const int k = j * nx + i;
  double off = 0, diag = 0;
  if (i > 0) { off += w * z[k - 1]; diag += w; } else diag += 2.0 * w;
  if (i < nx - 1) { off += e * z[k + 1]; diag += e; } else diag += 2.0 * e;
  if (j > 0) { off += so * z[k - nx]; diag += so; }
  if (j < ny - 1) { off += n * z[k + nx]; diag += n; } else diag += 2.0 * n;
  if (diag <= 0) return 0;
  return (off - diag * z[k]) / (double)s->lh2[0];
}

/* Krylov residual: r = b - B p64 from the saved b (the defect of
 * dc_defect(s, 0)), into rhs[0] with e = 0 as the next V-cycle's start;
 * ||r|| / ||b|| with the sums in the GPU's order */
static double kry_defect(S* s) {
  const int nx = s->nx, ny = s->ny, n = nx * ny;
  double* v = malloc(sizeof(double) * (size_t)n);
  double* w = malloc(sizeof(double) * (size_t)n);
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) {
      const int k = j * nx + i;
      const double b = (double)s->b0[k];
      const double r = apply64_open(s, i, j) ? b - apply64(s, s->p64, i, j) : 0.0;
      s->rhs[0][k] = (REAL)r;
      s->p[0][k] = 0;
      v[k] = r * r;
      w[k] = b * b;
    }
  const double rr = gpu_sum(v, n), bb = gpu_sum(w, n);
  set_sig(s, rr);
  free(v);
  free(w);
  return bb > 0 ? sqrt(rr / bb) : sqrt(rr);
}

/* Krylov start (schedule option krylov = 1), from the previous step's
 * pressure p rounded to FP32 (the direction's storage): a line search along
 * p from zero, x = a p with
 * a = (p, b) / (p, B p), which keeps the previous pressure where it fits
 * this step's right-hand side and scales it (or drops it, a = 0) where it
 * does not; p stays as the stored direction (r = b - a B p is orthogonal to
 * it), so the first FCG iteration is conjugated against it.  With p = 0 the
 * start is x = 0, r = b and no stored direction.  Returns ||r|| / ||b||. */
static double kry_start(S* s) {
  const int nx = s->nx, ny = s->ny, n = nx * ny;
  double* v = malloc(sizeof(double) * (size_t)n);
  double* bp = malloc(sizeof(double) * (size_t)n);
  double* d = malloc(sizeof(double) * (size_t)n);
  for (int k = 0; k < n; k++) {
    s->b0[k] = s->rhs[0][k];
    s->d32[k] = (float)s->p64[k];
    d[k] = (double)s->d32[k];
  }
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) bp[j * nx + i] = apply64(s, d, i, j);
  for (int k = 0; k < n; k++) v[k] = d[k] * (double)s->b0[k];
  const double pb = gpu_sum(v, n);
  for (int k = 0; k < n; k++) v[k] = d[k] * bp[k];
  const double pq = gpu_sum(v, n);
  for (int k = 0; k < n; k++) v[k] = (double)s->b0[k] * (double)s->b0[k];
  s->kbb = gpu_sum(v, n);
  const double a = pq != 0 ? pb / pq : 0.0;
  s->kdq = pq;
  for (int k = 0; k < n; k++) s->p64[k] = a * d[k];
  free(v);
  free(bp);
  free(d);
  const double rel = kry_defect(s);
  s->kbest = rel;
  s->kfall = 0;
  return rel;
}

/* One flexible-CG iteration after the V-cycle has left z = M r in p[0]
 * (Notay's FCG with one stored direction), with the residual recomputed
 * from the pressure each time rather than carried, so the direction can be
 * stored in FP32 (rounding it only perturbs the conjugacy):
 *   beta = (B z, d) / (d, B d)   (= (z, B d): B is symmetric),
 *   d <- z - beta d,  alpha = (r, z) / (d, B d),  p64 += alpha d,
 *   r = b - B p64,
 * where (d, B d) of the new direction expands to (z, B z) - beta (B z, d_old),
 * so one pass over z gives the three dot products; r enters them as the FP32
 * right-hand side the V-cycle solved for.  Returns ||r|| / ||b||.
 * A residual above twice the lowest of the solve so far ends the Krylov
 * iteration for this solve: the later cycles are plain defect correction
 * (alpha = 1, beta = 0: p64 += z) from where it stands.  A singular level-0
 * system (open cells closed off from every Dirichlet side) with a
 * right-hand side outside its range makes CG diverge once the compatible
 * part is solved, while defect correction stalls at the incompatible part. */
static double kry_step(S* s) {
  const int nx = s->nx, ny = s->ny, n = nx * ny;
  double alpha = 1.0, beta = 0.0;
  double* z = malloc(sizeof(double) * (size_t)n);
  for (int k = 0; k < n; k++) z[k] = (double)s->p[0][k];
  if (!s->kfall) {
    double* w = malloc(sizeof(double) * (size_t)n);
    double* v = malloc(sizeof(double) * (size_t)n);
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) w[j * nx + i] = apply64(s, z, i, j);
    for (int k = 0; k < n; k++) v[k] = w[k] * (double)s->d32[k];
    const double wd = gpu_sum(v, n);
    for (int k = 0; k < n; k++) v[k] = (double)s->rhs[0][k] * z[k];
    const double rz = gpu_sum(v, n);
    for (int k = 0; k < n; k++) v[k] = z[k] * w[k];
    const double zw = gpu_sum(v, n);
    beta = s->kdq != 0 ? wd / s->kdq : 0.0;
    const double dq = zw - beta * wd;
    alpha = dq != 0 ? rz / dq : 0.0;
    s->kdq = dq;
    free(w);
    free(v);
  }
  for (int k = 0; k < n; k++) {
    const float d = (float)(z[k] - beta * (double)s->d32[k]);
    s->d32[k] = d;
    s->p64[k] += alpha * (double)d;
  }
  free(z);
  const double rel = kry_defect(s);
  if (!s->kfall) {
    if (rel > 2.0 * s->kbest) s->kfall = 1;
    if (rel < s->kbest) s->kbest = rel;
  }
  return rel;
}

#ifndef CFO_DOUBLE
/* the first of the levels that the GPU's single-workgroup bottom solver
 * takes: the coarsest (<= 256 cells) and, when they fit, the one or two above
 * it (<= 1024 and <= 4096 cells).  The levels above it are smoothed by tiled
 * legs; with the FP16 smoother those keep p and the coefficients in FP16. */
static int mg_bottom_first(const S* s) {
  const int cap[3] = {256, 1024, 4096};
  int b = s->nlev - 1;
  for (int d = 1; d < 3 && b - 1 >= 0; d++) {
    if (s->lnx[b - 1] * s->lny[b - 1] > cap[d]) break;
    b--;
  }
  return b;
}

/* p of level l through its FP16 storage: p * sig rounded to FP16, read back
 * as h * isig */
static void p_through_half(S* s, int l) {
  const int n = s->lnx[l] * s->lny[l];
  REAL* p = s->p[l];
  for (int k = 0; k < n; k++) p[k] = cfo_h2f1(cfo_f2h1(p[k] * s->sig)) * s->isig;
}

static REAL* half_copy(const REAL* a, int n) {
  REAL* c = malloc(sizeof(REAL) * (size_t)n);
  for (int k = 0; k < n; k++) c[k] = cfo_h2f1(cfo_f2h1(a[k]));
  return c;
}
#endif

static void vcycle_ref(S* s, int nu);

/* the build's schedule option smoother_fp16 (with defect correction): the
 * V-cycle of js/fluid.js:294-312 where the levels above the bottom solver's
 * keep p and the face coefficients in FP16 between legs.  Arithmetic stays
 * FP32: each down leg sweeps, forms the residual and restricts it from the
 * FP32 values, then stores p; each up leg reads p and the coarse p, adds the
 * correction, sweeps and stores p. */
static void vcycle(S* s, int nu) {
#ifndef CFO_DOUBLE
  const int b = s->sch.smoother_fp16 ? mg_bottom_first(s) : 0;
  if (b > 0) {
    REAL *ax[MAXLEV], *ay[MAXLEV];
    for (int l = 0; l < b; l++) {
      const int nx = s->lnx[l], ny = s->lny[l];
      ax[l] = s->ax[l]; ay[l] = s->ay[l];
      s->ax[l] = half_copy(ax[l], (nx + 1) * ny);
      s->ay[l] = half_copy(ay[l], nx * (ny + 1));
    }
    const int L = s->nlev;
    for (int l = 0; l < L - 1; l++) {
      for (int q = 0; q < nu; q++) { k_relax(s, l, 0); k_relax(s, l, 1); }
      k_restrictRes(s, l);
      if (l < b) p_through_half(s, l);
    }
    if (s->sch.coarse_iters > 0) k_coarsest(s, L - 1, s->sch.coarse_iters);
    else coarse_direct(s, L - 1);
    for (int l = L - 2; l >= 0; l--) {
      k_prolong(s, l);
      for (int q = 0; q < nu; q++) { k_relax(s, l, 1); k_relax(s, l, 0); }
      if (l < b) p_through_half(s, l);
    }
    for (int l = 0; l < b; l++) {
      free(s->ax[l]); free(s->ay[l]);
      s->ax[l] = ax[l]; s->ay[l] = ay[l];
    }
    return;
  }
#endif
  vcycle_ref(s, nu);
}

/* vcycle (js/fluid.js:294-312) */
static void vcycle_ref(S* s, int nu) {
  const int L = s->nlev;
  const int dn = MUTANT(2) ? 1 : 0, up = MUTANT(1) ? 0 : 1;  /* first colour of a sweep, down and up */
  for (int l = 0; l < L - 1; l++) {
    for (int q = 0; q < nu; q++) { k_relax(s, l, dn); k_relax(s, l, 1 - dn); }
    k_restrictRes(s, l);
  }
  if (s->sch.coarse_iters > 0) k_coarsest(s, L - 1, s->sch.coarse_iters);
  else coarse_direct(s, L - 1);
  for (int l = L - 2; l >= 0; l--) {
    k_prolong(s, l);
    for (int q = 0; q < nu; q++) { k_relax(s, l, up); k_relax(s, l, 1 - up); }
  }
}

static void k_project(S* s) {                                                      /* :159-180 */
  const int nx = s->nx, ny = s->ny;
  const REAL dx = PP(dx), dt = PP(dt);
  const REAL* p = s->p[0];
  /* defect correction: differences of the FP64 pressure, rounded after the subtraction */
  const double* q = (s->sch.tol > 0 && s->sch.refine) ? s->p64 : NULL;
#pragma omp parallel for schedule(static)
  for (int j = 0; j <= ny; j++)
    for (int i = 0; i <= nx; i++) {
      if (i <= nx && j < ny) {
        int k = idxU(s, i, j);
        if (s->fx[k] > K(0)) {
          REAL dp;
          if (q) {
            double pl = i > 0 ? q[idxC(s, imax(i - 1, 0), j)] : -q[idxC(s, 0, j)];
            double pr = i < nx ? q[idxC(s, imin(i, nx - 1), j)] : -q[idxC(s, nx - 1, j)];
            dp = (REAL)(pr - pl);
          } else {
            REAL pl = i > 0 ? p[idxC(s, imax(i - 1, 0), j)] : -p[idxC(s, 0, j)];
            REAL pr = i < nx ? p[idxC(s, imin(i, nx - 1), j)] : -p[idxC(s, nx - 1, j)];
            dp = pr - pl;
          }
          REAL Tf = K(0.5) * (s->T[idxC(s, iclamp(i - 1, 0, nx - 1), j)] + s->T[idxC(s, imin(i, nx - 1), j)]);
          s->U[k] -= dt * invRho(s, Tf) * dp / dx;
        } else s->U[k] = K(0);
      }
      if (i < nx && j <= ny) {
        int k = idxV(s, i, j);
        if (j > 0 && s->fy[k] > K(0)) {
          REAL dp;
          if (q) {
            double pb = q[idxC(s, i, j - 1)];
            double pt = j < ny ? q[idxC(s, i, imin(j, ny - 1))] : -q[idxC(s, i, ny - 1)];
            dp = (REAL)(pt - pb);
          } else {
            REAL pb = p[idxC(s, i, j - 1)];
            REAL pt = j < ny ? p[idxC(s, i, imin(j, ny - 1))] : -p[idxC(s, i, ny - 1)];
            dp = pt - pb;
          }
          REAL Tf = K(0.5) * (s->T[idxC(s, i, j - 1)] + s->T[idxC(s, i, imin(j, ny - 1))]);
          s->V[k] -= dt * invRho(s, Tf) * dp / dx;
        } else s->V[k] = K(0);
      }
    }
  if (q) for (int k = 0; k < s->nC; k++) s->p[0][k] = (REAL)q[k];
}

/* ---------------- shaders/geometry.wgsl ---------------- */
static v2 nearest(const S* s, v2 p, uint32_t* bid_o) {                             /* :20-29 */
  REAL best = K(1e9);
  uint32_t bid = SOLID_NONE;
  for (int n = 0; n < s->P.nPieces; n++) {
    const float* pc = piece(s, (uint32_t)n);
    if (!(pc[7] > 0.0f)) continue;
    REAL d = sdBox(p, pc);
    if (d < best) { best = d; bid = (uint32_t)n * 4u + faceOf(p, pc); }
  }
  *bid_o = bid;
  v2 r = {best, 0};
  return r;
}
static void k_geometry(S* s) {                                                     /* :31-52 */
  const int nx = s->nx, ny = s->ny;
  const REAL dx = PP(dx);
#pragma omp parallel for schedule(static)
  for (int j = 0; j <= ny; j++)
    for (int i = 0; i <= nx; i++) {
      uint32_t id;
      if (i < nx && j < ny) {
        v2 c = {((REAL)i + K(0.5)) * dx, ((REAL)j + K(0.5)) * dx};
        v2 g = nearest(s, c, &id);
        int k = idxC(s, i, j);
        s->sdf[k] = g.x; s->sid[k] = id;
        if (id != SOLID_NONE && g.x >= K(0) && g.x < K(2.0) * dx) {
          REAL w = mx(K(0), K(1) - FABS(g.x - dx) / dx) / dx;
          int32_t add = iround(w * dx * dx * K(1.0e8));
#pragma omp atomic
          s->faceW[id] += add;
        }
      }
      if (i <= nx && j < ny) {
        v2 c = {(REAL)i * dx, ((REAL)j + K(0.5)) * dx};
        REAL d = nearest(s, c, &id).x;
        s->fx[idxU(s, i, j)] = clampr((MUTANT(10) ? K(0) : K(0.5)) + d / dx, K(0), K(1));
      }
      if (i < nx && j <= ny) {
        v2 c = {((REAL)i + K(0.5)) * dx, (REAL)j * dx};
        REAL d = nearest(s, c, &id).x;
        s->fy[idxV(s, i, j)] = j == 0 ? K(0) : clampr(K(0.5) + d / dx, K(0), K(1));
      }
    }
}

/* ---------------- shaders/radiation.wgsl ---------------- */
#define RPI K(3.14159265)
static inline REAL fractr(REAL x) { return x - FLOOR(x); }
static REAL planck580(REAL T) {                                                    /* :28-33 */
  REAL l = K(580e-9);
  REAL x = K(1.438777e-2) / (l * T);
  if (x > K(80.0)) return K(0);
  return K(1.191042e-16) / (l * l * l * l * l) / (EXP(x) - K(1)) * K(1e-9);
}
static REAL hitBox(v2 o, v2 d, const float* pc, REAL* face_o) {                    /* :36-51 */
  REAL c = COS((REAL)pc[2]), sn = SIN((REAL)pc[2]);
  REAL qx = o.x - (REAL)pc[0], qy = o.y - (REAL)pc[1];
  REAL lox = c * qx + sn * qy, loy = -sn * qx + c * qy;
  REAL ldx = c * d.x + sn * d.y, ldy = -sn * d.x + c * d.y;
  REAL h = (REAL)pc[3];
  REAL invx = K(1) / (FABS(ldx) < K(1e-9) ? K(1e-9) : ldx);
  REAL invy = K(1) / (FABS(ldy) < K(1e-9) ? K(1e-9) : ldy);
  REAL t1x = (-h - lox) * invx, t1y = (-h - loy) * invy;
  REAL t2x = (h - lox) * invx, t2y = (h - loy) * invy;
  REAL tminx = mn(t1x, t2x), tminy = mn(t1y, t2y), tmaxx = mx(t1x, t2x), tmaxy = mx(t1y, t2y);
  REAL tn = mx(tminx, tminy), tf = mn(tmaxx, tmaxy);
  if (tf < K(0) || tn > tf || tn < K(1e-5)) { *face_o = K(0); return K(-1); }
  REAL face;
  if (tminx > tminy) face = ldx < K(0) ? K(0) : K(2); else face = ldy < K(0) ? K(1) : K(3);
  if (MUTANT(12)) face = (REAL)(((int)face + 2) % 4);  /* the opposite face */
  *face_o = face;
  return tn;
}
static void gasAt(const S* s, v2 p, REAL* g0, REAL* g1) {                          /* :53-62 */
  int i = toint(p.x / PP(dx)), j = toint(p.y / PP(dx));
  if (i < 0 || j < 0 || i >= s->nx || j >= s->ny) { *g0 = K(0); *g1 = K(0); return; }
  int k = idxC(s, i, j);
  REAL Tg = s->T[k];
  v4 Yg = getY(s->Y, k);
  REAL fv = rhoOf(s, Tg) * Yg.w / PP(sootDensity);
  REAL kap = PP(kappaSootC) * fv * Tg + PP(kappaGas) * Yg.z;
  REAL T2 = Tg * Tg;
  *g0 = kap * PP(sigma) * (T2 * T2 - POW(PP(Tamb), K(4.0))) / RPI;
  *g1 = K(6.3) * fv / K(580e-9) * planck580(Tg);
}
static void k_radiation(S* s) {                                                    /* :64-158 */
  const int nFaces = s->P.nPieces * 4;
  const int ntgt = nFaces + CF_N_FLOOR;
  const REAL dx = PP(dx);
#pragma omp parallel for schedule(dynamic)
  for (int tgt = 0; tgt < ntgt; tgt++) {
    REAL sq[64], sv[64];
    int valid = 0;
    for (int li = 0; li < 64; li++) {
      v2 origin = {0, 0}, nrm = {0, 1};
      valid = 0;
      if (tgt < nFaces) {
        const float* pc = piece(s, (uint32_t)(tgt / 4));
        int f = tgt % 4;
        if (pc[7] > 0.0f) {
          REAL c = COS((REAL)pc[2]), sn = SIN((REAL)pc[2]);
          REAL lnx = 1, lny = 0;
          if (f == 1) { lnx = 0; lny = 1; } else if (f == 2) { lnx = -1; lny = 0; } else if (f == 3) { lnx = 0; lny = -1; }
          nrm.x = c * lnx - sn * lny; nrm.y = sn * lnx + c * lny;
          v2 tang = {-nrm.y, nrm.x};
          REAL u = (fractr((REAL)li * K(0.618034) + (REAL)s->P.frame * K(0.1234)) - K(0.5)) * K(1.6) * (REAL)pc[3];
          REAL hh = (REAL)pc[3] + K(1e-4);
          origin.x = (REAL)pc[0] + nrm.x * hh + tang.x * u;
          origin.y = (REAL)pc[1] + nrm.y * hh + tang.y * u;
          valid = 1;
        }
      } else {
        int fi = tgt - nFaces;
        if (fi < CF_N_FLOOR) {
          origin.x = ((REAL)fi + K(0.5)) / (REAL)CF_N_FLOOR * (REAL)s->nx * dx;
          origin.y = K(1e-4);
          nrm.x = 0; nrm.y = 1;
          valid = 1;
        }
      }
      REAL dq = K(0), dv = K(0);
      if (valid) {
        REAL jit = fractr((REAL)s->P.frame * K(0.7548777) + (REAL)tgt * K(0.5698403));
        REAL a = (((REAL)li + jit) / K(64) - K(0.5)) * RPI;
        REAL cosn = COS(a);
        REAL base = ATAN2(nrm.y, nrm.x);
        v2 d = {COS(base + a), SIN(base + a)};
        v2 o = origin;
        REAL J = K(0), Jv = K(0);
        REAL Twall = PP(Tamb);
        int bounced = 0;
        REAL ds = dx;
        for (int step = 0; step < 4096; step++) {
          REAL tHit = K(1e9), hitT = K(0);
          for (int n = 0; n < s->P.nPieces; n++) {
            const float* pc = piece(s, (uint32_t)n);
            if (!(pc[7] > 0.0f)) continue;
            REAL face;
            REAL hx = hitBox(o, d, pc, &face);
            if (hx > K(0) && hx < tHit) { tHit = hx; hitT = (REAL)pc[8 + (int)face]; }
          }
          REAL tFloor = K(1e9);
          if (d.y < K(-1e-6)) tFloor = -o.y / d.y;
          REAL tEnd = mn(tHit, tFloor);
          REAL W = (REAL)s->nx * dx, H = (REAL)s->ny * dx;
          REAL tx = K(1e9), ty = K(1e9);
          if (d.x > K(1e-6)) tx = (W - o.x) / d.x; else if (d.x < K(-1e-6)) tx = -o.x / d.x;
          if (d.y > K(1e-6)) ty = (H - o.y) / d.y;
          REAL tOut = mn(mn(tx, ty), tEnd);
          int ns = toint(CEIL(tOut / ds));
          for (int m = 0; m < ns; m++) {
            REAL t = ((REAL)m + K(0.5)) * tOut / (REAL)ns;
            v2 pp = {o.x + d.x * t, o.y + d.y * t};
            REAL g0, g1;
            gasAt(s, pp, &g0, &g1);
            J += g0 * tOut / (REAL)ns;
            Jv += g1 * tOut / (REAL)ns;
          }
          if (tHit < tFloor && tHit < K(1e8)) { Twall = hitT; break; }
          if (tFloor < K(1e8) && tFloor <= tHit && !bounced && !MUTANT(11)) {
            o.x = o.x + d.x * tFloor + K(0); o.y = o.y + d.y * tFloor + K(1e-5);
            d.y = -d.y;
            bounced = 1;
            continue;
          }
          Twall = PP(Tamb);
          break;
        }
        REAL w = RPI / K(64) * cosn;
        REAL T2 = Twall * Twall;
        dq = w * (K(0.5) * PP(sigma) * T2 * T2 + K(2.0) * J);
        dv = w * (K(2.0) * Jv);
      }
      sq[li] = dq; sv[li] = dv;
    }
    for (int st = 32; st > 0; st >>= 1)
      for (int li = 0; li < st; li++) { sq[li] += sq[li + st]; sv[li] += sv[li + st]; }
    if (valid) {  /* `valid` is uniform over the workgroup */
      REAL oq = s->rad[4 * tgt], ov = s->rad[4 * tgt + 1];
      REAL a = oq <= K(0) ? K(1.0) : K(0.35);
      s->rad[4 * tgt] = mixr(oq, sq[0], a);
      s->rad[4 * tgt + 1] = mixr(ov, sv[0], a);
      s->rad[4 * tgt + 2] = K(0); s->rad[4 * tgt + 3] = K(0);
    }
  }
}

/* ---------------- host side: js/fluid.js ---------------- */
static int param_index(const char* key) {
  for (int q = 0; q < CF_NPARAMS; q++) if (strcmp(PNAME[q], key) == 0) return q;
  return -1;
}

void* PFX(create)(int nx, int ny, double dx, double dt, int* err) {              /* js/fluid.js:27-91 */
  *err = 0;
  if (nx < 2 || ny < 2 || nx > 16384 || ny > 16384) { *err = CF_EINVAL; return NULL; }
  S* s = (S*)calloc(1, sizeof(S));
  s->nx = nx; s->ny = ny; s->dx = dx; s->dt = dt;
  s->nC = nx * ny; s->nU = (nx + 1) * ny; s->nV = nx * (ny + 1);
  /* levels: halve while even and > 256 cells (js/fluid.js:71-91, 6-level cap lifted) */
  int lx = nx, ly = ny; double h = dx;
  for (int l = 0; l < MAXLEV; l++) {
    s->lnx[l] = lx; s->lny[l] = ly; s->lh2[l] = (float)(h * h);
    s->nlev = l + 1;
    if (lx % 2 || ly % 2 || lx * ly <= 256) break;
    lx /= 2; ly /= 2; h *= 2;
  }
  if (s->lnx[s->nlev - 1] * s->lny[s->nlev - 1] > 256) { free(s); *err = CF_EBADSIZE; return NULL; }
  size_t nC = s->nC, nU = s->nU, nV = s->nV;
  s->U = calloc(nU, sizeof(REAL)); s->Uh = calloc(nU, sizeof(REAL)); s->U2 = calloc(nU, sizeof(REAL));
  s->V = calloc(nV, sizeof(REAL)); s->Vh = calloc(nV, sizeof(REAL)); s->V2 = calloc(nV, sizeof(REAL));
  s->T = calloc(nC, sizeof(REAL)); s->Th = calloc(nC, sizeof(REAL)); s->T2 = calloc(nC, sizeof(REAL));
  s->p64 = calloc(nC, sizeof(double)); s->b0 = calloc(nC, sizeof(REAL));
  s->d32 = calloc(nC, sizeof(float));
  s->Y = calloc(4 * nC, sizeof(REAL)); s->Yh = calloc(4 * nC, sizeof(REAL)); s->Y2 = calloc(4 * nC, sizeof(REAL));
  s->SQ = calloc(2 * nC, sizeof(REAL));
  s->sdf = calloc(nC, sizeof(REAL)); s->sid = calloc(nC, sizeof(uint32_t));
  s->fx = calloc(nU, sizeof(REAL)); s->fy = calloc(nV, sizeof(REAL));
  for (int l = 0; l < s->nlev; l++) {
    size_t a = (size_t)s->lnx[l] * s->lny[l];
    s->p[l] = calloc(a, sizeof(REAL)); s->rhs[l] = calloc(a, sizeof(REAL));
    s->ax[l] = calloc((size_t)(s->lnx[l] + 1) * s->lny[l], sizeof(REAL));
    s->ay[l] = calloc((size_t)s->lnx[l] * (s->lny[l] + 1), sizeof(REAL));
  }
  /* default params (js/fluid.js:35-45, constants js/config.js:12-47) */
  double Tamb = 293.15;
  double rhoAmb = 101325.0 * 0.0289 / (8.314462 * Tamb);
  double def[CF_NPARAMS] = {
    nx, ny, dx, dt, Tamb, rhoAmb, 0.232, 9.81,
    1300, 0.7, 1, 1, 2.0e10, 15000, 15.0e6, 15.0e6 / 13.1e6,
    6.0e6, 25000, 5.4e8, 19800,
    1862, 0.8, 0.05, 0.8,
    6.6, 5800, 0, 0,
    0, 0, 0, 0, 0, 0, 0, 600,
    1800, 5.670374e-8, 2.667, 0};
  memcpy(s->pv, def, sizeof(def));
  s->sch.nu = 2; s->sch.vcycles = 1; s->sch.tol = 0; s->sch.coarse_iters = 120; s->sch.smoother_fp16 = 0;
  s->sig = s->isig = 1.0f;
  s->sch.cfl = 0; s->sch.umin = 1.0f; s->sch.refine = 0; s->sch.krylov = 0;
  s->diffSub = 1;
  return s;
}

void PFX(destroy)(void* h) {
  S* s = (S*)h;
  if (!s) return;
  free(s->U); free(s->Uh); free(s->U2); free(s->V); free(s->Vh); free(s->V2);
  free(s->T); free(s->Th); free(s->T2); free(s->Y); free(s->Yh); free(s->Y2);
  free(s->p64); free(s->b0); free(s->d32);
  free(s->SQ); free(s->sdf); free(s->sid); free(s->fx); free(s->fy);
  for (int l = 0; l < s->nlev; l++) { free(s->p[l]); free(s->rhs[l]); free(s->ax[l]); free(s->ay[l]); }
  free(s);
}

/* the gas closures of shaders/common.wgsl:33-42, for closed-form checks */
double PFX(closure)(void* h, const char* which, double x) {
  const S* s = (const S*)h;
  if (!strcmp(which, "rhoOf")) return (double)rhoOf(s, (REAL)x);
  if (!strcmp(which, "muOf")) return (double)muOf(s, (REAL)x);
  if (!strcmp(which, "rhoDOf")) return (double)rhoDOf(s, (REAL)x);
  if (!strcmp(which, "invRho")) return (double)invRho(s, (REAL)x);
  if (!strcmp(which, "planck580")) return (double)planck580((REAL)x);
  return NAN;
}

int PFX(set_param)(void* h, const char* key, double v) {                          /* js/fluid.js:178 */
  S* s = (S*)h;
  if (!strcmp(key, "nPieces") && !(v >= 0 && v <= CF_MAX_PIECES)) return CF_EINVAL;
  if (strcmp(key, "time") == 0) { s->time = v; return 0; }
  if (strcmp(key, "frame") == 0) { s->frame = (uint32_t)(int64_t)v; return 0; }
  int q = param_index(key);
  if (q < 0) return CF_EINVAL;
  s->pv[q] = v;
  return 0;
}

int PFX(write_params)(void* h) {                                                  /* js/fluid.js:180-199 */
  S* s = (S*)h;
  s->pv[0] = s->nx; s->pv[1] = s->ny; s->pv[2] = s->dx; s->pv[3] = s->dt;
  s->pv[39] = s->time; s->pv[27] = (double)s->frame;
  uint32_t* raw = (uint32_t*)&s->P;
  for (int q = 0; q < CF_NPARAMS; q++) {
    if (PKIND[q] == 'f') { float f = (float)s->pv[q]; memcpy(&raw[q], &f, 4); }
    else if (PKIND[q] == 'i') { int32_t iv = (int32_t)s->pv[q]; memcpy(&raw[q], &iv, 4); }
    else { uint32_t uv = (uint32_t)(int64_t)s->pv[q]; raw[q] = uv; }
  }
  double Tmax = 2200, mu = 1.716e-5 * pow(Tmax / 273.15, 1.5) * 383.55 / (Tmax + 110.4);
  double rhoAmbP = 101325.0 * 0.0289 / (8.314462 * 293.15);
  double alpha = mu / 0.7 * s->pv[11] / (rhoAmbP * 293.15 / Tmax);
  /* with a CFL step the count is fixed by the largest dt the rule allows */
  const double dtmax = s->sch.cfl > 0 ? (double)s->sch.cfl * s->dx / (double)s->sch.umin : s->dt;
  int n = (int)ceil(dtmax / (s->dx * s->dx / (4.5 * alpha)));
  n = n % 2 ? n : n + 1;
  if (n > 5) n = 5;
  s->diffSub = n;
  s->dsub = (float)(s->dt / n);
  return 0;
}

int PFX(get_packed)(void* h, cf_params* out) { *out = ((S*)h)->P; return 0; }
int PFX(diff_sub)(void* h) { return ((S*)h)->diffSub; }
int PFX(nlevels)(void* h) { return ((S*)h)->nlev; }
double PFX(time)(void* h) { return ((S*)h)->time; }

int PFX(set_schedule)(void* h, const cf_schedule* sc) {
  S* s = (S*)h;
  if (sc->nu < 0 || sc->vcycles < 1 || sc->coarse_iters < 0) return CF_EINVAL;
  if (!(sc->cfl >= 0) || (sc->cfl > 0 && !(sc->umin > 0))) return CF_EINVAL;
  if (sc->refine && !(sc->tol > 0)) return CF_EINVAL;
  if (sc->krylov && !sc->refine) return CF_EINVAL;
#ifdef CFO_DOUBLE
  if (sc->smoother_fp16) return CF_EINVAL; /* an FP32-arithmetic option */
#else
  if (sc->smoother_fp16 && !sc->refine) return CF_EINVAL;
#endif
  if (sc->refine && !s->sch.refine) for (int k = 0; k < s->nC; k++) s->p64[k] = (double)s->p[0][k];
  const int cfl_changed = sc->cfl != s->sch.cfl || sc->umin != s->sch.umin;
  s->sch = *sc;
  if (cfl_changed) PFX(write_params)(h);
  return 0;
}
int PFX(last_vcycles)(void* h) { return ((S*)h)->last_vcycles; }
#ifndef CFO_DOUBLE
void cfo_f2h(const float* a, uint16_t* h, int n) { for (int k = 0; k < n; k++) h[k] = cfo_f2h1(a[k]); }
void cfo_h2f(const uint16_t* h, float* a, int n) { for (int k = 0; k < n; k++) a[k] = cfo_h2f1(h[k]); }
double cfo_gpu_sum(const double* v, int n) { return gpu_sum(v, n); }
#endif
double PFX(last_resid)(void* h) { return ((S*)h)->last_resid; }

int PFX(reset)(void* h) {                                                         /* js/fluid.js:162-176 */
  S* s = (S*)h;
  REAL T0 = (REAL)(float)293.15; /* new Float32Array(nC).fill(PHYS.Tamb) */
  for (int k = 0; k < s->nC; k++) {
    s->T[k] = s->T2[k] = s->Th[k] = T0;
    v4 y = V4(0, (REAL)(float)s->pv[6], 0, 0);
    putY(s->Y, k, y); putY(s->Y2, k, y); putY(s->Yh, k, y);
    s->SQ[2 * k] = s->SQ[2 * k + 1] = 0;
  }
  memset(s->U, 0, sizeof(REAL) * s->nU); memset(s->Uh, 0, sizeof(REAL) * s->nU); memset(s->U2, 0, sizeof(REAL) * s->nU);
  memset(s->V, 0, sizeof(REAL) * s->nV); memset(s->Vh, 0, sizeof(REAL) * s->nV); memset(s->V2, 0, sizeof(REAL) * s->nV);
  memset(s->acc, 0, sizeof(s->acc));
  memset(s->rad, 0, sizeof(s->rad));
  for (int l = 0; l < s->nlev; l++) memset(s->p[l], 0, sizeof(REAL) * s->lnx[l] * s->lny[l]);
  memset(s->p64, 0, sizeof(double) * s->nC);
  s->time = 0;
  return 0;
}

int PFX(upload_pieces)(void* h, const float* data, int n) {                       /* js/fluid.js:201-204 */
  S* s = (S*)h;
  if (n < 0 || n > CF_MAX_PIECES) return CF_EINVAL;
  s->pv[26] = n;
  memcpy(s->pieces, data, sizeof(s->pieces));
  return 0;
}

void PFX(clear_geometry)(void* h) { memset(((S*)h)->faceW, 0, sizeof(((S*)h)->faceW)); } /* js/fluid.js:213 */
void PFX(geometry)(void* h) { k_geometry((S*)h); }
void PFX(radiation)(void* h) { k_radiation((S*)h); }

/* the seven stages of encodeStepProfiled (js/fluid.js:256-292) */
int PFX(run_stage)(void* h, int stage) {
  S* s = (S*)h;
  switch (stage) {
    case 0: k_scalForward(s); k_scalCorrect(s); k_velForward(s); k_velCorrect(s); break;
    case 1:
      k_diffuse(s, s->T2, s->Y2, s->T, s->Y);
      for (int q = 1; q < s->diffSub; q += 2) {
        k_diffuse(s, s->T, s->Y, s->Th, s->Yh);
        k_diffuse(s, s->Th, s->Yh, s->T, s->Y);
      }
      break;
    case 2: k_sources(s); break;
    case 3: k_forces(s); break;
    case 4:
      k_coef(s); k_rhs0(s);
      for (int l = 0; l < s->nlev - 1; l++) k_coefRestrict(s, l);
      break;
    case 5: {
      int nv = 0;
      double r = -1;
      if (s->sch.tol > 0 && s->sch.refine) {
        r = s->sch.krylov ? kry_start(s) : dc_defect(s, 1);
        for (int c = 0; c < s->sch.vcycles && !(r < s->sch.tol); c++) {
          vcycle(s, s->sch.nu);
          nv++;
          if (s->sch.krylov) { r = kry_step(s); continue; }
          for (int k = 0; k < s->nC; k++) s->p64[k] += (double)s->p[0][k];
          r = dc_defect(s, 0);
        }
        s->last_vcycles = nv; s->last_resid = r;
        if (r >= s->sch.tol) return CF_ENOCONV;
        break;
      }
      for (int c = 0; c < s->sch.vcycles; c++) {
        vcycle(s, s->sch.nu);
        nv++;
        if (s->sch.tol > 0) { r = resid_norm(s); if (r < s->sch.tol) break; }
      }
      s->last_vcycles = nv; s->last_resid = r;
      if (s->sch.tol > 0 && r >= s->sch.tol) return CF_ENOCONV;
      break;
    }
    case 6: k_project(s); break;
    default: return CF_EINVAL;
  }
  return 0;
}

/* individual reference kernels, for per-kernel checks */
int PFX(run_kernel)(void* h, const char* name, int arg) {
  S* s = (S*)h;
  if (!strcmp(name, "scalForward")) k_scalForward(s);
  else if (!strcmp(name, "scalCorrect")) k_scalCorrect(s);
  else if (!strcmp(name, "velForward")) k_velForward(s);
  else if (!strcmp(name, "velCorrect")) k_velCorrect(s);
  else if (!strcmp(name, "diffuse")) k_diffuse(s, s->T2, s->Y2, s->T, s->Y);
  else if (!strcmp(name, "sources")) k_sources(s);
  else if (!strcmp(name, "forces")) k_forces(s);
  else if (!strcmp(name, "coef")) k_coef(s);
  else if (!strcmp(name, "rhs0")) k_rhs0(s);
  else if (!strcmp(name, "coefRestrict")) k_coefRestrict(s, arg);
  else if (!strcmp(name, "relax0")) k_relax(s, arg, 0);
  else if (!strcmp(name, "relax1")) k_relax(s, arg, 1);
  else if (!strcmp(name, "restrictRes")) k_restrictRes(s, arg);
  else if (!strcmp(name, "prolong")) k_prolong(s, arg);
  else if (!strcmp(name, "coarsest")) k_coarsest(s, s->nlev - 1, arg);
  else if (!strcmp(name, "project")) k_project(s);
  else if (!strcmp(name, "vcycle")) vcycle(s, arg);
  else return CF_EINVAL;
  return 0;
}

/* CFL time step (the build's schedule option, not in the reference):
 * dt = cfl * dx / max(max |u| over U and V faces, umin) in the field
 * precision, written into the parameter block; dsub = dt / diffSub. */
static void cfl_step(S* s) {
  REAL m = 0;
  for (size_t k = 0; k < s->nU; k++) m = FMAX(m, FABS(s->U[k]));
  for (size_t k = 0; k < s->nV; k++) m = FMAX(m, FABS(s->V[k]));
  const REAL dt = (REAL)s->sch.cfl * PP(dx) / FMAX(m, (REAL)s->sch.umin);
  s->P.dt = (float)dt;
  s->dt = (double)(float)dt;
  s->dsub = (float)((double)(float)dt / (double)s->diffSub);
}

int PFX(step)(void* h, int nsteps) {                                              /* js/fluid.js:225-253 */
  S* s = (S*)h;
  for (int q = 0; q < nsteps; q++) {
    if (s->sch.cfl > 0) cfl_step(s);
    for (int st = 0; st < 7; st++) {
      int rc = PFX(run_stage)(h, st);
      if (rc && rc != CF_ENOCONV) return rc;
    }
    s->time += s->dt;
  }
  return 0;
}

/* encodeReadback + afterSubmit: copy out, then clear acc (js/fluid.js:315-342) */
int PFX(readback)(void* h, int32_t* acc, float* rad, int32_t* faceW) {
  S* s = (S*)h;
  if (acc) memcpy(acc, s->acc, sizeof(s->acc));
  if (rad) for (int q = 0; q < CF_RAD_LEN * 4; q++) rad[q] = (float)s->rad[q];
  if (faceW) memcpy(faceW, s->faceW, sizeof(s->faceW));
  memset(s->acc, 0, sizeof(s->acc));
  return 0;
}

/* field table: name -> (REAL array, element count) */
static REAL* field(S* s, const char* name, size_t* count, int* lev_o) {
  *lev_o = -1;
  size_t nC = s->nC, nU = s->nU, nV = s->nV;
  struct { const char* n; REAL* p; size_t c; } t[] = {
    {"U", s->U, nU}, {"Uh", s->Uh, nU}, {"U2", s->U2, nU},
    {"V", s->V, nV}, {"Vh", s->Vh, nV}, {"V2", s->V2, nV},
    {"T", s->T, nC}, {"Th", s->Th, nC}, {"T2", s->T2, nC},
    {"Y", s->Y, 4 * nC}, {"Yh", s->Yh, 4 * nC}, {"Y2", s->Y2, 4 * nC},
    {"SQ", s->SQ, 2 * nC}, {"fx", s->fx, nU}, {"fy", s->fy, nV}};
  for (size_t q = 0; q < sizeof(t) / sizeof(t[0]); q++)
    if (!strcmp(t[q].n, name)) { *count = t[q].c; return t[q].p; }
  const char* pre[4] = {"p", "rhs", "ax", "ay"};
  for (int a = 0; a < 4; a++) {
    size_t L = strlen(pre[a]);
    if (strncmp(name, pre[a], L) == 0) {
      const char* rest = name + L;
      int l = 0;
      if (*rest) { char* e; l = (int)strtol(rest, &e, 10); if (*e) continue; }
      if (l < 0 || l >= s->nlev) return NULL;
      *lev_o = l;
      size_t lx = s->lnx[l], ly = s->lny[l];
      if (a == 0) { *count = lx * ly; return s->p[l]; }
      if (a == 1) { *count = lx * ly; return s->rhs[l]; }
      if (a == 2) { *count = (lx + 1) * ly; return s->ax[l]; }
      *count = lx * (ly + 1); return s->ay[l];
    }
  }
  return NULL;
}

/* read a field as float32 in the reference layout; cellG = (sdf, bitcast id) */
int PFX(read_field)(void* h, const char* name, void* dst, size_t bytes) {
  S* s = (S*)h;
  if (!strcmp(name, "cellG")) {
    if (bytes != (size_t)s->nC * 8) return CF_EINVAL;
    float* o = (float*)dst;
    for (int k = 0; k < s->nC; k++) { o[2 * k] = (float)s->sdf[k]; memcpy(&o[2 * k + 1], &s->sid[k], 4); }
    return 0;
  }
  if (!strcmp(name, "acc")) { if (bytes != sizeof(s->acc)) return CF_EINVAL; memcpy(dst, s->acc, bytes); return 0; }
  if (!strcmp(name, "faceW")) { if (bytes != sizeof(s->faceW)) return CF_EINVAL; memcpy(dst, s->faceW, bytes); return 0; }
  if (!strcmp(name, "pieces")) { if (bytes != sizeof(s->pieces)) return CF_EINVAL; memcpy(dst, s->pieces, bytes); return 0; }
  if (!strcmp(name, "rad")) {
    if (bytes != sizeof(float) * CF_RAD_LEN * 4) return CF_EINVAL;
    for (int q = 0; q < CF_RAD_LEN * 4; q++) ((float*)dst)[q] = (float)s->rad[q];
    return 0;
  }
  size_t c; int lev;
  REAL* f = field(s, name, &c, &lev);
  if (!f) return CF_EINVAL;
  if (bytes != c * 4) return CF_EINVAL;
  for (size_t q = 0; q < c; q++) ((float*)dst)[q] = (float)f[q];
  return 0;
}
/* same in double (FP64 oracle accuracy comparisons) */
int PFX(read_field64)(void* h, const char* name, double* dst, size_t count) {
  S* s = (S*)h;
  if (!strcmp(name, "sdf")) { if (count != (size_t)s->nC) return CF_EINVAL; for (int k = 0; k < s->nC; k++) dst[k] = s->sdf[k]; return 0; }
  if (!strcmp(name, "rad")) { if (count != CF_RAD_LEN * 4) return CF_EINVAL; for (int q = 0; q < CF_RAD_LEN * 4; q++) dst[q] = s->rad[q]; return 0; }
  if (!strcmp(name, "p64")) { if (count != (size_t)s->nC) return CF_EINVAL; memcpy(dst, s->p64, sizeof(double) * s->nC); return 0; }
  if (!strcmp(name, "d32")) { if (count != (size_t)s->nC) return CF_EINVAL; for (int k = 0; k < s->nC; k++) dst[k] = s->d32[k]; return 0; }
  size_t c; int lev;
  REAL* f = field(s, name, &c, &lev);
  if (!f || count != c) return CF_EINVAL;
  for (size_t q = 0; q < c; q++) dst[q] = (double)f[q];
  return 0;
}
int PFX(write_field)(void* h, const char* name, const void* src, size_t bytes) {
  S* s = (S*)h;
  if (!strcmp(name, "cellG")) {
    if (bytes != (size_t)s->nC * 8) return CF_EINVAL;
    const float* in = (const float*)src;
    for (int k = 0; k < s->nC; k++) { s->sdf[k] = (REAL)in[2 * k]; memcpy(&s->sid[k], &in[2 * k + 1], 4); }
    return 0;
  }
  if (!strcmp(name, "acc")) { if (bytes != sizeof(s->acc)) return CF_EINVAL; memcpy(s->acc, src, bytes); return 0; }
  if (!strcmp(name, "faceW")) { if (bytes != sizeof(s->faceW)) return CF_EINVAL; memcpy(s->faceW, src, bytes); return 0; }
  if (!strcmp(name, "pieces")) { if (bytes != sizeof(s->pieces)) return CF_EINVAL; memcpy(s->pieces, src, bytes); return 0; }
  if (!strcmp(name, "rad")) {
    if (bytes != sizeof(float) * CF_RAD_LEN * 4) return CF_EINVAL;
    for (int q = 0; q < CF_RAD_LEN * 4; q++) s->rad[q] = (REAL)((const float*)src)[q];
    return 0;
  }
  size_t c; int lev;
  REAL* f = field(s, name, &c, &lev);
  if (!f) return CF_EINVAL;
  if (bytes != c * 4) return CF_EINVAL;
  for (size_t q = 0; q < c; q++) f[q] = (REAL)((const float*)src)[q];
  return 0;
}
int PFX(write_field64)(void* h, const char* name, const double* src, size_t count) {
  S* s = (S*)h;
  if (!strcmp(name, "p64")) {
    if (count != (size_t)s->nC) return CF_EINVAL;
    memcpy(s->p64, src, sizeof(double) * s->nC);
    return 0;
  }
  size_t c; int lev;
  REAL* f = field(s, name, &c, &lev);
  if (!f || count != c) return CF_EINVAL;
  for (size_t q = 0; q < c; q++) f[q] = (REAL)src[q];
  return 0;
}
