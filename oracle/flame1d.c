/*
 * flame1d.c — restatement of tests/flame1d.mjs:13-71 (1D freely propagating
 * premixed flame in the lab frame with the simulator's gas closures).
 * TEST INFRASTRUCTURE ONLY (oracle).  One thread, the same operation order as
 * the JS, so the printed table of the reference script is reproduced
 * (tests/golden/flame1d.json).
 *
 * Compiled twice: in double (cfo_flame_speed, cfo_flame_speed_ex: the JS
 * numbers) and with -DFL_FLOAT in single precision (cfo_flame_speed32_ex: the
 * FP32 variant of the same algorithm).  The _ex forms take the ambient and
 * ignition temperatures, 293.15 K and 2000 K in the script (:6, :21).
 */
#include <math.h>
#include <stdlib.h>
#include "cf_oracle.h"

#ifdef FL_FLOAT
typedef float real;
#define R(x) ((float)(x))
#define POW powf
#define EXP expf
#define LOG logf
#define FABS fabsf
#define FMIN fminf
#define FMAX fmaxf
#define FLOOR floorf
#define FN_EX cfo_flame_speed32_ex
#else
typedef double real;
#define R(x) ((double)(x))
#define POW pow
#define EXP exp
#define LOG log
#define FABS fabs
#define FMIN fmin
#define FMAX fmax
#define FLOOR floor
#define FN_EX cfo_flame_speed_ex
#endif

/* tests/flame1d.mjs:9 — Sutherland viscosity */
static real mu1(real T) { return R(1.716e-5) * POW(T / R(273.15), R(1.5)) * R(383.55) / (T + R(110.4)); }

/* Math.round for the positive arguments used here (half rounds up) */
static int jsround(double x) { return (int)floor(x + 0.5); }

cfo_flame_result FN_EX(double B_, double Ta_, double dx_, double F_, double phi_, double L_, double tEnd_,
                       double Tamb_, double Tign_) {
  const real B = R(B_), Ta = R(Ta_), dx = R(dx_), F = R(F_), phi = R(phi_), L = R(L_), tEnd = R(tEnd_);
  /* tests/flame1d.mjs:6-8 */
  const real Tamb = R(Tamb_), rhoAmb = R(1.2), cp = R(1300), Pr = R(0.7), dHc = R(15e6), sO2 = R(15e6 / 13.1e6),
             YO = R(0.232);
  cfo_flame_result res = {0, 0, 0};
  const int n = jsround(L_ / dx_);                                        /* :15 */
  const real YFst = YO / sO2;                                             /* :16 */
  const real YF0 = phi * YFst / (R(1) + phi * YFst) * R(1), YO0 = YO / (R(1) + phi * YFst);
  real *T = malloc(sizeof(real) * n), *YF = malloc(sizeof(real) * n), *YOx = malloc(sizeof(real) * n);
  real *Tn = malloc(sizeof(real) * n), *Fn = malloc(sizeof(real) * n), *On = malloc(sizeof(real) * n);
  real *S = malloc(sizeof(real) * n), *u = calloc(n + 1, sizeof(real));
  for (int i = 0; i < n; i++) { T[i] = Tamb; YF[i] = YF0; YOx[i] = YO0; }
  const int nign = jsround(0.01 / dx_);                                   /* :21 */
  for (int i = 0; i < nign && i < n; i++) { T[i] = R(Tign_); YF[i] = 0; YOx[i] = FMAX(R(0), YO0 - sO2 * YF0); }
  const real alphaMax = mu1(R(2300)) / Pr * F / (rhoAmb * Tamb / R(2300)); /* :22-23 */
  const real dt = FMIN(R(0.2) * dx * dx / alphaMax, R(0.2) * dx / R(6));
  real t = 0, Tb = 0;
  int have_a = 0, have_b = 0, steps = 0;
  real at = 0, ax = 0, bt = 0, bx = 0, lt = 0, lx = 0;
  while (t < tEnd) {                                                      /* :28 */
    for (int i = 0; i < n; i++) { Tn[i] = T[i]; Fn[i] = YF[i]; On[i] = YOx[i]; }
    for (int i = 0; i < n; i++) {                                         /* :32-46 */
      const real rho = rhoAmb * Tamb / T[i];
      real dT = 0, dF = 0, dO = 0;
      for (int q = 0; q < 2; q++) {
        int j = q == 0 ? i - 1 : i + 1;
        if (j < 0 || j >= n) continue;
        const real a = mu1(R(0.5) * (T[i] + T[j])) / Pr * F;
        dT += a * (T[j] - T[i]); dF += a * (YF[j] - YF[i]); dO += a * (YOx[j] - YOx[i]);
      }
      const real c = dt / (rho * dx * dx);
      Tn[i] += c * dT; Fn[i] += c * dF; On[i] += c * dO;
      const real w = B * rho * Fn[i] * On[i] * EXP(-Ta / Tn[i]) / F;
      const real d = FMIN(FMIN(w * dt, Fn[i]), On[i] / sO2);
      Fn[i] -= d; On[i] -= sO2 * d; Tn[i] += d * dHc / cp;
      S[i] = LOG(Tn[i] / T[i]) / dt;
    }
    real* sw;
    sw = T; T = Tn; Tn = sw;
    sw = YF; YF = Fn; Fn = sw;
    sw = YOx; YOx = On; On = sw;
    for (int i = 0; i < n; i++) u[i + 1] = u[i] + S[i] * dx;              /* :49 */
    for (int i = 0; i < n; i++) { Tn[i] = T[i]; Fn[i] = YF[i]; On[i] = YOx[i]; }
    for (int i = 0; i < n; i++) {                                         /* :51-57 */
      const real uc = R(0.5) * (u[i] + u[i + 1]);
      const int j = uc > 0 ? i - 1 : i + 1;
      const int k = j < 0 ? 0 : (j > n - 1 ? n - 1 : j);
      const real cfl = FABS(uc) * dt / dx;
      Tn[i] = T[i] + cfl * (T[k] - T[i]); Fn[i] = YF[i] + cfl * (YF[k] - YF[i]); On[i] = YOx[i] + cfl * (YOx[k] - YOx[i]);
    }
    sw = T; T = Tn; Tn = sw;
    sw = YF; YF = Fn; Fn = sw;
    sw = YOx; YOx = On; On = sw;
    t += dt;
    steps++;
    real xf = 0;                                                          /* :61-64 */
    for (int i = 0; i < n; i++) if (T[i] > R(1200)) xf = ((real)i + R(0.5)) * dx;
    for (int i = 0; i < n; i++) Tb = FMAX(Tb, T[i]);
    if (!have_a && xf > R(0.3) * L) { have_a = 1; at = t; ax = xf; }
    if (!have_b && xf > R(0.75) * L) { have_b = 1; bt = t; bx = xf; }
    lt = t; lx = xf;
    if (xf > L * R(0.8)) break;
  }
  if (!have_b) { bt = lt; bx = lx; }                                       /* :68-70 */
  res.Tb = Tb; res.steps = steps;
  if (!have_a || steps == 0 || bt <= at) res.SL = 0;
  else res.SL = (bx - ax) / (bt - at) * Tamb / Tb;
  free(T); free(YF); free(YOx); free(Tn); free(Fn); free(On); free(S); free(u);
  return res;
}

#ifndef FL_FLOAT
cfo_flame_result cfo_flame_speed(double B, double Ta, double dx, double F, double phi, double L, double tEnd) {
  return cfo_flame_speed_ex(B, Ta, dx, F, phi, L, tEnd, 293.15, 2000.0);
}
#endif
