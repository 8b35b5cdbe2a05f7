/*
 * flame1d.c — restatement of tests/flame1d.mjs:13-71 (1D freely propagating
 * premixed flame in the lab frame with the simulator's gas closures).
 * TEST INFRASTRUCTURE ONLY (oracle).  Double precision, one thread, the same
 * operation order as the JS so the printed table of the reference script is
 * reproduced (tests/golden/flame1d.json).
 */
#include <math.h>
#include <stdlib.h>
#include "cf_oracle.h"

/* tests/flame1d.mjs:9 — Sutherland viscosity */
static double mu1(double T) { return 1.716e-5 * pow(T / 273.15, 1.5) * 383.55 / (T + 110.4); }

/* Math.round for the positive arguments used here (half rounds up) */
static int jsround(double x) { return (int)floor(x + 0.5); }

cfo_flame_result cfo_flame_speed(double B, double Ta, double dx, double F,
                                 double phi, double L, double tEnd) {
  /* tests/flame1d.mjs:6-8 */
  const double Tamb = 293.15, rhoAmb = 1.2, cp = 1300, Pr = 0.7, dHc = 15e6, sO2 = 15e6 / 13.1e6, YO = 0.232;
  cfo_flame_result res = {0, 0, 0};
  const int n = jsround(L / dx);                                          /* :15 */
  const double YFst = YO / sO2;                                           /* :16 */
  const double YF0 = phi * YFst / (1 + phi * YFst) * 1, YO0 = YO / (1 + phi * YFst);
  double *T = malloc(sizeof(double) * n), *YF = malloc(sizeof(double) * n), *YOx = malloc(sizeof(double) * n);
  double *Tn = malloc(sizeof(double) * n), *Fn = malloc(sizeof(double) * n), *On = malloc(sizeof(double) * n);
  double *S = malloc(sizeof(double) * n), *u = calloc(n + 1, sizeof(double));
  for (int i = 0; i < n; i++) { T[i] = Tamb; YF[i] = YF0; YOx[i] = YO0; }
  const int nign = jsround(0.01 / dx);                                    /* :21 */
  for (int i = 0; i < nign && i < n; i++) { T[i] = 2000; YF[i] = 0; YOx[i] = fmax(0, YO0 - sO2 * YF0); }
  const double alphaMax = mu1(2300) / Pr * F / (rhoAmb * Tamb / 2300);    /* :22-23 */
  const double dt = fmin(0.2 * dx * dx / alphaMax, 0.2 * dx / 6);
  double t = 0, Tb = 0;
  int have_a = 0, have_b = 0, steps = 0;
  double at = 0, ax = 0, bt = 0, bx = 0, lt = 0, lx = 0;
  while (t < tEnd) {                                                      /* :28 */
    for (int i = 0; i < n; i++) { Tn[i] = T[i]; Fn[i] = YF[i]; On[i] = YOx[i]; }
    for (int i = 0; i < n; i++) {                                         /* :32-46 */
      const double rho = rhoAmb * Tamb / T[i];
      double dT = 0, dF = 0, dO = 0;
      for (int q = 0; q < 2; q++) {
        int j = q == 0 ? i - 1 : i + 1;
        if (j < 0 || j >= n) continue;
        const double a = mu1(0.5 * (T[i] + T[j])) / Pr * F;
        dT += a * (T[j] - T[i]); dF += a * (YF[j] - YF[i]); dO += a * (YOx[j] - YOx[i]);
      }
      const double c = dt / (rho * dx * dx);
      Tn[i] += c * dT; Fn[i] += c * dF; On[i] += c * dO;
      const double w = B * rho * Fn[i] * On[i] * exp(-Ta / Tn[i]) / F;
      const double d = fmin(fmin(w * dt, Fn[i]), On[i] / sO2);
      Fn[i] -= d; On[i] -= sO2 * d; Tn[i] += d * dHc / cp;
      S[i] = log(Tn[i] / T[i]) / dt;
    }
    double* sw;
    sw = T; T = Tn; Tn = sw;
    sw = YF; YF = Fn; Fn = sw;
    sw = YOx; YOx = On; On = sw;
    for (int i = 0; i < n; i++) u[i + 1] = u[i] + S[i] * dx;              /* :49 */
    for (int i = 0; i < n; i++) { Tn[i] = T[i]; Fn[i] = YF[i]; On[i] = YOx[i]; }
    for (int i = 0; i < n; i++) {                                         /* :51-57 */
      const double uc = 0.5 * (u[i] + u[i + 1]);
      const int j = uc > 0 ? i - 1 : i + 1;
      const int k = j < 0 ? 0 : (j > n - 1 ? n - 1 : j);
      const double cfl = fabs(uc) * dt / dx;
      Tn[i] = T[i] + cfl * (T[k] - T[i]); Fn[i] = YF[i] + cfl * (YF[k] - YF[i]); On[i] = YOx[i] + cfl * (YOx[k] - YOx[i]);
    }
    sw = T; T = Tn; Tn = sw;
    sw = YF; YF = Fn; Fn = sw;
    sw = YOx; YOx = On; On = sw;
    t += dt;
    steps++;
    double xf = 0;                                                        /* :61-64 */
    for (int i = 0; i < n; i++) if (T[i] > 1200) xf = (i + 0.5) * dx;
    for (int i = 0; i < n; i++) Tb = fmax(Tb, T[i]);
    if (!have_a && xf > 0.3 * L) { have_a = 1; at = t; ax = xf; }
    if (!have_b && xf > 0.75 * L) { have_b = 1; bt = t; bx = xf; }
    lt = t; lx = xf;
    if (xf > L * 0.8) break;
  }
  if (!have_b) { bt = lt; bx = lx; }                                       /* :68-70 */
  res.Tb = Tb; res.steps = steps;
  if (!have_a || steps == 0 || bt <= at) res.SL = 0;
  else res.SL = (bx - ax) / (bt - at) * Tamb / Tb;
  free(T); free(YF); free(YOx); free(Tn); free(Fn); free(On); free(S); free(u);
  return res;
}
