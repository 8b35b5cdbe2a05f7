/*
 * cf_oracle.h — CPU oracle of campfire's gas step.  TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference algorithm (the WGSL kernels under
 * shaders/, the generated samplers of js/gpu.js:16-65, the schedule of
 * js/fluid.js and tests/flame1d.mjs).  It is compiled twice from cf_oracle.c:
 * with REAL=float (prefix cfo32_) for parity with the FP32 GPU path, and with
 * REAL=double (prefix cfo64_) as the FP64 accuracy reference.  Only the test
 * suite, the smoke check and the benchmark's CPU-baseline leg load it; the
 * product library never links or calls it.
 */
#ifndef CF_ORACLE_H
#define CF_ORACLE_H
#include <stddef.h>
#include <stdint.h>
#include "../include/campfire_hip.h"

/* flame1d.c: tests/flame1d.mjs:13-71 restated (double, single thread). */
typedef struct { double SL, Tb; int steps; } cfo_flame_result;
cfo_flame_result cfo_flame_speed(double B, double Ta, double dx, double F,
                                 double phi, double L, double tEnd);
/* The same with the ambient and ignition temperatures as arguments (293.15 K
 * and 2000 K in the script), in double and in single precision. */
cfo_flame_result cfo_flame_speed_ex(double B, double Ta, double dx, double F, double phi, double L, double tEnd,
                                    double Tamb, double Tign);
/* IEEE binary16 conversions of the oracle (round to nearest even, subnormals kept) */
void cfo_f2h(const float* a, uint16_t* h, int n);
void cfo_h2f(const uint16_t* h, float* a, int n);
/* sum of v[0..n) in the GPU's fixed reduction order (norms and dot products) */
double cfo_gpu_sum(const double* v, int n);
cfo_flame_result cfo_flame_speed32_ex(double B, double Ta, double dx, double F, double phi, double L, double tEnd,
                                      double Tamb, double Tign);

#endif
