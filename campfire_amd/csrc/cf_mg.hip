// cf_mg.hip — multigrid V-cycle of the pressure projection, fused per level.
//
// The reference runs the V-cycle of js/fluid.js:294-312 as one dispatch per
// red or black half-sweep, one per restriction and one per prolongation
// (shaders/pressure.wgsl:77-156): 61 dispatches at 1024^2.  Red-black
// Gauss-Seidel is a colour-ordered Jacobi, so a tile plus a halo of k cells
// reproduces k global half-sweeps exactly: every half-sweep invalidates one
// more ring from the halo edge and the tile interior stays exact.  That gives
//   k_mg_down  : nu sweeps (red, black) + residual restriction, halo 2nu+1
//   k_mg_up    : prolongation + nu sweeps (black, red), halo 2nu
//   k_mg_bottom: the last levels (<= 4096, 1024, 256 cells) in one LDS-resident workgroup,
//                including the coarsest level's red-black sweeps
// with the reference's arithmetic and operation order in every update.
#include "cf_kernels.h"

namespace cf {

// ---- tiled legs ----------------------------------------------------------------
// A tile of TX x TY fine cells plus a halo of H cells (the region) lives in
// LDS inside one more ring of zeros, so every cell's four neighbours sit at
// fixed offsets.  Each thread owns horizontal pairs of region cells, one of
// each colour, and keeps per cell its LDS index and flags, the four face
// coefficients (zero towards a missing neighbour, whose value is then taken
// times 0), the diagonal, its reciprocal and rhs*h2 in registers; the red
// and the black cells of its pairs are sorted apart at load, so a half-sweep
// updates one set without selects.  Cells on the region edge whose outward
// neighbour is a real cell are never updated; that error moves inward one
// ring per half-sweep and stays out of the tile interior.
struct LegCell {
  int ci;  // padded LDS index | LEG_UPD | LEG_IN
  float w, e, s, n, d, y, rh;
};
constexpr int LEG_UPD = 1 << 16, LEG_IN = 1 << 17, LEG_IDX = 0xffff;

__device__ __forceinline__ float ldc(const float* a, int k) { return a[k]; }
__device__ __forceinline__ float ldc(const __half* a, int k) { return __half2float(a[k]); }

// cell (x, y) at region position (rx, ry) without its rhs (rh = 0); CT:
// float, or __half for the FP16 smoother's coefficient copies
template <class CT>
__device__ __forceinline__ LegCell leg_cell_coef(int nx, int ny, const CT* __restrict__ ax,
                                                 const CT* __restrict__ ay, int x, int y, int rx, int ry, int RX,
                                                 int RY, int PX) {
  LegCell c;
  const bool in = x >= 0 && x < nx && y >= 0 && y < ny;
  const bool upd = in && !(rx == 0 && x > 0) && !(rx == RX - 1 && x < nx - 1) && !(ry == 0 && y > 0) &&
                   !(ry == RY - 1 && y < ny - 1);
  c.ci = ((ry + 1) * PX + rx + 1) | (upd ? LEG_UPD : 0) | (in ? LEG_IN : 0);
  // the loads are unconditional (a cell outside the domain reads cell 0's
  // faces and discards them): no branch separates them, so every load of a
  // thread's cells is in flight at once instead of one round trip per cell
  const int xx = in ? x : 0, yy = in ? y : 0;
  const float w = ldc(ax, yy * (nx + 1) + xx), e = ldc(ax, yy * (nx + 1) + xx + 1);
  const float s = ldc(ay, yy * nx + xx), n = ldc(ay, (yy + 1) * nx + xx);
  float diag = 0.f;  // shaders/pressure.wgsl:77-85, same accumulation order
  if (x > 0) diag += w; else diag += K(2.0) * w;
  if (x < nx - 1) diag += e; else diag += K(2.0) * e;
  if (y > 0) diag += s;
  if (y < ny - 1) diag += n; else diag += K(2.0) * n;
  c.d = in ? diag : K(0);
  c.y = in && diag > K(0) ? K(1) / diag : K(0);
  c.w = in && x > 0 ? w : 0.f;
  c.e = in && x < nx - 1 ? e : 0.f;
  c.s = in && y > 0 ? s : 0.f;
  c.n = in && y < ny - 1 ? n : 0.f;
  c.rh = 0.f;
  return c;
}

// the same with its rhs; rhs_out: the rhs (0 outside the domain)
template <class CT>
__device__ __forceinline__ LegCell leg_cell(int nx, int ny, const CT* __restrict__ ax, const CT* __restrict__ ay,
                                            const float* __restrict__ rhs, float h2, int x, int y, int rx, int ry,
                                            int RX, int RY, int PX, float& rhs_out) {
  LegCell c = leg_cell_coef(nx, ny, ax, ay, x, y, rx, ry, RX, RY, PX);
  const bool in = (c.ci & LEG_IN) != 0;
  const float r = rhs[in ? y * nx + x : 0];
  rhs_out = in ? r : 0.f;
  c.rh = in ? r * h2 : 0.f;
  return c;
}

// a level's value at cell (x, y) when `in` (else cell 0's, to be discarded)
__device__ __forceinline__ int leg_idx(bool in, int x, int y, int nx) { return in ? y * nx + x : 0; }

template <int M>
__device__ __forceinline__ LegCell leg_cell_m(const Level& f, const MgStore& st, int x, int y, int rx, int ry, int RX,
                                              int RY, int PX, float& rhs_out) {
  if constexpr ((M & 1) != 0)
    return leg_cell(f.nx, f.ny, st.f.ax, st.f.ay, f.rhs, f.h2, x, y, rx, ry, RX, RY, PX, rhs_out);
  else return leg_cell(f.nx, f.ny, f.ax, f.ay, f.rhs, f.h2, x, y, rx, ry, RX, RY, PX, rhs_out);
}

// a slot without a cell: reads around the first region cell, never writes
__device__ __forceinline__ LegCell leg_none(int PX) {
  LegCell c;
  c.ci = PX + 1;
  c.w = c.e = c.s = c.n = c.d = c.y = c.rh = 0.f;
  return c;
}

// A region's LDS array holds LEG_DUMMY words past its cells: lane l of a
// wave stores to word l there when its slot's cell is not updated.
constexpr int LEG_DUMMY = 64;
__device__ __forceinline__ float* leg_dummy(float* sp, int n) { return sp + n + (threadIdx.x & 63); }

// One half-sweep of a thread's cells of one colour, the red-black update of
// shaders/pressure.wgsl:87-94 (terms in its order).  The LDS reads of G
// cells at a time are issued before their arithmetic (G = Q: all of them;
// fewer where registers are short), and every update is stored without a
// branch: to its cell where the cell is updated, else to the lane's dummy
// word.  (The cells of one colour never neighbour each other, so no read of
// this half-sweep sees a store of it.)  A slot without a cell (leg_none)
// reads inside the region and stores to the dummy.
template <int PX, int G = 0, int Q>
__device__ __forceinline__ void leg_sweep(const LegCell (&c)[Q], float* sp, float* dummy) {
  constexpr int NG = G > 0 && G < Q ? G : Q;
#pragma unroll
  for (int g = 0; g < Q; g += NG) {
    float a0[NG], a1[NG], a2[NG], a3[NG];
#pragma unroll
    for (int q = 0; q < NG && g + q < Q; q++) {
      const int ci = c[g + q].ci & LEG_IDX;
      a0[q] = sp[ci - 1];
      a1[q] = sp[ci + 1];
      a2[q] = sp[ci - PX];
      a3[q] = sp[ci + PX];
    }
#pragma unroll
    for (int q = 0; q < NG && g + q < Q; q++) {
      const LegCell& x = c[g + q];
      const float off = x.w * a0[q] + x.e * a1[q] + x.s * a2[q] + x.n * a3[q];
      const float v = div_rcp_fast(off - x.rh, x.d, x.y);
      float* dst = (x.ci & LEG_UPD) ? sp + (x.ci & LEG_IDX) : dummy;
      *dst = x.d > K(0) ? v : K(0);
    }
  }
}

// p of a level: FP32, or FP16 holding p * sig (read back as h * isig, exact)
template <bool H>
__device__ __forceinline__ float ldp(const float* p, const __half* ph, int k, float isig) {
  if constexpr (H) return __half2float(ph[k]) * isig;
  else return p[k];
}
template <bool H>
__device__ __forceinline__ void stp(float* p, __half* ph, int k, float v, float sig) {
  if constexpr (H) ph[k] = __float2half(v * sig);
  else p[k] = v;
}

// M (storage, MgStore::mode): bit 0, the fine level's p and coefficients in
// FP16; bit 1, the coarse level's p in FP16.  Sweeps, residual and
// restriction are FP32 arithmetic either way.  The fine p is read from p and
// written to the leg buffer q.
template <int TX, int TY, int NU, int M>
struct MgDownOp {
  struct Args {
    Level f;
    Level c;
    MgStore st;
    const int* skip;
    int vg;      // tiles of this scene's launch
    int zero_p;  // start from p = 0 without reading the fine p
  };
  static constexpr int kThreads = TX == 64 ? (TX * TY / 4 + 63) / 64 * 64 : 256;  // about four cells a thread
  static constexpr int kTX = TX, kTY = TY;
  // the wide tiles (64x48, 64x32): two workgroups share a CU
  static constexpr int kMinWaves = TX * TY == 3072 ? 6 : (TX * TY == 2048 ? 4 : 1);
  static constexpr int H = 2 * NU + 1;
  static constexpr int RX = TX + 2 * H, RY = TY + 2 * H, PX = RX + 2, PY = RY + 2, NP = RX * RY / 2;
  static constexpr int PQ = (NP + kThreads - 1) / kThreads;
  static_assert(RX % 2 == 0, "pairs need an even region width");
  static __device__ __forceinline__ void run(const Args& a_, int vb) {
    const Level& f = a_.f;
    const Level& c = a_.c;
    if (a_.skip && *a_.skip) return;  // the step's residual target is met
    __shared__ float sp[PX * PY + LEG_DUMMY];
    __shared__ float sres[TX * TY];  // the interior's rhs, then its residual
    constexpr bool FH = (M & 1) != 0;
    const int nx = f.nx;
    const int gx = (nx + TX - 1) / TX;
    const int x0 = (vb % gx) * TX, y0 = (vb / gx) * TY;
    const int tid = threadIdx.x;
    const float h2 = f.h2;
    const float sig = M ? a_.st.dyn->sig : 1.f, isig = M ? a_.st.dyn->isig : 1.f;
    for (int k = tid; k < 2 * PX + 2 * RY; k += kThreads) {  // the ring of zeros
      const int idx = k < PX ? k : (k < 2 * PX ? (PY - 1) * PX + (k - PX) : ((k - 2 * PX) / 2 + 1) * PX +
                                                                               ((k & 1) ? PX - 1 : 0));
      sp[idx] = 0.f;
    }
    LegCell cR[PQ], cB[PQ];
    float pv[PQ][2], rv[PQ][2];
    // every global load first, branch-free (a slot without a pair loads pair 0's data)
#pragma unroll
    for (int q = 0; q < PQ; q++) {
      const int pi = tid + kThreads * q;
      const int pc = pi < NP ? pi : 0;
      const int ry = pc / (RX / 2), rx = 2 * (pc % (RX / 2));
      const int x = x0 - H + rx, y = y0 - H + ry;
      const LegCell c0 = leg_cell_m<M>(f, a_.st, x, y, rx, ry, RX, RY, PX, rv[q][0]);
      const LegCell c1 = leg_cell_m<M>(f, a_.st, x + 1, y, rx + 1, ry, RX, RY, PX, rv[q][1]);
      const bool in0 = (c0.ci & LEG_IN) != 0, in1 = (c1.ci & LEG_IN) != 0;
      float v0 = 0.f, v1 = 0.f;
      if (!a_.zero_p) {  // uniform over the launch's scene
        v0 = ldp<FH>(f.p, a_.st.f.p, leg_idx(in0, x, y, nx), isig);
        v1 = ldp<FH>(f.p, a_.st.f.p, leg_idx(in1, x + 1, y, nx), isig);
      }
      pv[q][0] = in0 ? v0 : 0.f;
      pv[q][1] = in1 ? v1 : 0.f;
      const bool red0 = ((x + y) & 1) == 0;
      cR[q] = red0 ? c0 : c1;
      cB[q] = red0 ? c1 : c0;
    }
#pragma unroll
    for (int q = 0; q < PQ; q++) {
      const int pi = tid + kThreads * q;
      if (pi < NP) {
        const int ry = pi / (RX / 2), rx = 2 * (pi % (RX / 2));
        sp[(ry + 1) * PX + rx + 1] = pv[q][0];
        sp[(ry + 1) * PX + rx + 2] = pv[q][1];
        if (ry >= H && ry < H + TY) {
          if (rx >= H && rx < H + TX) sres[(ry - H) * TX + (rx - H)] = rv[q][0];
          if (rx + 1 >= H && rx + 1 < H + TX) sres[(ry - H) * TX + (rx + 1 - H)] = rv[q][1];
        }
      } else {
        cR[q] = cB[q] = leg_none(PX);
      }
    }
    __syncthreads();
    for (int s = 0; s < NU; s++) {  // red then black (js/fluid.js:298-301)
      leg_sweep<PX>(cR, sp, leg_dummy(sp, PX * PY));
      __syncthreads();
      leg_sweep<PX>(cB, sp, leg_dummy(sp, PX * PY));
      __syncthreads();
    }
    // interior: p into the leg buffer, residual into sres (shaders/pressure.wgsl:96-101)
    auto finish = [&](const LegCell& a) {
      const int ci = a.ci & LEG_IDX;
      const int rx = ci % PX - 1, ry = ci / PX - 1;
      if ((a.ci & LEG_IN) && rx >= H && rx < H + TX && ry >= H && ry < H + TY) {
        const float pk = sp[ci];
        const float off = a.w * sp[ci - 1] + a.e * sp[ci + 1] + a.s * sp[ci - PX] + a.n * sp[ci + PX];
        const int r = (ry - H) * TX + (rx - H);
        sres[r] = a.d <= K(0) ? K(0) : sres[r] - (off - a.d * pk) / h2;
        stp<FH>(f.q, a_.st.f.q, (y0 - H + ry) * nx + (x0 - H + rx), pk, sig);
      }
    };
#pragma unroll
    for (int q = 0; q < PQ; q++) {
      finish(cR[q]);
      finish(cB[q]);
    }
    __syncthreads();
    // restriction to the coarse rhs, coarse p = 0 (:104-113)
    for (int k = tid; k < (TX / 2) * (TY / 2); k += kThreads) {
      const int li = k % (TX / 2), lj = k / (TX / 2);
      const int I = x0 / 2 + li, J = y0 / 2 + lj;
      if (I < c.nx && J < c.ny) {
        const float* r0 = sres + (2 * lj) * TX + 2 * li;
        const float r = r0[0] + r0[1] + r0[TX] + r0[TX + 1];
        c.rhs[J * c.nx + I] = K(0.25) * r;
        stp<(M & 2) != 0>(c.p, a_.st.c.p, J * c.nx + I, K(0), sig);
      }
    }
  }
};
template <int TX, int TY, int NU, int M>
__global__ __launch_bounds__((MgDownOp<TX, TY, NU, M>::kThreads), (MgDownOp<TX, TY, NU, M>::kMinWaves)) void k_mg_down(const typename MgDownOp<TX, TY, NU, M>::Args a_) {
  MgDownOp<TX, TY, NU, M>::run(a_, xcd_tile(blockIdx.x, gridDim.x));
}

// the fine p is read from the leg buffer q and written to p.  PROJ: the
// level-0 leg of a step's last V-cycle also projects the velocities of its
// tile's faces (shaders/pressure.wgsl's project, as k_project does it): one
// more ring of halo makes p exact one cell beyond the tile, where the faces on
// the tile's edges take their pressure differences.
template <int TX, int TY, int NU, int M, bool PROJ = false>
struct MgUpOp {
  struct ArgsLeg {
    Level f;
    Level c;
    MgStore st;
    const int* skip;
    int vg;  // tiles of this scene's launch
  };
  struct ArgsProj : ArgsLeg {
    Fields F;
  };
  using Args = std::conditional_t<PROJ, ArgsProj, ArgsLeg>;
  static_assert(!PROJ || M == 0, "the projection reads FP32 p");
  static constexpr int kThreads = TX == 64 ? (TX * TY / 4 + 63) / 64 * 64 : 256;  // about four cells a thread
  static constexpr int kTX = TX, kTY = TY;
  static constexpr int kMinWaves = TX == 64 && TY <= 32 ? 4 : 1;  // 64x32, 64x30: two workgroups share a CU
  static constexpr int H = 2 * NU + (PROJ ? 1 : 0);
  static constexpr int RX = TX + 2 * H, RY = TY + 2 * H, PX = RX + 2, PY = RY + 2, NP = RX * RY / 2;
  static constexpr int PQ = (NP + kThreads - 1) / kThreads;
  // cells whose sweep reads are in flight together: all, but one at a time in
  // the projecting leg, which holds the faces' velocities and fractions in
  // registers through the sweeps (more would spill)
  static constexpr int SG = PROJ ? 1 : 0;
  static_assert(RX % 2 == 0, "pairs need an even region width");
  static __device__ __forceinline__ void run(const Args& a_, int vb) {
    const Level& f = a_.f;
    const Level& c = a_.c;
    if (a_.skip && *a_.skip) return;
    __shared__ float sp[PX * PY + LEG_DUMMY];
    constexpr bool FH = (M & 1) != 0, CH = (M & 2) != 0;
    const int nx = f.nx;
    const int gx = (nx + TX - 1) / TX;
    const int x0 = (vb % gx) * TX, y0 = (vb / gx) * TY;
    const int tid = threadIdx.x;
    const int nxc = c.nx, nyc = c.ny;
    const float sig = M ? a_.st.dyn->sig : 1.f, isig = M ? a_.st.dyn->isig : 1.f;

    for (int k = tid; k < 2 * PX + 2 * RY; k += kThreads) {  // the ring of zeros
      const int idx = k < PX ? k : (k < 2 * PX ? (PY - 1) * PX + (k - PX) : ((k - 2 * PX) / 2 + 1) * PX +
                                                                               ((k & 1) ? PX - 1 : 0));
      sp[idx] = 0.f;
    }
    LegCell cR[PQ], cB[PQ];
    float pv[PQ][2], cv[PQ][2][4];
    // every global load first, branch-free (a slot without a pair, or a cell
    // outside the domain, loads cell 0's data and discards it); the values
    // combine in a second pass, so no select pulls a load into a branch
#pragma unroll
    for (int q = 0; q < PQ; q++) {
      const int pi = tid + kThreads * q;
      const int pc = pi < NP ? pi : 0;
      const int ry = pc / (RX / 2), rx = 2 * (pc % (RX / 2));
      const int y = y0 - H + ry;
      LegCell cc[2];
#pragma unroll
      for (int sub = 0; sub < 2; sub++) {
        const int x = x0 - H + rx + sub;
        float rr;
        cc[sub] = leg_cell_m<M>(f, a_.st, x, y, rx + sub, ry, RX, RY, PX, rr);
        const bool in = (cc[sub].ci & LEG_IN) != 0;
        const int xi = in ? x : 0, yi = in ? y : 0;
        // bilinear coarse correction, skipped in closed cells (:116-129)
        const int I = xi / 2, J = yi / 2;
        const int I2 = iclamp(I + ((xi & 1) == 1 ? 1 : -1), 0, nxc - 1);
        const int J2 = iclamp(J + ((yi & 1) == 1 ? 1 : -1), 0, nyc - 1);
        cv[q][sub][0] = ldp<CH>(c.p, a_.st.c.p, J * nxc + I, isig);
        cv[q][sub][1] = ldp<CH>(c.p, a_.st.c.p, J * nxc + I2, isig);
        cv[q][sub][2] = ldp<CH>(c.p, a_.st.c.p, J2 * nxc + I, isig);
        cv[q][sub][3] = ldp<CH>(c.p, a_.st.c.p, J2 * nxc + I2, isig);
        pv[q][sub] = ldp<FH>(f.q, a_.st.f.q, yi * nx + xi, isig);
      }
      const bool red0 = ((x0 - H + rx + y) & 1) == 0;
      cR[q] = red0 ? cc[0] : cc[1];
      cB[q] = red0 ? cc[1] : cc[0];
    }
#pragma unroll
    for (int q = 0; q < PQ; q++) {
      const int pi = tid + kThreads * q;
      if (pi < NP) {
        const int ry = pi / (RX / 2), rx = 2 * (pi % (RX / 2));
        const bool red0 = ((x0 - H + rx + y0 - H + ry) & 1) == 0;
#pragma unroll
        for (int sub = 0; sub < 2; sub++) {
          const LegCell& a = (sub == 0) == red0 ? cR[q] : cB[q];
          const float* cw = cv[q][sub];
          const float ccr = K(0.5625) * cw[0] + K(0.1875) * (cw[1] + cw[2]) + K(0.0625) * cw[3];
          sp[(ry + 1) * PX + rx + 1 + sub] = (a.ci & LEG_IN) ? pv[q][sub] + (a.d > K(0) ? ccr : K(0)) : 0.f;
        }
      } else {
        cR[q] = cB[q] = leg_none(PX);
      }
    }
    Pre pre;
    if constexpr (PROJ) pre_load(a_, x0, y0, pre);
    lds_sync();
    for (int s = 0; s < NU; s++) {  // black then red (js/fluid.js:307-310)
      leg_sweep<PX, SG>(cB, sp, leg_dummy(sp, PX * PY));
      lds_sync();
      leg_sweep<PX, SG>(cR, sp, leg_dummy(sp, PX * PY));
      lds_sync();
    }
    auto finish = [&](const LegCell& a) {
      const int ci = a.ci & LEG_IDX;
      const int rx = ci % PX - 1, ry = ci / PX - 1;
      if ((a.ci & LEG_IN) && rx >= H && rx < H + TX && ry >= H && ry < H + TY)
        stp<FH>(f.p, a_.st.f.p, (y0 - H + ry) * nx + (x0 - H + rx), sp[ci], sig);
    };
#pragma unroll
    for (int q = 0; q < PQ; q++) {
      finish(cR[q]);
      finish(cB[q]);
    }
    if constexpr (PROJ) project(a_, sp, x0, y0, pre);
  }
  // The faces of the tile: U faces x0 .. x0 + TX - 1 (and nx on the domain's
  // right edge), V faces y0 .. y0 + TY - 1 (and ny on its top).  Their
  // velocity and open fraction are loaded before the sweeps (they do not
  // depend on p), so those loads overlap the leg; the temperatures follow
  // the sweeps (holding them too would spill registers).
  static constexpr int QU = ((TX + 1) * TY + kThreads - 1) / kThreads, QV = (TX * (TY + 1) + kThreads - 1) / kThreads;
  struct Pre {
    float uf[QU], uu[QU];
    float vf[QV], vv[QV];
  };
  static __device__ __forceinline__ bool uface(int k, int x0, int y0, int nx, int ny, int& i, int& j) {
    const int a = k % (TX + 1), b = k / (TX + 1);
    i = x0 + a;
    j = y0 + b;
    return k < (TX + 1) * TY && j < ny && i <= nx && (a < TX || i == nx);
  }
  static __device__ __forceinline__ bool vface(int k, int x0, int y0, int nx, int ny, int& i, int& j) {
    const int a = k % TX, b = k / TX;
    i = x0 + a;
    j = y0 + b;
    return k < TX * (TY + 1) && i < nx && j <= ny && (b < TY || j == ny);
  }
  // (a slot without a face loads face 0's data and discards it)
  static __device__ __forceinline__ void pre_load(const Args& a_, int x0, int y0, Pre& r) {
    const Fields& F = a_.F;
    const int nx = a_.f.nx, ny = a_.f.ny;
#pragma unroll
    for (int q = 0; q < QU; q++) {
      int i, j;
      const bool ok = uface(threadIdx.x + kThreads * q, x0, y0, nx, ny, i, j);
      const int k = ok ? j * (nx + 1) + i : 0;
      r.uf[q] = F.fx[k];
      r.uu[q] = F.U[k];
    }
#pragma unroll
    for (int q = 0; q < QV; q++) {
      int i, j;
      const bool ok = vface(threadIdx.x + kThreads * q, x0, y0, nx, ny, i, j);
      const int k = ok ? j * nx + i : 0;
      r.vf[q] = F.fy[k];
      r.vv[q] = F.V[k];
    }
  }
  // k_project's update of each face (shaders/pressure.wgsl's project), the
  // pressure from the region (exact over the tile plus one ring)
  static __device__ __forceinline__ void project(const Args& a_, const float* sp, int x0, int y0, const Pre& r) {
    const Fields& F = a_.F;
    const cf_params P = load_params(F.P);
    const int nx = P.nx, ny = P.ny;
    const float dt = P.dt, dx = P.dx;
    auto cell = [&](int i, int j) { return (j - y0 + H + 1) * PX + (i - x0 + H + 1); };
#pragma unroll
    for (int q = 0; q < QU; q++) {
      int i, j;
      if (!uface(threadIdx.x + kThreads * q, x0, y0, nx, ny, i, j)) continue;
      const int kf = j * (nx + 1) + i;
      if (r.uf[q] > K(0)) {
        const int ci = cell(i, j);
        const float pl = i > 0 ? sp[ci - 1] : -sp[ci];
        const float pr = i < nx ? sp[ci] : -sp[ci - 1];
        const float dp = pr - pl;
        const float Tf = K(0.5) * (F.T[j * nx + iclamp(i - 1, 0, nx - 1)] + F.T[j * nx + min(i, nx - 1)]);
        F.U[kf] = r.uu[q] - dt * invRho(P, Tf) * dp / dx;
      } else F.U[kf] = K(0);
    }
#pragma unroll
    for (int q = 0; q < QV; q++) {
      int i, j;
      if (!vface(threadIdx.x + kThreads * q, x0, y0, nx, ny, i, j)) continue;
      const int kf = j * nx + i;
      if (j > 0 && r.vf[q] > K(0)) {
        const int ci = cell(i, j);
        const float pb = sp[ci - PX];
        const float pt = j < ny ? sp[ci] : -sp[ci - PX];
        const float dp = pt - pb;
        const float Tf = K(0.5) * (F.T[(j - 1) * nx + i] + F.T[min(j, ny - 1) * nx + i]);
        F.V[kf] = r.vv[q] - dt * invRho(P, Tf) * dp / dx;
      } else F.V[kf] = K(0);
    }
  }
};
template <int TX, int TY, int NU, int M>
__global__ __launch_bounds__((MgUpOp<TX, TY, NU, M>::kThreads), (MgUpOp<TX, TY, NU, M>::kMinWaves)) void k_mg_up(const typename MgUpOp<TX, TY, NU, M>::Args a_) {
  MgUpOp<TX, TY, NU, M>::run(a_, xcd_tile(blockIdx.x, gridDim.x));
}
template <int TX, int TY, int NU>
__global__ __launch_bounds__((MgUpOp<TX, TY, NU, 0, true>::kThreads), (MgUpOp<TX, TY, NU, 0, true>::kMinWaves)) void k_mg_up_project(
    const typename MgUpOp<TX, TY, NU, 0, true>::Args a_) {
  MgUpOp<TX, TY, NU, 0, true>::run(a_, xcd_tile(blockIdx.x, gridDim.x));
}

// ---- two levels per launch -------------------------------------------------------
// The legs of a level f and of the next coarser one c in one workgroup, for a
// tile of TC x TC cells of c (2TC x 2TC of f): c's rhs, its start from p = 0
// and its sweeps never leave LDS, and a V-cycle has one launch less per pair
// in each direction.  The halos compose level by level:
//   down: c's leg needs the tile plus 2NU+1 coarse cells (the lone leg's halo),
//         whose rhs is the restriction of the fine residual under them.  The
//         fine leg's region is those fine cells plus 2NU-1: the residual of
//         their two outer fine rings is wrong, but it reaches only the coarse
//         region's outer ring, whose cells the coarse leg never updates (that
//         ring never lies on the domain edge: the tile origin is even, the
//         halo odd and the level's size even).
//   up:   f's leg needs its tile plus 2NU fine cells, whose prolongation reads
//         coarse cells up to NU+1 beyond the coarse tile, so c's region is the
//         tile plus 3NU+1 coarse cells.
// Every update, residual, restriction and prolongation is the lone legs'
// arithmetic in their order, so the results are theirs to the last bit.  FP32
// storage only.

// the ring of zeros around an RX x RY region stored with row pitch PX = RX + 2
__device__ __forceinline__ void leg_ring(float* sp, int PX, int PY, int RY, int nt) {
  for (int k = threadIdx.x; k < 2 * PX + 2 * RY; k += nt) {
    const int idx = k < PX ? k : (k < 2 * PX ? (PY - 1) * PX + (k - PX) : ((k - 2 * PX) / 2 + 1) * PX +
                                                                             ((k & 1) ? PX - 1 : 0));
    sp[idx] = 0.f;
  }
}

// Levels f, c = f + 1, n = f + 2 (f >= 1): the down legs of f and c and the
// restriction into n.  Writes what the lone legs leave for the up legs: f.q,
// c.q and c.rhs over the tile, n.rhs and n.p = 0.  c.p is not written: no leg
// reads it before the up leg of c overwrites it.  f.p is not read either: it
// holds the zero that the restriction into f wrote.
template <int TC, int NU, int NT>
struct MgDown2Op {
  struct Args {
    Level f, c, n;
    const int* skip;
    int vg;  // tiles of this scene's launch
  };
  static constexpr int kThreads = NT;
  static constexpr int HC = 2 * NU + 1, HF = 2 * NU - 1;
  static constexpr int RC = TC + 2 * HC, RR = 2 * RC, RF = RR + 2 * HF;
  static constexpr int PC = RC + 2, PF = RF + 2;
  static constexpr int NPF = RF * RF / 2, NPC = RC * RC / 2;
  static constexpr int QF = (NPF + NT - 1) / NT, QC = (NPC + NT - 1) / NT;
  static_assert(RC % 2 == 0 && RF % 2 == 0 && TC % 2 == 0, "pairs need even region widths");
  static __device__ __forceinline__ void run(const Args& a_, int vb) {
    const Level& f = a_.f;
    const Level& c = a_.c;
    const Level& n = a_.n;
    if (a_.skip && *a_.skip) return;  // the step's residual target is met
    __shared__ float spf[PF * PF + LEG_DUMMY];
    __shared__ float sres[RR * RR];   // the fine rhs under the coarse region, then its residual
    __shared__ float spc[PC * PC + LEG_DUMMY];
    __shared__ float src[RC * RC];    // the coarse rhs over the coarse region
    __shared__ float sresc[TC * TC];  // the coarse residual over the tile
    const int tid = threadIdx.x;
    const int nx = f.nx, nxc = c.nx, nyc = c.ny;
    const int gx = (nxc + TC - 1) / TC;
    const int X0 = (vb % gx) * TC, Y0 = (vb / gx) * TC;  // the coarse tile
    const int xc0 = X0 - HC, yc0 = Y0 - HC;              // the coarse region
    const int xf0 = 2 * xc0 - HF, yf0 = 2 * yc0 - HF;    // the fine region
    leg_ring(spf, PF, PF, RF, NT);
    leg_ring(spc, PC, PC, RC, NT);
    // every global load before the first sweep: the fine cells, rhs and p,
    // the coarse cells' coefficients
    // (branch-free: a slot without a pair loads pair 0's data and discards it)
    LegCell fR[QF], fB[QF];
    float pv[QF][2], rv[QF][2];
#pragma unroll
    for (int q = 0; q < QF; q++) {
      const int pi = tid + NT * q;
      const int pc = pi < NPF ? pi : 0;
      const int ry = pc / (RF / 2), rx = 2 * (pc % (RF / 2));
      const int x = xf0 + rx, y = yf0 + ry;
      const LegCell c0 = leg_cell(f.nx, f.ny, f.ax, f.ay, f.rhs, f.h2, x, y, rx, ry, RF, RF, PF, rv[q][0]);
      const LegCell c1 = leg_cell(f.nx, f.ny, f.ax, f.ay, f.rhs, f.h2, x + 1, y, rx + 1, ry, RF, RF, PF, rv[q][1]);
      // f is a level above 0, whose V-cycle starts from the p = 0 that the
      // restriction into it wrote (shaders/pressure.wgsl:104-113): not read
      pv[q][0] = pv[q][1] = 0.f;
      const bool red0 = ((x + y) & 1) == 0;
      fR[q] = red0 ? c0 : c1;
      fB[q] = red0 ? c1 : c0;
    }
    LegCell cR[QC], cB[QC];
#pragma unroll
    for (int q = 0; q < QC; q++) {
      const int pi = tid + NT * q;
      const int pc = pi < NPC ? pi : 0;
      const int ry = pc / (RC / 2), rx = 2 * (pc % (RC / 2));
      const int X = xc0 + rx, Y = yc0 + ry;
      const LegCell c0 = leg_cell_coef(nxc, nyc, c.ax, c.ay, X, Y, rx, ry, RC, RC, PC);
      const LegCell c1 = leg_cell_coef(nxc, nyc, c.ax, c.ay, X + 1, Y, rx + 1, ry, RC, RC, PC);
      const bool red0 = ((X + Y) & 1) == 0;
      cR[q] = red0 ? c0 : c1;
      cB[q] = red0 ? c1 : c0;
    }
#pragma unroll
    for (int q = 0; q < QF; q++) {
      const int pi = tid + NT * q;
      if (pi < NPF) {
        const int ry = pi / (RF / 2), rx = 2 * (pi % (RF / 2));
        spf[(ry + 1) * PF + rx + 1] = pv[q][0];
        spf[(ry + 1) * PF + rx + 2] = pv[q][1];
        if (ry >= HF && ry < HF + RR) {
          if (rx >= HF && rx < HF + RR) sres[(ry - HF) * RR + (rx - HF)] = rv[q][0];
          if (rx + 1 >= HF && rx + 1 < HF + RR) sres[(ry - HF) * RR + (rx + 1 - HF)] = rv[q][1];
        }
      } else {
        fR[q] = fB[q] = leg_none(PF);
      }
    }
#pragma unroll
    for (int q = 0; q < QC; q++) {
      const int pi = tid + NT * q;
      if (pi < NPC) {
        const int ry = pi / (RC / 2), rx = 2 * (pi % (RC / 2));
        spc[(ry + 1) * PC + rx + 1] = 0.f;  // the coarse leg starts from p = 0 (shaders/pressure.wgsl:104-113)
        spc[(ry + 1) * PC + rx + 2] = 0.f;
      } else {
        cR[q] = cB[q] = leg_none(PC);
      }
    }
    __syncthreads();
    for (int s = 0; s < NU; s++) {  // red then black (js/fluid.js:298-301)
      leg_sweep<PF>(fR, spf, leg_dummy(spf, PF * PF));
      __syncthreads();
      leg_sweep<PF>(fB, spf, leg_dummy(spf, PF * PF));
      __syncthreads();
    }
    // the fine residual under the coarse region (:96-101); f's leg buffer over the tile
    auto finish_f = [&](const LegCell& a) {
      if (!(a.ci & LEG_IN)) return;
      const int ci = a.ci & LEG_IDX;
      const int rx = ci % PF - 1, ry = ci / PF - 1;
      const float pk = spf[ci];
      if (rx >= HF && rx < HF + RR && ry >= HF && ry < HF + RR) {
        const float off = a.w * spf[ci - 1] + a.e * spf[ci + 1] + a.s * spf[ci - PF] + a.n * spf[ci + PF];
        const int r = (ry - HF) * RR + (rx - HF);
        sres[r] = a.d <= K(0) ? K(0) : sres[r] - (off - a.d * pk) / f.h2;
      }
      constexpr int T0 = HF + 2 * HC;  // the fine tile's first region row and column
      if (rx >= T0 && rx < T0 + 2 * TC && ry >= T0 && ry < T0 + 2 * TC) f.q[(yf0 + ry) * nx + (xf0 + rx)] = pk;
    };
#pragma unroll
    for (int q = 0; q < QF; q++) {
      finish_f(fR[q]);
      finish_f(fB[q]);
    }
    __syncthreads();
    // restriction to the coarse rhs over the coarse region (:104-113); the tile's goes out for c's up leg
    for (int k = tid; k < RC * RC; k += NT) {
      const int li = k % RC, lj = k / RC;
      const int X = xc0 + li, Y = yc0 + lj;
      float v = 0.f;
      if (X >= 0 && X < nxc && Y >= 0 && Y < nyc) {
        const float* r0 = sres + (2 * lj) * RR + 2 * li;
        const float r = r0[0] + r0[1] + r0[RR] + r0[RR + 1];
        v = K(0.25) * r;
        if (li >= HC && li < HC + TC && lj >= HC && lj < HC + TC) c.rhs[Y * nxc + X] = v;
      }
      src[k] = v;
    }
    __syncthreads();
    auto set_rh = [&](LegCell& a) {  // rhs * h2, as leg_cell forms it
      if (a.ci & LEG_IN) {
        const int ci = a.ci & LEG_IDX;
        a.rh = src[(ci / PC - 1) * RC + (ci % PC - 1)] * c.h2;
      }
    };
#pragma unroll
    for (int q = 0; q < QC; q++) {
      set_rh(cR[q]);
      set_rh(cB[q]);
    }
    for (int s = 0; s < NU; s++) {
      leg_sweep<PC>(cR, spc, leg_dummy(spc, PC * PC));
      __syncthreads();
      leg_sweep<PC>(cB, spc, leg_dummy(spc, PC * PC));
      __syncthreads();
    }
    // the coarse residual over the tile; c's leg buffer
    auto finish_c = [&](const LegCell& a) {
      const int ci = a.ci & LEG_IDX;
      const int rx = ci % PC - 1, ry = ci / PC - 1;
      if ((a.ci & LEG_IN) && rx >= HC && rx < HC + TC && ry >= HC && ry < HC + TC) {
        const float pk = spc[ci];
        const float off = a.w * spc[ci - 1] + a.e * spc[ci + 1] + a.s * spc[ci - PC] + a.n * spc[ci + PC];
        const float rh = src[ry * RC + rx];
        sresc[(ry - HC) * TC + (rx - HC)] = a.d <= K(0) ? K(0) : rh - (off - a.d * pk) / c.h2;
        c.q[(yc0 + ry) * nxc + (xc0 + rx)] = pk;
      }
    };
#pragma unroll
    for (int q = 0; q < QC; q++) {
      finish_c(cR[q]);
      finish_c(cB[q]);
    }
    __syncthreads();
    // restriction to n's rhs, n's p = 0
    for (int k = tid; k < (TC / 2) * (TC / 2); k += NT) {
      const int li = k % (TC / 2), lj = k / (TC / 2);
      const int I = X0 / 2 + li, J = Y0 / 2 + lj;
      if (I < n.nx && J < n.ny) {
        const float* r0 = sresc + (2 * lj) * TC + 2 * li;
        const float r = r0[0] + r0[1] + r0[TC] + r0[TC + 1];
        n.rhs[J * n.nx + I] = K(0.25) * r;
        n.p[J * n.nx + I] = K(0);
      }
    }
  }
};
template <int TC, int NU, int NT>
__global__ __launch_bounds__(NT) void k_mg_down2(const typename MgDown2Op<TC, NU, NT>::Args a_) {
  MgDown2Op<TC, NU, NT>::run(a_, xcd_tile(blockIdx.x, gridDim.x));
}

// Levels f, c = f + 1, n = f + 2: the up legs of c and f, after n's.  Reads
// f.q, c.q, the rhs of both and n.p; writes c.p over the tile and f.p over the
// fine tile.
template <int TC, int NU, int NT>
struct MgUp2Op {
  struct Args {
    Level f, c, n;
    const int* skip;
    int vg;  // tiles of this scene's launch
  };
  static constexpr int kThreads = NT;
  static constexpr int HF = 2 * NU, HC = 3 * NU + 1;
  static constexpr int RF = 2 * TC + 2 * HF, RC = TC + 2 * HC;
  static constexpr int PF = RF + 2, PC = RC + 2;
  static constexpr int NPF = RF * RF / 2, NPC = RC * RC / 2;
  static constexpr int QF = (NPF + NT - 1) / NT, QC = (NPC + NT - 1) / NT;
  static_assert(RC % 2 == 0 && RF % 2 == 0, "pairs need even region widths");
  static __device__ __forceinline__ void run(const Args& a_, int vb) {
    const Level& f = a_.f;
    const Level& c = a_.c;
    const Level& n = a_.n;
    if (a_.skip && *a_.skip) return;
    __shared__ float spf[PF * PF + LEG_DUMMY];
    __shared__ float spc[PC * PC + LEG_DUMMY];
    const int tid = threadIdx.x;
    const int nx = f.nx, nxc = c.nx, nyc = c.ny, nxn = n.nx, nyn = n.ny;
    const int gx = (nxc + TC - 1) / TC;
    const int X0 = (vb % gx) * TC, Y0 = (vb / gx) * TC;  // the coarse tile
    const int xc0 = X0 - HC, yc0 = Y0 - HC;              // the coarse region
    const int xf0 = 2 * X0 - HF, yf0 = 2 * Y0 - HF;      // the fine region
    leg_ring(spf, PF, PF, RF, NT);
    leg_ring(spc, PC, PC, RC, NT);
    // the coarse cells: coefficients, rhs, start value q + the bilinear
    // correction from n, skipped in closed cells (shaders/pressure.wgsl:116-129)
    // (branch-free: a slot without a pair, or a cell outside the domain,
    // loads cell 0's data and discards it)
    LegCell cR[QC], cB[QC];
    float pv[QC][2];
#pragma unroll
    for (int q = 0; q < QC; q++) {
      const int pi = tid + NT * q;
      const int pc = pi < NPC ? pi : 0;
      const int ry = pc / (RC / 2), rx = 2 * (pc % (RC / 2));
      const int Y = yc0 + ry;
      LegCell cc[2];
#pragma unroll
      for (int sub = 0; sub < 2; sub++) {
        const int X = xc0 + rx + sub;
        float rr;
        cc[sub] = leg_cell(nxc, nyc, c.ax, c.ay, c.rhs, c.h2, X, Y, rx + sub, ry, RC, RC, PC, rr);
        const bool in = (cc[sub].ci & LEG_IN) != 0;
        const int Xi = in ? X : 0, Yi = in ? Y : 0;
        const int I = Xi / 2, J = Yi / 2;
        const int I2 = iclamp(I + ((Xi & 1) == 1 ? 1 : -1), 0, nxn - 1);
        const int J2 = iclamp(J + ((Yi & 1) == 1 ? 1 : -1), 0, nyn - 1);
        const float c00 = n.p[J * nxn + I], c10 = n.p[J * nxn + I2];
        const float c01 = n.p[J2 * nxn + I], c11 = n.p[J2 * nxn + I2];
        const float ccr = K(0.5625) * c00 + K(0.1875) * (c10 + c01) + K(0.0625) * c11;
        const float qc = c.q[Yi * nxc + Xi];
        pv[q][sub] = in ? qc + (cc[sub].d > K(0) ? ccr : K(0)) : 0.f;
      }
      const bool red0 = ((xc0 + rx + Y) & 1) == 0;
      cR[q] = red0 ? cc[0] : cc[1];
      cB[q] = red0 ? cc[1] : cc[0];
    }
    // the fine cells: coefficients, rhs and q (their correction waits for c's sweeps)
    LegCell fR[QF], fB[QF];
    float qR[QF], qB[QF];
#pragma unroll
    for (int q = 0; q < QF; q++) {
      const int pi = tid + NT * q;
      const int pc = pi < NPF ? pi : 0;
      const int ry = pc / (RF / 2), rx = 2 * (pc % (RF / 2));
      const int y = yf0 + ry;
      LegCell cc[2];
      float qv[2];
#pragma unroll
      for (int sub = 0; sub < 2; sub++) {
        const int x = xf0 + rx + sub;
        float rr;
        cc[sub] = leg_cell(f.nx, f.ny, f.ax, f.ay, f.rhs, f.h2, x, y, rx + sub, ry, RF, RF, PF, rr);
        const bool in = (cc[sub].ci & LEG_IN) != 0;
        const float v = f.q[leg_idx(in, x, y, nx)];
        qv[sub] = in ? v : 0.f;
      }
      const bool red0 = ((xf0 + rx + y) & 1) == 0;
      fR[q] = red0 ? cc[0] : cc[1];
      fB[q] = red0 ? cc[1] : cc[0];
      qR[q] = red0 ? qv[0] : qv[1];
      qB[q] = red0 ? qv[1] : qv[0];
    }
#pragma unroll
    for (int q = 0; q < QC; q++) {
      const int pi = tid + NT * q;
      if (pi < NPC) {
        const int ry = pi / (RC / 2), rx = 2 * (pi % (RC / 2));
        spc[(ry + 1) * PC + rx + 1] = pv[q][0];
        spc[(ry + 1) * PC + rx + 2] = pv[q][1];
      } else {
        cR[q] = cB[q] = leg_none(PC);
      }
    }
#pragma unroll
    for (int q = 0; q < QF; q++) {
      if (tid + NT * q >= NPF) {
        fR[q] = fB[q] = leg_none(PF);
        qR[q] = qB[q] = 0.f;
      }
    }
    __syncthreads();
    for (int s = 0; s < NU; s++) {  // black then red (js/fluid.js:307-310)
      leg_sweep<PC>(cB, spc, leg_dummy(spc, PC * PC));
      __syncthreads();
      leg_sweep<PC>(cR, spc, leg_dummy(spc, PC * PC));
      __syncthreads();
    }
    // c.p over the tile
    auto out_c = [&](const LegCell& a) {
      const int ci = a.ci & LEG_IDX;
      const int rx = ci % PC - 1, ry = ci / PC - 1;
      if ((a.ci & LEG_IN) && rx >= HC && rx < HC + TC && ry >= HC && ry < HC + TC)
        c.p[(yc0 + ry) * nxc + (xc0 + rx)] = spc[ci];
    };
#pragma unroll
    for (int q = 0; q < QC; q++) {
      out_c(cR[q]);
      out_c(cB[q]);
    }
    // the fine start values: q + the bilinear correction from c's region
    auto start = [&](const LegCell& a, float qv) {
      const int ci = a.ci & LEG_IDX;
      float v = 0.f;
      if (a.ci & LEG_IN) {
        const int x = xf0 + (ci % PF - 1), y = yf0 + (ci / PF - 1);
        const int I = x / 2, J = y / 2;
        const int I2 = iclamp(I + ((x & 1) == 1 ? 1 : -1), 0, nxc - 1);
        const int J2 = iclamp(J + ((y & 1) == 1 ? 1 : -1), 0, nyc - 1);
        const int r0 = (J - yc0 + 1) * PC - xc0 + 1, r2 = (J2 - yc0 + 1) * PC - xc0 + 1;
        const float c00 = spc[r0 + I], c10 = spc[r0 + I2], c01 = spc[r2 + I], c11 = spc[r2 + I2];
        const float ccr = K(0.5625) * c00 + K(0.1875) * (c10 + c01) + K(0.0625) * c11;
        v = qv + (a.d > K(0) ? ccr : K(0));
      }
      spf[ci] = v;
    };
#pragma unroll
    for (int q = 0; q < QF; q++) {
      if (tid + NT * q < NPF) {  // a slot without a cell must not write the first region cell
        start(fR[q], qR[q]);
        start(fB[q], qB[q]);
      }
    }
    __syncthreads();
    for (int s = 0; s < NU; s++) {
      leg_sweep<PF>(fB, spf, leg_dummy(spf, PF * PF));
      __syncthreads();
      leg_sweep<PF>(fR, spf, leg_dummy(spf, PF * PF));
      __syncthreads();
    }
    auto out_f = [&](const LegCell& a) {
      const int ci = a.ci & LEG_IDX;
      const int rx = ci % PF - 1, ry = ci / PF - 1;
      if ((a.ci & LEG_IN) && rx >= HF && rx < HF + 2 * TC && ry >= HF && ry < HF + 2 * TC)
        f.p[(yf0 + ry) * nx + (xf0 + rx)] = spf[ci];
    };
#pragma unroll
    for (int q = 0; q < QF; q++) {
      out_f(fR[q]);
      out_f(fB[q]);
    }
  }
};
template <int TC, int NU, int NT>
__global__ __launch_bounds__(NT) void k_mg_up2(const typename MgUp2Op<TC, NU, NT>::Args a_) {
  MgUp2Op<TC, NU, NT>::run(a_, xcd_tile(blockIdx.x, gridDim.x));
}

// ---- bottom levels in one workgroup ------------------------------------------------
// The last one to three levels of the hierarchy hold <= 4096, <= 1024 and
// <= 256 cells (each level above the coarsest has more than 256 cells and
// four times the cells of the next).  One workgroup of 1024 threads keeps p
// and rhs of all of them in LDS and each thread's cells' coefficients in
// registers: two horizontal pairs of the first level, one pair of the second
// (threads below 512), each pair one red and one black cell, sorted by colour
// at load, so every lane updates cells in every half-sweep; the coarsest
// level's two pairs per lane of the first wave.  Neighbours missing at the
// domain edge are read at the cell itself with a zero coefficient.  The
// coarsest level's face coefficients are restricted here from the level above
// when the caller has not produced them.
constexpr int BT = 1024;

struct Bottom {
  Level L[3];       // finest first; the last is the coarsest level
  const int* skip;  // may be null: nothing to do once *skip is non-zero
  double* band;     // coarse_iters = 0: scratch of the direct solve
  int n;            // 1..3
  int coarse_coef;  // 1: restrict the coarsest level's ax/ay from L[n-2] here
  int nu, coarse_iters;  // coarse_iters = 0: direct FP64 solve of the coarsest level
  int zero_p;            // the first level is above level 0: its p is the zero the restriction wrote, not read
};

template <int CPT>
struct CellSet {
  int k[CPT];   // linear cell index (or -1)
  int nb[CPT];  // bits 0..3: west, east, south, north neighbour exists; bit 4: colour
  float w[CPT], e[CPT], s[CPT], n[CPT], diag[CPT];
  float y[CPT];  // 1 / diag (0 for a closed cell)
};

// cell data of a level for the cells kidx[q] (ignored when outside the level);
// coefficient arrays may be in global memory (ax, ay) or in LDS.  Every
// load is issued before any arithmetic (a slot without a cell loads cell
// 0's coefficients and discards them), so they are in flight together.
template <int CPT>
__device__ __forceinline__ void cells_load(CellSet<CPT>& C, int nx, int ny, const float* ax, const float* ay,
                                           const int (&kidx)[CPT]) {
  float w[CPT], e[CPT], s[CPT], n[CPT];
  bool has[CPT];
  int ii[CPT], jj[CPT];
#pragma unroll
  for (int q = 0; q < CPT; q++) {
    const int k = kidx[q];
    has[q] = k >= 0 && k < nx * ny;
    const int kk = has[q] ? k : 0;
    const int i = kk % nx, j = kk / nx;
    ii[q] = i; jj[q] = j;
    w[q] = ax[j * (nx + 1) + i]; e[q] = ax[j * (nx + 1) + i + 1];
    s[q] = ay[j * nx + i]; n[q] = ay[(j + 1) * nx + i];
  }
#pragma unroll
  for (int q = 0; q < CPT; q++) {
    const int i = ii[q], j = jj[q];
    float diag = 0.f;  // the diagonal of shaders/pressure.wgsl:77-85, same accumulation order
    if (i > 0) diag += w[q]; else diag += K(2.0) * w[q];
    if (i < nx - 1) diag += e[q]; else diag += K(2.0) * e[q];
    if (j > 0) diag += s[q];
    if (j < ny - 1) diag += n[q]; else diag += K(2.0) * n[q];
    C.k[q] = has[q] ? kidx[q] : -1;
    C.diag[q] = has[q] ? diag : K(0);
    C.y[q] = has[q] && diag > K(0) ? K(1) / diag : K(0);
    C.w[q] = has[q] && i > 0 ? w[q] : 0.f;
    C.e[q] = has[q] && i < nx - 1 ? e[q] : 0.f;
    C.s[q] = has[q] && j > 0 ? s[q] : 0.f;
    C.n[q] = has[q] && j < ny - 1 ? n[q] : 0.f;
    C.nb[q] = has[q] ? ((i > 0 ? 1 : 0) | (i < nx - 1 ? 2 : 0) | (j > 0 ? 4 : 0) | (j < ny - 1 ? 8 : 0) |
                        (((i + j) & 1) << 4))
                     : 0;
  }
}

template <int CPT>
__device__ __forceinline__ float cells_off(const CellSet<CPT>& C, int q, const float* p, int nx) {
  const int k = C.k[q], b = C.nb[q];
  return C.w[q] * p[k - (b & 1)] + C.e[q] * p[k + ((b >> 1) & 1)] + C.s[q] * p[k - nx * ((b >> 2) & 1)] +
         C.n[q] * p[k + nx * ((b >> 3) & 1)];
}

// One cell of the coarsest level's sweeps with everything in registers: the
// LDS indices of its four neighbours (the cell itself where one is missing,
// with a zero coefficient), the coefficients, the diagonal, its reciprocal
// and rhs * h2.  A lane slot without a cell reads and writes index `dummy`
// (LDS that the sweeps do not otherwise use) with zero coefficients, so the
// update needs no branch.
struct SweepCell {
  int k, iw, ie, is, in;
  float w, e, s, n, d, y, r;
};

template <int CPT>
__device__ __forceinline__ SweepCell sweep_cell(const CellSet<CPT>& C, const float (&rh)[CPT], int q, int nx,
                                                int dummy) {
  SweepCell c;
  const int k = C.k[q], b = C.nb[q];
  const bool has = k >= 0;
  c.k = has ? k : dummy;
  c.iw = has ? k - (b & 1) : dummy;
  c.ie = has ? k + ((b >> 1) & 1) : dummy;
  c.is = has ? k - nx * ((b >> 2) & 1) : dummy;
  c.in = has ? k + nx * ((b >> 3) & 1) : dummy;
  c.w = C.w[q]; c.e = C.e[q]; c.s = C.s[q]; c.n = C.n[q];  // zero without a cell
  c.d = C.diag[q]; c.y = C.y[q]; c.r = has ? rh[q] : K(0);
  return c;
}

__device__ __forceinline__ SweepCell sweep_sel(bool a, const SweepCell& x, const SweepCell& z) {
  SweepCell c;
  c.k = a ? x.k : z.k; c.iw = a ? x.iw : z.iw; c.ie = a ? x.ie : z.ie; c.is = a ? x.is : z.is; c.in = a ? x.in : z.in;
  c.w = a ? x.w : z.w; c.e = a ? x.e : z.e; c.s = a ? x.s : z.s; c.n = a ? x.n : z.n;
  c.d = a ? x.d : z.d; c.y = a ? x.y : z.y; c.r = a ? x.r : z.r;
  return c;
}

// the red-black update of shaders/pressure.wgsl:87-94 (terms in its order)
// for two cells of one colour, which never neighbour each other: all eight
// reads are issued before either write.  A closed cell (d = 0, y = 0) gives
// a finite quotient that the select replaces by 0.
__device__ __forceinline__ void sweep_relax2(const SweepCell& a, const SweepCell& b, float* p) {
  const float a0 = p[a.iw], a1 = p[a.ie], a2 = p[a.is], a3 = p[a.in];
  const float b0 = p[b.iw], b1 = p[b.ie], b2 = p[b.is], b3 = p[b.in];
  const float offa = a.w * a0 + a.e * a1 + a.s * a2 + a.n * a3;
  const float offb = b.w * b0 + b.e * b1 + b.s * b2 + b.n * b3;
  const float va = div_rcp_fast(offa - a.r, a.d, a.y), vb = div_rcp_fast(offb - b.r, b.d, b.y);
  p[a.k] = a.d > K(0) ? va : K(0);
  p[b.k] = b.d > K(0) ? vb : K(0);
}

// ---- the coarsest sweeps of a 16 x 16 level in registers ----
// Lane (x, q) = (t mod 16, t / 16) of the first wave holds the cells
// (x, 4q .. 4q + 3), sorted by colour: R0 / R1 its red cells in rows
// 4q + {0, 1} / 4q + {2, 3}, K0 / K1 its black ones.  A cell's west and
// east neighbours are the same row in lanes x -+ 1, moved by a row shift
// inside the wave's 16-lane rows (the edge lane keeps its own value, which the
// zero coefficient multiplies, as the LDS form reads the cell itself); its
// vertical neighbours are in the lane's other registers or, past the ends of
// its four rows, in lane x of the next row of lanes below or above.  The
// arithmetic is sweep_relax2's, term by term, so p is bitwise the same.
struct Cell16 {
  float w, e, s, n, d, y, r;
};

__device__ __forceinline__ float relax16(const Cell16& c, float pw, float pe, float ps, float pn) {
  const float off = c.w * pw + c.e * pe + c.s * ps + c.n * pn;
  const float v = div_rcp_fast(off - c.r, c.d, c.y);
  return c.d > K(0) ? v : K(0);
}

// v of lane (x, q - 1) (below) or (x, q + 1) (above).  The 16-lane swap
// gives rows [v0 v0 v2 v2] and [v1 v1 v3 v3], the 32-lane swap of those
// [v0 v0 v1 v1] and [v2 v2 v3 v3]; each row of lanes takes its neighbour
// row from one of them (rows 0 below and 3 above get a value they ignore)
__device__ __forceinline__ float row_from(float v, int q, bool below) {
  const unsigned a = __float_as_uint(v);
  const auto p = __builtin_amdgcn_permlane16_swap(a, a, false, false);
  const auto r = __builtin_amdgcn_permlane32_swap(p[0], p[1], false, false);
  return __uint_as_float(below ? (q == 2 ? r[0] : p[0]) : (q == 1 ? r[1] : p[1]));
}
// v of lane x - 1 (west) / x + 1 (east) of the same 16-lane row; the lane at
// the row's end gets `self`
__device__ __forceinline__ float lane_west(float self, float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(self), __float_as_int(v), 0x111, 0xf, 0xf, false));
}
__device__ __forceinline__ float lane_east(float self, float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(self), __float_as_int(v), 0x101, 0xf, 0xf, false));
}

// `iters` red-black sweeps from zero of the 16 x 16 level whose face
// coefficients are axc, ayc and rhs sr2 (LDS); p to sp2.  The whole first
// wave calls it.
__device__ __forceinline__ void coarse_sweeps16(const float* axc, const float* ayc, const float* sr2, float* sp2,
                                             float h2, int iters, int t) {
  const int x = t & 15, q = t >> 4;
  const bool ev = (x & 1) == 0;  // red cells in rows 4q and 4q + 2 (else 4q + 1, 4q + 3)
  const int j0 = 4 * q;
  const int kk[4] = {(j0 + (ev ? 0 : 1)) * 16 + x, (j0 + (ev ? 2 : 3)) * 16 + x, (j0 + (ev ? 1 : 0)) * 16 + x,
                     (j0 + (ev ? 3 : 2)) * 16 + x};
  CellSet<4> C;
  cells_load<4>(C, 16, 16, axc, ayc, kk);
  Cell16 c[4];
#pragma unroll
  for (int m = 0; m < 4; m++) c[m] = {C.w[m], C.e[m], C.s[m], C.n[m], C.diag[m], C.y[m], sr2[kk[m]] * h2};
  const bool bottom = q == 0, top = q == 3;
  float R0 = K(0), R1 = K(0), K0 = K(0), K1 = K(0);
  for (int it = 0; it < iters; it++) {
    {  // red: the south neighbour of an even lane's R0 is K1 of the lanes below,
       // the north one of an odd lane's R1 is K0 of the lanes above
      const float X = row_from(ev ? K1 : K0, q, ev);
      const float s0 = ev ? (bottom ? R0 : X) : K0;
      const float n0 = ev ? K0 : K1;  // also R1's south neighbour
      const float n1 = ev ? K1 : (top ? R1 : X);
      const float a = relax16(c[0], lane_west(R0, K0), lane_east(R0, K0), s0, n0);
      const float b = relax16(c[1], lane_west(R1, K1), lane_east(R1, K1), n0, n1);
      R0 = a;
      R1 = b;
    }
    {  // black: the north neighbour of an even lane's K1 is R0 of the lanes
       // above, the south one of an odd lane's K0 is R1 of the lanes below
      const float X = row_from(ev ? R0 : R1, q, !ev);
      const float s0 = ev ? R0 : (bottom ? K0 : X);
      const float n0 = ev ? R1 : R0;  // also K1's south neighbour
      const float n1 = ev ? (top ? K1 : X) : R1;
      const float a = relax16(c[2], lane_west(K0, R0), lane_east(K0, R0), s0, n0);
      const float b = relax16(c[3], lane_west(K1, R1), lane_east(K1, R1), n0, n1);
      K0 = a;
      K1 = b;
    }
  }
  sp2[kk[0]] = R0;
  sp2[kk[1]] = R1;
  sp2[kk[2]] = K0;
  sp2[kk[3]] = K1;
}

// Entry q of a CellSet of slot 0 or 1 of the bottom solver (k = -1: none).
// Neighbours missing at the domain edge are read at the cell itself with a
// zero coefficient.
template <int CPT>
__device__ __forceinline__ float bot_off(const CellSet<CPT>& C, int q, const float* p, int nx) {
  const int k = C.k[q], b = C.nb[q];
  return C.w[q] * p[k - (b & 1)] + C.e[q] * p[k + ((b >> 1) & 1)] + C.s[q] * p[k - nx * ((b >> 2) & 1)] +
         C.n[q] * p[k + nx * ((b >> 3) & 1)];
}

// the red-black update of shaders/pressure.wgsl:87-94 (terms in its order;
// the quotient by the reciprocal-corrected division, as in the tiled legs)
template <int CPT>
__device__ __forceinline__ void bot_relax(const CellSet<CPT>& C, int q, int nx, float* p, const float* rhs, float h2) {
  if (C.k[q] >= 0) {
    const float off = bot_off<CPT>(C, q, p, nx);
    const float v = div_rcp_fast(off - rhs[C.k[q]] * h2, C.diag[q], C.y[q]);
    p[C.k[q]] = C.diag[q] > K(0) ? v : K(0);
  }
}

// the residual (shaders/pressure.wgsl:96-101)
template <int CPT>
__device__ __forceinline__ void bot_residual(const CellSet<CPT>& C, int q, int nx, const float* p, const float* rhs,
                                             float h2, float* res) {
  if (C.k[q] >= 0) {
    const float off = bot_off<CPT>(C, q, p, nx);
    res[C.k[q]] = C.diag[q] <= K(0) ? K(0) : rhs[C.k[q]] - (off - C.diag[q] * p[C.k[q]]) / h2;
  }
}

// bilinear coarse correction (shaders/pressure.wgsl:116-129)
template <int CPT>
__device__ __forceinline__ void bot_prolong(const CellSet<CPT>& C, int q, int nx, float* p, const float* pc, int nxc,
                                            int nyc) {
  const int k = C.k[q];
  if (k >= 0) {
    const int i = k % nx, j = k / nx;
    const int I = i / 2, J = j / 2;
    const int I2 = iclamp(I + ((i & 1) == 1 ? 1 : -1), 0, nxc - 1);
    const int J2 = iclamp(J + ((j & 1) == 1 ? 1 : -1), 0, nyc - 1);
    const float cc = K(0.5625) * pc[J * nxc + I] + K(0.1875) * (pc[J * nxc + I2] + pc[J2 * nxc + I]) +
                     K(0.0625) * pc[J2 * nxc + I2];
    p[k] += C.diag[q] > K(0) ? cc : K(0);
  }
}

// entry q of a CellSet in LDS, word m at st[(w + m) * 64 + lane] (8 words)
template <int CPT>
__device__ __forceinline__ void stash_put(float* st, int lane, int w, const CellSet<CPT>& C, int q) {
  float* o = st + w * 64 + lane;
  o[0] = __int_as_float(C.k[q]); o[64] = __int_as_float(C.nb[q]);
  o[128] = C.w[q]; o[192] = C.e[q]; o[256] = C.s[q]; o[320] = C.n[q]; o[384] = C.diag[q]; o[448] = C.y[q];
}
template <int CPT>
__device__ __forceinline__ void stash_get(const float* st, int lane, int w, CellSet<CPT>& C, int q) {
  const float* o = st + w * 64 + lane;
  C.k[q] = __float_as_int(o[0]); C.nb[q] = __float_as_int(o[64]);
  C.w[q] = o[128]; C.e[q] = o[192]; C.s[q] = o[256]; C.n[q] = o[320]; C.diag[q] = o[384]; C.y[q] = o[448];
}

// orders one wave's LDS traffic between half-sweeps without a workgroup barrier
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ double wave_sum64(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Direct FP64 solve of the coarsest level (schedule coarse_iters = 0):
// A p = -rhs h2, A the relaxation's operator (diag on the diagonal, minus the
// face coefficients to existing neighbours; a closed cell is an identity row),
// symmetric positive definite.  It is factored as L L^T inside its band of
// half-width ncx, band[k * W + d] = A[k][k - d] (W = ncx + 1, ncx <= 32), in
// the scene's scratch memory: one column per pair of workgroup barriers.  The
// first wave then solves L z = y and L^T p = z, each lane holding one of the
// last ncx unknowns, the products summed across the wave.
template <int CPT>
__device__ void coarse_direct(const CellSet<CPT>& C2, int ncx, int ncy, float h2, const float* sr2, float* sp2,
                              double* band, double* z) {
  const int t = threadIdx.x, N = ncx * ncy, W = ncx + 1;
  for (int q = t; q < N * W; q += BT) band[q] = 0.0;
  __syncthreads();
  if (t < 64) {
#pragma unroll
    for (int q = 0; q < CPT; q++) {
      const int k = C2.k[q];
      if (k < 0) continue;
      if (C2.diag[q] > K(0)) {
        band[k * W] = (double)C2.diag[q];
        if (C2.nb[q] & 1) band[k * W + 1] = -(double)C2.w[q];
        if (C2.nb[q] & 4) band[k * W + ncx] = -(double)C2.s[q];
        z[k] = -(double)sr2[k] * (double)h2;
      } else {
        band[k * W] = 1.0;
        z[k] = 0.0;
      }
    }
  }
  __syncthreads();
  for (int j = 0; j < N; j++) {
    const double ljj = sqrt(band[j * W]);
    if (t >= 1 && t <= ncx && j + t < N) band[(j + t) * W + t] /= ljj;
    __syncthreads();
    if (t == 0) band[j * W] = ljj;
    if (t < ncx * ncx) {
      const int a = t / ncx + 1, c = t % ncx + 1;  // row j + a, column j + c
      if (c <= a && j + a < N) band[(j + a) * W + (a - c)] -= band[(j + a) * W + a] * band[(j + c) * W + c];
    }
    __syncthreads();
  }
  if (t < 64) {
    const int l = t;  // lane l holds the unknown l rows back (forward) or ahead (backward)
    double win = 0.0;
    for (int j = 0; j < N; j++) {  // L z = y, y in z (LDS, read by every lane before lane 0 overwrites it)
      const double part = (l >= 1 && l <= ncx && j - l >= 0) ? band[j * W + l] * win : 0.0;
      const double zj = (z[j] - wave_sum64(part)) / band[j * W];
      const double up = __shfl_up(win, 1, 64);
      win = l == 1 ? zj : up;
      wave_sync();
      if (l == 0) z[j] = zj;
    }
    wave_sync();
    win = 0.0;
    for (int j = N - 1; j >= 0; j--) {  // L^T p = z
      const double part = (l >= 1 && l <= ncx && j + l < N) ? band[(j + l) * W + l] * win : 0.0;
      const double pj = (z[j] - wave_sum64(part)) / band[j * W];
      const double up = __shfl_up(win, 1, 64);
      win = l == 1 ? pj : up;
      if (l == 0) sp2[j] = (float)pj;
    }
  }
}


// coarse rhs = 1/4 of the 4 fine residuals (shaders/pressure.wgsl:104-113)
__device__ __forceinline__ void restrict_to(const float* res, int nxf, float* rhsc, float* pc, int nxc, int nyc) {
  for (int k = threadIdx.x; k < nxc * nyc; k += BT) {
    const int I = k % nxc, J = k / nxc;
    const float* r0 = res + (2 * J) * nxf + 2 * I;
    rhsc[k] = K(0.25) * (r0[0] + r0[1] + r0[nxf] + r0[nxf + 1]);
    pc[k] = K(0);
  }
}

// LDS of the bottom solver: p and rhs of the three slots, the residual scratch,
// the coarsest level's face coefficients (at most (nx+1)ny + nx(ny+1) <= 800),
// and the first
// wave's cell records while it sweeps the coarsest level (6 cells x 8 words)
constexpr int BOT_P0 = 0, BOT_R0 = 4096, BOT_P1 = 8192, BOT_R1 = 9216, BOT_P2 = 10240, BOT_R2 = 10496,
              BOT_RES = 10752, BOT_CAX = 14848, BOT_STASH = 15648,
              BOT_TOTAL = BOT_STASH + 6 * 8 * 64;
__device__ __forceinline__ int bot_p(int s) { return s == 0 ? BOT_P0 : (s == 1 ? BOT_P1 : BOT_P2); }
__device__ __forceinline__ int bot_r(int s) { return s == 0 ? BOT_R0 : (s == 1 ? BOT_R1 : BOT_R2); }

struct MgBottomOp {
  struct Args {
    Bottom B;
    int vg;  // tiles of this scene's launch
  };
  static constexpr int kThreads = 1024;
  static __device__ __forceinline__ void run(const Args& a_, int vb) {
    const Bottom& B = a_.B;
    if (B.skip && *B.skip) return;
    __shared__ float sm[BOT_TOTAL];
    float* sp2 = sm + BOT_P2;
    float* sr2 = sm + BOT_R2;
    float* sres = sm + BOT_RES;  // fine residuals
    float* stash = sm + BOT_STASH;
    auto P = [&](int s) { return sm + bot_p(s); };
    auto R = [&](int s) { return sm + bot_r(s); };
    const int nb = B.n, t = threadIdx.x;
    // slot s holds level B.L[s - s0]
    const int s0 = 3 - nb;
    auto lev = [&](int s) -> const Level& { return B.L[s - s0]; };
    const Level& Lc = B.L[nb - 1];
    const Level& Lf0 = B.L[0];
    float* axc = sm + BOT_CAX;
    float* ayc = axc + (Lc.nx + 1) * Lc.ny;
    const int nax = (Lc.nx + 1) * Lc.ny, nay = Lc.nx * (Lc.ny + 1), n0 = Lf0.nx * Lf0.ny;
    // ---- set-up: every global load is issued before the first store, so
    // they are in flight together.  The coarsest level's coefficients are
    // copied, or restricted here (:57-67); p (level 0's warm start; a level
    // above holds the restriction's zero, which is not read) and rhs of the
    // first level; the cell data of slots 0 and 1.
    // (a thread past the face count loads face 0 and discards it)
    float cax = 0.f, cay = 0.f;
    const int tu = t < nax ? t : 0, tv = t < nay ? t : 0;
    if (B.coarse_coef) {
      const Level& Lf = B.L[nb - 2];
      const int Iu = tu % (Lc.nx + 1), Ju = tu / (Lc.nx + 1), Iv = tv % Lc.nx, Jv = tv / Lc.nx;
      const float a0 = Lf.ax[(2 * Ju) * (Lf.nx + 1) + 2 * Iu], a1 = Lf.ax[(2 * Ju + 1) * (Lf.nx + 1) + 2 * Iu];
      const float b0 = Lf.ay[(2 * Jv) * Lf.nx + 2 * Iv], b1 = Lf.ay[(2 * Jv) * Lf.nx + 2 * Iv + 1];
      cax = K(0.5) * (a0 + a1);
      cay = K(0.5) * (b0 + b1);
    } else {
      cax = Lc.ax[tu];
      cay = Lc.ay[tv];
    }
    float pv[4], rv[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int k = t + BT * q;
      pv[q] = rv[q] = 0.f;
      if (k < n0) { pv[q] = B.zero_p ? K(0) : Lf0.p[k]; rv[q] = Lf0.rhs[k]; }
    }
    // Slot 0 (<= 4096 cells): thread t owns the horizontal pairs starting at
    // cells 2t and 2t + 2048; slot 1 (<= 1024): threads t < 512 own the pair
    // at 2t.  The levels above the coarsest have an even width, so each pair
    // is one red and one black cell, and every lane updates its cells of one
    // colour in each half-sweep.
    // A pair starting at an even column of row j has its red cell first
    // when j is even.  C0 holds the red cells of the two pairs, then the
    // black ones; C1 the red cell, then the black one.
    CellSet<4> C0;
    CellSet<2> C1;
    auto load0 = [&]() {
      const int nx = Lf0.nx, a = 2 * t, b = 2 * t + 2 * BT;
      const int ra = ((a / nx) & 1) ? a + 1 : a, rb = ((b / nx) & 1) ? b + 1 : b;
      const int k0[4] = {ra, rb, 2 * a + 1 - ra, 2 * b + 1 - rb};
      cells_load<4>(C0, nx, Lf0.ny, Lf0.ax, Lf0.ay, k0);
    };
    auto load1 = [&]() {
      const Level& L1 = lev(1);
      const int a = 2 * t, ra = ((a / L1.nx) & 1) ? a + 1 : a;
      const int k1[2] = {t < BT / 2 ? ra : -1, t < BT / 2 ? 2 * a + 1 - ra : -1};
      cells_load<2>(C1, L1.nx, L1.ny, L1.ax, L1.ay, k1);
    };
    if (nb == 3) load0();
    if (nb >= 2) load1();
    if (t < nax) { axc[t] = cax; if (B.coarse_coef) Lc.ax[t] = cax; }
    if (t < nay) { ayc[t] = cay; if (B.coarse_coef) Lc.ay[t] = cay; }
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int k = t + BT * q;
      if (k < n0) { P(s0)[k] = pv[q]; R(s0)[k] = rv[q]; }
    }
    __syncthreads();
    auto relax = [&](int s, int col) {  // col 0: red, 1: black
      const Level& L = lev(s);
      if (s == 0) {
        bot_relax<4>(C0, 2 * col, L.nx, P(0), R(0), L.h2);
        bot_relax<4>(C0, 2 * col + 1, L.nx, P(0), R(0), L.h2);
      } else {
        bot_relax<2>(C1, col, L.nx, P(1), R(1), L.h2);
      }
    };
    // ---- down (js/fluid.js:297-303)
    for (int s = s0; s < 2; s++) {
      const Level& L = lev(s);
      for (int q = 0; q < B.nu; q++) {
        relax(s, 0);
        __syncthreads();
        relax(s, 1);
        __syncthreads();
      }
      if (s == 0) {
#pragma unroll
        for (int q = 0; q < 4; q++) bot_residual<4>(C0, q, L.nx, P(0), R(0), L.h2, sres);
      } else {
#pragma unroll
        for (int q = 0; q < 2; q++) bot_residual<2>(C1, q, L.nx, P(1), R(1), L.h2, sres);
      }
      __syncthreads();
      const Level& Cn = lev(s + 1);
      restrict_to(sres, L.nx, R(s + 1), P(s + 1), Cn.nx, Cn.ny);
      __syncthreads();
    }
    // ---- coarsest: red-black sweeps from zero (shaders/pressure.wgsl:131-156)
    // by the first wave; the other waves wait at the next workgroup barrier
    for (int k = t; k < Lc.nx * Lc.ny; k += BT) sp2[k] = K(0);
    const int kw[4] = {t < 64 ? 2 * t : -1, t < 64 ? 2 * t + 1 : -1, t < 64 ? 2 * t + 128 : -1,
                       t < 64 ? 2 * t + 129 : -1};
    if (B.coarse_iters <= 0) {
      CellSet<4> C2;
      cells_load<4>(C2, Lc.nx, Lc.ny, axc, ayc, kw);
      __syncthreads();
      coarse_direct<4>(C2, Lc.nx, Lc.ny, Lc.h2, sr2, sp2, B.band, (double*)sres);
      // the cell data of slots 0 and 1 again, rather than held through the
      // factorization
      if (nb == 3) load0();
      if (nb >= 2) load1();
    } else {
      __syncthreads();
      // a wave-uniform branch (the first wave), so the registers of values
      // held only outside it are free inside it
      if (__builtin_amdgcn_readfirstlane(t) < 64) {
        // the wave's slot 0 and 1 cells wait in LDS, leaving their
        // registers to the coarsest level's cells
#pragma unroll
        for (int q = 0; q < 4; q++) stash_put<4>(stash, t, 8 * q, C0, q);
#pragma unroll
        for (int q = 0; q < 2; q++) stash_put<2>(stash, t, 8 * (4 + q), C1, q);
        if (Lc.nx == 16 && Lc.ny == 16) {
          coarse_sweeps16(axc, ayc, sr2, sp2, Lc.h2, B.coarse_iters, t);
        } else {
          CellSet<4> C2;
          cells_load<4>(C2, Lc.nx, Lc.ny, axc, ayc, kw);
          float rh[4];
#pragma unroll
          for (int q = 0; q < 4; q++) rh[q] = C2.k[q] >= 0 ? sr2[C2.k[q]] * Lc.h2 : K(0);
          // each lane's red and black cell of its two pairs; no cell: sres[0]
          SweepCell X[2][2];
          const int dummy = BOT_RES - BOT_P2;
#pragma unroll
          for (int a = 0; a < 2; a++) {
            const SweepCell c0 = sweep_cell<4>(C2, rh, 2 * a, Lc.nx, dummy), c1 = sweep_cell<4>(C2, rh, 2 * a + 1, Lc.nx, dummy);
            const bool first_red = ((C2.nb[2 * a] >> 4) & 1) == 0;
            X[0][a] = sweep_sel(first_red, c0, c1);
            X[1][a] = sweep_sel(first_red, c1, c0);
          }
          for (int it = 0; it < B.coarse_iters; it++) {
            sweep_relax2(X[0][0], X[0][1], sp2);
            wave_sync();
            sweep_relax2(X[1][0], X[1][1], sp2);
            wave_sync();
          }
        }
#pragma unroll
        for (int q = 0; q < 4; q++) stash_get<4>(stash, t, 8 * q, C0, q);
#pragma unroll
        for (int q = 0; q < 2; q++) stash_get<2>(stash, t, 8 * (4 + q), C1, q);
      }
    }
    __syncthreads();
    // ---- up (js/fluid.js:305-311)
    for (int s = 1; s >= s0; s--) {
      const Level& L = lev(s);
      const Level& Cn = lev(s + 1);
      if (s == 0) {
#pragma unroll
        for (int q = 0; q < 4; q++) bot_prolong<4>(C0, q, L.nx, P(0), P(1), Cn.nx, Cn.ny);
      } else {
#pragma unroll
        for (int q = 0; q < 2; q++) bot_prolong<2>(C1, q, L.nx, P(1), P(2), Cn.nx, Cn.ny);
      }
      __syncthreads();
      for (int q = 0; q < B.nu; q++) {  // black then red
        relax(s, 1);
        __syncthreads();
        relax(s, 0);
        __syncthreads();
      }
    }
    // write back p of every level and rhs of the levels below the first
    for (int s = s0; s < 3; s++) {
      const Level& L = lev(s);
      for (int k = t; k < L.nx * L.ny; k += BT) {
        L.p[k] = P(s)[k];
        if (s > s0) L.rhs[k] = R(s)[k];
      }
    }
  }
};
__global__ __launch_bounds__(1024) void k_mg_bottom(const MgBottomOp::Args a_) {
  MgBottomOp::run(a_, xcd_tile(blockIdx.x, gridDim.x));
}

// ---- launchers -------------------------------------------------------------------
template <int TX, int TY, int NU, int M>
static auto k_mg_down_of(MgDownOp<TX, TY, NU, M>) { return k_mg_down<TX, TY, NU, M>; }
template <int TX, int TY, int NU, int M>
static auto k_mg_up_of(MgUpOp<TX, TY, NU, M>) { return k_mg_up<TX, TY, NU, M>; }

static inline unsigned ntiles(int nx, int ny, int tx, int ty) {
  return (unsigned)(((nx + tx - 1) / tx) * ((ny + ty - 1) / ty));
}

// The wide tiles (code 48) where a level has at least 512 tiles of 64x64
// (more than one round of them: two workgroups per CU overlap loads and
// sweeps), 64x64 tiles of 1024 threads where it has at least 256 (the
// smoothing halo then adds 34 % to the tile's cells instead of 72 %), 32x32
// where it has at least 1024 of those, 16x16 below: a small level then has
// four times the workgroups, each with half the per-thread chain, and its
// legs are a shorter latency chain
int mg_tile_rule(const Level& f) {
  return mg_tile_for_count((int)ntiles(f.nx, f.ny, 64, 64), (int)ntiles(f.nx, f.ny, 32, 32));
}
// the same rule on the tile counts of a launch (n64 tiles of 64x64, n32 of
// 32x32), which for a pool's phase sum over its scenes: many small levels
// together fill the device with large tiles, whose halo costs less
int mg_tile_for_count(int n64, int n32) {
  if (n64 >= 512) return 48;  // more than one round of 64x64 workgroups: the wide tiles
  if (n64 >= 256) return 64;
  return n32 >= 1024 ? 32 : 16;
}
// a tile chosen per level (the autotuner's) overrides the rule; every tile
// gives the same results, the legs being exact at any tile size
int mg_tile(const Level& f) { return f.tile ? f.tile : mg_tile_rule(f); }

// dispatch on (tile, nu, storage mode) to the instantiation
// Tile code 48 is the wide tile: 64 cells wide, short enough that two
// workgroups share a CU (their registers fit twice), so one workgroup's loads
// overlap the other's sweeps where a launch has several rounds of tiles:
// 64x48 down legs (768 threads) at nu <= 2, 64x32 at nu = 3, 64x32 up legs
// (512 threads), 64x30 projecting up legs in pools (512 threads: the one more
// ring of halo of 64x32 would give each thread a fourth pair of cells); a lone
// scene's projecting leg keeps 64x64 tiles.
template <template <int, int, int, int> class Op>
struct WideRows {
  static constexpr int rows(int) { return 32; }
};
template <>
struct WideRows<MgDownOp> {
  static constexpr int rows(int nu) { return nu <= 2 ? 48 : 32; }
};
template <template <int, int, int, int> class Op, class F>
static bool mg_dispatch(int tile, int nu, int mode, F&& f) {
#define CF_MG_CASE(T, NU, M)                                                               \
  if (tile == T && nu == NU && mode == M) {                                                \
    f(Op<(T == 48 ? 64 : T), (T == 48 ? WideRows<Op>::rows(NU) : T), NU, M>{});          \
    return true;                                                                           \
  }
#define CF_MG_TILE(T) \
  CF_MG_CASE(T, 1, 0) CF_MG_CASE(T, 2, 0) CF_MG_CASE(T, 3, 0) \
  CF_MG_CASE(T, 1, 1) CF_MG_CASE(T, 2, 1) CF_MG_CASE(T, 3, 1) \
  CF_MG_CASE(T, 1, 3) CF_MG_CASE(T, 2, 3) CF_MG_CASE(T, 3, 3)
  CF_MG_TILE(64)
  CF_MG_TILE(48)
  CF_MG_TILE(32)
  CF_MG_TILE(16)
#undef CF_MG_TILE
#undef CF_MG_CASE
  return false;
}

bool launch_mg_down(const Level& f, const Level& c, const MgStore& st, int nu, const int* skip, hipStream_t s,
                    int zero_p) {
  const int t = mg_tile(f);
  return mg_dispatch<MgDownOp>(t, nu, st.mode, [&](auto op) {
    using Op = decltype(op);
    const int nb = (int)ntiles(f.nx, f.ny, Op::kTX, Op::kTY);
    launch(k_mg_down_of(op), dim3(nb), dim3(Op::kThreads), 0, s, typename Op::Args{f, c, st, skip, nb, zero_p});
  });
}

bool launch_mg_up(const Level& f, const Level& c, const MgStore& st, int nu, const int* skip, hipStream_t s) {
  const int t = mg_tile(f);
  return mg_dispatch<MgUpOp>(t, nu, st.mode, [&](auto op) {
    using Op = decltype(op);
    const int nb = (int)ntiles(f.nx, f.ny, Op::kTX, Op::kTY);
    launch(k_mg_up_of(op), dim3(nb), dim3(Op::kThreads), 0, s, typename Op::Args{f, c, st, skip, nb});
  });
}

// two-level legs: 16 x 16 coarse cells per workgroup where that still gives
// the launch 128 workgroups, else 8 x 8
int mg_pair_tile(const Level& c) { return ntiles(c.nx, c.ny, 16, 16) >= 128 ? 16 : 8; }

template <int TC, int NU, int NT>
static auto k_mg_down2_of(MgDown2Op<TC, NU, NT>) { return k_mg_down2<TC, NU, NT>; }
template <int TC, int NU, int NT>
static auto k_mg_up2_of(MgUp2Op<TC, NU, NT>) { return k_mg_up2<TC, NU, NT>; }

// threads per workgroup: about one pair of fine cells per thread and sweep
// (down: (2TC + 6NU + 4)^2 / 2 cells, up: (2TC + 4NU)^2 / 2)
template <template <int, int, int> class Op, int NT16, int NT8, class F>
static bool pair_dispatch(int tc, int nu, F&& f) {
#define CF_PAIR_CASE(T, NU, NT) \
  if (tc == T && nu == NU) { f(Op<T, NU, NT>{}); return true; }
  CF_PAIR_CASE(16, 1, NT16) CF_PAIR_CASE(16, 2, NT16) CF_PAIR_CASE(16, 3, NT16)
  CF_PAIR_CASE(8, 1, NT8) CF_PAIR_CASE(8, 2, NT8) CF_PAIR_CASE(8, 3, NT8)
#undef CF_PAIR_CASE
  return false;
}

bool launch_mg_down2(const Level& f, const Level& c, const Level& n, int nu, int tile, const int* skip,
                     hipStream_t s) {
  const int tc = tile ? tile : mg_pair_tile(c), nb = (int)ntiles(c.nx, c.ny, tc, tc);
  return pair_dispatch<MgDown2Op, 1024, 512>(tc, nu, [&](auto op) {
    using Op = decltype(op);
    launch(k_mg_down2_of(op), dim3(nb), dim3(Op::kThreads), 0, s, typename Op::Args{f, c, n, skip, nb});
  });
}

bool launch_mg_up2(const Level& f, const Level& c, const Level& n, int nu, int tile, const int* skip, hipStream_t s) {
  const int tc = tile ? tile : mg_pair_tile(c), nb = (int)ntiles(c.nx, c.ny, tc, tc);
  return pair_dispatch<MgUp2Op, 1024, 256>(tc, nu, [&](auto op) {
    using Op = decltype(op);
    launch(k_mg_up2_of(op), dim3(nb), dim3(Op::kThreads), 0, s, typename Op::Args{f, c, n, skip, nb});
  });
}

bool launch_mg_up_project(const Level& f, const Level& c, const Fields& F, int nu, const int* skip, hipStream_t s) {
  const int t = mg_tile(f) == 48 ? 64 : mg_tile(f), nb = (int)ntiles(f.nx, f.ny, t, t);
#define CF_UPP_CASE(T, NU)                                                                          \
  if (t == T && nu == NU) {                                                                         \
    using Op = MgUpOp<T, T, NU, 0, true>;                                                           \
    Op::Args a{};                                                                                   \
    a.f = f; a.c = c; a.st = MgStore{}; a.skip = skip; a.vg = nb; a.F = F;                          \
    launch(k_mg_up_project<T, T, NU>, dim3(nb), dim3(Op::kThreads), 0, s, a);                       \
    return true;                                                                                    \
  }
  CF_UPP_CASE(64, 1) CF_UPP_CASE(64, 2) CF_UPP_CASE(64, 3)
  CF_UPP_CASE(32, 1) CF_UPP_CASE(32, 2) CF_UPP_CASE(32, 3)
  CF_UPP_CASE(16, 1) CF_UPP_CASE(16, 2) CF_UPP_CASE(16, 3)
#undef CF_UPP_CASE
  return false;
}

bool phase_mg_down(BatchPhase& ph, const Level& f, const Level& c, const MgStore& st, int nu, const int* skip,
                   int t, int zero_p) {
  return mg_dispatch<MgDownOp>(t, nu, st.mode, [&](auto op) {
    using Op = decltype(op);
    const int nb = (int)ntiles(f.nx, f.ny, Op::kTX, Op::kTY);
    phase_add<Op>(ph, typename Op::Args{f, c, st, skip, nb, zero_p}, nb, dim3(Op::kThreads), "mg_down");
  });
}

bool phase_mg_up(BatchPhase& ph, const Level& f, const Level& c, const MgStore& st, int nu, const int* skip,
                 int t) {
  return mg_dispatch<MgUpOp>(t, nu, st.mode, [&](auto op) {
    using Op = decltype(op);
    const int nb = (int)ntiles(f.nx, f.ny, Op::kTX, Op::kTY);
    phase_add<Op>(ph, typename Op::Args{f, c, st, skip, nb}, nb, dim3(Op::kThreads), "mg_up");
  });
}

bool phase_mg_up_project(BatchPhase& ph, const Level& f, const Level& c, const Fields& F, int nu, const int* skip,
                         int tile) {
  const int t = tile;
#define CF_UPP_PHASE(T, NU)                                                                \
  if (t == T && nu == NU) {                                                                \
    using Op = MgUpOp<(T == 48 ? 64 : T), (T == 48 ? 30 : T), NU, 0, true>;                \
    const int nb = (int)ntiles(f.nx, f.ny, Op::kTX, Op::kTY);                              \
    typename Op::Args a{};                                                                 \
    a.f = f; a.c = c; a.st = MgStore{}; a.skip = skip; a.vg = nb; a.F = F;                 \
    phase_add<Op>(ph, a, nb, dim3(Op::kThreads), "mg_up_project");                        \
    return true;                                                                           \
  }
  CF_UPP_PHASE(64, 1) CF_UPP_PHASE(64, 2) CF_UPP_PHASE(64, 3)
  CF_UPP_PHASE(48, 1) CF_UPP_PHASE(48, 2) CF_UPP_PHASE(48, 3)
  CF_UPP_PHASE(32, 1) CF_UPP_PHASE(32, 2) CF_UPP_PHASE(32, 3)
  CF_UPP_PHASE(16, 1) CF_UPP_PHASE(16, 2) CF_UPP_PHASE(16, 3)
#undef CF_UPP_PHASE
  return false;
}

// the bottom kernel takes the last min(3, nlev) levels when each fits its slot
int mg_bottom_first(const Level* L, int nlev) {
  const int cap[3] = {256, 1024, 4096};
  int b = nlev - 1;
  for (int d = 1; d < 3 && b - 1 >= 0; d++) {
    if (L[b - 1].nx * L[b - 1].ny > cap[d]) break;
    b--;
  }
  return b;
}

static bool bottom_args(const Level* L, int nlev, int first, bool coarse_coef, int nu, int coarse_iters, const int* skip,
                        double* band, Bottom& B) {
  const int n = nlev - first;
  if (n < 1 || n > 3 || L[nlev - 1].nx * L[nlev - 1].ny > 256) return false;
  if (coarse_coef && n < 2) return false;
  if (coarse_iters <= 0 && (!band || L[nlev - 1].nx > 32)) return false;
  for (int q = 0; q < n; q++) B.L[q] = L[first + q];
  B.n = n; B.coarse_coef = coarse_coef ? 1 : 0; B.nu = nu; B.coarse_iters = coarse_iters; B.skip = skip;
  B.zero_p = first >= 1 ? 1 : 0;
  B.band = band;
  return true;
}

bool phase_mg_bottom(BatchPhase& ph, const Level* L, int nlev, int first, bool coarse_coef, int nu, int coarse_iters,
                     const int* skip, double* band) {
  Bottom B;
  if (!bottom_args(L, nlev, first, coarse_coef, nu, coarse_iters, skip, band, B)) return false;
  phase_add<MgBottomOp>(ph, MgBottomOp::Args{B, 1}, 1, dim3(BT), "mg_bottom");
  return true;
}

bool launch_mg_bottom(const Level* L, int nlev, int first, bool coarse_coef, int nu, int coarse_iters, const int* skip,
                      double* band, hipStream_t s) {
  Bottom B;
  if (!bottom_args(L, nlev, first, coarse_coef, nu, coarse_iters, skip, band, B)) return false;
  launch(k_mg_bottom, dim3(1), dim3(BT), 0, s, MgBottomOp::Args{B, 1});
  return true;
}

}  // namespace cf
