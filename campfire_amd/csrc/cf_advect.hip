// cf_advect.hip — MacCormack advection of one step in a single pass.
//
// The reference runs it as two dispatch pairs (shaders/advect.wgsl):
// scalForward + velForward write Th, Yh, Uh, Vh; scalCorrect + velCorrect
// read them back at the backward-traced points and write T2, Y2, U2, V2.
// Here one workgroup of 512 threads owns a 32x32 tile of cells and keeps
// everything of the tile's neighbourhood in LDS:
//   A. U, V, T, Y over the tile + 3 cells, as the values the boundary-aware
//      samplers of js/gpu.js:16-65 return for each index (ambient outside the
//      sides and top, floor clamp, no-slip ghost for U), so sampling LDS is
//      sampling the field;
//   B. the forward pass (Th, Yh, Uh, Vh) over the tile + 1 cell, by the
//      same rules;
//   C. the correction and limiter for the tile's cells and faces.
// These halos hold every footprint of a displacement below one cell per
// step.  A bilinear footprint that leaves a window reads the global field, or
// evaluates the forward pass at that index, with the same arithmetic, so the
// result never depends on the window sizes.  Two 73 KB workgroups fit in a
// CU's LDS (16 waves); the forward pass over the ring around the tile adds
// 13 % to the tile's own.  Th, Yh, Uh, Vh are not written to memory.
#include "cf_kernels.h"

namespace cf {

namespace {

constexpr int AX = 32, AY = 32;  // output tile (cells)
constexpr int NT = 512;          // threads: two tile cells each
constexpr int AH = 1;            // halo of the forward-pass window
constexpr int AR = 3;            // halo of the staged input window
constexpr int A_CX = AX + 2 * AR, A_CY = AY + 2 * AR;  // T, Y
constexpr int A_UX = A_CX + 1, A_UY = A_CY;            // U faces
constexpr int A_VX = A_CX, A_VY = A_CY + 1;            // V faces
constexpr int B_CX = AX + 2 * AH, B_CY = AY + 2 * AH;  // Th, Yh
constexpr int B_UX = B_CX + 1, B_UY = B_CY;            // Uh
constexpr int B_VX = B_CX, B_VY = B_CY + 1;            // Vh

__device__ __forceinline__ float mixv(float a, float b, float t) { return mixf(a, b, t); }
__device__ __forceinline__ float4 mixv(float4 a, float4 b, float t) { return mix4(a, b, t); }

template <class T>
using lds_ptr = __attribute__((address_space(3))) const T*;

// A window of an index space held in LDS, element (i, j) at
// s[(j - oy) * W + (i - ox)].  CK: reads flag `ok` false when a bilinear
// footprint leaves the window; the caller then redoes the point from global
// memory.  Without CK the caller has shown that every footprint lies inside
// (a workgroup whose velocities move no point by a cell), and `ok` is never
// touched.
template <class T, int W, int H, bool CK>
struct Win {
  const T* s;
  int ox, oy;
  float fox, foy;  // ox, oy as floats
  float cb;        // LDS byte address of index (0, 0): s - (oy * W + ox) elements (set by init)
  // On grids up to 16384 cells a side every term of the byte address
  // by * W * sizeof(T) + bx * sizeof(T) + cb is an integer below 2^24, so it
  // is formed exactly in float with two fused multiply-adds and converted once.
  __device__ __forceinline__ void init(const T* s_, int ox_, int oy_) {
    s = s_; ox = ox_; oy = oy_;
    fox = (float)ox_; foy = (float)oy_;
    const float base = (float)(unsigned)(uintptr_t)(lds_ptr<T>)s_;
    cb = base - (foy * (float)W + fox) * (float)sizeof(T);
  }
  // the footprint whose lower-left corner is at the floored position (bx, by).
  // CK: tested in float, exact for values in the window (small integers) and
  // false for NaN or huge ones, so the conversion needs no clamp first; the
  // element index b * W + a is formed in float too
  __device__ __forceinline__ const T* at(float bx, float by, bool& ok) const {
    if constexpr (!CK) {
      const unsigned byte = (unsigned)(int)fmaf(by, (float)(W * sizeof(T)), fmaf(bx, (float)sizeof(T), cb));
      return (const T*)(lds_ptr<T>)byte;
    } else {
      const float a = bx - fox, b = by - foy;
      const bool in = a >= K(0) && a < (float)(W - 1) && b >= K(0) && b < (float)(H - 1);
      ok = ok && in;
      return s + (in ? (unsigned)(int)fmaf(b, (float)W, a) : 0u);
    }
  }
  __device__ __forceinline__ T own(int i, int j) const { return s[(j - oy) * W + (i - ox)]; }
  // bilinear sample at (fx, fy) = grid position minus the field's stagger
  __device__ __forceinline__ T bil(float fx, float fy, bool& ok) const {
    const float bx = floorf(fx), by = floorf(fy);
    const float tx = fx - bx, ty = fy - by;
    const T* p = at(bx, by, ok);
    return mixv(mixv(p[0], p[1], tx), mixv(p[W], p[W + 1], tx), ty);
  }
};

// x / d for the traces: div_rcp_fast without the sign fix-up, so a zero
// quotient may carry the wrong sign; every quotient here is subtracted from a
// positive position (or +0), where either zero gives the same result
__device__ __forceinline__ float div_trace(float x, float d, float y) {
  const float q = x * y;
  return fmaf(fmaf(-q, d, x), y, q);
}

__device__ __forceinline__ float2 mm4v(float a, float c, float d, float e) {
  return make_float2(fminf(fminf(a, c), fminf(d, e)), fmaxf(fmaxf(a, c), fmaxf(d, e)));
}

// what a tile's passes read (idx: 1/dx correctly rounded, for div_rcp)
template <bool CK>
struct Tile {
  Win<float, A_CX, A_CY, CK> T;
  Win<float4, A_CX, A_CY, CK> Y;
  Win<float, A_UX, A_UY, CK> U;
  Win<float, A_VX, A_VY, CK> V;
  Win<float, B_CX, B_CY, CK> Th;
  Win<float4, B_CX, B_CY, CK> Yh;
  Win<float, B_UX, B_UY, CK> Uh;
  Win<float, B_VX, B_VY, CK> Vh;
  float dx, idx;
};

// RK2 departure point (shaders/advect.wgsl:28-34) from the velocity windows,
// the divisions by dx as correctly rounded quotients (div_rcp_fast).  Traces
// start at a cell centre or a face, whose first velocity samples have
// integer footprints and weights 0.5 or 0: they are read at fixed offsets
// and mixed with constant weights, the arithmetic of bil() on those points.
// Every B and C point's first footprint lies inside the A windows (their
// halo is AR, the points are within AH of the tile, the footprint reaches 1
// more), so `ok` is unchanged by these samples.  A point's forward (+dt) and
// backward (-dt) traces start from the same first velocity (in cells per
// unit time, w1), which the forward pass keeps for the correction.
template <class TL>
__device__ __forceinline__ float2 trace_from(const TL& t, float2 g, float2 w1, float dt, bool& ok) {
  const float gx = g.x, gy = g.y;
  const float hdt = K(0.5) * dt;
  const float mx_ = gx - hdt * w1.x, my_ = gy - hdt * w1.y;
  const float wx = div_trace(t.U.bil(mx_ - K(0), my_ - K(0.5), ok), t.dx, t.idx);
  const float wy = div_trace(t.V.bil(mx_ - K(0.5), my_ - K(0), ok), t.dx, t.idx);
  return make_float2(gx - dt * wx, gy - dt * wy);
}
template <class TL>
__device__ __forceinline__ float2 first_vel(const TL& t, float u, float v) {
  return make_float2(div_trace(u, t.dx, t.idx), div_trace(v, t.dx, t.idx));
}
// from cell centre (i + 1/2, j + 1/2): U footprint (i, j) with (0.5, 0), V (i, j) with (0, 0.5)
template <class TL>
__device__ __forceinline__ float2 vel_cell(const TL& t, int i, int j) {
  const float u = mixf(mixf(t.U.own(i, j), t.U.own(i + 1, j), K(0.5)), mixf(t.U.own(i, j + 1), t.U.own(i + 1, j + 1), K(0.5)),
                       K(0));
  const float v = mixf(mixf(t.V.own(i, j), t.V.own(i + 1, j), K(0)), mixf(t.V.own(i, j + 1), t.V.own(i + 1, j + 1), K(0)),
                       K(0.5));
  return first_vel(t, u, v);
}
// from U face (i, j + 1/2): U footprint (i, j) with (0, 0), V (i - 1, j) with (0.5, 0.5)
template <class TL>
__device__ __forceinline__ float2 vel_uface(const TL& t, int i, int j) {
  const float u = mixf(mixf(t.U.own(i, j), t.U.own(i + 1, j), K(0)), mixf(t.U.own(i, j + 1), t.U.own(i + 1, j + 1), K(0)),
                       K(0));
  const float v = mixf(mixf(t.V.own(i - 1, j), t.V.own(i, j), K(0.5)), mixf(t.V.own(i - 1, j + 1), t.V.own(i, j + 1), K(0.5)),
                       K(0.5));
  return first_vel(t, u, v);
}
// from V face (i + 1/2, j): U footprint (i, j - 1) with (0.5, 0.5), V (i, j) with (0, 0)
template <class TL>
__device__ __forceinline__ float2 vel_vface(const TL& t, int i, int j) {
  const float u = mixf(mixf(t.U.own(i, j - 1), t.U.own(i + 1, j - 1), K(0.5)), mixf(t.U.own(i, j), t.U.own(i + 1, j), K(0.5)),
                       K(0.5));
  const float v = mixf(mixf(t.V.own(i, j), t.V.own(i + 1, j), K(0)), mixf(t.V.own(i, j + 1), t.V.own(i + 1, j + 1), K(0)),
                       K(0));
  return first_vel(t, u, v);
}
// the start points of the three kinds
__device__ __forceinline__ float2 at_cell(int i, int j) { return make_float2((float)i + K(0.5), (float)j + K(0.5)); }
__device__ __forceinline__ float2 at_uface(int i, int j) { return make_float2((float)i, (float)j + K(0.5)); }
__device__ __forceinline__ float2 at_vface(int i, int j) { return make_float2((float)i + K(0.5), (float)j); }

// ---- global-memory forms (points whose footprints leave the windows) ----
// forward pass at any index, as the corrector's samplers see Th/Yh, Uh, Vh
__device__ __forceinline__ void g_fwd_cell(const Fields& F, const cf_params& P, int i, int j, float& th, float4& yh) {
  const int nx = P.nx, ny = P.ny;
  if (i < 0 || i >= nx || j >= ny) { th = P.Tamb; yh = make_float4(0.f, P.YOamb, 0.f, 0.f); return; }
  j = max(j, 0);
  const float2 d = trace(F.U, F.V, nx, ny, P.dx, (float)i + K(0.5), (float)j + K(0.5), P.dt);
  th = sC(F.T, nx, ny, P.Tamb, d.x, d.y);
  yh = sC4(F.Y, nx, ny, make_float4(0.f, P.YOamb, 0.f, 0.f), d.x, d.y);
}
__device__ __forceinline__ float g_fwd_u(const Fields& F, const cf_params& P, int i, int j) {
  const int nx = P.nx, ny = P.ny;
  if (i < 0 || i > nx || j >= ny) return K(0);
  const int jj = max(j, 0);
  const float2 d = trace(F.U, F.V, nx, ny, P.dx, (float)i, (float)jj + K(0.5), P.dt);
  const float u = sU(F.U, nx, ny, d.x, d.y);
  return j < 0 ? -u : u;
}
__device__ __forceinline__ float g_fwd_v(const Fields& F, const cf_params& P, int i, int j) {
  const int nx = P.nx, ny = P.ny;
  if (i < 0 || i >= nx || j > ny) return K(0);
  j = max(j, 0);
  const float2 d = trace(F.U, F.V, nx, ny, P.dx, (float)i + K(0.5), (float)j, P.dt);
  return sV(F.V, nx, ny, d.x, d.y);
}

// the corrector of one cell / U face / V face from global memory only
// (shaders/advect.wgsl:49-66, 78-103)
__device__ __forceinline__ void g_corr_cell(const Fields& F, const cf_params& P, int i, int j, float& To, float4& Yo) {
  const int nx = P.nx, ny = P.ny;
  const float4 aY = make_float4(0.f, P.YOamb, 0.f, 0.f);
  const float gx = (float)i + K(0.5), gy = (float)j + K(0.5);
  const float2 d = trace(F.U, F.V, nx, ny, P.dx, gx, gy, P.dt);
  const float2 f = trace(F.U, F.V, nx, ny, P.dx, gx, gy, -P.dt);
  const float2 mT = mmC(F.T, nx, ny, P.Tamb, d.x, d.y);
  float4 lo, hi;
  mmC4(F.Y, nx, ny, aY, d.x, d.y, lo, hi);
  float Thk;
  float4 Yhk;
  g_fwd_cell(F, P, i, j, Thk, Yhk);
  const float4 Yk = F.Y[j * nx + i];
  // sC / sC4 of the forward fields at f
  const float fx = f.x - K(0.5), fy = f.y - K(0.5);
  const float bx = floorf(fx), by = floorf(fy);
  const float tx = fx - bx, ty = fy - by;
  const int bi = toint(bx), bj = toint(by);
  float t00, t10, t01, t11;
  float4 y00, y10, y01, y11;
  g_fwd_cell(F, P, bi, bj, t00, y00);
  g_fwd_cell(F, P, bi + 1, bj, t10, y10);
  g_fwd_cell(F, P, bi, bj + 1, t01, y01);
  g_fwd_cell(F, P, bi + 1, bj + 1, t11, y11);
  const float sbT = mixf(mixf(t00, t10, tx), mixf(t01, t11, tx), ty);
  const float4 sb = mix4(mix4(y00, y10, tx), mix4(y01, y11, tx), ty);
  const float Tc = Thk + K(0.5) * (F.T[j * nx + i] - sbT);
  const float4 Yc = make_float4(Yhk.x + K(0.5) * (Yk.x - sb.x), Yhk.y + K(0.5) * (Yk.y - sb.y),
                                Yhk.z + K(0.5) * (Yk.z - sb.z), Yhk.w + K(0.5) * (Yk.w - sb.w));
  const bool inT = Tc >= mT.x && Tc <= mT.y;
  const bool inY = Yc.x >= lo.x && Yc.y >= lo.y && Yc.z >= lo.z && Yc.w >= lo.w &&
                   Yc.x <= hi.x && Yc.y <= hi.y && Yc.z <= hi.z && Yc.w <= hi.w;
  const float mixC = fminf(Yc.x, Yc.y / P.sO2);
  const float mixL = fminf(Yhk.x, Yhk.y / P.sO2);
  const bool take = inT && inY && mixC <= mixL + K(1e-5);
  To = take ? Tc : Thk;
  Yo = take ? Yc : Yhk;
}
__device__ __forceinline__ float g_corr_u(const Fields& F, const cf_params& P, int i, int j) {
  const int nx = P.nx, ny = P.ny;
  const float gx = (float)i, gy = (float)j + K(0.5);
  const float2 b = trace(F.U, F.V, nx, ny, P.dx, gx, gy, -P.dt);
  const float fx = b.x - K(0), fy = b.y - K(0.5);
  const float bx = floorf(fx), by = floorf(fy);
  const float tx = fx - bx, ty = fy - by;
  const int bi = toint(bx), bj = toint(by);
  const float back = mixf(mixf(g_fwd_u(F, P, bi, bj), g_fwd_u(F, P, bi + 1, bj), tx),
                          mixf(g_fwd_u(F, P, bi, bj + 1), g_fwd_u(F, P, bi + 1, bj + 1), tx), ty);
  const float2 d = trace(F.U, F.V, nx, ny, P.dx, gx, gy, P.dt);
  const float2 m = mmU(F.U, nx, ny, d.x, d.y);
  return clampf(g_fwd_u(F, P, i, j) + K(0.5) * (F.U[j * (nx + 1) + i] - back), m.x, m.y);
}
__device__ __forceinline__ float g_corr_v(const Fields& F, const cf_params& P, int i, int j) {
  const int nx = P.nx, ny = P.ny;
  const float gx = (float)i + K(0.5), gy = (float)j;
  const float2 b = trace(F.U, F.V, nx, ny, P.dx, gx, gy, -P.dt);
  const float fx = b.x - K(0.5), fy = b.y - K(0);
  const float bx = floorf(fx), by = floorf(fy);
  const float tx = fx - bx, ty = fy - by;
  const int bi = toint(bx), bj = toint(by);
  const float back = mixf(mixf(g_fwd_v(F, P, bi, bj), g_fwd_v(F, P, bi + 1, bj), tx),
                          mixf(g_fwd_v(F, P, bi, bj + 1), g_fwd_v(F, P, bi + 1, bj + 1), tx), ty);
  const float2 d = trace(F.U, F.V, nx, ny, P.dx, gx, gy, P.dt);
  const float2 m = mmV(F.V, nx, ny, d.x, d.y);
  return clampf(g_fwd_v(F, P, i, j) + K(0.5) * (F.V[j * nx + i] - back), m.x, m.y);
}

}  // namespace

struct AdvectOp {
  using Args = Fields;
  static constexpr int kThreads = NT;
  static constexpr int kMinWaves = 4;  // two workgroups per CU, as the LDS allows
  static __device__ __forceinline__ void run(const Fields& F, int tile);
  static int tiles(int nx, int ny) { return ((nx + AX - 1) / AX) * ((ny + AY - 1) / AY); }
};

__device__ __forceinline__ void AdvectOp::run(const Fields& F, int tile) {
  __shared__ float sT[A_CX * A_CY];
  __shared__ float4 sY[A_CX * A_CY];
  __shared__ float sU[A_UX * A_UY];
  __shared__ float sV[A_VX * A_VY];
  __shared__ float bT[B_CX * B_CY];
  __shared__ float4 bY[B_CX * B_CY];
  __shared__ float bU[B_UX * B_UY];
  __shared__ float bV[B_VX * B_VY];
  const cf_params P = load_params(F.P);
  const int nx = P.nx, ny = P.ny;
  const int gx = (nx + AX - 1) / AX;
  const int x0 = (tile % gx) * AX, y0 = (tile / gx) * AY;
  const int tid = threadIdx.x;
  const float dt = P.dt;
  const float4 aY = make_float4(0.f, P.YOamb, 0.f, 0.f);
  const int ax0 = x0 - AR, ay0 = y0 - AR, bx0 = x0 - AH, by0 = y0 - AH;
  // a staged velocity below vlim moves no trace of this workgroup by a cell
  // (below), so every bilinear footprint of its passes lies in the windows
  // (grids up to 2^14 cells a side keep the rounding of a position under
  // 2^-10 cell, and the untested windows' byte addresses exact in float)
  const float vlim = K(0.99) * P.dx / fabsf(dt);
  bool fastv = nx <= 16384 && ny <= 16384;  // false: also a staged |U| or |V| >= vlim, or NaN

  // A. inputs as the samplers see them (fC, fC4, fU, fV).  Every load is
  // issued before the first store, branch-free (an index outside the field
  // loads element 0, and the sampler's boundary value replaces it), so the
  // thread's loads are in flight together instead of one round trip per
  // element
  {
    constexpr int QC = (A_CX * A_CY + NT - 1) / NT, QU = (A_UX * A_UY + NT - 1) / NT,
                  QV = (A_VX * A_VY + NT - 1) / NT;
    float tv[QC], uv[QU], vv[QV];
    float4 yv[QC];
#pragma unroll
    for (int q = 0; q < QC; q++) {
      const int k = min(tid + NT * q, A_CX * A_CY - 1);
      const int i = ax0 + k % A_CX, j = ay0 + k / A_CX;
      const bool in = i >= 0 && i < nx && j < ny;
      const int idx = in ? max(j, 0) * nx + i : 0;
      tv[q] = F.T[idx];
      yv[q] = F.Y[idx];
    }
#pragma unroll
    for (int q = 0; q < QU; q++) {
      const int k = min(tid + NT * q, A_UX * A_UY - 1);
      const int i = ax0 + k % A_UX, j = ay0 + k / A_UX;
      const bool in = i >= 0 && i <= nx && j < ny;
      uv[q] = F.U[in ? max(j, 0) * (nx + 1) + i : 0];
    }
#pragma unroll
    for (int q = 0; q < QV; q++) {
      const int k = min(tid + NT * q, A_VX * A_VY - 1);
      const int i = ax0 + k % A_VX, j = ay0 + k / A_VX;
      const bool in = i >= 0 && i < nx && j <= ny;
      vv[q] = F.V[in ? max(j, 0) * nx + i : 0];
    }
#pragma unroll
    for (int q = 0; q < QC; q++) {
      const int k = tid + NT * q;
      if (k < A_CX * A_CY) {
        const int i = ax0 + k % A_CX, j = ay0 + k / A_CX;
        const bool in = i >= 0 && i < nx && j < ny;
        sT[k] = in ? tv[q] : P.Tamb;
        // per component: a select of two float4 values would go through scratch memory
        sY[k] = make_float4(in ? yv[q].x : aY.x, in ? yv[q].y : aY.y, in ? yv[q].z : aY.z, in ? yv[q].w : aY.w);
      }
    }
#pragma unroll
    for (int q = 0; q < QU; q++) {
      const int k = tid + NT * q;
      if (k < A_UX * A_UY) {  // outside: 0; below the floor: the no-slip ghost -U(i, 0)
        const int i = ax0 + k % A_UX, j = ay0 + k / A_UX;
        const bool in = i >= 0 && i <= nx && j < ny;
        const float su = in ? (j < 0 ? -uv[q] : uv[q]) : K(0);
        sU[k] = su;
        fastv = fastv && fabsf(su) < vlim;
      }
    }
#pragma unroll
    for (int q = 0; q < QV; q++) {
      const int k = tid + NT * q;
      if (k < A_VX * A_VY) {
        const int i = ax0 + k % A_VX, j = ay0 + k / A_VX;
        const bool in = i >= 0 && i < nx && j <= ny;
        const float sv = in ? vv[q] : K(0);
        sV[k] = sv;
        fastv = fastv && fabsf(sv) < vlim;
      }
    }
  }
  // every wave writes its own slot, so the slots need no clearing
  __shared__ int sSlow[NT / 64];
  {
    const unsigned long long slow = __ballot(!fastv);
    if ((tid & 63) == 0) sSlow[tid >> 6] = slow != 0ull;
  }
  __syncthreads();
  int slow = 0;
#pragma unroll
  for (int w = 0; w < NT / 64; w++) slow |= sSlow[w];

  // The displacement bound.  A trace's velocity samples interpolate staged
  // values, so |u|, |v| < vlim, and the trace moves its point by at most
  // dt*vlim/dx = 0.99 cell (plus a few ulps of rounding, far below the 1 %
  // margin), in x and in y.  With the window halos (AR = 3 staged, AH = 1
  // forward) a point of B lies within a cell of the tile and its footprints
  // within 1 + 1 + 1 cells; a point of C lies in the tile and its backward
  // footprint within 1 + 1 cells of it, inside the forward window.  So the
  // workgroup runs B and C without window tests (CK = false), and a
  // workgroup with a faster velocity (or a NaN) runs them tested, with the
  // global-memory redo.  Both forms do the same arithmetic on the same values.
  auto passes = [&](auto ck) {
    constexpr bool CK = decltype(ck)::value;
    Tile<CK> t;
    t.T.init(sT, ax0, ay0); t.Y.init(sY, ax0, ay0); t.U.init(sU, ax0, ay0); t.V.init(sV, ax0, ay0);
    t.Th.init(bT, bx0, by0); t.Yh.init(bY, bx0, by0); t.Uh.init(bU, bx0, by0); t.Vh.init(bV, bx0, by0);
    t.dx = P.dx;
    t.idx = K(1) / P.dx;

    // B. forward pass over the tile + AH, as the corrector's samplers see it.
    // The corrector's departure trace of a cell or face (+dt) is the forward
    // pass's trace of the same point, so each thread first does the forward
    // pass of the points it will correct in C and keeps their traces in
    // registers (B1); the rest of the window follows (B2).
    const float sO2i = K(1) / P.sO2;
    const bool ex = x0 + AX == nx, ey = y0 + AY == ny;
    constexpr int NQ = AX * AY + AY + AX;  // C's work items: the tile's cells (+ faces), then edge faces
    constexpr int NS = (NQ + NT - 1) / NT;
    // the points C corrects: item q -> (i, j) and which of its cell / U / V it owns
    auto item = [&](int q, int& i, int& j, bool& oc, bool& ou, bool& ov) {
      const bool tp = q < AX * AY, eu = q >= AX * AY && q < AX * AY + AY, ev = q >= AX * AY + AY && q < NQ;
      i = tp ? x0 + q % AX : (eu ? nx : x0 + (q - AX * AY - AY));
      j = tp ? y0 + q / AX : (eu ? y0 + (q - AX * AY) : ny);
      oc = tp && i < nx && j < ny;
      ou = (tp && i <= nx && j < ny) || (eu && ex && j < ny);
      ov = (tp && i < nx && j <= ny) || (ev && ey && i < nx);
    };
    // forward values of one point (i, j >= 0 inside the domain); d, okd: its trace
    auto fwd_cell = [&](int i, int j, float& th, float4& yh, float2& d, float2& w1, bool& okd) {
      okd = true;
      w1 = vel_cell(t, i, j);
      d = trace_from(t, at_cell(i, j), w1, dt, okd);
      bool ok = okd;
      th = t.T.bil(d.x - K(0.5), d.y - K(0.5), ok);
      yh = t.Y.bil(d.x - K(0.5), d.y - K(0.5), ok);
      if (!ok) g_fwd_cell(F, P, i, j, th, yh);
    };
    auto fwd_u = [&](int i, int j, float2& d, float2& w1, bool& okd) {
      okd = true;
      w1 = vel_uface(t, i, j);
      d = trace_from(t, at_uface(i, j), w1, dt, okd);
      bool ok = okd;
      float u = t.U.bil(d.x - K(0), d.y - K(0.5), ok);
      if (!ok) u = g_fwd_u(F, P, i, j);
      return u;
    };
    auto fwd_v = [&](int i, int j, float2& d, float2& w1, bool& okd) {
      okd = true;
      w1 = vel_vface(t, i, j);
      d = trace_from(t, at_vface(i, j), w1, dt, okd);
      bool ok = okd;
      float v = t.V.bil(d.x - K(0.5), d.y - K(0), ok);
      if (!ok) v = g_fwd_v(F, P, i, j);
      return v;
    };
    // the traces of one work item's cell and faces and their first
    // velocities, kept from B1 for C
    struct Slot {
      float2 dC, dU, dV;
      float2 wC, wU, wV;
      bool okC, okU, okV;
    };
    static_assert(NS == 3 && AX * AY == 2 * NT && AX + AY <= NT, "two tile slots and one edge slot per thread");
    Slot S0, S1;
    // B1: the tile rectangle's points of work items tid and tid + NT: the
    // forward pass of the points C corrects (traces kept), the outside values
    // for the rest (beyond the domain's right or top edge)
    auto b1 = [&](int q, Slot& S) {
      int i, j;
      bool oc, ou, ov;
      item(q, i, j, oc, ou, ov);
      S.dC = S.dU = S.dV = S.wC = S.wU = S.wV = make_float2(0.f, 0.f);
      S.okC = S.okU = S.okV = false;
      float th = P.Tamb;
      float4 yh = aY;
      if (oc) fwd_cell(i, j, th, yh, S.dC, S.wC, S.okC);
      const int qb = (j - by0) * B_CX + (i - bx0);
      bT[qb] = th;
      bY[qb] = yh;
      bU[(j - by0) * B_UX + (i - bx0)] = ou ? fwd_u(i, j, S.dU, S.wU, S.okU) : K(0);
      bV[(j - by0) * B_VX + (i - bx0)] = ov ? fwd_v(i, j, S.dV, S.wV, S.okV) : K(0);
    };
    b1(tid, S0);
    b1(tid + NT, S1);
    // B2: the ring of the window around the tile rectangle, enumerated without
    // gaps: r -> window (a, b) for a W x H window whose rectangle starts at
    // (AH, AH) with size AX x AY
    auto ring = [&](int r, int W, int& a, int& b) {
      const int top = AH * W, mid = AY * (W - AX);
      if (r < top) { a = r % W; b = r / W; }
      else if (r < top + mid) {
        const int c = (r - top) % (W - AX);
        a = c < AH ? c : c + AX;
        b = AH + (r - top) / (W - AX);
      } else { a = (r - top - mid) % W; b = AH + AY + (r - top - mid) / W; }
    };
    for (int r = tid; r < B_CX * B_CY - AX * AY; r += NT) {
      int a, b;
      ring(r, B_CX, a, b);
      const int i = bx0 + a, j = by0 + b;
      float th = P.Tamb;
      float4 yh = aY;
      if (!(i < 0 || i >= nx || j >= ny)) {
        float2 d, w1;
        bool okd;
        fwd_cell(i, max(j, 0), th, yh, d, w1, okd);
      }
      bT[b * B_CX + a] = th;
      bY[b * B_CX + a] = yh;
    }
    for (int r = tid; r < B_UX * B_UY - AX * AY; r += NT) {
      int a, b;
      ring(r, B_UX, a, b);
      const int i = bx0 + a, j = by0 + b;
      float u = K(0);
      if (!(i < 0 || i > nx || j >= ny)) {
        float2 d, w1;
        bool okd;
        u = fwd_u(i, max(j, 0), d, w1, okd);
        if (j < 0) u = -u;
      }
      bU[b * B_UX + a] = u;
    }
    for (int r = tid; r < B_VX * B_VY - AX * AY; r += NT) {
      int a, b;
      ring(r, B_VX, a, b);
      const int i = bx0 + a, j = by0 + b;
      float v = K(0);
      if (!(i < 0 || i >= nx || j > ny)) {
        float2 d, w1;
        bool okd;
        v = fwd_v(i, max(j, 0), d, w1, okd);
      }
      bV[b * B_VX + a] = v;
    }
    __syncthreads();

    // C. correction of the tile's cells and faces; the tile at the right (top)
    // domain edge also owns the U (V) faces on that edge
    auto c1 = [&](int q, const Slot& S) {
      int i, j;
      bool oc, ou, ov;
      item(q, i, j, oc, ou, ov);
      if (oc) {  // scalCorrect (:49-66)
        const int k = j * nx + i;
        bool ok = CK ? S.okC : true;
        const float2 d = S.dC;
        const float2 f = trace_from(t, at_cell(i, j), S.wC, -dt, ok);
        const float* pT = t.T.at(floorf(d.x - K(0.5)), floorf(d.y - K(0.5)), ok);
        const float4* pY = t.Y.at(floorf(d.x - K(0.5)), floorf(d.y - K(0.5)), ok);
        const float sbT = t.Th.bil(f.x - K(0.5), f.y - K(0.5), ok);
        const float4 sb = t.Yh.bil(f.x - K(0.5), f.y - K(0.5), ok);
        float To;
        float4 Yo;
        if (ok) {
          const float2 mT = mm4v(pT[0], pT[1], pT[A_CX], pT[A_CX + 1]);
          const float4 a = pY[0], c = pY[1], dd = pY[A_CX], e = pY[A_CX + 1];
          const float4 lo = make_float4(fminf(fminf(a.x, c.x), fminf(dd.x, e.x)), fminf(fminf(a.y, c.y), fminf(dd.y, e.y)),
                                        fminf(fminf(a.z, c.z), fminf(dd.z, e.z)), fminf(fminf(a.w, c.w), fminf(dd.w, e.w)));
          const float4 hi = make_float4(fmaxf(fmaxf(a.x, c.x), fmaxf(dd.x, e.x)), fmaxf(fmaxf(a.y, c.y), fmaxf(dd.y, e.y)),
                                        fmaxf(fmaxf(a.z, c.z), fmaxf(dd.z, e.z)), fmaxf(fmaxf(a.w, c.w), fmaxf(dd.w, e.w)));
          const float Thk = t.Th.own(i, j);
          const float4 Yhk = t.Yh.own(i, j), Yk = t.Y.own(i, j);
          const float Tc = Thk + K(0.5) * (t.T.own(i, j) - sbT);
          const float4 Yc = make_float4(Yhk.x + K(0.5) * (Yk.x - sb.x), Yhk.y + K(0.5) * (Yk.y - sb.y),
                                        Yhk.z + K(0.5) * (Yk.z - sb.z), Yhk.w + K(0.5) * (Yk.w - sb.w));
          const bool inT = Tc >= mT.x && Tc <= mT.y;
          const bool inY = Yc.x >= lo.x && Yc.y >= lo.y && Yc.z >= lo.z && Yc.w >= lo.w &&
                           Yc.x <= hi.x && Yc.y <= hi.y && Yc.z <= hi.z && Yc.w <= hi.w;
          const float mixC = fminf(Yc.x, div_rcp(Yc.y, P.sO2, sO2i));
          const float mixL = fminf(Yhk.x, div_rcp(Yhk.y, P.sO2, sO2i));
          const bool take = inT && inY && mixC <= mixL + K(1e-5);
          To = take ? Tc : Thk;
          Yo = take ? Yc : Yhk;
        } else {
          g_corr_cell(F, P, i, j, To, Yo);
        }
        F.T2[k] = To;
        F.Y2[k] = Yo;
      }
      if (ou) {  // velCorrect, U (:78-90)
        bool ok = CK ? S.okU : true;
        const float2 b = trace_from(t, at_uface(i, j), S.wU, -dt, ok);
        const float back = t.Uh.bil(b.x - K(0), b.y - K(0.5), ok);
        const float2 d = S.dU;
        const float* pU = t.U.at(floorf(d.x - K(0)), floorf(d.y - K(0.5)), ok);
        float u;
        if (ok) {
          const float2 mm = mm4v(pU[0], pU[1], pU[A_UX], pU[A_UX + 1]);
          u = clampf(t.Uh.own(i, j) + K(0.5) * (t.U.own(i, j) - back), mm.x, mm.y);
        } else {
          u = g_corr_u(F, P, i, j);
        }
        F.U2[j * (nx + 1) + i] = u;
      }
      if (ov) {  // velCorrect, V (:91-103)
        bool ok = CK ? S.okV : true;
        const float2 b = trace_from(t, at_vface(i, j), S.wV, -dt, ok);
        const float back = t.Vh.bil(b.x - K(0.5), b.y - K(0), ok);
        const float2 d = S.dV;
        const float* pV = t.V.at(floorf(d.x - K(0.5)), floorf(d.y - K(0)), ok);
        float v;
        if (ok) {
          const float2 mm = mm4v(pV[0], pV[1], pV[A_VX], pV[A_VX + 1]);
          v = clampf(t.Vh.own(i, j) + K(0.5) * (t.V.own(i, j) - back), mm.x, mm.y);
        } else {
          v = g_corr_v(F, P, i, j);
        }
        F.V2[j * nx + i] = v;
      }
    };
    c1(tid, S0);
    c1(tid + NT, S1);
    {  // the domain-edge faces: traced here
      int i, j;
      bool oc, ou, ov;
      item(tid + 2 * NT, i, j, oc, ou, ov);
      Slot S2;
      S2.dC = S2.dU = S2.dV = S2.wC = S2.wU = S2.wV = make_float2(0.f, 0.f);
      S2.okC = S2.okU = S2.okV = false;
      if (ou) { S2.okU = true; S2.wU = vel_uface(t, i, j); S2.dU = trace_from(t, at_uface(i, j), S2.wU, dt, S2.okU); }
      if (ov) { S2.okV = true; S2.wV = vel_vface(t, i, j); S2.dV = trace_from(t, at_vface(i, j), S2.wV, dt, S2.okV); }
      c1(tid + 2 * NT, S2);
    }
  };
  if (slow) passes(std::true_type{});
  else passes(std::false_type{});
}

__global__ __launch_bounds__(NT) void k_advect(Fields F) { AdvectOp::run(F, xcd_tile(blockIdx.x, gridDim.x)); }

void launch_advect(const Fields& F, hipStream_t s) {
  launch(k_advect, dim3(AdvectOp::tiles(F.nx, F.ny)), dim3(NT), 0, s, F);
}
void phase_advect(BatchPhase& ph, const Fields& F) {
  phase_add<AdvectOp>(ph, F, AdvectOp::tiles(F.nx, F.ny), dim3(NT), "advect");
}

}  // namespace cf
