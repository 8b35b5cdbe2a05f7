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
//   k_mg_bottom: every level of <= 1024 cells in one LDS-resident workgroup,
//                including the coarsest level's red-black sweeps
// with the reference's arithmetic and operation order in every update.
#include "cf_kernels.h"

namespace cf {

// ---- tiled down leg ------------------------------------------------------------
// Tile TX x TY fine cells per workgroup (TX, TY even), region = tile + halo.
template <int TX, int TY, int NU>
__global__ __launch_bounds__(256) void k_mg_down(Level f, Level c) {
  constexpr int H = 2 * NU + 1;
  constexpr int RX = TX + 2 * H, RY = TY + 2 * H, NR = RX * RY;
  constexpr int Q = (NR + 255) / 256;
  __shared__ float sp[NR];
  __shared__ float sres[TX * TY];
  const int nx = f.nx, ny = f.ny;
  const int gx = (nx + TX - 1) / TX, gy = (ny + TY - 1) / TY;
  const int t = xcd_tile(blockIdx.x, gx * gy);
  const int x0 = (t % gx) * TX, y0 = (t / gx) * TY;
  const int tid = threadIdx.x;
  const float h2 = f.h2;
  float cw[Q], ce[Q], cs[Q], cn[Q], cr[Q];
  int xs[Q], ys[Q];
  bool in[Q], upd[Q];
#pragma unroll
  for (int q = 0; q < Q; q++) {
    const int cidx = tid + 256 * q;
    const int rx = cidx % RX, ry = cidx / RX;
    const int x = x0 - H + rx, y = y0 - H + ry;
    xs[q] = x; ys[q] = y;
    in[q] = cidx < NR && x >= 0 && x < nx && y >= 0 && y < ny;
    // a cell on the region edge whose outward neighbour is a real cell cannot be updated
    upd[q] = in[q] && !(rx == 0 && x > 0) && !(rx == RX - 1 && x < nx - 1) && !(ry == 0 && y > 0) &&
             !(ry == RY - 1 && y < ny - 1);
    cw[q] = ce[q] = cs[q] = cn[q] = cr[q] = 0.f;
    if (in[q]) {
      cw[q] = f.ax[y * (nx + 1) + x]; ce[q] = f.ax[y * (nx + 1) + x + 1];
      cs[q] = f.ay[y * nx + x]; cn[q] = f.ay[(y + 1) * nx + x];
      cr[q] = f.rhs[y * nx + x];
      sp[cidx] = f.p[y * nx + x];
    } else if (cidx < NR) {
      sp[cidx] = 0.f;
    }
  }
  __syncthreads();
  for (int hs = 0; hs < 2 * NU; hs++) {
    const int col = hs & 1;  // red (0) then black (1), js/fluid.js:298-301
#pragma unroll
    for (int q = 0; q < Q; q++) {
      const int x = xs[q], y = ys[q];
      if (upd[q] && ((x + y) & 1) == col) {
        const int cidx = tid + 256 * q;
        float off = 0.f, diag = 0.f;
        if (x > 0) { off += cw[q] * sp[cidx - 1]; diag += cw[q]; } else { diag += K(2.0) * cw[q]; }
        if (x < nx - 1) { off += ce[q] * sp[cidx + 1]; diag += ce[q]; } else { diag += K(2.0) * ce[q]; }
        if (y > 0) { off += cs[q] * sp[cidx - RX]; diag += cs[q]; }
        if (y < ny - 1) { off += cn[q] * sp[cidx + RX]; diag += cn[q]; } else { diag += K(2.0) * cn[q]; }
        sp[cidx] = diag > K(0) ? (off - cr[q] * h2) / diag : K(0);
      }
    }
    __syncthreads();
  }
  // interior: write p, residual into sres (shaders/pressure.wgsl:96-101)
#pragma unroll
  for (int q = 0; q < Q; q++) {
    const int cidx = tid + 256 * q;
    const int rx = cidx % RX, ry = cidx / RX;
    if (in[q] && rx >= H && rx < H + TX && ry >= H && ry < H + TY) {
      const int x = xs[q], y = ys[q];
      float off = 0.f, diag = 0.f;
      if (x > 0) { off += cw[q] * sp[cidx - 1]; diag += cw[q]; } else { diag += K(2.0) * cw[q]; }
      if (x < nx - 1) { off += ce[q] * sp[cidx + 1]; diag += ce[q]; } else { diag += K(2.0) * ce[q]; }
      if (y > 0) { off += cs[q] * sp[cidx - RX]; diag += cs[q]; }
      if (y < ny - 1) { off += cn[q] * sp[cidx + RX]; diag += cn[q]; } else { diag += K(2.0) * cn[q]; }
      const float pk = sp[cidx];
      sres[(ry - H) * TX + (rx - H)] = diag <= K(0) ? K(0) : cr[q] - (off - diag * pk) / h2;
      f.p[y * nx + x] = pk;
    }
  }
  __syncthreads();
  // restriction to the coarse rhs, coarse p = 0 (:104-113)
  for (int k = tid; k < (TX / 2) * (TY / 2); k += 256) {
    const int li = k % (TX / 2), lj = k / (TX / 2);
    const int I = x0 / 2 + li, J = y0 / 2 + lj;
    if (I < c.nx && J < c.ny) {
      const float* r0 = sres + (2 * lj) * TX + 2 * li;
      const float r = r0[0] + r0[1] + r0[TX] + r0[TX + 1];
      c.rhs[J * c.nx + I] = K(0.25) * r;
      c.p[J * c.nx + I] = K(0);
    }
  }
}

// ---- tiled up leg ----------------------------------------------------------------
template <int TX, int TY, int NU>
__global__ __launch_bounds__(256) void k_mg_up(Level f, Level c) {
  constexpr int H = 2 * NU;
  constexpr int RX = TX + 2 * H, RY = TY + 2 * H, NR = RX * RY;
  constexpr int Q = (NR + 255) / 256;
  __shared__ float sp[NR];
  const int nx = f.nx, ny = f.ny;
  const int gx = (nx + TX - 1) / TX, gy = (ny + TY - 1) / TY;
  const int t = xcd_tile(blockIdx.x, gx * gy);
  const int x0 = (t % gx) * TX, y0 = (t / gx) * TY;
  const int tid = threadIdx.x;
  const float h2 = f.h2;
  const int nxc = c.nx, nyc = c.ny;
  float cw[Q], ce[Q], cs[Q], cn[Q], cr[Q];
  int xs[Q], ys[Q];
  bool in[Q], upd[Q];
#pragma unroll
  for (int q = 0; q < Q; q++) {
    const int cidx = tid + 256 * q;
    const int rx = cidx % RX, ry = cidx / RX;
    const int x = x0 - H + rx, y = y0 - H + ry;
    xs[q] = x; ys[q] = y;
    in[q] = cidx < NR && x >= 0 && x < nx && y >= 0 && y < ny;
    upd[q] = in[q] && !(rx == 0 && x > 0) && !(rx == RX - 1 && x < nx - 1) && !(ry == 0 && y > 0) &&
             !(ry == RY - 1 && y < ny - 1);
    cw[q] = ce[q] = cs[q] = cn[q] = cr[q] = 0.f;
    if (in[q]) {
      cw[q] = f.ax[y * (nx + 1) + x]; ce[q] = f.ax[y * (nx + 1) + x + 1];
      cs[q] = f.ay[y * nx + x]; cn[q] = f.ay[(y + 1) * nx + x];
      cr[q] = f.rhs[y * nx + x];
      // bilinear coarse correction, skipped in closed cells (:116-129)
      const int I = x / 2, J = y / 2;
      const int I2 = iclamp(I + ((x & 1) == 1 ? 1 : -1), 0, nxc - 1);
      const int J2 = iclamp(J + ((y & 1) == 1 ? 1 : -1), 0, nyc - 1);
      const float cc = K(0.5625) * c.p[J * nxc + I] + K(0.1875) * (c.p[J * nxc + I2] + c.p[J2 * nxc + I]) +
                       K(0.0625) * c.p[J2 * nxc + I2];
      float diag = 0.f;
      if (x > 0) diag += cw[q]; else diag += K(2.0) * cw[q];
      if (x < nx - 1) diag += ce[q]; else diag += K(2.0) * ce[q];
      if (y > 0) diag += cs[q];
      if (y < ny - 1) diag += cn[q]; else diag += K(2.0) * cn[q];
      sp[cidx] = f.p[y * nx + x] + (diag > K(0) ? cc : K(0));
    } else if (cidx < NR) {
      sp[cidx] = 0.f;
    }
  }
  __syncthreads();
  for (int hs = 0; hs < 2 * NU; hs++) {
    const int col = (hs & 1) ? 0 : 1;  // black (1) then red (0), js/fluid.js:307-310
#pragma unroll
    for (int q = 0; q < Q; q++) {
      const int x = xs[q], y = ys[q];
      if (upd[q] && ((x + y) & 1) == col) {
        const int cidx = tid + 256 * q;
        float off = 0.f, diag = 0.f;
        if (x > 0) { off += cw[q] * sp[cidx - 1]; diag += cw[q]; } else { diag += K(2.0) * cw[q]; }
        if (x < nx - 1) { off += ce[q] * sp[cidx + 1]; diag += ce[q]; } else { diag += K(2.0) * ce[q]; }
        if (y > 0) { off += cs[q] * sp[cidx - RX]; diag += cs[q]; }
        if (y < ny - 1) { off += cn[q] * sp[cidx + RX]; diag += cn[q]; } else { diag += K(2.0) * cn[q]; }
        sp[cidx] = diag > K(0) ? (off - cr[q] * h2) / diag : K(0);
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int q = 0; q < Q; q++) {
    const int cidx = tid + 256 * q;
    const int rx = cidx % RX, ry = cidx / RX;
    if (in[q] && rx >= H && rx < H + TX && ry >= H && ry < H + TY) f.p[ys[q] * nx + xs[q]] = sp[cidx];
  }
}

// ---- bottom levels in one workgroup ------------------------------------------------
// LDS layout per level: p[n], rhs[n], ax[(nx+1)ny], ay[nx(ny+1)].
struct BottomLevels {
  Level L[4];
  int off[4];  // float offset of each level's block in LDS
  int nlev;    // levels handled here (the last one is the coarsest)
  int nu, coarse_iters, first_is_level0;
};

// one half-sweep of a bottom level held in LDS; 16x16 threads walk rows and
// same-colour columns without integer division
__device__ __forceinline__ void b_relax(const Level& L, float* p, const float* rhs, const float* ax, const float* ay,
                                        int col) {
  const int nx = L.nx, ny = L.ny;
  const int half = (nx + 1) / 2;
  const float h2 = L.h2;
  for (int j = threadIdx.y; j < ny; j += 16) {
    const int par = (j + col) & 1;
    for (int ii = threadIdx.x; ii < half; ii += 16) {
      const int i = 2 * ii + par;
      if (i >= nx) continue;
      const float w = ax[j * (nx + 1) + i], e = ax[j * (nx + 1) + i + 1], s = ay[j * nx + i], n = ay[(j + 1) * nx + i];
      float off = 0.f, diag = 0.f;
      if (i > 0) { off += w * p[j * nx + i - 1]; diag += w; } else { diag += K(2.0) * w; }
      if (i < nx - 1) { off += e * p[j * nx + i + 1]; diag += e; } else { diag += K(2.0) * e; }
      if (j > 0) { off += s * p[(j - 1) * nx + i]; diag += s; }
      if (j < ny - 1) { off += n * p[(j + 1) * nx + i]; diag += n; } else { diag += K(2.0) * n; }
      const int k2 = j * nx + i;
      p[k2] = diag > K(0) ? (off - rhs[k2] * h2) / diag : K(0);
    }
  }
}

__device__ __forceinline__ float b_residual(const Level& L, const float* p, const float* rhs, const float* ax,
                                            const float* ay, int i, int j) {
  const int nx = L.nx, ny = L.ny;
  const float w = ax[j * (nx + 1) + i], e = ax[j * (nx + 1) + i + 1], s = ay[j * nx + i], n = ay[(j + 1) * nx + i];
  float off = 0.f, diag = 0.f;
  if (i > 0) { off += w * p[j * nx + i - 1]; diag += w; } else { diag += K(2.0) * w; }
  if (i < nx - 1) { off += e * p[j * nx + i + 1]; diag += e; } else { diag += K(2.0) * e; }
  if (j > 0) { off += s * p[(j - 1) * nx + i]; diag += s; }
  if (j < ny - 1) { off += n * p[(j + 1) * nx + i]; diag += n; } else { diag += K(2.0) * n; }
  if (diag <= K(0)) return K(0);
  return rhs[j * nx + i] - (off - diag * p[j * nx + i]) / L.h2;
}

__global__ __launch_bounds__(256) void k_mg_bottom(BottomLevels B) {
  extern __shared__ float lds[];
  const int T = 256;  // 16 x 16 threads
  const int tid = threadIdx.y * 16 + threadIdx.x;
  // load coefficients of every level and p, rhs of the first
  for (int l = 0; l < B.nlev; l++) {
    const Level& L = B.L[l];
    const int n = L.nx * L.ny, nax = (L.nx + 1) * L.ny, nay = L.nx * (L.ny + 1);
    float* base = lds + B.off[l];
    float *p = base, *rhs = base + n, *ax = rhs + n, *ay = ax + nax;
    for (int k = tid; k < nax; k += T) ax[k] = L.ax[k];
    for (int k = tid; k < nay; k += T) ay[k] = L.ay[k];
    if (l == 0)
      for (int k = tid; k < n; k += T) { p[k] = L.p[k]; rhs[k] = L.rhs[k]; }
  }
  __syncthreads();
  auto P = [&](int l) { return lds + B.off[l]; };
  auto R = [&](int l) { return lds + B.off[l] + B.L[l].nx * B.L[l].ny; };
  auto AX = [&](int l) { return R(l) + B.L[l].nx * B.L[l].ny; };
  auto AY = [&](int l) { return AX(l) + (B.L[l].nx + 1) * B.L[l].ny; };
  const int last = B.nlev - 1;
  // down (js/fluid.js:297-303)
  for (int l = 0; l < last; l++) {
    const Level& L = B.L[l];
    for (int q = 0; q < B.nu; q++) {
      b_relax(L, P(l), R(l), AX(l), AY(l), 0); __syncthreads();
      b_relax(L, P(l), R(l), AX(l), AY(l), 1); __syncthreads();
    }
    const Level& C = B.L[l + 1];
    float *pc = P(l + 1), *rc = R(l + 1);
    for (int k = tid; k < C.nx * C.ny; k += T) {
      const int I = k % C.nx, J = k / C.nx;
      const float r = b_residual(L, P(l), R(l), AX(l), AY(l), 2 * I, 2 * J) +
                      b_residual(L, P(l), R(l), AX(l), AY(l), 2 * I + 1, 2 * J) +
                      b_residual(L, P(l), R(l), AX(l), AY(l), 2 * I, 2 * J + 1) +
                      b_residual(L, P(l), R(l), AX(l), AY(l), 2 * I + 1, 2 * J + 1);
      rc[k] = K(0.25) * r;
      pc[k] = K(0);
    }
    __syncthreads();
  }
  // coarsest: sweeps from zero (shaders/pressure.wgsl:131-156)
  {
    const Level& L = B.L[last];
    float* p = P(last);
    for (int k = tid; k < L.nx * L.ny; k += T) p[k] = K(0);
    __syncthreads();
    for (int it = 0; it < B.coarse_iters; it++) {
      b_relax(L, p, R(last), AX(last), AY(last), 0); __syncthreads();
      b_relax(L, p, R(last), AX(last), AY(last), 1); __syncthreads();
    }
  }
  // up (js/fluid.js:305-311)
  for (int l = last - 1; l >= 0; l--) {
    const Level& L = B.L[l];
    const Level& C = B.L[l + 1];
    float *p = P(l), *pc = P(l + 1);
    const float *ax = AX(l), *ay = AY(l);
    for (int k = tid; k < L.nx * L.ny; k += T) {
      const int i = k % L.nx, j = k / L.nx;
      const int I = i / 2, J = j / 2;
      const int I2 = iclamp(I + ((i & 1) == 1 ? 1 : -1), 0, C.nx - 1);
      const int J2 = iclamp(J + ((j & 1) == 1 ? 1 : -1), 0, C.ny - 1);
      const float cc = K(0.5625) * pc[J * C.nx + I] + K(0.1875) * (pc[J * C.nx + I2] + pc[J2 * C.nx + I]) +
                       K(0.0625) * pc[J2 * C.nx + I2];
      float diag = 0.f;
      const int nx = L.nx;
      if (i > 0) diag += ax[j * (nx + 1) + i]; else diag += K(2.0) * ax[j * (nx + 1) + i];
      if (i < nx - 1) diag += ax[j * (nx + 1) + i + 1]; else diag += K(2.0) * ax[j * (nx + 1) + i + 1];
      if (j > 0) diag += ay[j * nx + i];
      if (j < L.ny - 1) diag += ay[(j + 1) * nx + i]; else diag += K(2.0) * ay[(j + 1) * nx + i];
      p[k] += diag > K(0) ? cc : K(0);
    }
    __syncthreads();
    for (int q = 0; q < B.nu; q++) {
      b_relax(L, p, R(l), ax, ay, 1); __syncthreads();
      b_relax(L, p, R(l), ax, ay, 0); __syncthreads();
    }
  }
  // write back p and rhs of every bottom level (the first level's rhs is unchanged)
  for (int l = 0; l < B.nlev; l++) {
    const Level& L = B.L[l];
    const int n = L.nx * L.ny;
    for (int k = tid; k < n; k += T) {
      L.p[k] = P(l)[k];
      if (l > 0) L.rhs[k] = R(l)[k];
    }
  }
}

// ---- launchers -------------------------------------------------------------------
static constexpr int MG_TX = 32, MG_TY = 32;

static inline unsigned ntiles(int nx, int ny, int tx, int ty) {
  return (unsigned)(((nx + tx - 1) / tx) * ((ny + ty - 1) / ty));
}

bool launch_mg_down(const Level& f, const Level& c, int nu, hipStream_t s) {
  const unsigned nb = ntiles(f.nx, f.ny, MG_TX, MG_TY);
  switch (nu) {
    case 1: hipLaunchKernelGGL((k_mg_down<MG_TX, MG_TY, 1>), dim3(nb), dim3(256), 0, s, f, c); return true;
    case 2: hipLaunchKernelGGL((k_mg_down<MG_TX, MG_TY, 2>), dim3(nb), dim3(256), 0, s, f, c); return true;
    case 3: hipLaunchKernelGGL((k_mg_down<MG_TX, MG_TY, 3>), dim3(nb), dim3(256), 0, s, f, c); return true;
    default: return false;
  }
}

bool launch_mg_up(const Level& f, const Level& c, int nu, hipStream_t s) {
  const unsigned nb = ntiles(f.nx, f.ny, MG_TX, MG_TY);
  switch (nu) {
    case 1: hipLaunchKernelGGL((k_mg_up<MG_TX, MG_TY, 1>), dim3(nb), dim3(256), 0, s, f, c); return true;
    case 2: hipLaunchKernelGGL((k_mg_up<MG_TX, MG_TY, 2>), dim3(nb), dim3(256), 0, s, f, c); return true;
    case 3: hipLaunchKernelGGL((k_mg_up<MG_TX, MG_TY, 3>), dim3(nb), dim3(256), 0, s, f, c); return true;
    default: return false;
  }
}

size_t mg_bottom_lds_bytes(const Level* L, int nlev) {
  size_t fl = 0;
  for (int l = 0; l < nlev; l++) fl += 2 * (size_t)L[l].nx * L[l].ny + (size_t)(L[l].nx + 1) * L[l].ny + (size_t)L[l].nx * (L[l].ny + 1);
  return fl * 4;
}

bool launch_mg_bottom(const Level* L, int nlev, int nu, int coarse_iters, hipStream_t s) {
  if (nlev < 1 || nlev > 4) return false;
  BottomLevels B;
  int off = 0;
  for (int l = 0; l < nlev; l++) {
    B.L[l] = L[l];
    B.off[l] = off;
    off += 2 * L[l].nx * L[l].ny + (L[l].nx + 1) * L[l].ny + L[l].nx * (L[l].ny + 1);
  }
  B.nlev = nlev; B.nu = nu; B.coarse_iters = coarse_iters; B.first_is_level0 = 0;
  const size_t bytes = (size_t)off * 4;
  if (bytes > 64 * 1024) return false;
  hipLaunchKernelGGL(k_mg_bottom, dim3(1), dim3(16, 16), bytes, s, B);
  return true;
}

}  // namespace cf
