// gfx950 kernels of the momentum, continuity, projection and energy phases
// of the LESwHT time step. Arithmetic follows the reference expression by
// expression (built with -ffp-contract=off) so the HIP results agree with the
// FP64 reference to rounding; the reference file:line is on each kernel.
#include "lw_kernels.h"
#include "lw_scan.h"

namespace lw {

#define BLK3 dim3(64, 4, 1)
// interior points, the unit of the per-launch byte counts below
static inline double pts(const Geo &g) { return (double)g.n1m * g.n2m * g.n3m; }
static inline dim3 grid3(int ni, int nj, int nk) {
  return dim3((unsigned)((ni + 63) / 64), (unsigned)((nj + 3) / 4), (unsigned)(nk > 0 ? nk : 1));
}

// ---------------------------------------------------------------------------
// reductions: wave64 shuffles, then one LDS pass; fixed shapes -> deterministic
// ---------------------------------------------------------------------------
__device__ inline double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

template <int NT>
__device__ inline double block_sum(double v, double *sh) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double r = 0.0;
  if (w == 0) {
    r = lane < NT / 64 ? sh[lane] : 0.0;
    r = wave_sum(r);
  }
  __syncthreads();
  return r;  // valid in thread 0
}

__device__ inline void atomic_max_nonneg(unsigned long long *a, double v) {
  atomicMax(a, (unsigned long long)__double_as_longlong(v));
}

// ---------------------------------------------------------------------------
// momentum right-hand side, slv_mmtm.f90:315-567, with the LES cross terms of
// rhssgs (sgs.f90:1136-1200) computed in place of the rhs1 round trip
// ---------------------------------------------------------------------------
// the right-hand side of a component at a cell and its new explicit term
// (the next substep's rk3o); the old rk3o stays where it is, and the launch
// functions swap rk3o and rk3oo after writing the new term into rk3oo
struct RhsOut {
  double rhs, rk3;
};

__device__ inline RhsOut rhs_u_at(const Geo &g, const Sub &sb, const Fld &f, int les, int i, int j, int k) {
  const double *u = f.u, *v = f.v, *w = f.w, *p = f.p;
  const long c = P(g, i, j, k);
  const long dx = 1, dy = g.NX, dz = g.NX * g.NY;
  const double fju = g.fixju[j], fjl = g.fixjl[j], fku = g.fixku[k], fkl = g.fixkl[k];
  const double un = 0.5 * (g.f2fy[j + 1] * u[c] + g.f2fy[j] * u[c + dy]) * g.c2cyi[j + 1] * (1. - fju) +
                    u[P(g, i, g.n2, k)] * fju;
  const double us = 0.5 * (g.f2fy[j] * u[c - dy] + g.f2fy[j - 1] * u[c]) * g.c2cyi[j] * (1. - fjl) +
                    u[P(g, i, 0, k)] * fjl;
  const double uc = 0.5 * (g.f2fz[k + 1] * u[c] + g.f2fz[k] * u[c + dz]) * g.c2czi[k + 1] * (1. - fku) +
                    u[P(g, i, j, g.n3)] * fku;
  const double uf = 0.5 * (g.f2fz[k] * u[c - dz] + g.f2fz[k - 1] * u[c]) * g.c2czi[k] * (1. - fkl) +
                    u[P(g, i, j, 0)] * fkl;
  const double vnx = 0.5 * (g.f2fx[i - 1] * v[c + dy] + g.f2fx[i] * v[c - dx + dy]) * g.c2cxi[i];
  const double vsx = 0.5 * (g.f2fx[i - 1] * v[c] + g.f2fx[i] * v[c - dx]) * g.c2cxi[i];
  const double wcx = 0.5 * (g.f2fx[i - 1] * w[c + dz] + g.f2fx[i] * w[c - dx + dz]) * g.c2cxi[i];
  const double wfx = 0.5 * (g.f2fx[i - 1] * w[c] + g.f2fx[i] * w[c - dx]) * g.c2cxi[i];
  const double anxe = 0.5 * (u[c + dx] + u[c]);
  const double anxw = 0.5 * (u[c] + u[c - dx]);
  const double anxx = (anxe * anxe - anxw * anxw) * g.c2cxi[i];
  const double anx1 = (un * vnx - us * vsx) * g.f2fyi[j];
  const double anx2 = (uc * wcx - uf * wfx) * g.f2fzi[k];
  double rk3x = -anxx - anx1 - anx2;
  const double alx1 = (u[c + dx] - u[c]) * g.f2fxi[i];
  const double alx2 = (u[c] - u[c - dx]) * g.f2fxi[i - 1];
  const double alx3 = (u[c + dy] - u[c]) * g.c2cyi[j + 1];
  const double alx4 = (u[c] - u[c - dy]) * g.c2cyi[j];
  const double alx5 = (u[c + dz] - u[c]) * g.c2czi[k + 1];
  const double alx6 = (u[c] - u[c - dz]) * g.c2czi[k];
  const double alxx = (alx1 - alx2) * g.c2cxi[i];
  const double alxy = (alx3 - alx4) * g.f2fyi[j];
  const double alxz = (alx5 - alx6) * g.f2fzi[k];
  double alx = 1. / sb.re * (alxx + alxy + alxz);
  if (les) {
    const double *nu = f.nu, *n2 = f.nu1[1], *n3 = f.nu1[2];
    const double talxx = (nu[c] * alx1 - nu[c - dx] * alx2) * g.c2cxi[i];
    const double talxy = (n3[c + dy] * alx3 - n3[c] * alx4) * g.f2fyi[j];
    const double talxz = (n2[c + dz] * alx5 - n2[c] * alx6) * g.f2fzi[k];
    alx = alx + 1.0 * (2. * talxx + talxy + talxz);
    const double rhsu2 = g.f2fyi[j] * g.c2cxi[i] *
                         (n3[c + dy] * (v[c + dy] - v[c - dx + dy]) - n3[c] * (v[c] - v[c - dx]));
    const double rhsu3 = g.f2fzi[k] * g.c2cxi[i] *
                         (n2[c + dz] * (w[c + dz] - w[c - dx + dz]) - n2[c] * (w[c] - w[c - dx]));
    rk3x = rk3x + 1.0 * (rhsu2 + rhsu3);
  }
  const double ro = f.rk3o[0][c];
  // the channel's mean pressure gradient pmi(0) (meanpg) when ICH != 0
  const double pmi0 = g.ich != 0 ? f.chan[0] : 0.0;
  return {sb.dt * (sb.gam * rk3x + sb.ro * ro + 2. * sb.alpha * alx -
                   2. * sb.alpha * (p[c] - p[c - dx]) * g.c2cxi[i] -
                   2. * sb.alpha * pmi0 * (double)(g.ich != 0 ? 1 : 0)),
          rk3x};
}

__device__ inline RhsOut rhs_v_at(const Geo &g, const Sub &sb, const Fld &f, int les, int i, int j, int k) {
  const double *u = f.u, *v = f.v, *w = f.w, *p = f.p;
  const long c = P(g, i, j, k);
  const long dx = 1, dy = g.NX, dz = g.NX * g.NY;
  const double fiu = g.fixiu[i], fil = g.fixil[i], fku = g.fixku[k], fkl = g.fixkl[k];
  const double ve = (0.5 * (g.f2fx[i + 1] * v[c] + g.f2fx[i] * v[c + dx]) * g.c2cxi[i + 1]) * (1. - fiu) +
                    v[P(g, g.n1, j, k)] * fiu;
  const double vw = (0.5 * (g.f2fx[i] * v[c - dx] + g.f2fx[i - 1] * v[c]) * g.c2cxi[i]) * (1. - fil) +
                    v[P(g, 0, j, k)] * fil;
  const double vc = (0.5 * (g.f2fz[k + 1] * v[c] + g.f2fz[k] * v[c + dz]) * g.c2czi[k + 1]) * (1. - fku) +
                    v[P(g, i, j, g.n3)] * fku;
  const double vf = (0.5 * (g.f2fz[k] * v[c - dz] + g.f2fz[k - 1] * v[c]) * g.c2czi[k]) * (1. - fkl) +
                    v[P(g, i, j, 0)] * fkl;
  const double uey = 0.5 * (g.f2fy[j - 1] * u[c + dx] + g.f2fy[j] * u[c + dx - dy]) * g.c2cyi[j];
  const double uwy = 0.5 * (g.f2fy[j - 1] * u[c] + g.f2fy[j] * u[c - dy]) * g.c2cyi[j];
  const double wcy = 0.5 * (g.f2fy[j - 1] * w[c + dz] + g.f2fy[j] * w[c - dy + dz]) * g.c2cyi[j];
  const double wfy = 0.5 * (g.f2fy[j - 1] * w[c] + g.f2fy[j] * w[c - dy]) * g.c2cyi[j];
  const double anyn = 0.5 * (v[c + dy] + v[c]);
  const double anys = 0.5 * (v[c] + v[c - dy]);
  const double anyy = g.c2cyi[j] * (anyn * anyn - anys * anys);
  const double any1 = (uey * ve - uwy * vw) * g.f2fxi[i];
  const double any2 = (vc * wcy - vf * wfy) * g.f2fzi[k];
  double rk3y = -anyy - any1 - any2;
  const double aly1 = (v[c + dx] - v[c]) * g.c2cxi[i + 1];
  const double aly2 = (v[c] - v[c - dx]) * g.c2cxi[i];
  const double aly3 = (v[c + dy] - v[c]) * g.f2fyi[j];
  const double aly4 = (v[c] - v[c - dy]) * g.f2fyi[j - 1];
  const double aly5 = (v[c + dz] - v[c]) * g.c2czi[k + 1];
  const double aly6 = (v[c] - v[c - dz]) * g.c2czi[k];
  const double alyx = (aly1 - aly2) * g.f2fxi[i];
  const double alyy = (aly3 - aly4) * g.c2cyi[j];
  const double alyz = (aly5 - aly6) * g.f2fzi[k];
  double aly = 1. / sb.re * (alyx + alyy + alyz);
  if (les) {
    const double *nu = f.nu, *n1 = f.nu1[0], *n3 = f.nu1[2];
    const double talyx = (n3[c + dx] * aly1 - n3[c] * aly2) * g.f2fxi[i];
    const double talyy = (nu[c] * aly3 - nu[c - dy] * aly4) * g.c2cyi[j];
    const double talyz = (n1[c + dz] * aly5 - n1[c] * aly6) * g.f2fzi[k];
    aly = aly + 1.0 * (talyx + 2. * talyy + talyz);
    const double rhsv1 = g.f2fxi[i] * g.c2cyi[j] *
                         (n3[c + dx] * (u[c + dx] - u[c + dx - dy]) - n3[c] * (u[c] - u[c - dy]));
    const double rhsv3 = g.f2fzi[k] * g.c2cyi[j] *
                         (n1[c + dz] * (w[c + dz] - w[c - dy + dz]) - n1[c] * (w[c] - w[c - dy]));
    rk3y = rk3y + 1.0 * (rhsv1 + rhsv3);
  }
  const double ro = f.rk3o[1][c];
  return {sb.dt * (sb.gam * rk3y + sb.ro * ro + 2. * sb.alpha * aly - 2. * sb.alpha * (p[c] - p[c - dy]) * g.c2cyi[j]),
          rk3y};
}

__device__ inline RhsOut rhs_w_at(const Geo &g, const Sub &sb, const Fld &f, int les, int i, int j, int k) {
  const double *u = f.u, *v = f.v, *w = f.w, *p = f.p;
  const long c = P(g, i, j, k);
  const long dx = 1, dy = g.NX, dz = g.NX * g.NY;
  const double fiu = g.fixiu[i], fil = g.fixil[i], fju = g.fixju[j], fjl = g.fixjl[j];
  const double we = 0.5 * (g.f2fx[i + 1] * w[c] + g.f2fx[i] * w[c + dx]) * g.c2cxi[i + 1] * (1. - fiu) +
                    fiu * w[P(g, g.n1, j, k)];
  const double ww = 0.5 * (g.f2fx[i] * w[c - dx] + g.f2fx[i - 1] * w[c]) * g.c2cxi[i] * (1. - fil) +
                    fil * w[P(g, 0, j, k)];
  const double wn = (0.5 * (g.f2fy[j + 1] * w[c] + g.f2fy[j] * w[c + dy]) * g.c2cyi[j + 1]) * (1. - fju) +
                    fju * w[P(g, i, g.n2, k)];
  const double ws = (0.5 * (g.f2fy[j] * w[c - dy] + g.f2fy[j - 1] * w[c]) * g.c2cyi[j]) * (1. - fjl) +
                    fjl * w[P(g, i, 0, k)];
  const double uez = 0.5 * (g.f2fz[k - 1] * u[c + dx] + g.f2fz[k] * u[c + dx - dz]) * g.c2czi[k];
  const double uwz = 0.5 * (g.f2fz[k - 1] * u[c] + g.f2fz[k] * u[c - dz]) * g.c2czi[k];
  const double vnz = 0.5 * (g.f2fz[k - 1] * v[c + dy] + g.f2fz[k] * v[c + dy - dz]) * g.c2czi[k];
  const double vsz = 0.5 * (g.f2fz[k - 1] * v[c] + g.f2fz[k] * v[c - dz]) * g.c2czi[k];
  const double anzc = 0.5 * (w[c + dz] + w[c]);
  const double anzf = 0.5 * (w[c] + w[c - dz]);
  const double anzz = (anzc * anzc - anzf * anzf) * g.c2czi[k];
  const double anz1 = (uez * we - uwz * ww) * g.f2fxi[i];
  const double anz2 = (vnz * wn - vsz * ws) * g.f2fyi[j];
  double rk3z = -anzz - anz1 - anz2;
  const double alz1 = (w[c + dx] - w[c]) * g.c2cxi[i + 1];
  const double alz2 = (w[c] - w[c - dx]) * g.c2cxi[i];
  const double alz3 = (w[c + dy] - w[c]) * g.c2cyi[j + 1];
  const double alz4 = (w[c] - w[c - dy]) * g.c2cyi[j];
  const double alz5 = (w[c + dz] - w[c]) * g.f2fzi[k];
  const double alz6 = (w[c] - w[c - dz]) * g.f2fzi[k - 1];
  const double alzx = (alz1 - alz2) * g.f2fxi[i];
  const double alzy = (alz3 - alz4) * g.f2fyi[j];
  const double alzz = (alz5 - alz6) * g.c2czi[k];
  double alz = 1. / sb.re * (alzx + alzy + alzz);
  if (les) {
    const double *nu = f.nu, *n1 = f.nu1[0], *n2 = f.nu1[1];
    const double talzx = (n2[c + dx] * alz1 - n2[c] * alz2) * g.f2fxi[i];
    const double talzy = (n1[c + dy] * alz3 - n1[c] * alz4) * g.f2fyi[j];
    const double talzz = (nu[c] * alz5 - nu[c - dz] * alz6) * g.c2czi[k];
    alz = alz + 1.0 * (talzx + talzy + 2. * talzz);
    const double rhsw1 = g.f2fxi[i] * g.c2czi[k] *
                         (n2[c + dx] * (u[c + dx] - u[c + dx - dz]) - n2[c] * (u[c] - u[c - dz]));
    const double rhsw2 = g.f2fyi[j] * g.c2czi[k] *
                         (n1[c + dy] * (v[c + dy] - v[c + dy - dz]) - n1[c] * (v[c] - v[c - dz]));
    rk3z = rk3z + 1.0 * (rhsw1 + rhsw2);
  }
  const double ro = f.rk3o[2][c];
  return {sb.dt * (sb.gam * rk3z + sb.ro * ro + 2. * sb.alpha * alz - 2. * sb.alpha * (p[c] - p[c - dz]) * g.c2czi[k]),
          rk3z};
}

// one component per launch (LW_RHS_FUSED=0)
LW_KERNEL void k_rhs_u(LW_KP(1024, 0), const Geo &g, const Sub &sb, const Fld &f, int les) {
  const int i = g.ib + blockIdx.x * 64 + threadIdx.x;
  const int j = 1 + blockIdx.y * 4 + threadIdx.y;
  const int k = 1 + blockIdx.z;
  if (i > g.n1m || j > g.n2m) return;
  const RhsOut o = rhs_u_at(g, sb, f, les, i, j, k);
  const long c = P(g, i, j, k);
  f.rhs[0][c] = o.rhs;
  f.rk3oo[0][c] = o.rk3;
}

LW_KERNEL void k_rhs_v(LW_KP(1024, 0), const Geo &g, const Sub &sb, const Fld &f, int les) {
  const int i = 1 + blockIdx.x * 64 + threadIdx.x;
  const int j = g.jb + blockIdx.y * 4 + threadIdx.y;
  const int k = 1 + blockIdx.z;
  if (i > g.n1m || j > g.n2m) return;
  const RhsOut o = rhs_v_at(g, sb, f, les, i, j, k);
  const long c = P(g, i, j, k);
  f.rhs[1][c] = o.rhs;
  f.rk3oo[1][c] = o.rk3;
}

LW_KERNEL void k_rhs_w(LW_KP(1024, 0), const Geo &g, const Sub &sb, const Fld &f, int les) {
  const int i = 1 + blockIdx.x * 64 + threadIdx.x;
  const int j = 1 + blockIdx.y * 4 + threadIdx.y;
  const int k = g.kb + blockIdx.z;
  if (i > g.n1m || j > g.n2m) return;
  const RhsOut o = rhs_w_at(g, sb, f, les, i, j, k);
  const long c = P(g, i, j, k);
  f.rhs[2][c] = o.rhs;
  f.rk3oo[2][c] = o.rk3;
}

// the three components in one pass over the cells: every load of u, v, w,
// p and the viscosities serves all three. Each component is evaluated at
// every interior cell, also at i = 1 < ib, j = 1 < jb and k = 1 < kb, where
// its stencil stays inside the arrays (ghost layer 0 and metric index 0, as
// on a periodic grid), and stored only where its own kernel would store it.
// 150 VGPRs, three waves per SIMD: at 384x256x64 LES 0.36 ms per launch
// against 0.46 for the three launches (limiting it to 128 registers spills
// and takes 0.54 ms).
LW_KERNEL void k_rhs_uvw(LW_KP(256, 0), const Geo &g, const Sub &sb, const Fld &f, int les) {
  const int i = 1 + blockIdx.x * 64 + threadIdx.x;
  const int j = 1 + blockIdx.y * 4 + threadIdx.y;
  const int k = 1 + blockIdx.z;
  if (i > g.n1m || j > g.n2m) return;
  const RhsOut ou = rhs_u_at(g, sb, f, les, i, j, k);
  const RhsOut ov = rhs_v_at(g, sb, f, les, i, j, k);
  const RhsOut ow = rhs_w_at(g, sb, f, les, i, j, k);
  const long c = P(g, i, j, k);
  if (i >= g.ib) {
    f.rhs[0][c] = ou.rhs;
    f.rk3oo[0][c] = ou.rk3;
  }
  if (j >= g.jb) {
    f.rhs[1][c] = ov.rhs;
    f.rk3oo[1][c] = ov.rk3;
  }
  if (k >= g.kb) {
    f.rhs[2][c] = ow.rhs;
    f.rk3oo[2][c] = ow.rk3;
  }
}

static bool rhs_fused_enabled() {
  const char *e = getenv("LW_RHS_FUSED");
  return !(e && atoi(e) == 0);
}

void launch_rhs_mom(hipStream_t s, const Geo &g, const Sub &sb, Fld &f, int les) {
  if (g.n3m != 1 && rhs_fused_enabled()) {
    // reads u v w p, three rk3o (+ four viscosities with LES), writes three
    // rhs and three new terms
    LW_LAUNCH("k_rhs_uvw", pts(g) * 8.0 * (les ? 17 : 13), 0, k_rhs_uvw, grid3(g.n1m, g.n2m, g.n3m), BLK3, 0, s, g,
              sb, f, les);
  } else {
    // reads u v w p rk3o (+ three viscosities with LES), writes rhs and the
    // new term
    const double b = pts(g) * 8.0 * (les ? 10 : 7);
    LW_LAUNCH("k_rhs_u", b, 0, k_rhs_u, grid3(g.n1m - g.ib + 1, g.n2m, g.n3m), BLK3, 0, s, g, sb, f, les);
    LW_LAUNCH("k_rhs_v", b, 0, k_rhs_v, grid3(g.n1m, g.n2m - g.jb + 1, g.n3m), BLK3, 0, s, g, sb, f, les);
    if (g.n3m != 1)
      LW_LAUNCH("k_rhs_w", b, 0, k_rhs_w, grid3(g.n1m, g.n2m, g.n3m - g.kb + 1), BLK3, 0, s, g, sb, f, les);
  }
  for (int c = 0; c < 3; ++c) std::swap(f.rk3o[c], f.rk3oo[c]);
}

// ---------------------------------------------------------------------------
// immersed-boundary forcing, slv_mmtm.f90:569-819 and slv_engy.f90:201-306:
// phase 1 gathers the 27-point targets from the unmodified rhs1 (the
// reference's rhstmp copy), phase 2 scatters; inner points force to zero.
// ---------------------------------------------------------------------------
__device__ inline void reindex_dev(int n, int nm, int &ip, int &ipp) {
  if (ipp == n) ipp = 1;
  else if (ipp == n + 1) { ip = 1; ipp = 2; }
  else if (ipp == 0) ipp = nm;
  else if (ipp == -1) { ip = nm; ipp = nm - 1; }
}

__device__ inline void ibm_gather_at(const Geo &g, const IBComp &ib, const double *fld, const double *rhs, int n) {
  const int ii = ib.ifc[n], jj = ib.jfc[n], kk = ib.kfc[n];
  const int *ix = ib.indx + 3 * n;
  int ip = ii + ix[0], jp = jj + ix[1], kp = kk + ix[2];
  int ipp = ii + ix[0] * 2, jpp = jj + ix[1] * 2, kpp = kk + ix[2] * 2;
  if (ipp >= g.n1 || ipp <= 0) reindex_dev(g.n1, g.n1m, ip, ipp);
  if (kpp >= g.n3 || kpp <= 0) reindex_dev(g.n3, g.n3m, kp, kpp);
  const int is[3] = {ii, ip, ipp}, js[3] = {jj, jp, jpp}, ks[3] = {kk, kp, kpp};
  const double *gf = ib.geomfac + 27 * (long)n;
  double targ = gf[0] * 0.0;
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b)
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        if (a == 0 && b == 0 && d == 0) continue;
        const long q = P(g, is[a], js[b], ks[d]);
        targ = targ + gf[9 * a + 3 * b + d] * (fld[q] + rhs[q]);
      }
  ib.targ[n] = targ;
}

__device__ inline void ibm_scatter_at(const Geo &g, const IBComp &ib, const double *fld, double *rhs, double dti,
                                      int has_dudtr, int n) {
  const long q = P(g, ib.ifc[n], ib.jfc[n], ib.kfc[n]);
  const double targ = n < ib.nintp ? ib.targ[n] : 0.0;
  const double fv = fld[q];
  const double fc = (targ - fv - rhs[q]) * dti;
  ib.fcv[n] = fc;
  if (has_dudtr) ib.dudtr[n] = (targ - fv) * dti;
  rhs[q] = targ - fv;
  ib.fcvavg[n] = ib.fcvavg[n] + fc;
}

LW_KERNEL void k_ibm_gather(LW_KP(1024, 0), const Geo &g, const IBComp &ib, const double *fld, const double *rhs) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n < ib.nintp) ibm_gather_at(g, ib, fld, rhs, n);
}

LW_KERNEL void k_ibm_scatter(LW_KP(1024, 0), const Geo &g, const IBComp &ib, const double *fld, double *rhs, double dti,
                             int has_dudtr) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n < ib.nbody) ibm_scatter_at(g, ib, fld, rhs, dti, has_dudtr, n);
}

// the three momentum components in one launch per phase (blockIdx.y is the
// component): a component's gather reads only its own field and rhs, and its
// scatter writes only its own rhs, so gathering all three before scattering
// any gives the per-component order's results
LW_KERNEL void k_ibm_gather3(LW_KP(256, 0), const Geo &g, const IBComp &i0, const IBComp &i1, const IBComp &i2,
                             const Fld &f) {
  // a switch with constant component indices: selecting among the
  // arguments at run time would copy them to scratch memory
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  switch (blockIdx.y) {
    case 0: if (n < i0.nintp) ibm_gather_at(g, i0, f.u, f.rhs[0], n); break;
    case 1: if (n < i1.nintp) ibm_gather_at(g, i1, f.v, f.rhs[1], n); break;
    default: if (n < i2.nintp) ibm_gather_at(g, i2, f.w, f.rhs[2], n); break;
  }
}

LW_KERNEL void k_ibm_scatter3(LW_KP(256, 0), const Geo &g, const IBComp &i0, const IBComp &i1, const IBComp &i2,
                              const Fld &f, double dti) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  switch (blockIdx.y) {
    case 0: if (n < i0.nbody) ibm_scatter_at(g, i0, f.u, f.rhs[0], dti, 1, n); break;
    case 1: if (n < i1.nbody) ibm_scatter_at(g, i1, f.v, f.rhs[1], dti, 1, n); break;
    default: if (n < i2.nbody) ibm_scatter_at(g, i2, f.w, f.rhs[2], dti, 1, n); break;
  }
}

void launch_ibm_mom(hipStream_t s, const Geo &g, const Sub &sb, const Fld &f, IBComp *ib) {
  const double dti = 1. / sb.dt;
  int ni = 0, nb = 0;
  double bi = 0, bb = 0;
  for (int c = 0; c < 3; ++c) {
    ni = std::max(ni, ib[c].nintp);
    nb = std::max(nb, ib[c].nbody);
    bi += ib[c].nintp * (27.0 * 16 + 3 * 4 + 8 + 8);
    bb += ib[c].nbody * (3 * 4 + 8 + 8 + 8 + 8);
  }
  if (ni > 0)
    LW_LAUNCH("k_ibm_gather", bi, 0, k_ibm_gather3, dim3((ni + 255) / 256, 3), dim3(256), 0, s, g, ib[0], ib[1], ib[2],
              f);
  if (nb > 0)
    LW_LAUNCH("k_ibm_scatter", bb, 0, k_ibm_scatter3, dim3((nb + 255) / 256, 3), dim3(256), 0, s, g, ib[0], ib[1], ib[2],
              f, dti);
}

void launch_ibm(hipStream_t s, const Geo &g, const Sub &sb, const Fld &f, IBComp *ib, int comp) {
  const double *fld = comp == 0 ? f.u : comp == 1 ? f.v : comp == 2 ? f.w : f.t;
  const double dti = 1. / sb.dt;
  IBComp &c = ib[comp];
  if (c.nintp > 0)
    LW_LAUNCH("k_ibm_gather", c.nintp * (27.0 * 16 + 3 * 4 + 8 + 8), 0, k_ibm_gather, dim3((c.nintp + 255) / 256),
              dim3(256), 0, s, g, c, fld, f.rhs[comp]);
  if (c.nbody > 0)
    LW_LAUNCH("k_ibm_scatter", c.nbody * (3 * 4 + 8 + 8 + 8 + 8), 0, k_ibm_scatter, dim3((c.nbody + 255) / 256),
              dim3(256), 0, s, g, c, fld, f.rhs[comp], dti, comp < 3 ? 1 : 0);
}

// ---------------------------------------------------------------------------
// convective outlet, slv_mmtm.f90:873-988 (convbc) + 991-1041 (rhsincorpbc).
// Three dependent reductions over boundary planes: each reduction kernel
// writes kObc block partials to f.obc, and every workgroup of the kernel that
// needs the total re-sums the partials in one fixed order, so all of them hold
// bit-identical totals without a one-workgroup launch in between. The
// reference sums each plane serially; the fixed tree differs from it only by
// rounding.
// ---------------------------------------------------------------------------
constexpr int kObc = 64;  // workgroups of the outlet-plane kernels
constexpr int kObcT = 256;
// f.obc layout: [0..3][kObc] flux partials (u inlet, v walls, w ends, u
// outlet), [4][kObc] outflow partials

__device__ inline double obc_total(const double *part, double *sh) {
  const double v = threadIdx.x < kObc ? part[threadIdx.x] : 0.0;
  const double r = block_sum<kObcT>(v, sh);
  if (threadIdx.x == 0) sh[kObcT / 64] = r;
  __syncthreads();
  const double out = sh[kObcT / 64];
  __syncthreads();
  return out;
}

// boundary fluxes: qin = inlet + (bottom - top) + (front - back), and the
// outlet-plane mean of u
LW_KERNEL void k_obc_flux(LW_KP(kObcT, 0), const Geo &g, const Fld &f) {
  __shared__ double sh[kObcT / 64 + 1];
  const int n1 = g.n1, n1m = g.n1m, n2m = g.n2m, n3m = g.n3m, n2 = g.n2, n3 = g.n3;
  const int str = kObc * kObcT, q0 = blockIdx.x * kObcT + threadIdx.x;
  const double *u = f.u, *v = f.v, *w = f.w;
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  for (int q = q0; q < n2m * n3m; q += str) {
    const int j = 1 + q % n2m, k = 1 + q / n2m;
    a0 += u[P(g, 1, j, k)] * g.f2fy[j] * g.f2fz[k];
    a3 += u[P(g, n1, j, k)] * g.f2fy[j] * g.f2fz[k];
  }
  for (int q = q0; q < n1m * n3m; q += str) {
    const int i = 1 + q % n1m, k = 1 + q / n1m;
    a1 += (v[P(g, i, 1, k)] - v[P(g, i, n2, k)]) * g.f2fx[i] * g.f2fz[k];
  }
  for (int q = q0; q < n1m * n2m; q += str) {
    const int i = 1 + q % n1m, j = 1 + q / n1m;
    a2 += (w[P(g, i, j, 1)] - w[P(g, i, j, n3)]) * g.f2fx[i] * g.f2fy[j];
  }
  const double s0 = block_sum<kObcT>(a0, sh), s1 = block_sum<kObcT>(a1, sh), s2 = block_sum<kObcT>(a2, sh),
               s3 = block_sum<kObcT>(a3, sh);
  if (threadIdx.x == 0) {
    f.obc[0 * kObc + blockIdx.x] = s0;
    f.obc[1 * kObc + blockIdx.x] = s1;
    f.obc[2 * kObc + blockIdx.x] = s2;
    f.obc[3 * kObc + blockIdx.x] = s3;
  }
}

struct ObcIn {
  double qin, ubar;
};
__device__ inline ObcIn obc_in(const Geo &g, const Fld &f, double *sh) {
  ObcIn r;
  const double s0 = obc_total(f.obc, sh), s1 = obc_total(f.obc + kObc, sh), s2 = obc_total(f.obc + 2 * kObc, sh),
               s3 = obc_total(f.obc + 3 * kObc, sh);
  r.qin = s0;
  r.qin = r.qin + s1;
  r.qin = r.qin + s2;
  r.ubar = s3 / (g.yl * g.zl);
  return r;
}

// outlet velocities before the mass correction, and the outflow partials
LW_KERNEL void k_obc_out_mom(LW_KP(kObcT, 0), const Geo &g, const Sub &sb, const Fld &f) {
  __shared__ double sh[kObcT / 64 + 1];
  const ObcIn in = obc_in(g, f, sh);
  const int n1 = g.n1, n1m = g.n1m, n2m = g.n2m, n3m = g.n3m;
  const double *u = f.u, *v = f.v, *w = f.w;
  const double ucoef = in.ubar * sb.dtconst * g.f2fxi[n1m];
  const double vwcoef = in.ubar * sb.dtconst * g.c2cxi[n1];
  double a = 0.0;
  for (int q = blockIdx.x * kObcT + threadIdx.x; q < n2m * n3m; q += kObc * kObcT) {
    const int j = 1 + q % n2m, k = 1 + q / n2m;
    const long pb = j + g.NY * k;
    double uo, vo, wo;
    if (f.mask[P(g, n1m, j, k)] & 64) {  // outlet cell solid: zero gradient
      uo = u[P(g, n1m, j, k)];
      vo = v[P(g, n1m, j, k)];
      wo = w[P(g, n1m, j, k)];
    } else {
      uo = u[P(g, n1, j, k)] - ucoef * (u[P(g, n1, j, k)] - u[P(g, n1m, j, k)]);
      vo = v[P(g, n1, j, k)] - vwcoef * (v[P(g, n1, j, k)] - v[P(g, n1m, j, k)]);
      wo = w[P(g, n1, j, k)] - vwcoef * (w[P(g, n1, j, k)] - w[P(g, n1m, j, k)]);
    }
    f.uout[pb] = uo;
    f.vout[pb] = vo;
    f.wout[pb] = wo;
    a += uo * g.f2fy[j] * g.f2fz[k];
  }
  const double s = block_sum<kObcT>(a, sh);
  if (threadIdx.x == 0) f.obc[4 * kObc + blockIdx.x] = s;
  (void)n3m;
}

// mass correction qin/qout, the outlet deltas, and rhsincorpbc (ICH = 0)
LW_KERNEL void k_obc_apply_mom(LW_KP(kObcT, 0), const Geo &g, const Sub &sb, const Fld &f, int les) {
  __shared__ double sh[kObcT / 64 + 1];
  const ObcIn in = obc_in(g, f, sh);
  const double qout = obc_total(f.obc + 4 * kObc, sh);
  const double qratio = in.qin / qout;
  const int n1 = g.n1, n1m = g.n1m, n2m = g.n2m, n3m = g.n3m;
  const double *u = f.u, *v = f.v, *w = f.w;
  const double cre = les ? sb.re : 0., cre2 = les ? 2. * sb.re : 0.;
  for (int q = blockIdx.x * kObcT + threadIdx.x; q < n2m * n3m; q += kObc * kObcT) {
    const int j = 1 + q % n2m, k = 1 + q / n2m;
    const long pb = j + g.NY * k;
    double uo = f.uout[pb], vo = f.vout[pb], wo = f.wout[pb];
    if (!(f.mask[P(g, n1m, j, k)] & 64)) {
      uo = uo * qratio;
      vo = vo * qratio;
      wo = wo * qratio;
      f.uout[pb] = uo;
      f.vout[pb] = vo;
      f.wout[pb] = wo;
    }
    const double du = uo - u[P(g, n1, j, k)], dv = vo - v[P(g, n1, j, k)], dw = wo - w[P(g, n1, j, k)];
    f.duout[pb] = du;
    f.dvout[pb] = dv;
    f.dwout[pb] = dw;
    const long c = P(g, n1m, j, k);
    f.rhs[0][c] = f.rhs[0][c] - sb.acoef * g.ciu[n1m] * (1. + cre2 * f.nu[c]) * du;
    if (j >= 2)
      f.rhs[1][c] = f.rhs[1][c] - sb.acoef * g.civw[n1m] * (1. + cre * f.nu1[2][P(g, n1, j, k)]) * dv;
    if (k >= g.kb)
      f.rhs[2][c] = f.rhs[2][c] - sb.acoef * g.civw[n1m] * (1. + cre * f.nu1[1][P(g, n1, j, k)]) * dw;
  }
  (void)n3m;
}

// z periodic ghosts of the outlet deltas
LW_KERNEL void k_obc_ghost(LW_KP(1024, 0), const Geo &g, double *d0, double *d1, double *d2) {
  const int j = blockIdx.x * 64 + threadIdx.x;
  if (j > g.n2) return;
  double *d[3] = {d0, d1, d2};
  for (int c = 0; c < 3; ++c) {
    if (!d[c]) continue;
    d[c][j + g.NY * g.n3] = d[c][j + g.NY * 1];
    d[c][j] = d[c][j + g.NY * g.n3m];
  }
}

static void launch_obc_flux(hipStream_t s, const Geo &g, const Fld &f) {
  const double b = ((double)g.n2m * g.n3m * 2 + (double)g.n1m * g.n3m * 2 + (double)g.n1m * g.n2m * 2) * 8.0;
  LW_LAUNCH("k_obc_flux", b, 0, k_obc_flux, dim3(kObc), dim3(kObcT), 0, s, g, f);
}

// constant flow rate (ICH = 1): no outlet, its velocities are zero
// (slv_mmtm.f90:980-984)
LW_KERNEL void k_obc_zero(LW_KP(1024, 0), const Geo &g, const Fld &f) {
  const long q = blockIdx.x * 256L + threadIdx.x;
  if (q >= (long)g.NY * g.NZ) return;
  f.uout[q] = 0.0;
  f.vout[q] = 0.0;
  f.wout[q] = 0.0;
}

void launch_convbc_mom(hipStream_t s, const Geo &g, const Sub &sb, const Fld &f, const Body &, int les) {
  const double pl = (double)g.n2m * g.n3m;
  if (g.ich == 1) {
    const long n = (long)g.NY * g.NZ;
    LW_LAUNCH("k_obc_zero", n * 24.0, 0, k_obc_zero, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, g, f);
    return;
  }
  launch_obc_flux(s, g, f);
  LW_LAUNCH("k_obc_out_mom", pl * 8.0 * 10, 0, k_obc_out_mom, dim3(kObc), dim3(kObcT), 0, s, g, sb, f);
  LW_LAUNCH("k_obc_apply_mom", pl * 8.0 * 20, 0, k_obc_apply_mom, dim3(kObc), dim3(kObcT), 0, s, g, sb, f, les);
  LW_LAUNCH("k_obc_ghost", (double)(g.n2 + 1) * 8.0 * 12, 0, k_obc_ghost, dim3((g.n2 + 64) / 64), dim3(64), 0, s, g,
            f.duout, f.dvout, f.dwout);
}

// ---------------------------------------------------------------------------
// ADI line solves, slv_mmtm.f90:1044-1641 and slv_engy.f90:460-609.
// One thread per line; for the z and y sweeps neighbouring threads hold
// neighbouring i. C = 0,1,2 momentum component, 3 temperature.
// ---------------------------------------------------------------------------
// Coefficients of a line (the LHS assembly of slv_mmtm.f90:1044-1362 and
// slv_engy.f90:460-609): a = am*(1 + cc*vis(lower face)),
// c = cm*(1 + cc*vis(upper face)), b = bd - a - c, right-hand side bd*rhs,
// with bd = 1/acoef (times Pr for temperature). LineCf names, per direction
// D (0 x, 1 y, 2 z) and component C, the viscosity-like array, the offset
// `lo` (-1 or 0, in points along the line) of its lower-face value, the
// metric arrays and the multipliers.
struct LineCf {
  const double *vis, *am, *cm;
  int lo;
  double cc, bd;
};

template <int D, int C>
__device__ inline LineCf line_cf(const Geo &g, const Fld &f, const Sub &sb, double cre, double cre2) {
  LineCf l;
  l.bd = C == 3 ? sb.acoefi * sb.pr : sb.acoefi;
  l.lo = 0;
  l.cc = cre;
  if (D == 0) {
    l.am = C == 0 ? g.aiu : g.aivw;
    l.cm = C == 0 ? g.ciu : g.civw;
    if (C == 0) {
      l.vis = f.nu;
      l.lo = -1;
      l.cc = cre2;
    } else {
      l.vis = C == 1 ? f.nu1[2] : C == 2 ? f.nu1[1] : f.al1[0];
    }
  } else if (D == 1) {
    l.am = C == 1 ? g.ajv : g.ajuw;
    l.cm = C == 1 ? g.cjv : g.cjuw;
    if (C == 1) {
      l.vis = f.nu;
      l.lo = -1;
      l.cc = cre2;
    } else {
      l.vis = C == 0 ? f.nu1[2] : C == 2 ? f.nu1[0] : f.al1[2];
    }
  } else {
    l.am = C == 2 ? g.akw : g.akuv;
    l.cm = C == 2 ? g.ckw : g.ckuv;
    if (C == 2) {
      l.vis = f.nu;
      l.lo = -1;
      l.cc = cre2;
    } else {
      l.vis = C == 0 ? f.nu1[1] : C == 1 ? f.nu1[0] : f.al1[1];
    }
  }
  return l;
}

// Each thread solves one line. The loads of a chunk of kAdiCh points are
// issued together ahead of the recurrence over them (the recurrence itself is
// the reference's, operation for operation), so a line waits for memory once
// per chunk instead of once per point.
constexpr int kAdiCh = 8;

// non-periodic line along stride `st` from index lo_ to n (trdiag1/trdiag2,
// slv_mmtm.f90:1365-1455): forward elimination into gm/r; the coefficient of
// index l uses metrics am[l], cm[l]
__device__ inline double line_forward(const LineCf &l, double *r, double *gm, long q0, long st, int lo_, int n) {
  double bet, up, cprev;
  {
    const long q = q0 + lo_ * st;
    const double a = l.am[lo_] * (1. + l.cc * l.vis[q + l.lo * st]);
    const double c = l.cm[lo_] * (1. + l.cc * l.vis[q + (l.lo + 1) * st]);
    const double b = l.bd - a - c;
    bet = 1. / b;
    up = (l.bd * r[q]) * bet;
    cprev = c;
    r[q] = up;
  }
  // software pipeline: the next chunk's loads are in flight while this
  // chunk's recurrence runs. The loads run past the line's end (values
  // unused): into the next plane of the padded arrays, whose planes k + 1 exist
  // for every line (k <= n3m), and into the padded metric arrays. Addresses
  // advance by pointer steps, so no per-point index arithmetic is needed.
  const double *pv = l.vis + q0 + (long)(lo_ + 1 + l.lo) * st;
  const double *pam = l.am + lo_ + 1, *pcm = l.cm + lo_ + 1;
  double *pr = r + q0 + (long)(lo_ + 1) * st, *pg = gm + q0 + (long)(lo_ + 1) * st;
  double v[kAdiCh + 1], rv[kAdiCh], am[kAdiCh], cm[kAdiCh];
  auto load = [&](const double *lv, const double *lr, const double *la, const double *lc) {
#pragma unroll
    for (int u = 0; u <= kAdiCh; ++u) v[u] = lv[u * st];
#pragma unroll
    for (int u = 0; u < kAdiCh; ++u) {
      rv[u] = lr[u * st];
      am[u] = la[u];
      cm[u] = lc[u];
    }
  };
  load(pv, pr, pam, pcm);
  for (int ib = lo_ + 1; ib <= n; ib += kAdiCh) {
    double cv[kAdiCh + 1], crv[kAdiCh], cam[kAdiCh], ccm[kAdiCh];
#pragma unroll
    for (int u = 0; u <= kAdiCh; ++u) cv[u] = v[u];
#pragma unroll
    for (int u = 0; u < kAdiCh; ++u) {
      crv[u] = rv[u];
      cam[u] = am[u];
      ccm[u] = cm[u];
    }
    const long sk = kAdiCh * st;
    if (ib + kAdiCh <= n) load(pv + sk, pr + sk, pam + kAdiCh, pcm + kAdiCh);
#pragma unroll
    for (int u = 0; u < kAdiCh; ++u) {
      if (ib + u > n) break;
      const double a = cam[u] * (1. + l.cc * cv[u]);
      const double c = ccm[u] * (1. + l.cc * cv[u + 1]);
      const double b = l.bd - a - c;
      const double gv = l.bd * crv[u];
      const double gam = cprev * bet;
      bet = 1. / (b - a * gam);
      up = (gv - a * up) * bet;
      pg[u * st] = gam;
      pr[u * st] = up;
      cprev = c;
    }
    pv += sk;
    pr += sk;
    pg += sk;
    pam += kAdiCh;
    pcm += kAdiCh;
  }
  return up;
}

// periodic z line (trdiag3p, slv_mmtm.f90:1572-1641); result into rhs
template <int C>
__device__ inline void adi_z(LW_KB, const Geo &g, const Sub &sb, const Fld &f, double cre, double cre2) {
  const int ilo = C == 0 ? g.ib : 1, jlo = C == 1 ? g.jb : 1;
  const int i = ilo + blockIdx.x * 64 + threadIdx.x;
  const int j = jlo + blockIdx.y * 4 + threadIdx.y;
  if (i > g.n1m || j > g.n2m) return;
  const int n = g.n3m;
  const LineCf l = line_cf<2, C>(g, f, sb, cre, cre2);
  const long st = (long)g.NX * g.NY, q0 = P(g, i, j, 0);
  double *r = f.rhs[C];
  double *sq = f.sa[C % 3], *ss = f.sb[C % 3];
  const double fn = l.bd * r[q0 + n * st];
  double qp, sp, fp;
  {
    const long q = q0 + st;
    const double a = l.am[1] * (1. + l.cc * l.vis[q + l.lo * st]);
    const double c = l.cm[1] * (1. + l.cc * l.vis[q + (l.lo + 1) * st]);
    const double b = l.bd - a - c;
    const double binv = 1. / b;
    qp = -c * binv;
    sp = -a * binv;
    fp = (l.bd * r[q]) * binv;
    sq[q] = qp;
    ss[q] = sp;
    r[q] = fp;
  }
  double an = 0, bn = 0, cn = 0;
  for (int kb = 2; kb <= n; kb += kAdiCh) {
    double v[kAdiCh + 1], rv[kAdiCh], am[kAdiCh], cm[kAdiCh];
#pragma unroll
    for (int u = 0; u <= kAdiCh; ++u)
      v[u] = kb + u <= n + 1 ? l.vis[q0 + (long)(kb + u + l.lo) * st] : 0.0;
#pragma unroll
    for (int u = 0; u < kAdiCh; ++u) {
      const bool ok = kb + u <= n;
      rv[u] = ok ? r[q0 + (long)(kb + u) * st] : 0.0;
      am[u] = ok ? l.am[kb + u] : 0.0;
      cm[u] = ok ? l.cm[kb + u] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < kAdiCh; ++u) {
      const int k = kb + u;
      if (k > n) break;
      const double a = am[u] * (1. + l.cc * v[u]);
      const double c = cm[u] * (1. + l.cc * v[u + 1]);
      const double b = l.bd - a - c;
      const double gv = l.bd * rv[u];
      const double pn = 1. / (b + a * qp);
      qp = -c * pn;
      sp = -a * sp * pn;
      fp = (gv - a * fp) * pn;
      const long q = q0 + (long)k * st;
      sq[q] = qp;
      ss[q] = sp;
      r[q] = fp;
      an = a;
      bn = b;
      cn = c;
    }
  }
  // backward pass: s(n) = 1, qe(n) = 0; qe overwrites f in place
  double snext = 1.0, qenext = 0.0;
  for (int kb = n - 1; kb >= 1; kb -= kAdiCh) {
    double qv[kAdiCh], sv[kAdiCh], rv[kAdiCh];
#pragma unroll
    for (int u = 0; u < kAdiCh; ++u) {
      const long q = q0 + (long)(kb - u) * st;
      qv[u] = kb - u >= 1 ? sq[q] : 0.0;
      sv[u] = kb - u >= 1 ? ss[q] : 0.0;
      rv[u] = kb - u >= 1 ? r[q] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < kAdiCh; ++u) {
      if (kb - u < 1) break;
      const long q = q0 + (long)(kb - u) * st;
      const double sk = sv[u] + qv[u] * snext;
      const double qek = rv[u] + qv[u] * qenext;
      ss[q] = sk;
      r[q] = qek;
      snext = sk;
      qenext = qek;
    }
  }
  const long q1 = q0 + st, qn1 = q0 + (long)(n - 1) * st;
  const double qe1 = r[q1], s1 = ss[q1];
  const double qenm = r[qn1], snm = ss[qn1];
  const double fnn = (fn - cn * qe1 - an * qenm) / (cn * s1 + an * snm + bn);
  r[q0 + (long)n * st] = fnn;
  for (int kb = n - 1; kb >= 1; kb -= kAdiCh) {
    double sv[kAdiCh], rv[kAdiCh];
#pragma unroll
    for (int u = 0; u < kAdiCh; ++u) {
      const long q = q0 + (long)(kb - u) * st;
      sv[u] = kb - u >= 1 ? ss[q] : 0.0;
      rv[u] = kb - u >= 1 ? r[q] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < kAdiCh; ++u) {
      if (kb - u < 1) break;
      r[q0 + (long)(kb - u) * st] = fnn * sv[u] + rv[u];
    }
  }
}

// periodic x line (trdiag1p, slv_mmtm.f90:1501-1570: the recurrences of
// trdiag3p along x) of the x-periodic channel; field += solution. One thread
// per (j, k) line in the reference's order, q and s in the scratch arrays,
// qe in place of the right-hand side. x is the contiguous direction, so the
// threads' accesses are a row apart: this path is not tuned.
template <int C>
__device__ inline void adi_xp(LW_KB, const Geo &g, const Sub &sb, const Fld &f, double cre, double cre2) {
  const int jlo = C == 1 ? g.jb : 1, klo = C == 2 ? g.kb : 1;
  const int nj = g.n2m - jlo + 1;
  const int line = blockIdx.x * 64 + threadIdx.x;
  if (line >= nj * (g.n3m - klo + 1)) return;
  const int j = jlo + line % nj, k = klo + line / nj;
  const int n = g.n1m;
  const LineCf l = line_cf<0, C>(g, f, sb, cre, cre2);
  const long q0 = P(g, 0, j, k);
  double *r = f.rhs[C], *fld = C == 0 ? f.u : C == 1 ? f.v : f.w;
  double *sq = f.sa[C % 3], *ss = f.sb[C % 3];
  const double fn = l.bd * r[q0 + n];
  double qp, sp, fp, an = 0, bn = 0, cn = 0;
  {
    const long q = q0 + 1;
    const double a = l.am[1] * (1. + l.cc * l.vis[q + l.lo]);
    const double c = l.cm[1] * (1. + l.cc * l.vis[q + l.lo + 1]);
    const double b = l.bd - a - c;
    const double binv = 1. / b;
    qp = -c * binv;
    sp = -a * binv;
    fp = (l.bd * r[q]) * binv;
    sq[q] = qp;
    ss[q] = sp;
    r[q] = fp;
  }
  for (int i = 2; i <= n; ++i) {
    const long q = q0 + i;
    const double a = l.am[i] * (1. + l.cc * l.vis[q + l.lo]);
    const double c = l.cm[i] * (1. + l.cc * l.vis[q + l.lo + 1]);
    const double b = l.bd - a - c;
    const double pn = 1. / (b + a * qp);
    qp = -c * pn;
    sp = -a * sp * pn;
    fp = (l.bd * r[q] - a * fp) * pn;
    sq[q] = qp;
    ss[q] = sp;
    r[q] = fp;
    an = a;
    bn = b;
    cn = c;
  }
  double snext = 1.0, qenext = 0.0;
  for (int i = n - 1; i >= 1; --i) {
    const long q = q0 + i;
    const double sk = ss[q] + sq[q] * snext;
    const double qek = r[q] + sq[q] * qenext;
    ss[q] = sk;
    r[q] = qek;
    snext = sk;
    qenext = qek;
  }
  const double fnn = (fn - cn * r[q0 + 1] - an * r[q0 + n - 1]) / (cn * ss[q0 + 1] + an * ss[q0 + n - 1] + bn);
  fld[q0 + n] = fnn + fld[q0 + n];
  for (int i = n - 1; i >= 1; --i) fld[q0 + i] = (fnn * ss[q0 + i] + r[q0 + i]) + fld[q0 + i];
}

template <int C0>
LW_KERNEL void k_adi_xp(LW_KP(64, 0), const Geo &g, const Sub &sb, const Fld &f, double cre, double cre2) {
  switch (blockIdx.y) {
    case 0: adi_xp<0>(LW_KBA, g, sb, f, cre, cre2); break;
    case 1: adi_xp<1>(LW_KBA, g, sb, f, cre, cre2); break;
    default: adi_xp<2>(LW_KBA, g, sb, f, cre, cre2); break;
  }
}

// non-periodic y line (trdiag2, slv_mmtm.f90:1415-1455); result into rhs
template <int C>
__device__ inline void adi_y(LW_KB, const Geo &g, const Sub &sb, const Fld &f, double cre, double cre2) {
  const int ilo = C == 0 ? g.ib : 1, jlo = C == 1 ? g.jb : 1, klo = C == 2 ? g.kb : 1;
  const int i = ilo + blockIdx.x * 64 + threadIdx.x;
  const int k = klo + blockIdx.y * 4 + threadIdx.y;
  if (i > g.n1m || k > g.n3m) return;
  const LineCf l = line_cf<1, C>(g, f, sb, cre, cre2);
  const long st = g.NX, q0 = P(g, i, 0, k);
  double *r = f.rhs[C];
  double *gm = f.sa[C % 3];
  double unext = line_forward(l, r, gm, q0, st, jlo, g.n2m);
  // back substitution, pipelined like the forward pass; the loads run below
  // the line's start into plane k-1 (k >= 1), values unused
  double *pr = r + q0 + (long)(g.n2m - 1) * st;
  const double *pg = gm + q0 + (long)g.n2m * st;
  double rv[kAdiCh], gv[kAdiCh];
  auto load = [&](const double *lr, const double *lg) {
#pragma unroll
    for (int u = 0; u < kAdiCh; ++u) {
      rv[u] = lr[-u * st];
      gv[u] = lg[-u * st];
    }
  };
  load(pr, pg);
  for (int jb = g.n2m - 1; jb >= jlo; jb -= kAdiCh) {
    double crv[kAdiCh], cgv[kAdiCh];
#pragma unroll
    for (int u = 0; u < kAdiCh; ++u) {
      crv[u] = rv[u];
      cgv[u] = gv[u];
    }
    const long sk = kAdiCh * st;
    if (jb - kAdiCh >= jlo) load(pr - sk, pg - sk);
#pragma unroll
    for (int u = 0; u < kAdiCh; ++u) {
      if (jb - u < jlo) break;
      const double uu = crv[u] - cgv[u] * unext;
      pr[-u * st] = uu;
      unext = uu;
    }
    pr -= sk;
    pg -= sk;
  }
}

// non-periodic x line (trdiag1, slv_mmtm.f90:1365-1412); field += solution.
// x is the contiguous direction, so a thread per line would read addresses a
// row apart. One wavefront takes 64 lines (consecutive (j, k)) and walks them
// in chunks of kXch points: each chunk is loaded through LDS two lines at a
// time (each half of the wavefront reads 32 consecutive doubles of one line),
// every lane then runs the reference recurrence along its own line out of
// LDS, and the results leave through LDS the same way.
constexpr int kXch = 31;        // points per chunk
constexpr int kXld = kXch + 2;  // LDS row stride (odd, conflict-free columns)

template <int C>
__device__ inline void adi_x(LW_KB, const Geo &g, const Sub &sb, const Fld &f, double cre, double cre2, double *V, double *R,
                             double *G, double *AM, double *CM) {
  const int ilo = C == 0 ? g.ib : 1, jlo = C == 1 ? g.jb : 1, klo = C == 2 ? g.kb : 1;
  const int lane = threadIdx.x;
  const int nj = g.n2m - jlo + 1;
  const int nlines = nj * (g.n3m - klo + 1);
  const int line0 = blockIdx.x * 64;
  if (line0 >= nlines) return;  // the grid covers the component with the most lines
  const int nl = nlines - line0 < 64 ? nlines - line0 : 64;  // lines of this wavefront
  const bool mine = lane < nl;
  const LineCf l = line_cf<0, C>(g, f, sb, cre, cre2);
  double *fld = C == 0 ? f.u : C == 1 ? f.v : C == 2 ? f.w : f.t;
  double *r = f.rhs[C];
  double *gm = f.sa[C % 3];
  const int n1m = g.n1m;
  // row offset of this lane's line; line L's offset is lane L's (broadcast)
  const long q0 = mine ? P(g, 0, jlo + (line0 + lane) % nj, klo + (line0 + lane) / nj) : 0;
  // the row offsets of the lines this lane stages: line 2t + (lane >> 5)
  long qs[32];
#pragma unroll
  for (int t = 0; t < 32; ++t) qs[t] = __shfl(q0, 2 * t + (lane >> 5), 64);
  const int hl = lane >> 5, o = lane & 31;
  double bet = 0.0, up = 0.0, cprev = 0.0;
  if (mine) {  // first point, i = ilo
    const long q = q0 + ilo;
    const double a = l.am[ilo] * (1. + l.cc * l.vis[q + l.lo]);
    const double c = l.cm[ilo] * (1. + l.cc * l.vis[q + l.lo + 1]);
    const double b = l.bd - a - c;
    bet = 1. / b;
    up = (l.bd * r[q]) * bet;
    cprev = c;
    r[q] = up;
  }
  // forward elimination, chunks [ib, ib + cnt)
  for (int ib = ilo + 1; ib <= n1m; ib += kXch) {
    const int cnt = n1m - ib + 1 < kXch ? n1m - ib + 1 : kXch;
    if (lane < cnt) {
      AM[lane] = l.am[ib + lane];
      CM[lane] = l.cm[ib + lane];
    }
    // unconditional loads, 16 line pairs in flight at a time: lanes past the
    // last line read row 0 of the first plane and points past the chunk read
    // the next row, all inside the padded array; nothing beyond cnt is used
#pragma unroll
    for (int t0 = 0; t0 < 32; t0 += 16) {
      double tv[16], tr[16];
#pragma unroll
      for (int t = 0; t < 16; ++t) {
        tv[t] = l.vis[qs[t0 + t] + ib + l.lo + o];
        tr[t] = r[qs[t0 + t] + ib + o];
      }
#pragma unroll
      for (int t = 0; t < 16; ++t) {
        const int L = 2 * (t0 + t) + hl;
        V[L * kXld + o] = tv[t];
        R[L * kXld + o] = tr[t];
      }
    }
    __syncthreads();
    if (mine) {
      const double *v = V + lane * kXld;
      double *rr = R + lane * kXld, *gg = G + lane * kXld;
#pragma unroll 4
      for (int o = 0; o < cnt; ++o) {
        const double a = AM[o] * (1. + l.cc * v[o]);
        const double c = CM[o] * (1. + l.cc * v[o + 1]);
        const double b = l.bd - a - c;
        const double gv = l.bd * rr[o];
        const double gam = cprev * bet;
        bet = 1. / (b - a * gam);
        up = (gv - a * up) * bet;
        gg[o] = gam;
        rr[o] = up;
        cprev = c;
      }
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < 32; ++t) {
      const int L = 2 * t + hl;
      if (L < nl && o < cnt) {
        gm[qs[t] + ib + o] = G[L * kXld + o];
        r[qs[t] + ib + o] = R[L * kXld + o];
      }
    }
    __syncthreads();
  }
  // back substitution, chunks [bot, top] from the top down; V holds the field
  double unext = up;
  if (mine) fld[q0 + n1m] = unext + fld[q0 + n1m];
  for (int top = n1m - 1; top >= ilo; top -= kXch) {
    const int bot = top - kXch + 1 > ilo ? top - kXch + 1 : ilo;
    const int cnt = top - bot + 1;
#pragma unroll
    for (int t0 = 0; t0 < 32; t0 += 8) {
      double tr[8], tg[8], tf[8];
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        const long qL = qs[t0 + t] + bot;
        tr[t] = r[qL + o];
        tg[t] = gm[qL + o + 1];
        tf[t] = fld[qL + o];
      }
#pragma unroll
      for (int t = 0; t < 8; ++t) {
        const int L = 2 * (t0 + t) + hl;
        R[L * kXld + o] = tr[t];
        G[L * kXld + o] = tg[t];
        V[L * kXld + o] = tf[t];
      }
    }
    __syncthreads();
    if (mine) {
      const double *rr = R + lane * kXld, *gg = G + lane * kXld;
      double *ff = V + lane * kXld;
#pragma unroll 4
      for (int o = cnt - 1; o >= 0; --o) {
        const double uu = rr[o] - gg[o] * unext;
        ff[o] = uu + ff[o];
        unext = uu;
      }
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < 32; ++t) {
      const int L = 2 * t + hl;
      if (L < nl && o < cnt) fld[qs[t] + bot + o] = V[L * kXld + o];
    }
    __syncthreads();
  }
}

// y sweep with one half-wavefront per line (tri_line, lines of 32 lanes): a
// workgroup stages the viscosity and right-hand side of kYL lines of
// consecutive i (one plane k) through LDS, read along i; each half-wavefront
// solves one line out of LDS with the line held in registers, and the
// solutions leave through LDS the same way. No per-point scratch goes to HBM.
constexpr int kYL = 16;

__host__ __device__ inline int adi_y_ld(int n2) { return n2 + 1; }

template <int R, bool FULL, int C>
__device__ inline void adi_y_wave(LW_KB, const Geo &g, const Sub &sb, const Fld &f, double cre, double cre2, double *tv,
                                  double *tr) {
  const LineCf l = line_cf<1, C>(g, f, sb, cre, cre2);
  const int ilo = C == 0 ? g.ib : 1, jlo = C == 1 ? g.jb : 1, klo = C == 2 ? g.kb : 1;
  const int i0 = ilo + blockIdx.x * kYL, k = klo + blockIdx.y;
  if (k > g.n3m) return;  // uniform: the grid covers the widest component
  const int n1m = g.n1m, n2m = g.n2m, N = n2m - jlo + 1, LD = adi_y_ld(g.n2);
  const int tid = threadIdx.x + 64 * threadIdx.y;
  double *r = f.rhs[C];
  for (int q = tid; q < kYL * (g.n2 + 1); q += 32 * kYL) {
    const int li = q % kYL, j = q / kYL, i = i0 + li;
    const long p = P(g, i <= n1m ? i : n1m, j, k);
    tv[li * LD + j] = l.vis[p];
    tr[li * LD + j] = r[p];
  }
  __syncthreads();
  // two lines per wavefront, one per half (lines of 32 lanes)
  const int lane = threadIdx.x, ll = lane & 31, li = 2 * threadIdx.y + (lane >> 5);
  const double *v = tv + li * LD;
  double *rr = tr + li * LD;
  double a[R], b[R], c[R], ff[R], x[R];
#pragma unroll
  for (int u = 0; u < R; ++u) {
    const int t = 1 + ll * R + u;
    const int j = (FULL || t <= N) ? jlo + t - 1 : n2m;
    a[u] = l.am[j] * (1. + l.cc * v[j + l.lo]);
    c[u] = l.cm[j] * (1. + l.cc * v[j + l.lo + 1]);
    b[u] = l.bd - a[u] - c[u];
    ff[u] = l.bd * rr[j];
  }
  tri_line<R, FULL, 32>(a, b, c, ff, N, lane, x);
  __syncthreads();
#pragma unroll
  for (int u = 0; u < R; ++u) {
    const int t = 1 + ll * R + u;
    if (FULL || t <= N) rr[jlo + t - 1] = x[u];
  }
  __syncthreads();
  for (int q = tid; q < kYL * N; q += 32 * kYL) {
    const int li2 = q % kYL, j = jlo + q / kYL, i = i0 + li2;
    if (i <= n1m) r[P(g, i, j, k)] = tr[li2 * LD + j];
  }
}

template <int R, int C>
__device__ inline void adi_y_wave_c(LW_KB, const Geo &g, const Sub &sb, const Fld &f, double cre, double cre2, double *tv,
                                    double *tr) {
  const int jlo = C == 1 ? g.jb : 1;
  if (g.n2m - jlo + 1 == 32 * R) adi_y_wave<R, true, C>(LW_KBA, g, sb, f, cre, cre2, tv, tr);
  else adi_y_wave<R, false, C>(LW_KBA, g, sb, f, cre, cre2, tv, tr);
}

template <int R, int C0>
LW_KERNEL void k_adi_yw(LW_KP(32 * kYL, 0), const Geo &g, const Sub &sb, const Fld &f, double cre, double cre2) {
  extern __shared__ double yl[];
  double *tv = yl, *tr = yl + kYL * adi_y_ld(g.n2);
  if constexpr (C0 == 3) {
    adi_y_wave_c<R, 3>(LW_KBA, g, sb, f, cre, cre2, tv, tr);
  } else {
    switch (blockIdx.z) {
      case 0: adi_y_wave_c<R, 0>(LW_KBA, g, sb, f, cre, cre2, tv, tr); break;
      case 1: adi_y_wave_c<R, 1>(LW_KBA, g, sb, f, cre, cre2, tv, tr); break;
      default: adi_y_wave_c<R, 2>(LW_KBA, g, sb, f, cre, cre2, tv, tr); break;
    }
  }
}

typedef ::lw::KFn<Geo, Sub, Fld, double, double> AdiWaveKernel;
template <int C0>
static AdiWaveKernel pick_adi_yw(int n2m) {  // rows per lane on lines of 32 lanes
  if (n2m <= 32) return LW_KFN((k_adi_yw<1, C0>));
  if (n2m <= 64) return LW_KFN((k_adi_yw<2, C0>));
  if (n2m <= 128) return LW_KFN((k_adi_yw<4, C0>));
  if (n2m <= 256) return LW_KFN((k_adi_yw<8, C0>));
  return {};
}
static inline unsigned adi_yw_lds(const Geo &g) { return (unsigned)(2 * kYL * adi_y_ld(g.n2) * sizeof(double)); }
// LDS limit set on the instance pick_adi_yw chooses for n2m: the size its
// longest line needs (n2m = 32 R rows). The attribute belongs to the kernel
// instance, which every case with rows in the same range shares, so it is
// set to the range's largest size and never lowered under a running case.
static inline unsigned adi_yw_lds_bucket(int n2m) {
  const int r = n2m <= 32 ? 1 : n2m <= 64 ? 2 : n2m <= 128 ? 4 : 8;
  return (unsigned)(2 * kYL * adi_y_ld(32 * r + 1) * sizeof(double));
}

// periodic z sweep with one half-wavefront per line (tri_per_line32, N =
// n3m = 32*R rows): a workgroup stages the viscosity and right-hand side of
// kZL lines of consecutive i (one row j) through LDS, plane by plane, and the
// solutions go back to rhs the same way; no per-point scratch goes to HBM.
constexpr int kZL = 32;

template <int R, int C>
__device__ inline void adi_z_wave(LW_KB, const Geo &g, const Sub &sb, const Fld &f, double cre, double cre2, double *tv,
                                  double *tr) {
  constexpr int N = 32 * R, LD = N + 1;
  const LineCf l = line_cf<2, C>(g, f, sb, cre, cre2);
  const int ilo = C == 0 ? g.ib : 1, jlo = C == 1 ? g.jb : 1;
  const int i0 = ilo + blockIdx.x * kZL, j = jlo + blockIdx.y;
  if (j > g.n2m) return;  // uniform: the grid covers the widest component
  const int n1m = g.n1m;
  const long st = (long)g.NX * g.NY;
  const int tid = threadIdx.x + 64 * threadIdx.y;
  double *r = f.rhs[C];
  // tv[s] = vis at plane 1 + lo + s (s = 0..N): a_t reads tv[t-1], c_t tv[t]
  for (int q = tid; q < kZL * (N + 1); q += 32 * kZL) {
    const int li = q % kZL, s = q / kZL, i = i0 + li;
    const long p0 = P(g, i <= n1m ? i : n1m, j, 0);
    tv[li * LD + s] = l.vis[p0 + (long)(1 + l.lo + s) * st];
    if (s < N) tr[li * LD + s] = r[p0 + (long)(1 + s) * st];
  }
  __syncthreads();
  const int lane = threadIdx.x, ll = lane & 31, li = 2 * threadIdx.y + (lane >> 5);
  const double *v = tv + li * LD;
  double *rr = tr + li * LD;
  double a[R], b[R], c[R], ff[R], x[R];
#pragma unroll
  for (int u = 0; u < R; ++u) {
    const int t = 1 + ll * R + u;
    a[u] = l.am[t] * (1. + l.cc * v[t - 1]);
    c[u] = l.cm[t] * (1. + l.cc * v[t]);
    b[u] = l.bd - a[u] - c[u];
    ff[u] = l.bd * rr[t - 1];
  }
  tri_per_line32<R>(a, b, c, ff, lane, x);
  __syncthreads();
#pragma unroll
  for (int u = 0; u < R; ++u) rr[ll * R + u] = x[u];
  __syncthreads();
  for (int q = tid; q < kZL * N; q += 32 * kZL) {
    const int li2 = q % kZL, s = q / kZL, i = i0 + li2;
    if (i <= n1m) r[P(g, i, j, 0) + (long)(1 + s) * st] = tr[li2 * LD + s];
  }
}

template <int R, int C0>
LW_KERNEL void k_adi_zw(LW_KP(32 * kZL, 0), const Geo &g, const Sub &sb, const Fld &f, double cre, double cre2) {
  __shared__ double zl[2 * kZL * (32 * R + 1)];
  double *tv = zl, *tr = zl + kZL * (32 * R + 1);
  if constexpr (C0 == 3) {
    adi_z_wave<R, 3>(LW_KBA, g, sb, f, cre, cre2, tv, tr);
  } else {
    switch (blockIdx.z) {
      case 0: adi_z_wave<R, 0>(LW_KBA, g, sb, f, cre, cre2, tv, tr); break;
      case 1: adi_z_wave<R, 1>(LW_KBA, g, sb, f, cre, cre2, tv, tr); break;
      default: adi_z_wave<R, 2>(LW_KBA, g, sb, f, cre, cre2, tv, tr); break;
    }
  }
}

template <int C0>
static AdiWaveKernel pick_adi_zw(int n3m) {  // N = 32*R planes exactly
  if (n3m == 32) return LW_KFN((k_adi_zw<1, C0>));
  if (n3m == 64) return LW_KFN((k_adi_zw<2, C0>));
  if (n3m == 128) return LW_KFN((k_adi_zw<4, C0>));
  return {};
}

// x sweep with one wavefront per line (tri_line, lines of 64 lanes, R rows
// per lane): the line's viscosity and right-hand side are read along i
// (contiguous) into LDS, each lane takes its R consecutive rows from there,
// and the solution goes back through LDS to be added to the field along i.
// Unlike the serial sweep, nothing but the field is written to HBM. LDS
// index s + s / R spreads a lane's rows over the banks.
constexpr int kXWL = 4;  // lines (wavefronts) per workgroup
template <int R>
__host__ __device__ constexpr int xw_ld() {
  return 64 * (R + 1) + 2;
}
template <int R>
__device__ inline int xw_p(int s) {
  return s + s / R;
}

template <int R, bool FULL, int C>
__device__ inline void adi_x_wave(LW_KB, const Geo &g, const Sub &sb, const Fld &f, double cre, double cre2, double *lds) {
  const int ilo = C == 0 ? g.ib : 1, jlo = C == 1 ? g.jb : 1, klo = C == 2 ? g.kb : 1;
  const int lane = threadIdx.x, w = threadIdx.y;
  const int nj = g.n2m - jlo + 1, nlines = nj * (g.n3m - klo + 1);
  if (blockIdx.x * kXWL >= nlines) return;  // uniform: the grid covers the component with the most lines
  const int line = blockIdx.x * kXWL + w;
  const bool mine = line < nlines;
  const int ln = mine ? line : nlines - 1;  // a wavefront past the last line reads that line, stores nothing
  const int N = g.n1m - ilo + 1;
  const LineCf l = line_cf<0, C>(g, f, sb, cre, cre2);
  double *fld = C == 0 ? f.u : C == 1 ? f.v : C == 2 ? f.w : f.t;
  const double *r = f.rhs[C];
  const long q0 = P(g, 0, jlo + ln % nj, klo + ln / nj) + ilo;  // row t = 1 of the line
  double *V = lds + (long)w * 2 * xw_ld<R>(), *X = V + xw_ld<R>();
  // V[s] = vis at row s + 1 + lo (s = 0..N): a_t reads V[t-1], c_t V[t]
  for (int s = lane; s <= N; s += 64) V[xw_p<R>(s)] = l.vis[q0 + l.lo + s];
  for (int s = lane; s < N; s += 64) X[xw_p<R>(s)] = r[q0 + s];
  __syncthreads();
  double a[R], b[R], c[R], ff[R], x[R];
#pragma unroll
  for (int u = 0; u < R; ++u) {
    const int t = 1 + lane * R + u;
    const int tc = (FULL || t <= N) ? t : N;
    const int i = ilo + tc - 1;
    a[u] = l.am[i] * (1. + l.cc * V[xw_p<R>(tc - 1)]);
    c[u] = l.cm[i] * (1. + l.cc * V[xw_p<R>(tc)]);
    b[u] = l.bd - a[u] - c[u];
    ff[u] = l.bd * X[xw_p<R>(tc - 1)];
  }
  tri_line<R, FULL, 64>(a, b, c, ff, N, lane, x);
  __syncthreads();
#pragma unroll
  for (int u = 0; u < R; ++u) {
    const int t = 1 + lane * R + u;
    if (FULL || t <= N) X[xw_p<R>(t - 1)] = x[u];
  }
  __syncthreads();
  if (mine)
    for (int s = lane; s < N; s += 64) fld[q0 + s] = X[xw_p<R>(s)] + fld[q0 + s];
}

template <int R, int C>
__device__ inline void adi_x_wave_c(LW_KB, const Geo &g, const Sub &sb, const Fld &f, double cre, double cre2, double *lds) {
  const int ilo = C == 0 ? g.ib : 1;
  if (g.n1m - ilo + 1 == 64 * R) adi_x_wave<R, true, C>(LW_KBA, g, sb, f, cre, cre2, lds);
  else adi_x_wave<R, false, C>(LW_KBA, g, sb, f, cre, cre2, lds);
}

template <int R, int C0>
LW_KERNEL void k_adi_xw(LW_KP(64 * kXWL, 0), const Geo &g, const Sub &sb, const Fld &f, double cre, double cre2) {
  __shared__ double xl[kXWL * 2 * xw_ld<R>()];
  if constexpr (C0 == 3) {
    adi_x_wave_c<R, 3>(LW_KBA, g, sb, f, cre, cre2, xl);
  } else {
    switch (blockIdx.y) {
      case 0: adi_x_wave_c<R, 0>(LW_KBA, g, sb, f, cre, cre2, xl); break;
      case 1: adi_x_wave_c<R, 1>(LW_KBA, g, sb, f, cre, cre2, xl); break;
      default: adi_x_wave_c<R, 2>(LW_KBA, g, sb, f, cre, cre2, xl); break;
    }
  }
}

template <int C0>
static AdiWaveKernel pick_adi_xw(int n1m) {  // rows per lane on lines of 64 lanes
  if (n1m <= 64) return LW_KFN((k_adi_xw<1, C0>));
  if (n1m <= 128) return LW_KFN((k_adi_xw<2, C0>));
  if (n1m <= 192) return LW_KFN((k_adi_xw<3, C0>));
  if (n1m <= 256) return LW_KFN((k_adi_xw<4, C0>));
  if (n1m <= 384) return LW_KFN((k_adi_xw<6, C0>));
  if (n1m <= 512) return LW_KFN((k_adi_xw<8, C0>));
  return {};
}

// once per handle: which sweeps run with whole lines in registers (bit 1:
// y, bit 2: x, bit 4: z; LW_ADI_WAVE sets the mask, default 7). The y kernel
// needs more than 64 KiB of LDS at 256 rows; lines longer than 256 (y) or
// 512 (x) rows, and z lines of other than 32, 64 or 128 planes, keep the
// serial sweeps.
hipError_t prepare_adi_kernels(const Geo &g, int *use_wave) {
  int mask = 7;
  if (const char *e = getenv("LW_ADI_WAVE")) mask = atoi(e);
  *use_wave = 0;
  if ((mask & 2) && pick_adi_xw<0>(g.n1m)) *use_wave |= 2;
  if ((mask & 4) && pick_adi_zw<0>(g.n3m)) *use_wave |= 4;
  if (!(mask & 1) || !pick_adi_yw<0>(g.n2m) || adi_yw_lds_bucket(g.n2m) > 160 * 1024) return hipSuccess;
  const int lds = (int)adi_yw_lds_bucket(g.n2m);
  hipError_t e = ::lw::set_max_lds(pick_adi_yw<0>(g.n2m), lds);
  if (e == hipSuccess)
    e = ::lw::set_max_lds(pick_adi_yw<3>(g.n2m), lds);
  if (e == hipSuccess) *use_wave |= 1;
  return e;
}

// One launch per direction for all momentum components at once (their
// solves are independent: each has its own rhs, coefficients and scratch),
// so the serial line recurrences of u, v and w overlap on the device.
// blockIdx.z (z, y sweeps) or blockIdx.y (x sweep) is the component offset
// from C0; C0 = 3 is the temperature alone.
template <int C0>
LW_KERNEL void k_adi_z(LW_KP(256, 0), const Geo &g, const Sub &sb, const Fld &f, double cre, double cre2) {
  if constexpr (C0 == 3) {
    adi_z<3>(LW_KBA, g, sb, f, cre, cre2);
  } else {
    switch (blockIdx.z) {
      case 0: adi_z<0>(LW_KBA, g, sb, f, cre, cre2); break;
      case 1: adi_z<1>(LW_KBA, g, sb, f, cre, cre2); break;
      default: adi_z<2>(LW_KBA, g, sb, f, cre, cre2); break;
    }
  }
}

template <int C0>
LW_KERNEL void k_adi_y(LW_KP(256, 0), const Geo &g, const Sub &sb, const Fld &f, double cre, double cre2) {
  if constexpr (C0 == 3) {
    adi_y<3>(LW_KBA, g, sb, f, cre, cre2);
  } else {
    switch (blockIdx.z) {
      case 0: adi_y<0>(LW_KBA, g, sb, f, cre, cre2); break;
      case 1: adi_y<1>(LW_KBA, g, sb, f, cre, cre2); break;
      default: adi_y<2>(LW_KBA, g, sb, f, cre, cre2); break;
    }
  }
}

template <int C0>
LW_KERNEL void k_adi_x(LW_KP(64, 0), const Geo &g, const Sub &sb, const Fld &f, double cre, double cre2) {
  __shared__ double V[64 * kXld], R[64 * kXld], G[64 * kXld], AM[kXch], CM[kXch];
  if constexpr (C0 == 3) {
    adi_x<3>(LW_KBA, g, sb, f, cre, cre2, V, R, G, AM, CM);
  } else {
    switch (blockIdx.y) {
      case 0: adi_x<0>(LW_KBA, g, sb, f, cre, cre2, V, R, G, AM, CM); break;
      case 1: adi_x<1>(LW_KBA, g, sb, f, cre, cre2, V, R, G, AM, CM); break;
      default: adi_x<2>(LW_KBA, g, sb, f, cre, cre2, V, R, G, AM, CM); break;
    }
  }
}

// z, y, x sweeps of nc components from c0 (momentum: c0 = 0, nc = 3, or 2
// when n3m = 1 and w has no equation; temperature: c0 = 3, nc = 1)
template <int C0>
static void adi3(hipStream_t s, const Geo &g, const Sub &sb, const Fld &f, double cre, double cre2, int nc) {
  static const char *nz = C0 == 3 ? "k_adi_z<t>" : "k_adi_z<uvw>";
  static const char *ny = C0 == 3 ? "k_adi_y<t>" : "k_adi_y<uvw>";
  static const char *nx = C0 == 3 ? "k_adi_x<t>" : "k_adi_x<uvw>";
  // the grids cover the widest component; z and y sweeps: viscosity and rhs
  // in, rhs out; x sweep: also the field in and out (per component)
  if (g.adi_wave & 4)
    LW_LAUNCH_K(nz, pts(g) * 24.0 * nc, 0, pick_adi_zw<C0>(g.n3m),
              dim3((unsigned)((g.n1m + kZL - 1) / kZL), (unsigned)g.n2m, (unsigned)nc), dim3(64, kZL / 2), 0, s, g, sb,
              f, cre, cre2);
  else if (g.n3m != 1)
    LW_LAUNCH(nz, pts(g) * 24.0 * nc, 0, k_adi_z<C0>, grid3(g.n1m, g.n2m, nc), BLK3, 0, s, g, sb, f, cre, cre2);
  if (g.adi_wave & 1)
    LW_LAUNCH_K(ny, pts(g) * 24.0 * nc, 0, pick_adi_yw<C0>(g.n2m),
              dim3((unsigned)((g.n1m + kYL - 1) / kYL), (unsigned)g.n3m, (unsigned)nc), dim3(64, kYL / 2), adi_yw_lds(g), s,
              g, sb, f, cre, cre2);
  else
    LW_LAUNCH(ny, pts(g) * 24.0 * nc, 0, k_adi_y<C0>, grid3(g.n1m, g.n3m, nc), BLK3, 0, s, g, sb, f, cre, cre2);
  const int xlines = g.n2m * g.n3m;
  if (g.xprdic && C0 != 3)
    LW_LAUNCH(nx, pts(g) * 32.0 * nc, 0, k_adi_xp<C0>, dim3((unsigned)((xlines + 63) / 64), (unsigned)nc), dim3(64),
              0, s, g, sb, f, cre, cre2);
  else if (g.adi_wave & 2)
    LW_LAUNCH_K(nx, pts(g) * 32.0 * nc, 0, pick_adi_xw<C0>(g.n1m),
              dim3((unsigned)((xlines + kXWL - 1) / kXWL), (unsigned)nc), dim3(64, kXWL), 0, s, g, sb, f, cre, cre2);
  else
    LW_LAUNCH(nx, pts(g) * 32.0 * nc, 0, k_adi_x<C0>, dim3((unsigned)((xlines + 63) / 64), (unsigned)nc), dim3(64),
              0, s, g, sb, f, cre, cre2);
}

void launch_adi_mom(hipStream_t s, const Geo &g, const Sub &sb, const Fld &f, int les) {
  const double cre = (double)les * sb.re, cre2 = 2. * (double)les * sb.re;
  adi3<0>(s, g, sb, f, cre, cre2, g.n3m != 1 ? 3 : 2);
}

void launch_adi_t(hipStream_t s, const Geo &g, const Sub &sb, const Fld &f, int les) {
  const double cre = (double)les * sb.re * sb.pr;
  adi3<3>(s, g, sb, f, cre, 0.0, 1);
}

// retrv_uvw + prdic_adj_uvw(0) for periodic z (slv_mmtm.f90:1710-1857)
LW_KERNEL void k_retrv_mom(LW_KP(1024, 0), const Geo &g, const Fld &f) {
  const int j = blockIdx.x * 64 + threadIdx.x;  // 0..n2
  const int k = blockIdx.y;                     // 0..n3
  if (j > g.n2) return;
  if (k >= 1 && k <= g.n3m && j >= 1 && j <= g.n2m) {
    const long pb = j + g.NY * k;
    f.u[P(g, g.n1, j, k)] = f.uout[pb];
    if (j >= 2) f.v[P(g, g.n1, j, k)] = f.vout[pb];
    if (k >= g.kb) f.w[P(g, g.n1, j, k)] = f.wout[pb];
  }
}

LW_KERNEL void k_zghost_uvw(LW_KP(1024, 0), const Geo &g, const Fld &f) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  const int j = blockIdx.y;
  if (i > g.n1) return;
  const long lo = P(g, i, j, 0), hi = P(g, i, j, g.n3), a = P(g, i, j, g.n3m), b = P(g, i, j, 1);
  if (i >= 1) {
    f.u[lo] = f.u[a];
    f.u[hi] = f.u[b];
  }
  if (j >= 1) {
    f.v[lo] = f.v[a];
    f.v[hi] = f.v[b];
  }
  f.w[lo] = f.w[a];
  f.w[hi] = f.w[b];
}

// x-periodic ghost planes (prdic_adj_uvw, slv_mmtm.f90:1755-1786; ucalc,
// slv_auxi.f90:388-412): u over j 0..n2, v over j 1..n2 (k 0..n3), w over
// k 1..n3 (j 0..n2)
LW_KERNEL void k_xghost_uvw(LW_KP(1024, 0), const Geo &g, const Fld &f) {
  const int j = blockIdx.x * 64 + threadIdx.x;  // 0..n2
  const int k = blockIdx.y;                     // 0..n3
  if (j > g.n2) return;
  const long lo = P(g, 0, j, k), hi = P(g, g.n1, j, k), a = P(g, g.n1m, j, k), b = P(g, 1, j, k);
  f.u[lo] = f.u[a];
  f.u[hi] = f.u[b];
  if (j >= 1) {
    f.v[lo] = f.v[a];
    f.v[hi] = f.v[b];
  }
  if (k >= 1) {
    f.w[lo] = f.w[a];
    f.w[hi] = f.w[b];
  }
}

static void launch_xghost_uvw(hipStream_t s, const Geo &g, const Fld &f) {
  LW_LAUNCH("k_xghost_uvw", (double)(g.n2 + 1) * (g.n3 + 1) * 8.0 * 12, 0, k_xghost_uvw,
            dim3((g.n2 + 64) / 64, g.n3 + 1), dim3(64), 0, s, g, f);
}

// pcalc's x-periodic ghosts of p (slv_auxi.f90:480-489)
LW_KERNEL void k_xghost_p(LW_KP(1024, 0), const Geo &g, const Fld &f) {
  const int j = 1 + blockIdx.x * 64 + threadIdx.x;
  const int k = 1 + blockIdx.y;
  if (j > g.n2m) return;
  f.p[P(g, 0, j, k)] = f.p[P(g, g.n1m, j, k)];
  f.p[P(g, g.n1, j, k)] = f.p[P(g, 1, j, k)];
}

void launch_prdic_uvw(hipStream_t s, const Geo &g, const Fld &f) {
  if (g.xprdic) launch_xghost_uvw(s, g, f);
  LW_LAUNCH("k_zghost_uvw", (double)(g.n1 + 1) * (g.n2 + 1) * 8.0 * 12, 0, k_zghost_uvw,
            dim3((g.n1 + 64) / 64, g.n2 + 1), dim3(64), 0, s, g, f);
}

void launch_retrv_mom(hipStream_t s, const Geo &g, const Fld &f) {
  LW_LAUNCH("k_retrv_mom", (double)(g.n2 + 1) * (g.n3 + 1) * 8.0 * 9, 0, k_retrv_mom,
            dim3((g.n2 + 64) / 64, g.n3 + 1), dim3(64), 0, s, g, f);
  if (g.xprdic) launch_xghost_uvw(s, g, f);
  LW_LAUNCH("k_zghost_uvw", (double)(g.n1 + 1) * (g.n2 + 1) * 8.0 * 12, 0, k_zghost_uvw,
            dim3((g.n1 + 64) / 64, g.n2 + 1), dim3(64), 0, s, g, f);
}

// ---------------------------------------------------------------------------
// divgs, slv_cont.f90:2-208. The IB mass source is applied per cell in the
// reference's loop order (u own, u from i+1, v own, v from j+1, w own, w from
// k+1), which reproduces its sequential accumulation without atomics.
// ---------------------------------------------------------------------------
LW_KERNEL void k_divgs(LW_KP(1024, 0), const Geo &g, const Sub &sb, const Fld &f, int masson) {
  const int i = 1 + blockIdx.x * 64 + threadIdx.x;
  const int j = 1 + blockIdx.y * 4 + threadIdx.y;
  const int k = 1 + blockIdx.z;
  if (i > g.n1m || j > g.n2m) return;
  const long c = P(g, i, j, k);
  const long dy = g.NX, dz = g.NX * g.NY;
  const double d1 = (f.u[c + 1] - f.u[c]) * g.f2fxi[i];
  const double d2 = (f.v[c + dy] - f.v[c]) * g.f2fyi[j];
  const double d3 = (f.w[c + dz] - f.w[c]) * g.f2fzi[k];
  double d = (d1 + d2 + d3);
  if (masson) {
    double qm = 0.0;
    const uint8_t m0 = f.mask[c];
    // u: own point, then point i+1 whose imv is i (non-periodic x: i+1 <= n1m)
    if (m0 & M_BODY_U) {
      const double val = f.u[c] * g.f2fxi[i];
      qm = qm - val;
      d = d + val;
    }
    if (i + 1 <= g.n1m && g.imv[i + 1] == i && (f.mask[c + 1] & M_BODY_U)) {
      const double val = f.u[c + 1] * g.f2fxi[i];
      qm = qm + val;
      d = d - val;
    }
    if (m0 & M_BODY_V) {
      const double val = f.v[c] * g.f2fyi[j];
      qm = qm - val;
      d = d + val;
    }
    if (j + 1 <= g.n2m && g.jmv[j + 1] == j && (f.mask[c + dy] & M_BODY_V)) {
      const double val = f.v[c + dy] * g.f2fyi[j];
      qm = qm + val;
      d = d - val;
    }
    if (m0 & M_BODY_W) {
      const double val = f.w[c] * g.f2fzi[k];
      qm = qm - val;
      d = d + val;
    }
    {
      const int kp = g.kpv[k];  // the point whose kmv is k
      const long cp = P(g, i, j, kp);
      if (g.kmv[kp] == k && (f.mask[cp] & M_BODY_W)) {
        const double val = f.w[cp] * g.f2fzi[k];
        qm = qm + val;
        d = d - val;
      }
    }
    f.qmass[c] = qm;
  }
  f.div[c] = d * sb.dtconsti;
}

void launch_divgs(hipStream_t s, const Geo &g, const Sub &sb, const Fld &f, int masson) {
  // u v w in, div out (+ qmass out with the mass source)
  LW_LAUNCH("k_divgs", pts(g) * 8.0 * (masson ? 5 : 4), 0, k_divgs, grid3(g.n1m, g.n2m, g.n3m), BLK3, 0, s, g,
            sb, f, masson);
}

// ---------------------------------------------------------------------------
// ucalc (slv_auxi.f90:233-416) and pcalc (419-493) fused; the outlet-plane
// mean of phi is reduced first (block partials into f.obc[5 * kObc..], then
// one fixed-order sum)
// ---------------------------------------------------------------------------
LW_KERNEL void k_phiref(LW_KP(kObcT, 0), const Geo &g, const Fld &f) {
  __shared__ double sh[kObcT / 64 + 1];
  double a = 0.0;
  for (int q = blockIdx.x * kObcT + threadIdx.x; q < g.n2m * g.n3m; q += kObc * kObcT) {
    const int j = 1 + q % g.n2m, k = 1 + q / g.n2m;
    a += f.phi[P(g, g.n1m, j, k)] * g.f2fy[j] * g.f2fz[k];
  }
  const double s = block_sum<kObcT>(a, sh);
  if (threadIdx.x == 0) f.obc[5 * kObc + blockIdx.x] = s;
}

LW_KERNEL void k_phiref_sum(LW_KP(kObcT, 0), const Geo &g, const Fld &f) {
  __shared__ double sh[kObcT / 64 + 1];
  const double s = obc_total(f.obc + 5 * kObc, sh);
  if (threadIdx.x == 0) f.red[0] = s / (g.yl * g.zl);
}

LW_KERNEL void k_ucalc_pcalc(LW_KP(1024, 0), const Geo &g, const Sub &sb, const Fld &f) {
  const int i = 1 + blockIdx.x * 64 + threadIdx.x;
  const int j = 1 + blockIdx.y * 4 + threadIdx.y;
  const int k = 1 + blockIdx.z;
  if (i > g.n1m || j > g.n2m) return;
  const long c = P(g, i, j, k);
  const double *phi = f.phi;
  const double dc = sb.dtconst;
  if (i >= g.ib) {
    const int im = g.imv[i];
    f.u[c] = f.u[c] - dc * (phi[c] - phi[P(g, im, j, k)]) * g.c2cxi[i];
  }
  if (j >= 2) f.v[c] = f.v[c] - dc * (phi[c] - phi[c - g.NX]) * g.c2cyi[j];
  if (k >= g.kb) {
    const int km = g.kmv[k];
    f.w[c] = f.w[c] - dc * (phi[c] - phi[P(g, i, j, km)]) * g.c2czi[k];
  }
  const double pn = f.p[c] + phi[c] - 0.5 * sb.dtconst * f.div[c] * 1. / sb.re;
  f.p[c] = pn - f.red[0];
}

LW_KERNEL void k_zghost_p(LW_KP(1024, 0), const Geo &g, const Fld &f) {
  const int i = 1 + blockIdx.x * 64 + threadIdx.x;
  const int j = 1 + blockIdx.y;
  if (i > g.n1m) return;
  f.p[P(g, i, j, 0)] = f.p[P(g, i, j, g.n3m)];
  f.p[P(g, i, j, g.n3)] = f.p[P(g, i, j, 1)];
}

void launch_projection(hipStream_t s, const Geo &g, const Sub &sb, const Fld &f) {
  LW_LAUNCH("k_phiref", (double)g.n2m * g.n3m * 8.0, 0, k_phiref, dim3(kObc), dim3(kObcT), 0, s, g, f);
  LW_LAUNCH("k_phiref_sum", kObc * 8.0, 0, k_phiref_sum, dim3(1), dim3(kObcT), 0, s, g, f);
  // phi div in, u v w p in and out
  LW_LAUNCH("k_ucalc_pcalc", pts(g) * 8.0 * 10, 0, k_ucalc_pcalc, grid3(g.n1m, g.n2m, g.n3m), BLK3, 0, s, g, sb,
            f);
  LW_LAUNCH("k_zghost_uvw", (double)(g.n1 + 1) * (g.n2 + 1) * 8.0 * 12, 0, k_zghost_uvw,
            dim3((g.n1 + 64) / 64, g.n2 + 1), dim3(64), 0, s, g, f);
  if (g.xprdic) launch_xghost_uvw(s, g, f);
  LW_LAUNCH("k_zghost_p", (double)g.n1m * g.n2m * 8.0 * 4, 0, k_zghost_p, dim3((g.n1m + 63) / 64, g.n2m), dim3(64),
            0, s, g, f);
  if (g.xprdic)
    LW_LAUNCH("k_xghost_p", (double)g.n2m * g.n3m * 8.0 * 4, 0, k_xghost_p, dim3((g.n2m + 63) / 64, g.n3m), dim3(64),
              0, s, g, f);
}

// ---------------------------------------------------------------------------
// energy right-hand side, slv_engy.f90:29-198 (ICONJG = 0, IDECOMP = 0)
// ---------------------------------------------------------------------------
// energy right-hand side, slv_engy.f90:29-198 (ICONJG = 0, IDECOMP = 0)
// ---------------------------------------------------------------------------
LW_KERNEL void k_rhs_t(LW_KP(1024, 0), const Geo &g, const Sub &sb, const Fld &f, int les) {
  const int i = 1 + blockIdx.x * 64 + threadIdx.x;
  const int j = 1 + blockIdx.y * 4 + threadIdx.y;
  const int k = 1 + blockIdx.z;
  if (i > g.n1m || j > g.n2m) return;
  const double *u = f.u, *v = f.v, *w = f.w, *t = f.t;
  const int kp = g.kpv[k], km = g.kmv[k], jp = g.jpv[j], jm = g.jmv[j], ip = g.ipv[i], im = g.imv[i];
  const double fiu = g.fixiu[i], fil = g.fixil[i], fju = g.fixju[j], fjl = g.fixjl[j], fku = g.fixku[k],
               fkl = g.fixkl[k];
  const long c = P(g, i, j, k);
  const long cip = P(g, ip, j, k), cim = P(g, im, j, k), cjp = P(g, i, jp, k), cjm = P(g, i, jm, k),
             ckp = P(g, i, j, kp), ckm = P(g, i, j, km);
  const double te = 0.5 * (g.f2fx[i] * t[cip] + g.f2fx[ip] * t[c]) * g.c2cxi[ip] * (1. - fiu) + t[cip] * fiu;
  const double tw = 0.5 * (g.f2fx[im] * t[c] + g.f2fx[i] * t[cim]) * g.c2cxi[i] * (1. - fil) + t[cim] * fil;
  const double tn = 0.5 * (g.f2fy[j] * t[cjp] + g.f2fy[jp] * t[c]) * g.c2cyi[jp] * (1. - fju) + t[cjp] * fju;
  const double ts = 0.5 * (g.f2fy[jm] * t[c] + g.f2fy[j] * t[cjm]) * g.c2cyi[j] * (1. - fjl) + t[cjm] * fjl;
  const double tc = 0.5 * (g.f2fz[k] * t[ckp] + g.f2fz[kp] * t[c]) * g.c2czi[kp] * (1. - fku) + t[ckp] * fku;
  const double tf = 0.5 * (g.f2fz[km] * t[c] + g.f2fz[k] * t[ckm]) * g.c2czi[k] * (1. - fkl) + t[ckm] * fkl;
  const double ant1 = (te * u[cip] - tw * u[c]) * g.f2fxi[i];
  const double ant2 = (tn * v[cjp] - ts * v[c]) * g.f2fyi[j];
  const double ant3 = (tc * w[ckp] - tf * w[c]) * g.f2fzi[k];
  const double omega = 1.0;
  double rk3t = -omega * (ant1 + ant2 + ant3) + 0.0;
  const double alt1 = (t[cip] - t[c]) * g.c2cxi[ip];
  const double alt2 = (t[c] - t[cim]) * g.c2cxi[i];
  const double alt3 = (t[cjp] - t[c]) * g.c2cyi[jp];
  const double alt4 = (t[c] - t[cjm]) * g.c2cyi[j];
  const double alt5 = (t[ckp] - t[c]) * g.c2czi[kp];
  const double alt6 = (t[c] - t[ckm]) * g.c2czi[k];
  const double altx = (alt1 - alt2) * g.f2fxi[i];
  const double alty = (alt3 - alt4) * g.f2fyi[j];
  const double altz = (alt5 - alt6) * g.f2fzi[k];
  double alt = 1. / (sb.re * sb.pr) * (altx + alty + altz);
  if (les) {
    const double *a1 = f.al1[0], *a2 = f.al1[1], *a3 = f.al1[2];
    const double taltx = g.f2fxi[i] * (a1[cip] * g.c2cxi[ip] * (t[cip] - t[c]) -
                                       a1[c] * g.c2cxi[i] * (t[c] - t[cim]));
    const double talty = g.f2fyi[j] * (a2[cjp] * g.c2cyi[jp] * (t[cjp] - t[c]) -
                                       a2[c] * g.c2cyi[j] * (t[c] - t[cjm]));
    const double taltz = g.f2fzi[k] * (a3[ckp] * g.c2czi[kp] * (t[ckp] - t[c]) -
                                       a3[c] * g.c2czi[k] * (t[c] - t[ckm]));
    alt = alt + 1.0 * omega * (taltx + talty + taltz);
    // rhssgs_t leaves zero in rhs1(:,:,:,4) (sgs.f90:1776-1803)
    rk3t = rk3t + 1.0 * omega * (0.0 + 0.0 + 0.0);
  }
  const double ro = f.rk3o[3][c];
  f.rhs[3][c] = sb.dt * (sb.gam * rk3t + sb.ro * ro + 2. * sb.alpha * alt);
  // the new term into the spare buffer; the host swaps rk3o and rk3oo
  f.rk3oo[3][c] = rk3t;
}

void launch_rhs_t(hipStream_t s, const Geo &g, const Sub &sb, Fld &f, int les) {
  // reads t u v w rk3o (+ three diffusivities with LES), writes rhs and the
  // new term
  LW_LAUNCH("k_rhs_t", pts(g) * 8.0 * (les ? 10 : 7), 0, k_rhs_t, grid3(g.n1m, g.n2m, g.n3m), BLK3, 0, s, g, sb, f,
            les);
  std::swap(f.rk3o[3], f.rk3oo[3]);
}


// ---------------------------------------------------------------------------
// convbc_t (slv_engy.f90:309-413) + rhsincorpbc_t (416-457), ICONJG = 0,
// with the plane reductions done as in the momentum outlet
// ---------------------------------------------------------------------------
LW_KERNEL void k_obc_out_t(LW_KP(kObcT, 0), const Geo &g, const Sub &sb, const Fld &f) {
  __shared__ double sh[kObcT / 64 + 1];
  const ObcIn in = obc_in(g, f, sh);
  const int n1 = g.n1, n1m = g.n1m, n2m = g.n2m, n3m = g.n3m;
  const double *u = f.u, *t = f.t;
  const double ucoef = in.ubar * sb.dtconst * g.f2fxi[n1m];
  const double vwcoef = in.ubar * sb.dtconst * g.c2cxi[n1];
  double a = 0.0;
  for (int q = blockIdx.x * kObcT + threadIdx.x; q < n2m * n3m; q += kObc * kObcT) {
    const int j = 1 + q % n2m, k = 1 + q / n2m;
    const long pb = j + g.NY * k;
    const double uo = u[P(g, n1, j, k)] - ucoef * (u[P(g, n1, j, k)] - u[P(g, n1m, j, k)]);
    f.uout[pb] = uo;
    if (f.mask[P(g, n1m, j, k)] & 64)
      f.tout[pb] = t[P(g, n1m, j, k)];
    else
      f.tout[pb] = t[P(g, n1, j, k)] - vwcoef * (t[P(g, n1, j, k)] - t[P(g, n1m, j, k)]);
    a += uo * g.f2fy[j] * g.f2fz[k];
  }
  const double s = block_sum<kObcT>(a, sh);
  if (threadIdx.x == 0) f.obc[4 * kObc + blockIdx.x] = s;
  (void)n3m;
}

LW_KERNEL void k_obc_apply_t(LW_KP(kObcT, 0), const Geo &g, const Sub &sb, const Fld &f, int les) {
  __shared__ double sh[kObcT / 64 + 1];
  const ObcIn in = obc_in(g, f, sh);
  const double qout = obc_total(f.obc + 4 * kObc, sh);
  const double qratio = fabs(qout) > 1.0e-14 ? in.qin / qout : 1.0;
  const int n1 = g.n1, n1m = g.n1m, n2m = g.n2m;
  const double *t = f.t;
  const double cre2 = les ? 2. * sb.re * sb.pr : 0.;
  for (int q = blockIdx.x * kObcT + threadIdx.x; q < n2m * g.n3m; q += kObc * kObcT) {
    const int j = 1 + q % n2m, k = 1 + q / n2m;
    const long pb = j + g.NY * k;
    double to = f.tout[pb];
    if (!(f.mask[P(g, n1m, j, k)] & 64)) {
      to = to * qratio;
      f.tout[pb] = to;
    }
    const double dt = to - t[P(g, n1, j, k)];
    f.dtout[pb] = dt;
    const long c = P(g, n1m, j, k);
    f.rhs[3][c] = f.rhs[3][c] - sb.acoef * 1. * 1. * g.ciu[n1m] * (1. + cre2 * 1.0 * f.al[c]) * dt;
  }
}

void launch_convbc_t(hipStream_t s, const Geo &g, const Sub &sb, const Fld &f, const Body &, int les,
                     const lw_settings &) {
  const double pl = (double)g.n2m * g.n3m;
  launch_obc_flux(s, g, f);
  LW_LAUNCH("k_obc_out_t", pl * 8.0 * 7, 0, k_obc_out_t, dim3(kObc), dim3(kObcT), 0, s, g, sb, f);
  LW_LAUNCH("k_obc_apply_t", pl * 8.0 * 7, 0, k_obc_apply_t, dim3(kObc), dim3(kObcT), 0, s, g, sb, f, les);
  LW_LAUNCH("k_obc_ghost", (double)(g.n2 + 1) * 8.0 * 4, 0, k_obc_ghost, dim3((g.n2 + 64) / 64), dim3(64), 0, s, g,
            f.dtout, (double *)nullptr, (double *)nullptr);
}

// retrv_t, prdic_adj_t (periodic z) and wallbc_t (slv_engy.f90:612-812)
LW_KERNEL void k_retrv_t(LW_KP(1024, 0), const Geo &g, const Fld &f) {
  const int j = 1 + blockIdx.x * 64 + threadIdx.x;
  const int k = 1 + blockIdx.y;
  if (j > g.n2m) return;
  f.t[P(g, g.n1, j, k)] = f.tout[j + g.NY * k];
}

LW_KERNEL void k_zghost_t(LW_KP(1024, 0), const Geo &g, const Fld &f) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  const int j = blockIdx.y;
  if (i > g.n1) return;
  f.t[P(g, i, j, 0)] = f.t[P(g, i, j, g.n3m)];
  f.t[P(g, i, j, g.n3)] = f.t[P(g, i, j, 1)];
}

LW_KERNEL void k_inlet_t(LW_KP(1024, 0), const Geo &g, const Fld &f, double tsol, double tflu) {
  const int j = blockIdx.x * 64 + threadIdx.x;
  const int k = blockIdx.y;
  if (j > g.n2) return;
  const long c = P(g, 0, j, k);
  f.t[c] = (f.mask[c] & 128) ? tsol : tflu;
}

LW_KERNEL void k_ywall_t(LW_KP(1024, 0), const Geo &g, const Fld &f, int bb, int bt, double vb, double vt) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  const int k = blockIdx.y;
  if (i > g.n1) return;
  const long c0 = P(g, i, 0, k), c1 = P(g, i, 1, k), cn = P(g, i, g.n2, k), cnm = P(g, i, g.n2m, k);
  if (bb == 0) f.t[c0] = vb;
  else if (bb == 1) f.t[c0] = f.t[c1] - vb / g.c2cyi[1];
  if (bt == 0) f.t[cn] = vt;
  else if (bt == 1) f.t[cn] = f.t[cnm] + vt / g.c2cyi[g.n2];
}

void launch_energy_bcs(hipStream_t s, const Geo &g, const Fld &f, const Body &, const lw_settings &st) {
  LW_LAUNCH("k_retrv_t", (double)g.n2m * g.n3m * 8.0 * 3, 0, k_retrv_t, dim3((g.n2m + 63) / 64, g.n3m), dim3(64),
            0, s, g, f);
  LW_LAUNCH("k_zghost_t", (double)(g.n1 + 1) * (g.n2 + 1) * 8.0 * 4, 0, k_zghost_t, dim3((g.n1 + 64) / 64, g.n2 + 1),
            dim3(64), 0, s, g, f);
  LW_LAUNCH("k_inlet_t", (double)(g.n2 + 1) * (g.n3 + 1) * 8.0, 0, k_inlet_t, dim3((g.n2 + 64) / 64, g.n3 + 1),
            dim3(64), 0, s, g, f, st.tsol_i, st.tflu_i);
  LW_LAUNCH("k_ywall_t", (double)(g.n1 + 1) * (g.n3 + 1) * 8.0 * 4, 0, k_ywall_t, dim3((g.n1 + 64) / 64, g.n3 + 1), dim3(64), 0, s, g, f, (int)st.bc_t_ybtm,
                     (int)st.bc_t_ytop, st.val_t_ybtm, st.val_t_ytop);
}

// ---------------------------------------------------------------------------
// CFL (slv_auxi.f90:53-102), convergence_check / masscheck
// (misc_init.f90:1047-1104), lagforce (slv_auxi.f90:519-570)
// ---------------------------------------------------------------------------
// block-wide max of non-negative values, then one atomic per workgroup
__device__ inline void block_max_to(double v, unsigned long long *dst, double *sh) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_down(v, o, 64));
  const int t = threadIdx.y * 64 + threadIdx.x;
  if ((t & 63) == 0) sh[t >> 6] = v;
  __syncthreads();
  if (t == 0) atomic_max_nonneg(dst, fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3])));
}

// each thread walks k for one (i, j); one workgroup covers 64 x 4 columns
LW_KERNEL void k_cfl(LW_KP(256, 0), const Geo &g, const Fld &f, double dt) {
  __shared__ double sh[4];
  const int i = 1 + blockIdx.x * 64 + threadIdx.x;
  const int j = 1 + blockIdx.y * 4 + threadIdx.y;
  double c0 = 0.0;
  if (i <= g.n1m && j <= g.n2m) {
    for (int k = 1; k <= g.n3m; ++k) {
      const long c = P(g, i, j, k);
      const double c1 = fabs(f.u[c] + f.u[c + 1]) * 0.5 * g.f2fxi[i];
      const double c2 = fabs(f.v[c] + f.v[c + g.NX]) * 0.5 * g.f2fyi[j];
      const double c3 = fabs(f.w[c] + f.w[c + g.NX * g.NY]) * 0.5 * g.f2fzi[k];
      c0 = fmax(c0, (c1 + c2 + c3) * dt);
    }
  }
  block_max_to(c0, &f.redmax[0], sh);
}

void launch_cfl(hipStream_t s, const Geo &g, const Fld &f, double dt) {
  LW_LAUNCH("k_cfl", pts(g) * 24.0, 0, k_cfl, grid3(g.n1m, g.n2m, 1), BLK3, 0, s, g, f, dt);
}

LW_KERNEL void k_diag(LW_KP(256, 0), const Geo &g, const Fld &f, int masson) {
  __shared__ double sh[4];
  const int i = 1 + blockIdx.x * 64 + threadIdx.x;
  const int j = 1 + blockIdx.y * 4 + threadIdx.y;
  double dv = 0.0, qm = 0.0;
  if (i <= g.n1m && j <= g.n2m) {
    for (int k = 1; k <= g.n3m; ++k) {
      const long c = P(g, i, j, k);
      const double d1 = (f.u[c + 1] - f.u[c]) * g.f2fxi[i];
      const double d2 = (f.v[c + g.NX] - f.v[c]) * g.f2fyi[j];
      const double d3 = (f.w[c + g.NX * g.NY] - f.w[c]) * g.f2fzi[k];
      const double dd = masson ? d1 + d2 + d3 - f.qmass[c] : d1 + d2 + d3;
      dv = fmax(dv, fabs(dd));
      if (masson) qm = fmax(qm, fabs(f.qmass[c]));
    }
  }
  block_max_to(dv, &f.redmax[1], sh);
  __syncthreads();
  block_max_to(qm, &f.redmax[2], sh);
}

void launch_diag(hipStream_t s, const Geo &g, const Fld &f, int masson) {
  LW_LAUNCH("k_diag", pts(g) * 8.0 * (masson ? 4 : 3), 0, k_diag, grid3(g.n1m, g.n2m, 1), BLK3, 0, s, g, f,
            masson);
}

// lagforce (slv_auxi.f90:519-570): volume-weighted sums of the IB forcing of
// u, v, w over the forcing points; block partials, then one fixed-order sum
LW_KERNEL void k_lagforce(LW_KP(256, 0), const Geo &g, const Sub &sb, const Fld &f, const IBComp &u, const IBComp &v, const IBComp &w) {
  __shared__ double sh[4];
  const IBComp *cs[3] = {&u, &v, &w};
  // unrolled: constant component indices keep the arguments out of scratch
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const IBComp &ib = *cs[c];
    double acc = 0.0;
    for (int n = blockIdx.x * 256 + threadIdx.x; n < ib.nbody; n += gridDim.x * 256) {
      const int ii = ib.ifc[n], jj = ib.jfc[n], kk = ib.kfc[n];
      const long q = P(g, ii, jj, kk);
      const double rr = ib.dudtr[n] - (sb.gam * f.rk3o[c][q] + sb.ro * f.rk3oo[c][q]);
      double vol;
      if (c == 0) vol = g.c2cx[ii] * g.f2fy[jj] * g.f2fz[kk];
      else if (c == 1) vol = g.f2fx[ii] * g.c2cy[jj] * g.f2fz[kk];
      else vol = g.f2fx[ii] * g.f2fy[jj] * g.c2cz[kk];
      acc += vol * rr;
    }
    const double s = block_sum<256>(acc, sh);
    if (threadIdx.x == 0) f.partial[c * kRedBlocks + blockIdx.x] = s;
  }
}

LW_KERNEL void k_lagforce_sum(LW_KP(256, 0), const Sub &sb, const Fld &f) {
  __shared__ double sh[4];
  for (int c = 0; c < 3; ++c) {
    double a = 0.0;
    for (int q = threadIdx.x; q < kRedBlocks; q += 256) a += f.partial[c * kRedBlocks + q];
    const double s = block_sum<256>(a, sh);
    if (threadIdx.x == 0) f.red[8 + c] = (sb.msub == 1 ? 0.0 : f.red[8 + c]) + s;
  }
}

void launch_lagforce(hipStream_t s, const Geo &g, const Sub &sb, const Fld &f, IBComp *ib) {
  LW_LAUNCH("k_lagforce", (double)(ib[0].nbody + ib[1].nbody + ib[2].nbody) * (3 * 4 + 8 * 5), 0, k_lagforce,
            dim3(kRedBlocks), dim3(256), 0, s, g, sb, f, ib[0], ib[1], ib[2]);
  LW_LAUNCH("k_lagforce_sum", 3.0 * kRedBlocks * 8, 0, k_lagforce_sum, dim3(1), dim3(256), 0, s, sb, f);
}


// ---------------------------------------------------------------------------
// per-step diagnostics of the driver loop (solver.f90:158-172): the IB and
// wall-shear parts of draglift (slv_auxi.f90:669-840), the boundary fluxes of
// heatflux (slv_auxi.f90:573-666) and the bulk temperatures of writehistory
// (misc_hist.f90:73-181). Each sum s is reduced by kDiagB workgroups into
// block partials, then one workgroup sums every s in one fixed order into
// red[kDiagRed + s]. The fluid tests of the boundary planes come from masks
// computed on the host with the reference's funcbody test.
// ---------------------------------------------------------------------------
constexpr int kDiagB = 128;

__device__ inline double diag_term(const Geo &g, const Fld &f, const DiagPlanes &dp, const IBComp *ib, int s,
                                   long q, double re) {
  const int n1 = g.n1, n2 = g.n2, n3 = g.n3, n1m = g.n1m, n2m = g.n2m, n3m = g.n3m;
  if (s <= D_IBW) {  // -fcvavg * u-cell volume (the reference uses c2cx for all three)
    const IBComp &c = ib[s];
    const int ii = c.ifc[q], jj = c.jfc[q], kk = c.kfc[q];
    return c.fcvavg[q] * (g.c2cx[ii] * g.f2fy[jj] * g.f2fz[kk]);
  }
  if (s <= D_YT_W) {  // y walls: one-sided second-order gradient at the wall
    const bool top = s >= D_YT_U;
    const int i = 1 + (int)(q % n1m), k = 1 + (int)(q / n1m);
    if (!(top ? dp.wyt : dp.wyb)[i + (long)(n1 + 1) * k]) return 0.0;
    const double dy1 = top ? 1.0 / g.c2cyi[n2] : 1.0 / g.c2cyi[1];
    const double dy2 = top ? dy1 + 1.0 / g.c2cyi[n2m] : dy1 + 1.0 / g.c2cyi[2];
    const int j1 = top ? n2m : 1, j2 = top ? n2m - 1 : 2;
    const double *a = (s == D_YB_U || s == D_YT_U) ? f.u : f.w;
    const double grad = (a[P(g, i, j1, k)] * (dy2 * dy2) - a[P(g, i, j2, k)] * (dy1 * dy1)) / (dy1 * dy2 * (dy2 - dy1));
    const double tmp = grad * g.c2cx[i] * g.f2fz[k];
    return tmp / re;
  }
  if (s <= D_ZT_V) {  // z walls
    const bool top = s >= D_ZT_U;
    const int i = 1 + (int)(q % n1m), j = 1 + (int)(q / n1m);
    if (!(top ? dp.wzt : dp.wzb)[i + (long)(n1 + 1) * j]) return 0.0;
    const double dz1 = top ? 1.0 / g.c2czi[n3] : 1.0 / g.c2czi[1];
    const double dz2 = top ? dz1 + 1.0 / g.c2czi[n3m] : dz1 + 1.0 / g.c2czi[2];
    const int k1 = top ? n3m : 1, k2 = top ? n3m - 1 : 2;
    const double *a = (s == D_ZB_U || s == D_ZT_U) ? f.u : f.v;
    const double grad = (a[P(g, i, j, k1)] * (dz2 * dz2) - a[P(g, i, j, k2)] * (dz1 * dz1)) / (dz1 * dz2 * (dz2 - dz1));
    const double tmp = grad * g.c2cx[i] * g.f2fy[j];
    return tmp / re;
  }
  if (s <= D_HXP) {  // x faces: u * T at the face * area
    const bool hi = s == D_HXP;
    const int j = 1 + (int)(q % n2m), k = 1 + (int)(q / n2m);
    if (!(hi ? dp.hxp : dp.hxm)[j + (long)(n2 + 1) * k]) return 0.0;
    const double area = g.f2fy[j] * g.f2fz[k];
    const double tface = hi ? 0.5 * (f.t[P(g, n1m, j, k)] + f.t[P(g, n1, j, k)])
                            : 0.5 * (f.t[P(g, 1, j, k)] + f.t[P(g, 0, j, k)]);
    return f.u[P(g, hi ? n1 : 1, j, k)] * tface * area;
  }
  if (s <= D_HYP) {  // y faces
    const bool hi = s == D_HYP;
    const int i = 1 + (int)(q % n1m), k = 1 + (int)(q / n1m);
    if (!(hi ? dp.hyp : dp.hym)[i + (long)(n1 + 1) * k]) return 0.0;
    const double area = g.f2fx[i] * g.f2fz[k];
    const double tface = hi ? 0.5 * (f.t[P(g, i, n2m, k)] + f.t[P(g, i, n2, k)])
                            : 0.5 * (f.t[P(g, i, 1, k)] + f.t[P(g, i, 0, k)]);
    return f.v[P(g, i, hi ? n2 : 1, k)] * tface * area;
  }
  if (s <= D_HZP) {  // z faces
    const bool hi = s == D_HZP;
    const int i = 1 + (int)(q % n1m), j = 1 + (int)(q / n1m);
    if (!(hi ? dp.hzp : dp.hzm)[i + (long)(n1 + 1) * j]) return 0.0;
    const double area = g.f2fx[i] * g.f2fy[j];
    const double tface = hi ? 0.5 * (f.t[P(g, i, j, n3m)] + f.t[P(g, i, j, n3)])
                            : 0.5 * (f.t[P(g, i, j, 1)] + f.t[P(g, i, j, 0)]);
    return f.w[P(g, i, j, hi ? n3 : 1)] * tface * area;
  }
  // bulk temperatures: fluid / solid by the cell-centre test, u-cell volume
  const int i = 1 + (int)(q % n1m), j = 1 + (int)((q / n1m) % n2m), k = 1 + (int)(q / ((long)n1m * n2m));
  const long c = P(g, i, j, k);
  const bool solid = f.mask[c] & M_SOLID_C;
  const double cellvol = g.c2cx[i] * g.f2fy[j] * g.f2fz[k];
  if (s == D_TFV) return solid ? 0.0 : cellvol;
  if (s == D_TFI) return solid ? 0.0 : f.t[c] * cellvol;
  if (s == D_TSV) return solid ? cellvol : 0.0;
  return solid ? f.t[c] * cellvol : 0.0;
}

__device__ inline long diag_size(const Geo &g, const IBComp *ib, int s) {
  if (s <= D_IBW) return ib[s].nbody;
  if (s <= D_ZT_V) return (long)g.n1m * (s <= D_YT_W ? g.n3m : g.n2m);
  if (s <= D_HXP) return (long)g.n2m * g.n3m;
  if (s <= D_HYP) return (long)g.n1m * g.n3m;
  if (s <= D_HZP) return (long)g.n1m * g.n2m;
  return (long)g.n1m * g.n2m * g.n3m;
}

// The four bulk-temperature sums over the volume share one pass: the
// workgroups of their four rows split the volume into 4*kDiagB chunks, and
// each forms the partials of all four sums for its chunk (kDiagP partials
// per sum; the other sums use the first kDiagB).
constexpr int kDiagP = 4 * kDiagB;

LW_KERNEL void k_step_diag(LW_KP(256, 0), const Geo &g, const Fld &f, const IBComp &u, const IBComp &v, const IBComp &w, const DiagPlanes &dp,
                                                   unsigned active, double re) {
  __shared__ double sh[4];
  const int s = blockIdx.y;
  if (!((active >> s) & 1u)) return;
  const IBComp ib[3] = {u, v, w};
  const long n = diag_size(g, ib, s);
  if (s >= D_TFV) {
    const int c = (s - D_TFV) * kDiagB + blockIdx.x;  // chunk 0 .. kDiagP-1
    double a[4] = {0.0, 0.0, 0.0, 0.0};
    for (long q = c * 256L + threadIdx.x; q < n; q += (long)kDiagP * 256) {
      const int i = 1 + (int)(q % g.n1m), j = 1 + (int)((q / g.n1m) % g.n2m), k = 1 + (int)(q / ((long)g.n1m * g.n2m));
      const long pc = P(g, i, j, k);
      const bool solid = f.mask[pc] & M_SOLID_C;
      const double cellvol = g.c2cx[i] * g.f2fy[j] * g.f2fz[k];
      const double tv = f.t[pc] * cellvol;
      a[0] += solid ? 0.0 : cellvol;  // D_TFV
      a[1] += solid ? 0.0 : tv;       // D_TFI
      a[2] += solid ? cellvol : 0.0;  // D_TSV
      a[3] += solid ? tv : 0.0;       // D_TSI
    }
    for (int t = 0; t < 4; ++t) {
      const double r = block_sum<256>(a[t], sh);
      if (threadIdx.x == 0) f.dpart[(D_TFV + t) * kDiagP + c] = r;
    }
    return;
  }
  double a = 0.0;
  for (long q = blockIdx.x * 256L + threadIdx.x; q < n; q += (long)kDiagB * 256) a += diag_term(g, f, dp, ib, s, q, re);
  const double r = block_sum<256>(a, sh);
  if (threadIdx.x == 0) f.dpart[s * kDiagP + blockIdx.x] = r;
}

LW_KERNEL void k_step_diag_sum(LW_KP(256, 0), const Fld &f, unsigned active) {
  __shared__ double sh[4];
  for (int s = 0; s < D_NSUM; ++s) {
    if (!((active >> s) & 1u)) continue;
    const int np = s >= D_TFV ? kDiagP : kDiagB;
    double v = 0.0;
    for (int q = threadIdx.x; q < np; q += 256) v += f.dpart[s * kDiagP + q];
    const double r = block_sum<256>(v, sh);
    if (threadIdx.x == 0) f.red[kDiagRed + s] = r;
  }
}

void launch_step_diag(hipStream_t s, const Geo &g, const Fld &f, IBComp *ib, const DiagPlanes &dp, unsigned active,
                      double re) {
  if (!active) return;
  double bytes = 0.0;
  if (g_prof)
    for (int q = 0; q < D_NSUM; ++q)
      if ((active >> q) & 1u)  // the four bulk-temperature sums: mask and T once (9 B per cell)
        bytes += q <= D_IBW ? 36.0 * (double)ib[q].nbody : q == D_TFV ? 9.0 * pts(g) : q > D_TFV ? 0.0 : 32.0 * 1e4;
  LW_LAUNCH("k_step_diag", bytes, 0, k_step_diag, dim3(kDiagB, D_NSUM), dim3(256), 0, s, g, f, ib[0], ib[1], ib[2], dp,
            active, re);
  LW_LAUNCH("k_step_diag_sum", D_NSUM * kDiagP * 8.0, 0, k_step_diag_sum, dim3(1), dim3(256), 0, s, f, active);
}


// the step's history values from the sums (draglift, slv_auxi.f90:700-840;
// heatflux, 636-666; writehistory, misc_hist.f90:145-178) and their running
// time averages from avg_tst on. hist: [0..2] cd, [3..5] cdavg, [6..8] heat
// flux, [9..11] its average, [12..13] bulk T fluid/solid, [14..15] their
// averages, [16..19] cd integral and duration, [20..23] heat-flux integral and
// duration, [24..26] bulk-T integral and duration.
LW_KERNEL void k_hist_update(LW_KP(1024, 0), const Fld &f, double *hist, unsigned active, int ibmon, int heat, int hstep, double time,
                              double dt, double avg_tst, double vol) {
  if (threadIdx.x != 0) return;
  const double *S = f.red + kDiagRed;
  const bool avg = time >= avg_tst;
  if (ibmon) {
    double cd[3] = {-S[D_IBU], -S[D_IBV], -S[D_IBW]};
    for (int l = 0; l < 3; ++l) cd[l] = cd[l] + f.red[8 + l];
    if ((active >> D_YB_U) & 1u) {
      cd[0] = cd[0] + S[D_YB_U];
      cd[2] = cd[2] + S[D_YB_W];
    }
    if ((active >> D_YT_U) & 1u) {
      cd[0] = cd[0] + S[D_YT_U];
      cd[2] = cd[2] + S[D_YT_W];
    }
    if ((active >> D_ZB_U) & 1u) {
      cd[0] = cd[0] + S[D_ZB_U];
      cd[1] = cd[1] + S[D_ZB_V];
    }
    if ((active >> D_ZT_U) & 1u) {
      cd[0] = cd[0] + S[D_ZT_U];
      cd[1] = cd[1] + S[D_ZT_V];
    }
    for (int l = 0; l < 3; ++l) {
      cd[l] = cd[l] / vol;
      hist[l] = cd[l];
      hist[3 + l] = 0.0;
    }
    if (avg) {
      for (int l = 0; l < 3; ++l) hist[16 + l] = hist[16 + l] + cd[l] * dt;
      hist[19] = hist[19] + dt;
      if (hist[19] > 0.0)
        for (int l = 0; l < 3; ++l) hist[3 + l] = hist[16 + l] / hist[19];
    }
  }
  if (heat) {
    const double hf[3] = {S[D_HXP] - S[D_HXM], S[D_HYP] - S[D_HYM], S[D_HZP] - S[D_HZM]};
    for (int l = 0; l < 3; ++l) {
      hist[6 + l] = hf[l];
      hist[9 + l] = 0.0;
    }
    if (avg) {
      for (int l = 0; l < 3; ++l) hist[20 + l] = hist[20 + l] + hf[l] * dt;
      hist[23] = hist[23] + dt;
      if (hist[23] > 0.0)
        for (int l = 0; l < 3; ++l) hist[9 + l] = hist[20 + l] / hist[23];
    }
    if (hstep) {
      const double tf = S[D_TFV] > 0.0 ? S[D_TFI] / S[D_TFV] : 0.0;
      const double ts = S[D_TSV] > 0.0 ? S[D_TSI] / S[D_TSV] : 0.0;
      hist[12] = tf;
      hist[13] = ts;
      hist[14] = 0.0;
      hist[15] = 0.0;
      if (avg) {
        hist[24] = hist[24] + tf * dt;
        hist[25] = hist[25] + ts * dt;
        hist[26] = hist[26] + dt;
        if (hist[26] > 0.0) {
          hist[14] = hist[24] / hist[26];
          hist[15] = hist[25] / hist[26];
        }
      }
    }
  }
}

void launch_hist_update(hipStream_t s, const Fld &f, double *hist, unsigned active, int ibmon, int heat, int hstep,
                        double time, double dt, double avg_tst, double vol) {
  LW_LAUNCH("k_hist_update", 0, 0, k_hist_update, dim3(1), dim3(64), 0, s, f, hist, active, ibmon, heat, hstep, time,
            dt, avg_tst, vol);
}

}  // namespace lw
