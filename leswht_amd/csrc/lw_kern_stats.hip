// Time-averaged statistics of the flow field, field_avg (misc_hist.f90:287-472):
// every step after AVG_TST, each cell centre accumulates dt-weighted sums of
// the cell-centre velocities, their products, p, p^2, T, T^2, the vorticity,
// its products and S:S. One thread per cell, i fastest, so the 23 running sums
// (interior arrays, i fastest) are read and written along i in one pass; the
// velocity stencil reads hit L2 for the neighbouring rows.
#include "lw_kernels.h"

namespace lw {
namespace {

__device__ inline long I3s(const Geo &g, int i, int j, int k) {
  return (long)(i - 1) + (long)g.n1m * ((long)(j - 1) + (long)g.n2m * (long)(k - 1));
}

// misc_hist.f90:339-364 (centre values, products) and 383-466 (gradients with
// literal i+-1, j+-1, k+-1 neighbours and the fix* one-sided switches,
// vorticity, strain), in the reference's operation order
LW_KERNEL void k_field_avg(LW_KP(256, 0), const Geo &g, const Fld &f, const AvgBuf &b, double dt, int heat) {
  const int i = blockIdx.x * 256 + threadIdx.x + 1, j = blockIdx.y + 1, k = blockIdx.z + 1;
  if (i > g.n1m) return;
  const double *u = f.u, *v = f.v, *w = f.w;
  const int n1 = g.n1, n2 = g.n2, n3 = g.n3;
#define U(a_, b_, c_) u[P(g, a_, b_, c_)]
#define V(a_, b_, c_) v[P(g, a_, b_, c_)]
#define W(a_, b_, c_) w[P(g, a_, b_, c_)]
  const long q = I3s(g, i, j, k);
  const double ucc = 0.5 * (U(i, j, k) + U(i + 1, j, k));
  const double vcc = 0.5 * (V(i, j, k) + V(i, j + 1, k));
  const double wcc = 0.5 * (W(i, j, k) + W(i, j, k + 1));
  double *const *a = b.a;
  a[0][q] = a[0][q] + ucc * dt;
  a[1][q] = a[1][q] + vcc * dt;
  a[2][q] = a[2][q] + wcc * dt;
  a[3][q] = a[3][q] + (ucc * ucc) * dt;
  a[4][q] = a[4][q] + ucc * vcc * dt;
  a[5][q] = a[5][q] + ucc * wcc * dt;
  a[6][q] = a[6][q] + (vcc * vcc) * dt;
  a[7][q] = a[7][q] + vcc * wcc * dt;
  a[8][q] = a[8][q] + (wcc * wcc) * dt;
  if (heat) {
    const double tt = f.t[P(g, i, j, k)];
    a[21][q] = a[21][q] + tt * dt;
    a[22][q] = a[22][q] + (tt * tt) * dt;
  }
  const double pp = f.p[P(g, i, j, k)];
  a[9][q] = a[9][q] + pp * dt;
  a[10][q] = a[10][q] + (pp * pp) * dt;

  const double fju = g.fixju[j], fjl = g.fixjl[j], fku = g.fixku[k], fkl = g.fixkl[k], fiu = g.fixiu[i],
               fil = g.fixil[i];
  const double vg11 = g.f2fxi[i] * (U(i + 1, j, k) - U(i, j, k));
  double up = g.c2cyi[j + 1] * 0.25 *
                  (g.f2fy[j + 1] * (U(i, j, k) + U(i + 1, j, k)) + g.f2fy[j] * (U(i, j + 1, k) + U(i + 1, j + 1, k))) *
                  (1. - fju) +
              0.5 * (U(i, n2, k) + U(i + 1, n2, k)) * fju;
  double um = g.c2cyi[j] * 0.25 *
                  (g.f2fy[j] * (U(i, j - 1, k) + U(i + 1, j - 1, k)) + g.f2fy[j - 1] * (U(i, j, k) + U(i + 1, j, k))) *
                  (1. - fjl) +
              0.5 * (U(i, 0, k) + U(i + 1, 0, k)) * fjl;
  const double vg12 = g.f2fyi[j] * (up - um);
  up = g.c2czi[k + 1] * 0.25 *
           (g.f2fz[k + 1] * (U(i, j, k) + U(i + 1, j, k)) + g.f2fz[k] * (U(i, j, k + 1) + U(i + 1, j, k + 1))) *
           (1. - fku) +
       0.5 * (U(i, j, n3) + U(i + 1, j, n3)) * fku;
  um = g.c2czi[k] * 0.25 *
           (g.f2fz[k] * (U(i, j, k - 1) + U(i + 1, j, k - 1)) + g.f2fz[k - 1] * (U(i, j, k) + U(i + 1, j, k))) *
           (1. - fkl) +
       0.5 * (U(i, j, 0) + U(i + 1, j, 0)) * fkl;
  const double vg13 = g.f2fzi[k] * (up - um);

  double vp = g.c2cxi[i + 1] * 0.25 *
                  (g.f2fx[i + 1] * (V(i, j, k) + V(i, j + 1, k)) + g.f2fx[i] * (V(i + 1, j, k) + V(i + 1, j + 1, k))) *
                  (1. - fiu) +
              0.5 * (V(n1, j, k) + V(n1, j + 1, k)) * fiu;
  double vm = g.c2cxi[i] * 0.25 *
                  (g.f2fx[i] * (V(i - 1, j, k) + V(i - 1, j + 1, k)) + g.f2fx[i - 1] * (V(i, j, k) + V(i, j + 1, k))) *
                  (1. - fil) +
              0.5 * (V(0, j, k) + V(0, j + 1, k)) * fil;
  const double vg21 = g.f2fxi[i] * (vp - vm);
  const double vg22 = g.f2fyi[j] * (V(i, j + 1, k) - V(i, j, k));
  vp = g.c2czi[k + 1] * 0.25 *
           (g.f2fz[k + 1] * (V(i, j, k) + V(i, j + 1, k)) + g.f2fz[k] * (V(i, j, k + 1) + V(i, j + 1, k + 1))) *
           (1. - fku) +
       0.5 * (V(i, j, n3) + V(i, j + 1, n3)) * fku;
  vm = g.c2czi[k] * 0.25 *
           (g.f2fz[k] * (V(i, j, k - 1) + V(i, j + 1, k - 1)) + g.f2fz[k - 1] * (V(i, j, k) + V(i, j + 1, k))) *
           (1. - fkl) +
       0.5 * (V(i, j, 0) + V(i, j + 1, 0)) * fkl;
  const double vg23 = g.f2fzi[k] * (vp - vm);

  double wp = g.c2cxi[i + 1] * 0.25 *
                  (g.f2fx[i + 1] * (W(i, j, k) + W(i, j, k + 1)) + g.f2fx[i] * (W(i + 1, j, k) + W(i + 1, j, k + 1))) *
                  (1. - fiu) +
              0.5 * (W(n1, j, k) + W(n1, j, k + 1)) * fiu;
  double wm = g.c2cxi[i] * 0.25 *
                  (g.f2fx[i] * (W(i - 1, j, k) + W(i - 1, j, k + 1)) + g.f2fx[i - 1] * (W(i, j, k) + W(i, j, k + 1))) *
                  (1. - fil) +
              0.5 * (W(0, j, k) + W(0, j, k + 1)) * fil;
  const double vg31 = g.f2fxi[i] * (wp - wm);
  wp = g.c2cyi[j + 1] * 0.25 *
           (g.f2fy[j + 1] * (W(i, j, k) + W(i, j, k + 1)) + g.f2fy[j] * (W(i, j + 1, k) + W(i, j + 1, k + 1))) *
           (1. - fju) +
       0.5 * (W(i, n2, k) + W(i, n2, k + 1)) * fju;
  wm = g.c2cyi[j] * 0.25 *
           (g.f2fy[j] * (W(i, j - 1, k) + W(i, j - 1, k + 1)) + g.f2fy[j - 1] * (W(i, j, k) + W(i, j, k + 1))) *
           (1. - fjl) +
       0.5 * (W(i, 0, k) + W(i, 0, k + 1)) * fjl;
  const double vg32 = g.f2fyi[j] * (wp - wm);
  const double vg33 = g.f2fzi[k] * (W(i, j, k + 1) - W(i, j, k));
#undef U
#undef V
#undef W
  const double sr1 = vg11, sr2 = 0.5 * (vg12 + vg21), sr3 = 0.5 * (vg13 + vg31), sr4 = vg22,
               sr5 = 0.5 * (vg23 + vg32), sr6 = vg33;
  const double wx = vg32 - vg23, wy = vg13 - vg31, wz = vg21 - vg12;
  a[11][q] = a[11][q] + wx * dt;
  a[12][q] = a[12][q] + wy * dt;
  a[13][q] = a[13][q] + wz * dt;
  a[14][q] = a[14][q] + wx * wx * dt;
  a[15][q] = a[15][q] + wx * wy * dt;
  a[16][q] = a[16][q] + wx * wz * dt;
  a[17][q] = a[17][q] + wy * wy * dt;
  a[18][q] = a[18][q] + wy * wz * dt;
  a[19][q] = a[19][q] + wz * wz * dt;
  const double srsr = 2. * (sr2 * sr2) + (sr1 * sr1) + 2. * (sr3 * sr3) + (sr4 * sr4) + 2. * (sr5 * sr5) + (sr6 * sr6);
  a[20][q] = a[20][q] + srsr * dt;
}

// tracer (misc_hist.f90:184-209): the cell-centre velocities and the
// pressure of each traced cell (i, j, k) of n: out[c n + q] for c = u, v, w, p
LW_KERNEL void k_trace(LW_KP(1024, 0), const Geo &g, const Fld &f, const int *pts, int n, double *out) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  const int i = pts[3 * q], j = pts[3 * q + 1], k = pts[3 * q + 2];
  out[q] = 0.5 * (f.u[P(g, i, j, k)] + f.u[P(g, i + 1, j, k)]);
  out[n + q] = 0.5 * (f.v[P(g, i, j, k)] + f.v[P(g, i, j + 1, k)]);
  out[2 * n + q] = 0.5 * (f.w[P(g, i, j, k)] + f.w[P(g, i, j, k + 1)]);
  out[3 * n + q] = f.p[P(g, i, j, k)];
}

}  // namespace

void launch_trace(hipStream_t s, const Geo &g, const Fld &f, const int *pts, int n, double *out) {
  if (n <= 0) return;
  LW_LAUNCH("k_trace", 0, 0, k_trace, dim3(1), dim3(64), 0, s, g, f, pts, n, out);
}

void launch_field_avg(hipStream_t s, const Geo &g, const Fld &f, const AvgBuf &b, double dt, int heat) {
  const double cells = (double)g.n1m * g.n2m * g.n3m;
  const int nacc = heat ? kAvgArrays : kAvgArrays - 2;
  // u, v, w, p (and T) read once; every running sum read and written
  LW_LAUNCH("k_field_avg", cells * 8.0 * ((heat ? 5 : 4) + 2 * nacc), 0, k_field_avg,
            dim3((g.n1m + 255) / 256, g.n2m, g.n3m), dim3(256), 0, s, g, f, b, dt, heat);
}

}  // namespace lw
