// gfx950 kernels of the subgrid-scale models (sgs.f90): dynamic Vreman with a
// global Germano coefficient (sgs_dvmg, 405-638), constant Vreman (sgs_cvm,
// 335-402), dynamic and constant eddy diffusivity (sgs_edm 1297-1478,
// sgs_cedm 1250-1294), the Simpson box test filter (test_filter 641-769),
// zeroing in the body (nutzero/alpzero) and edge interpolation
// (nutinterpol 941-1133, alpinterpol 1591-1773).
#include <stdlib.h>

#include "lw_kernels.h"

namespace lw {

#define BLK3 dim3(64, 4, 1)
static inline dim3 grid3(int ni, int nj, int nk) {
  return dim3((unsigned)((ni + 63) / 64), (unsigned)((nj + 3) / 4), (unsigned)(nk > 0 ? nk : 1));
}

__device__ inline long I3(const Geo &g, int i, int j, int k) {
  return (long)(i - 1) + (long)g.n1m * ((long)(j - 1) + (long)g.n2m * (long)(k - 1));
}

// velgrad (sgs.f90:772-875): a = du_j/dx_i in (11,21,31,12,22,32,13,23,33), sr = S_ij
__device__ void velgrad(const Geo &g, const Fld &f, int i, int j, int k, double a[9], double sr[6]) {
  const double *u = f.u, *v = f.v, *w = f.w;
  const int n1 = g.n1, n2 = g.n2, n3 = g.n3;
  const int kp = g.kpv[k], km = g.kmv[k], jp = g.jpv[j], jm = g.jmv[j], ip = g.ipv[i], im = g.imv[i];
  const double fju = g.fixju[j], fjl = g.fixjl[j], fku = g.fixku[k], fkl = g.fixkl[k], fiu = g.fixiu[i],
               fil = g.fixil[i];
#define U(a_, b_, c_) u[P(g, a_, b_, c_)]
#define V(a_, b_, c_) v[P(g, a_, b_, c_)]
#define W(a_, b_, c_) w[P(g, a_, b_, c_)]
  const double vg11 = g.f2fxi[i] * (U(ip, j, k) - U(i, j, k));
  double up = g.c2cyi[jp] * 0.25 * (g.f2fy[jp] * (U(i, j, k) + U(ip, j, k)) + g.f2fy[j] * (U(i, jp, k) + U(ip, jp, k))) *
                  (1. - fju) +
              0.5 * (U(i, n2, k) + U(ip, n2, k)) * fju;
  double um = g.c2cyi[j] * 0.25 * (g.f2fy[j] * (U(i, jm, k) + U(ip, jm, k)) + g.f2fy[jm] * (U(i, j, k) + U(ip, j, k))) *
                  (1. - fjl) +
              0.5 * (U(i, 0, k) + U(ip, 0, k)) * fjl;
  const double vg12 = g.f2fyi[j] * (up - um);
  up = g.c2czi[kp] * 0.25 * (g.f2fz[kp] * (U(i, j, k) + U(ip, j, k)) + g.f2fz[k] * (U(i, j, kp) + U(ip, j, kp))) *
           (1. - fku) +
       0.5 * (U(i, j, n3) + U(ip, j, n3)) * fku;
  um = g.c2czi[k] * 0.25 * (g.f2fz[k] * (U(i, j, km) + U(ip, j, km)) + g.f2fz[km] * (U(i, j, k) + U(ip, j, k))) *
           (1. - fkl) +
       0.5 * (U(i, j, 0) + U(ip, j, 0)) * fkl;
  const double vg13 = g.f2fzi[k] * (up - um);
  double vp = g.c2cxi[ip] * 0.25 * (g.f2fx[ip] * (V(i, j, k) + V(i, jp, k)) + g.f2fx[i] * (V(ip, j, k) + V(ip, jp, k))) *
                  (1. - fiu) +
              0.5 * (V(n1, j, k) + V(n1, jp, k)) * fiu;
  double vm = g.c2cxi[i] * 0.25 * (g.f2fx[i] * (V(im, j, k) + V(im, jp, k)) + g.f2fx[im] * (V(i, j, k) + V(i, jp, k))) *
                  (1. - fil) +
              0.5 * (V(0, j, k) + V(0, jp, k)) * fil;
  const double vg21 = g.f2fxi[i] * (vp - vm);
  const double vg22 = g.f2fyi[j] * (V(i, jp, k) - V(i, j, k));
  vp = g.c2czi[kp] * 0.25 * (g.f2fz[kp] * (V(i, j, k) + V(i, jp, k)) + g.f2fz[k] * (V(i, j, kp) + V(i, jp, kp))) *
           (1. - fku) +
       0.5 * (V(i, j, n3) + V(i, jp, n3)) * fku;
  vm = g.c2czi[k] * 0.25 * (g.f2fz[k] * (V(i, j, km) + V(i, jp, km)) + g.f2fz[km] * (V(i, j, k) + V(i, jp, k))) *
           (1. - fkl) +
       0.5 * (V(i, j, 0) + V(i, jp, 0)) * fkl;
  const double vg23 = g.f2fzi[k] * (vp - vm);
  double wp = g.c2cxi[ip] * 0.25 * (g.f2fx[ip] * (W(i, j, k) + W(i, j, kp)) + g.f2fx[i] * (W(ip, j, k) + W(ip, j, kp))) *
                  (1. - fiu) +
              0.5 * (W(n1, j, k) + W(n1, j, kp)) * fiu;
  double wm = g.c2cxi[i] * 0.25 * (g.f2fx[i] * (W(im, j, k) + W(im, j, kp)) + g.f2fx[im] * (W(i, j, k) + W(i, j, kp))) *
                  (1. - fil) +
              0.5 * (W(0, j, k) + W(0, j, kp)) * fil;
  const double vg31 = g.f2fxi[i] * (wp - wm);
  wp = g.c2cyi[jp] * 0.25 * (g.f2fy[jp] * (W(i, j, k) + W(i, j, kp)) + g.f2fy[j] * (W(i, jp, k) + W(i, jp, kp))) *
           (1. - fju) +
       0.5 * (W(i, n2, k) + W(i, n2, kp)) * fju;
  wm = g.c2cyi[j] * 0.25 * (g.f2fy[j] * (W(i, jm, k) + W(i, jm, kp)) + g.f2fy[jm] * (W(i, j, k) + W(i, j, kp))) *
           (1. - fjl) +
       0.5 * (W(i, 0, k) + W(i, 0, kp)) * fjl;
  const double vg32 = g.f2fyi[j] * (wp - wm);
  const double vg33 = g.f2fzi[k] * (W(i, j, kp) - W(i, j, k));
#undef U
#undef V
#undef W
  a[0] = vg11; a[1] = vg21; a[2] = vg31;
  a[3] = vg12; a[4] = vg22; a[5] = vg32;
  a[6] = vg13; a[7] = vg23; a[8] = vg33;
  sr[0] = vg11;
  sr[1] = 0.5 * (vg12 + vg21);
  sr[2] = 0.5 * (vg13 + vg31);
  sr[3] = vg22;
  sr[4] = 0.5 * (vg23 + vg32);
  sr[5] = vg33;
}

// Vreman sqrt(B_beta / alpha:alpha) (sgs.f90:442-470)
__device__ inline double vreman(const double alp[9], double sx, double sy, double sz) {
  double b[6];
  b[0] = sx * alp[0] * alp[0] + sy * alp[3] * alp[3] + sz * alp[6] * alp[6];
  b[1] = sx * alp[0] * alp[1] + sy * alp[3] * alp[4] + sz * alp[6] * alp[7];
  b[2] = sx * alp[0] * alp[2] + sy * alp[3] * alp[5] + sz * alp[6] * alp[8];
  b[3] = sx * alp[1] * alp[1] + sy * alp[4] * alp[4] + sz * alp[7] * alp[7];
  b[4] = sx * alp[1] * alp[2] + sy * alp[4] * alp[5] + sz * alp[7] * alp[8];
  b[5] = sx * alp[2] * alp[2] + sy * alp[5] * alp[5] + sz * alp[8] * alp[8];
  const double aaa = alp[0] * alp[0] + alp[1] * alp[1] + alp[2] * alp[2] + alp[3] * alp[3] +
                     alp[4] * alp[4] + alp[5] * alp[5] + alp[6] * alp[6] + alp[7] * alp[7] + alp[8] * alp[8];
  double bbb = b[0] * b[3] + b[3] * b[5] + b[5] * b[0] - b[1] * b[1] - b[2] * b[2] - b[4] * b[4];
  bbb = fmax(bbb, 1.0e-16);
  if (aaa == 0.) return 0.0;
  return sqrt(bbb / aaa);
}

__device__ inline double fwidth(double fm, double f0, double fp, int fixu, int fixl) {
  return 0.5 * fm * (1. - (double)fixu) + f0 + 0.5 * fp * (1. - (double)fixl);
}

// temgrad (sgs.f90:1480-1525)
__device__ void temgrad(const Geo &g, const Fld &f, int i, int j, int k, double a[3]) {
  const double *t = f.t;
  const int kp = g.kpv[k], km = g.kmv[k], jp = g.jpv[j], jm = g.jmv[j], ip = g.ipv[i], im = g.imv[i];
  const double fiu = g.fixiu[i], fil = g.fixil[i], fju = g.fixju[j], fjl = g.fixjl[j], fku = g.fixku[k],
               fkl = g.fixkl[k];
  const long c = P(g, i, j, k);
  double tp = 0.5 * (g.f2fx[i] * t[P(g, ip, j, k)] + g.f2fx[ip] * t[c]) * g.c2cxi[ip] * (1. - fiu) +
              t[P(g, ip, j, k)] * fiu;
  double tm = 0.5 * (g.f2fx[im] * t[c] + g.f2fx[i] * t[P(g, im, j, k)]) * g.c2cxi[i] * (1. - fil) +
              t[P(g, im, j, k)] * fil;
  a[0] = g.f2fxi[i] * (tp - tm);
  tp = 0.5 * (g.f2fy[j] * t[P(g, i, jp, k)] + g.f2fy[jp] * t[c]) * g.c2cyi[jp] * (1. - fju) + t[P(g, i, jp, k)] * fju;
  tm = 0.5 * (g.f2fy[jm] * t[c] + g.f2fy[j] * t[P(g, i, jm, k)]) * g.c2cyi[j] * (1. - fjl) + t[P(g, i, jm, k)] * fjl;
  a[1] = g.f2fyi[j] * (tp - tm);
  tp = 0.5 * (g.f2fz[k] * t[P(g, i, j, kp)] + g.f2fz[kp] * t[c]) * g.c2czi[kp] * (1. - fku) + t[P(g, i, j, kp)] * fku;
  tm = 0.5 * (g.f2fz[km] * t[c] + g.f2fz[k] * t[P(g, i, j, km)]) * g.c2czi[k] * (1. - fkl) + t[P(g, i, j, km)] * fkl;
  a[2] = g.f2fzi[k] * (tp - tm);
}

// ---------------------------------------------------------------------------
// Vreman, dynamic global coefficient (sgs_dvmg)
// buffers: a[0..8] alpha_ij, a[9..14] M_ij, a[15..17] u_i, a[18..23] L_ij
// ---------------------------------------------------------------------------
LW_KERNEL void k_dvmg1(LW_KP(1024, 0), const Geo &g, const Fld &f, const SgsBuf &b) {
  const int i = 1 + blockIdx.x * 64 + threadIdx.x;
  const int j = 1 + blockIdx.y * 4 + threadIdx.y;
  const int k = 1 + blockIdx.z;
  if (i > g.n1m || j > g.n2m) return;
  double alp[9], str[6];
  velgrad(g, f, i, j, k, alp, str);
  const double r = vreman(alp, g.f2fx[i] * g.f2fx[i], g.f2fy[j] * g.f2fy[j], g.f2fz[k] * g.f2fz[k]);
  f.nu[P(g, i, j, k)] = r;
  const long q = I3(g, i, j, k);
#pragma unroll
  for (int l = 0; l < 9; ++l) b.a[l][q] = alp[l];
#pragma unroll
  for (int l = 0; l < 6; ++l) b.a[9 + l][q] = r * str[l];
}

LW_KERNEL void k_dvmg2(LW_KP(1024, 0), const Geo &g, const Fld &f, const SgsBuf &b) {
  const int i = 1 + blockIdx.x * 64 + threadIdx.x;
  const int j = 1 + blockIdx.y * 4 + threadIdx.y;
  const int k = 1 + blockIdx.z;
  if (i > g.n1m || j > g.n2m) return;
  const long q = I3(g, i, j, k);
  double alp[9], str[6];
#pragma unroll
  for (int l = 0; l < 9; ++l) alp[l] = b.a[l][q];
  str[0] = alp[0];
  str[1] = 0.5 * (alp[3] + alp[1]);
  str[2] = 0.5 * (alp[6] + alp[2]);
  str[3] = alp[4];
  str[4] = 0.5 * (alp[7] + alp[5]);
  str[5] = alp[8];
  const double dx = fwidth(g.f2fx[i - 1], g.f2fx[i], g.f2fx[i + 1], g.fixiu[i], g.fixil[i]);
  const double dy = fwidth(g.f2fy[j - 1], g.f2fy[j], g.f2fy[j + 1], g.fixju[j], g.fixjl[j]);
  const double dz = fwidth(g.f2fz[k - 1], g.f2fz[k], g.f2fz[k + 1], g.fixku[k], g.fixkl[k]);
  const double ami = vreman(alp, dx * dx, dy * dy, dz * dz);
#pragma unroll
  for (int l = 0; l < 6; ++l) b.a[9 + l][q] = ami * str[l] - b.a[9 + l][q];
  const long c = P(g, i, j, k);
  const double u1 = 0.5 * (f.u[c] + f.u[c + 1]);
  const double u2 = 0.5 * (f.v[c] + f.v[c + g.NX]);
  const double u3 = 0.5 * (f.w[c] + f.w[c + g.NX * g.NY]);
  b.a[15][q] = u1;
  b.a[16][q] = u2;
  b.a[17][q] = u3;
  b.a[18][q] = u1 * u1;
  b.a[19][q] = u1 * u2;
  b.a[20][q] = u1 * u3;
  b.a[21][q] = u2 * u2;
  b.a[22][q] = u2 * u3;
  b.a[23][q] = u3 * u3;
}

// per-cell Germano products, summed per block into partial[2*block + 0/1]
template <int NT>
__device__ inline double bsum(double v, double *sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double r = 0.0;
  if (w == 0) {
    r = lane < NT / 64 ? sh[lane] : 0.0;
    for (int o = 32; o > 0; o >>= 1) r += __shfl_down(r, o, 64);
  }
  __syncthreads();
  return r;
}

LW_KERNEL void k_dvmg3(LW_KP(kRedThreads, 0), const Geo &g, const Fld &f, const SgsBuf &b, const int64_t *nwall) {
  __shared__ double sh[kRedThreads / 64];
  const long n = (long)g.n1m * g.n2m * g.n3m;
  double lm = 0.0, mm = 0.0;
  for (long q = blockIdx.x * (long)kRedThreads + threadIdx.x; q < n; q += (long)gridDim.x * kRedThreads) {
    const int i = 1 + (int)(q % g.n1m);
    const int j = 1 + (int)((q / g.n1m) % g.n2m);
    const int k = 1 + (int)(q / ((long)g.n1m * g.n2m));
    const double volume = g.f2fx[i] * g.f2fy[j] * g.f2fz[k];
    const double U1 = b.a[15][q], U2 = b.a[16][q], U3 = b.a[17][q];
    const double l1 = b.a[18][q] - U1 * U1, l2 = b.a[19][q] - U1 * U2, l3 = b.a[20][q] - U1 * U3,
                 l4 = b.a[21][q] - U2 * U2, l5 = b.a[22][q] - U2 * U3, l6 = b.a[23][q] - U3 * U3;
    const double m1 = b.a[9][q], m2 = b.a[10][q], m3 = b.a[11][q], m4 = b.a[12][q], m5 = b.a[13][q],
                 m6 = b.a[14][q];
    const double nw = (double)nwall[q];
    lm += (2. * l2 * m2 + l1 * m1 + 2. * l3 * m3 + l4 * m4 + 2. * l5 * m5 + l6 * m6) * nw * volume;
    mm += (2. * (m2 * m2) + m1 * m1 + 2. * (m3 * m3) + m4 * m4 + 2. * (m5 * m5) + m6 * m6) * nw * volume;
  }
  const double s1 = bsum<kRedThreads>(lm, sh);
  const double s2 = bsum<kRedThreads>(mm, sh);
  if (threadIdx.x == 0) {
    f.partial[2 * blockIdx.x] = s1;
    f.partial[2 * blockIdx.x + 1] = s2;
  }
}

// final sum of the block partials and the clipped coefficient into red[slot]
LW_KERNEL void k_coef(LW_KP(kRedThreads, 0), const Fld &f, int nblk, double scale, int slot) {
  __shared__ double sh[kRedThreads / 64];
  double a = 0.0, c = 0.0;
  for (int q = threadIdx.x; q < nblk; q += kRedThreads) {
    a += f.partial[2 * q];
    c += f.partial[2 * q + 1];
  }
  const double lm = bsum<kRedThreads>(a, sh);
  const double mm = bsum<kRedThreads>(c, sh);
  if (threadIdx.x == 0) f.red[slot] = fmax(0., scale * lm / mm);
}

// 3-point Simpson box test filter in x (i = 2..n1m-1, non-periodic) and z
// (periodic), in that order (test_filter, sgs.f90:641-769). in -> out.
// A thread walks kFiltK planes of its (i, j) column up the periodic
// z-direction and keeps the x-filtered values of planes k-1, k, k+1 in
// registers, so an input value is read (kFiltK+2)/kFiltK times per launch
// instead of three times. Whether the column is x-filtered and its three
// weights are decided once per thread; the column's loads are unconditional
// (an unfiltered column reads its own point for both neighbours), so the
// loads of later planes can be issued ahead of the arithmetic.
constexpr int kFiltK = 16;

LW_KERNEL void k_filter(LW_KP(256, 0), const Geo &g, const double *__restrict__ in, double *__restrict__ out,
                                                int fx, int fz) {
  const int i = 1 + blockIdx.x * 64 + threadIdx.x;
  const int j = 1 + blockIdx.y * 4 + threadIdx.y;
  const int n3m = g.n3m;
  const int k0 = 1 + blockIdx.z * kFiltK, k1 = k0 + kFiltK - 1 < n3m ? k0 + kFiltK - 1 : n3m;
  if (i > g.n1m || j > g.n2m) return;
  // x periodic: the end cells wrap (sgs.f90:687-700); otherwise they stay
  const bool xi = fx && (g.xprdic ? true : (i >= 2 && i <= g.n1m - 1));
  const int ic = xi ? i : 2;  // a weight row that exists
  const double c0 = g.cfx1[3 * ic], c1 = g.cfx1[3 * ic + 1], c2 = g.cfx1[3 * ic + 2];
  const long dm = !xi ? 0 : i == 1 ? g.n1m - 1 : -1, dp = !xi ? 0 : i == g.n1m ? 1 - g.n1m : 1;
  const long q0 = I3(g, i, j, 1), st = (long)g.n1m * g.n2m;
  auto X = [&](int k) {
    const double *a = in + q0 + (long)(k - 1) * st;
    const double am = a[dm], a0 = a[0], ap = a[dp];
    return xi ? c0 * am + c1 * a0 + c2 * ap : a0;
  };
  if (!fz) {
#pragma unroll 4
    for (int k = k0; k <= k1; ++k) out[q0 + (long)(k - 1) * st] = X(k);
    return;
  }
  double xm = X(k0 == 1 ? n3m : k0 - 1), x0 = X(k0);
#pragma unroll 4
  for (int k = k0; k <= k1; ++k) {
    const double xp = X(k == n3m ? 1 : k + 1);
    out[q0 + (long)(k - 1) * st] = g.cfz1[3 * k] * xm + g.cfz1[3 * k + 1] * x0 + g.cfz1[3 * k + 2] * xp;
    xm = x0;
    x0 = xp;
  }
}

static inline double pts(const Geo &g) { return (double)g.n1m * g.n2m * g.n3m; }

// filters a[lo..hi] out of place; the filtered buffer takes the array's slot
// and the old one becomes the spare
static void filter_many(hipStream_t s, const Geo &g, SgsBuf &b, int lo, int hi, int filter) {
  const int fx = (filter == 1 || filter == 3) ? 1 : 0;
  const int fz = (filter == 2 || filter == 3) ? 1 : 0;
  for (int l = lo; l <= hi; ++l) {
    LW_LAUNCH("k_filter", pts(g) * 16.0, 0, k_filter, grid3(g.n1m, g.n2m, (g.n3m + kFiltK - 1) / kFiltK), BLK3, 0, s,
              g, b.a[l], b.tmp, fx, fz);
    double *t = b.a[l];
    b.a[l] = b.tmp;
    b.tmp = t;
  }
}

// ---------------------------------------------------------------------------
// The dynamic models in one pass each. The quantities the reference filters
// (sgs.f90:405-638 Vreman: alpha_ij, nu S_ij, u_i, u_i u_j; 1297-1478 eddy
// diffusivity: S_ij, dT/dx_i, alpha dT/dx_i, T, u_i, u_i T) are formed per
// cell in registers and filtered on the fly instead of going through 24
// interior arrays and one filter launch per array: a wavefront owns 62
// consecutive cells of a row j (lanes 1..62; lanes 0 and 63 compute the x
// neighbours), takes each cell's quantities from the lanes next to it for
// the x filter and marches kSgsK planes up the periodic z-direction, adding
// each plane's x-filtered values into the z-filtered sums of the planes
// below, above and at it in the order test_filter adds them (641-769). The
// Germano products of each cell (593-618, 1447-1462) go into the thread's
// sums; workgroup sums go to f.partial for k_coef. The filtered values are
// the reference's to the bit; the sums differ from a serial loop's by the
// order of the additions only.
// ---------------------------------------------------------------------------
constexpr int kSgsK = 16;   // planes a wavefront marches
constexpr int kSgsOwn = 62; // cells of a row a wavefront owns

struct SgsTile {
  int i, ic, j, k0, k1, lane;
  bool own, xi;
  double cx0, cx1, cx2;
};

__device__ inline bool sgs_tile(const Geo &g, long t, int fx, SgsTile &T) {
  const int nti = (g.n1m + kSgsOwn - 1) / kSgsOwn, ntk = (g.n3m + kSgsK - 1) / kSgsK;
  if (t >= (long)nti * g.n2m * ntk) return false;
  const int ti = (int)(t % nti);
  const long rest = t / nti;
  T.j = 1 + (int)(rest % g.n2m);
  const int tk = (int)(rest / g.n2m);
  T.lane = threadIdx.x & 63;
  T.i = ti * kSgsOwn + T.lane;
  // the x neighbours of the row's end cells: clamped (their filter is off),
  // or on an x-periodic grid the cells across the period
  if (g.xprdic) T.ic = T.i < 1 ? g.n1m : (T.i > g.n1m ? 1 : T.i);
  else T.ic = T.i < 1 ? 1 : (T.i > g.n1m ? g.n1m : T.i);
  T.own = T.lane >= 1 && T.lane <= kSgsOwn && T.i <= g.n1m;
  T.xi = fx && (g.xprdic ? (T.i >= 1 && T.i <= g.n1m) : (T.i >= 2 && T.i <= g.n1m - 1));
  const int ir = T.xi ? T.i : 2;
  T.cx0 = g.cfx1[3 * ir];
  T.cx1 = g.cfx1[3 * ir + 1];
  T.cx2 = g.cfx1[3 * ir + 2];
  T.k0 = 1 + tk * kSgsK;
  T.k1 = T.k0 + kSgsK - 1 < g.n3m ? T.k0 + kSgsK - 1 : g.n3m;
  return true;
}

// the x filter, in place, of NQ quantities of the lanes' cells (every lane
// of the wavefront calls it); FT is the filter's arithmetic (double, or float
// in the mixed-precision form)
template <int NQ, typename FT>
__device__ inline void sgs_xfilter(const SgsTile &T, FT (&q)[NQ]) {
  const FT c0 = (FT)T.cx0, c1 = (FT)T.cx1, c2 = (FT)T.cx2;
#pragma unroll
  for (int l = 0; l < NQ; ++l) {
    const FT am = __shfl_up(q[l], 1, 64), ap = __shfl_down(q[l], 1, 64);
    q[l] = T.xi ? c0 * am + c1 * q[l] + c2 * ap : q[l];
  }
}

// Vreman quantities of a cell: alpha_ij (9), nu S_ij (6), u_i (3), u_i u_j
// (6) (sgs.f90:434-470, 538-566); returns nu
template <typename FT>
__device__ inline double dvmg_cell(const Geo &g, const Fld &f, int i, int j, int k, FT (&q)[24]) {
  double alp[9], str[6];
  velgrad(g, f, i, j, k, alp, str);
  const double r = vreman(alp, g.f2fx[i] * g.f2fx[i], g.f2fy[j] * g.f2fy[j], g.f2fz[k] * g.f2fz[k]);
#pragma unroll
  for (int l = 0; l < 9; ++l) q[l] = (FT)alp[l];
#pragma unroll
  for (int l = 0; l < 6; ++l) q[9 + l] = (FT)(r * str[l]);
  const long c = P(g, i, j, k);
  const double u1 = 0.5 * (f.u[c] + f.u[c + 1]);
  const double u2 = 0.5 * (f.v[c] + f.v[c + g.NX]);
  const double u3 = 0.5 * (f.w[c] + f.w[c + g.NX * g.NY]);
  q[15] = (FT)u1; q[16] = (FT)u2; q[17] = (FT)u3;
  q[18] = (FT)(u1 * u1); q[19] = (FT)(u1 * u2); q[20] = (FT)(u1 * u3);
  q[21] = (FT)(u2 * u2); q[22] = (FT)(u2 * u3); q[23] = (FT)(u3 * u3);
  return r;
}

// the Germano products of a cell from its filtered quantities (k_dvmg2's M,
// k_dvmg3's L and sums)
template <typename FT>
__device__ inline void dvmg_germano(const Geo &g, int i, int j, int k, const FT (&Ff)[24], const int64_t *nwall,
                                    double &lm, double &mm) {
  double F[24];
#pragma unroll
  for (int l = 0; l < 24; ++l) F[l] = (double)Ff[l];
  double alp[9], str[6];
#pragma unroll
  for (int l = 0; l < 9; ++l) alp[l] = F[l];
  str[0] = alp[0];
  str[1] = 0.5 * (alp[3] + alp[1]);
  str[2] = 0.5 * (alp[6] + alp[2]);
  str[3] = alp[4];
  str[4] = 0.5 * (alp[7] + alp[5]);
  str[5] = alp[8];
  const double dx = fwidth(g.f2fx[i - 1], g.f2fx[i], g.f2fx[i + 1], g.fixiu[i], g.fixil[i]);
  const double dy = fwidth(g.f2fy[j - 1], g.f2fy[j], g.f2fy[j + 1], g.fixju[j], g.fixjl[j]);
  const double dz = fwidth(g.f2fz[k - 1], g.f2fz[k], g.f2fz[k + 1], g.fixku[k], g.fixkl[k]);
  const double ami = vreman(alp, dx * dx, dy * dy, dz * dz);
  const double m1 = ami * str[0] - F[9], m2 = ami * str[1] - F[10], m3 = ami * str[2] - F[11],
               m4 = ami * str[3] - F[12], m5 = ami * str[4] - F[13], m6 = ami * str[5] - F[14];
  const double U1 = F[15], U2 = F[16], U3 = F[17];
  const double l1 = F[18] - U1 * U1, l2 = F[19] - U1 * U2, l3 = F[20] - U1 * U3, l4 = F[21] - U2 * U2,
               l5 = F[22] - U2 * U3, l6 = F[23] - U3 * U3;
  const double volume = g.f2fx[i] * g.f2fy[j] * g.f2fz[k];
  const double nw = (double)nwall[I3(g, i, j, k)];
  lm += (2. * l2 * m2 + l1 * m1 + 2. * l3 * m3 + l4 * m4 + 2. * l5 * m5 + l6 * m6) * nw * volume;
  mm += (2. * (m2 * m2) + m1 * m1 + 2. * (m3 * m3) + m4 * m4 + 2. * (m5 * m5) + m6 * m6) * nw * volume;
}

// eddy-diffusivity quantities of a cell: S_ij (6), dT/dx_i (3), alpha
// dT/dx_i (3), T, u_i (3), u_i T (3) (sgs.f90:1325-1345, 1395-1430);
// returns alpha
template <typename FT>
__device__ inline double edm_cell(const Geo &g, const Fld &f, int i, int j, int k, FT (&q)[19]) {
  double d[9], s[6], alp[3];
  velgrad(g, f, i, j, k, d, s);
  temgrad(g, f, i, j, k, alp);
  const double del = pow(g.f2fx[i] * g.f2fy[j] * g.f2fz[k], 1. / 3.);
  double sss = s[0] * s[0] + 2 * (s[1] * s[1]) + 2 * (s[2] * s[2]) + s[3] * s[3] + 2 * (s[4] * s[4]) + s[5] * s[5];
  sss = fmax(sss, 1.0e-16);
  const double al = (del * del) * sqrt(2. * sss);
#pragma unroll
  for (int l = 0; l < 6; ++l) q[l] = (FT)s[l];
#pragma unroll
  for (int l = 0; l < 3; ++l) {
    q[6 + l] = (FT)alp[l];
    q[9 + l] = (FT)(al * alp[l]);
  }
  const long c = P(g, i, j, k);
  const double u1 = 0.5 * (f.u[c] + f.u[c + 1]);
  const double u2 = 0.5 * (f.v[c] + f.v[c + g.NX]);
  const double u3 = 0.5 * (f.w[c] + f.w[c + g.NX * g.NY]);
  const double tc = f.t[c];
  q[12] = (FT)tc;
  q[13] = (FT)u1; q[14] = (FT)u2; q[15] = (FT)u3;
  q[16] = (FT)(u1 * tc); q[17] = (FT)(u2 * tc); q[18] = (FT)(u3 * tc);
  return al;
}

template <typename FT>
__device__ inline void edm_germano(const Geo &g, int i, int j, int k, const FT (&Ff)[19], const int64_t *nwall,
                                   double &lm, double &mm) {
  double F[19];
#pragma unroll
  for (int l = 0; l < 19; ++l) F[l] = (double)Ff[l];
  const double d1 = fwidth(g.f2fx[i - 1], g.f2fx[i], g.f2fx[i + 1], g.fixiu[i], g.fixil[i]);
  const double d2 = fwidth(g.f2fy[j - 1], g.f2fy[j], g.f2fy[j + 1], g.fixju[j], g.fixjl[j]);
  const double d3 = fwidth(g.f2fz[k - 1], g.f2fz[k], g.f2fz[k + 1], g.fixku[k], g.fixkl[k]);
  const double del = pow(d1 * d2 * d3, 1. / 3.);
  const double a1 = F[0], a2 = F[1], a3 = F[2], a4 = F[3], a5 = F[4], a6 = F[5];
  double sss = a1 * a1 + 2 * (a2 * a2) + 2 * (a3 * a3) + a4 * a4 + 2 * (a5 * a5) + a6 * a6;
  sss = fmax(sss, 1.0e-16);
  const double ami = (del * del) * sqrt(2. * sss);
  const double m1 = ami * F[6] - F[9], m2 = ami * F[7] - F[10], m3 = ami * F[8] - F[11];
  const double l1 = F[16] - F[13] * F[12], l2 = F[17] - F[14] * F[12], l3 = F[18] - F[15] * F[12];
  const double volume = g.f2fx[i] * g.f2fy[j] * g.f2fz[k];
  const double nw = (double)nwall[I3(g, i, j, k)];
  lm += (l1 * m1 + l2 * m2 + l3 * m3) * nw * volume;
  mm += (m1 * m1 + m2 * m2 + m3 * m3) * nw * volume;
}

// MODEL 0: dynamic Vreman (writes nu), 1: dynamic eddy diffusivity (writes
// alpha). Workgroups of 4 wavefronts walk the tiles; each workgroup's two
// sums go to f.partial[2 b], f.partial[2 b + 1].
// FT: the test filter's arithmetic, double (the reference's) or float (the
// mixed-precision form: the filtered quantities feed only the Germano sums,
// which stay FP64, and the model coefficient they give). The Vreman
// instances ask for two waves per SIMD (FP64: 39 spilled registers, FP32:
// 4); the eddy-diffusivity ones keep one (two spilled 151 registers in FP64,
// and in FP32 an earlier form spilled 122 and ran 1.7 ms).
template <int MODEL, int NQ, typename FT>
LW_KERNEL void k_sgs_dyn(LW_KP(256, MODEL == 0 ? 2 : 0), const Geo &g, const Fld &f, const int64_t *nwall, int fx, int fz) {
  __shared__ double sh[4];
  // the first term of the plane above for each quantity: for the Vreman
  // model in LDS, one slot per lane (each lane reads back only what it
  // wrote), which leaves the registers for two waves per SIMD (FP64 1.12
  // against 1.19 ms, FP32 0.65 against 0.82 at 384x256x64); the eddy-
  // diffusivity model keeps it in registers (in LDS it ran 1.06 against
  // 1.03 ms at one wave per SIMD either way)
  constexpr bool kLdsB = MODEL == 0;
  __shared__ FT Bs[4][kLdsB ? NQ : 1][64];
  double lm = 0.0, mm = 0.0;
  const int w = threadIdx.x >> 6;
  SgsTile T;
  for (long t = (long)blockIdx.x * 4 + w; sgs_tile(g, t, fx, T); t += (long)gridDim.x * 4) {
    const int n3m = g.n3m;
    // z sums: A of the plane awaiting its last term, B of the plane awaiting
    // its middle one
    FT A[NQ], B[kLdsB ? 1 : NQ];
    FT x[NQ];
#pragma unroll
    for (int l = 0; l < NQ; ++l) A[l] = (FT)0.0;
    // B is written at kk = k0 - 1 before its first read at kk = k0 (fz), and
    // not used without the z filter
    const int kb = fz ? T.k0 - 1 : T.k0, ke = fz ? T.k1 + 1 : T.k1;
    for (int kk = kb; kk <= ke; ++kk) {
      const int k = kk < 1 ? n3m : (kk > n3m ? 1 : kk);
      double v;
      if constexpr (MODEL == 0) v = dvmg_cell(g, f, T.ic, T.j, k, x);
      else v = edm_cell(g, f, T.ic, T.j, k, x);
      if (T.own && kk >= T.k0 && kk <= T.k1) (MODEL == 0 ? f.nu : f.al)[P(g, T.i, T.j, k)] = v;
      sgs_xfilter<NQ>(T, x);
      if (!fz) {
        if (T.own) {
          if constexpr (MODEL == 0) dvmg_germano(g, T.i, T.j, k, x, nwall, lm, mm);
          else edm_germano(g, T.i, T.j, k, x, nwall, lm, mm);
        }
        continue;
      }
      // the plane below (kk - 1) is complete with this plane's term
      if (kk >= T.k0 + 1) {
        const int km = kk - 1;
        const FT cz = (FT)g.cfz1[3 * km + 2];
#pragma unroll
        for (int l = 0; l < NQ; ++l) A[l] = A[l] + cz * x[l];
        if (T.own) {
          if constexpr (MODEL == 0) dvmg_germano(g, T.i, T.j, km, A, nwall, lm, mm);
          else edm_germano(g, T.i, T.j, km, A, nwall, lm, mm);
        }
      }
      // this plane's middle term; it then awaits its last
      if (kk >= T.k0 && kk <= T.k1) {
        const FT cz = (FT)g.cfz1[3 * k + 1];
#pragma unroll
        for (int l = 0; l < NQ; ++l) {
          if constexpr (kLdsB) A[l] = Bs[w][l][T.lane] + cz * x[l];
          else A[l] = B[l] + cz * x[l];
        }
      }
      // the first term of the plane above
      if (kk + 1 <= T.k1) {
        const int kp = kk + 1;
        const FT cz = (FT)g.cfz1[3 * kp];
#pragma unroll
        for (int l = 0; l < NQ; ++l) {
          if constexpr (kLdsB) Bs[w][l][T.lane] = cz * x[l];
          else B[l] = cz * x[l];
        }
      }
    }
  }
  // workgroup sums (fixed shape)
  for (int o = 32; o > 0; o >>= 1) {
    lm += __shfl_down(lm, o, 64);
    mm += __shfl_down(mm, o, 64);
  }
  const int lane = threadIdx.x & 63;
  if (lane == 0) sh[w] = lm;
  __syncthreads();
  if (threadIdx.x == 0) f.partial[2 * blockIdx.x] = ((sh[0] + sh[1]) + sh[2]) + sh[3];
  __syncthreads();
  if (lane == 0) sh[w] = mm;
  __syncthreads();
  if (threadIdx.x == 0) f.partial[2 * blockIdx.x + 1] = ((sh[0] + sh[1]) + sh[2]) + sh[3];
}

// read at every handle's creation, so that cases of one process may differ
bool sgs_fused_enabled() {
  const char *e = getenv("LW_SGS_FUSED");
  return !(e && atoi(e) == 0);
}

static void launch_sgs_dyn(hipStream_t s, const Geo &g, const Fld &f, const int64_t *nwall, int filter, int model,
                           double scale, int slot, int fp32) {
  const int fx = (filter == 1 || filter == 3) ? 1 : 0;
  const int fz = (filter == 2 || filter == 3) ? 1 : 0;
  const double cells = pts(g);
  // u v w read (with their x and z neighbour planes), nu and nwall: about 6
  // arrays (7 with T)
  if (model == 0 && !fp32)
    LW_LAUNCH("k_sgs_dvmg", cells * 8.0 * 6, 0, (k_sgs_dyn<0, 24, double>), dim3(kRedBlocks), dim3(256), 0, s, g, f,
              nwall, fx, fz);
  else if (model == 0)
    LW_LAUNCH("k_sgs_dvmg", cells * 8.0 * 6, 0, (k_sgs_dyn<0, 24, float>), dim3(kRedBlocks), dim3(256), 0, s, g, f,
              nwall, fx, fz);
  else if (!fp32)
    LW_LAUNCH("k_sgs_edm", cells * 8.0 * 7, 0, (k_sgs_dyn<1, 19, double>), dim3(kRedBlocks), dim3(256), 0, s, g, f,
              nwall, fx, fz);
  else
    LW_LAUNCH("k_sgs_edm", cells * 8.0 * 7, 0, (k_sgs_dyn<1, 19, float>), dim3(kRedBlocks), dim3(256), 0, s, g, f,
              nwall, fx, fz);
  LW_LAUNCH("k_coef", kRedBlocks * 16.0, 0, k_coef, dim3(1), dim3(kRedThreads), 0, s, f, kRedBlocks, scale, slot);
}

void launch_sgs_dvmg_fused(hipStream_t s, const Geo &g, const Fld &f, const int64_t *nwall, int filter, int fp32) {
  launch_sgs_dyn(s, g, f, nwall, filter, 0, -0.5, 4, fp32);
}
void launch_sgs_edm_fused(hipStream_t s, const Geo &g, const Fld &f, const int64_t *nwall, int filter, int fp32) {
  launch_sgs_dyn(s, g, f, nwall, filter, 1, -1.0, 5, fp32);
}

// red slots: 4 = csgsts, 5 = csgshf
void launch_sgs_dvmg(hipStream_t s, const Geo &g, const Fld &f, SgsBuf &b, int filter) {
  // u v w in; nu, alpha_ij (9) and the Vreman products (6) out
  LW_LAUNCH("k_dvmg1", pts(g) * 8.0 * 19, 0, k_dvmg1, grid3(g.n1m, g.n2m, g.n3m), BLK3, 0, s, g, f, b);
  filter_many(s, g, b, 9, 14, filter);
  filter_many(s, g, b, 0, 8, filter);
  // 15 filtered arrays and u v w in; M_ij (6), u_i (3) and u_i u_j (6) out
  LW_LAUNCH("k_dvmg2", pts(g) * 8.0 * 33, 0, k_dvmg2, grid3(g.n1m, g.n2m, g.n3m), BLK3, 0, s, g, f, b);
  filter_many(s, g, b, 15, 23, filter);
}

void launch_sgs_dvmg_coef(hipStream_t s, const Geo &g, const Fld &f, const SgsBuf &b, const int64_t *nwall) {
  LW_LAUNCH("k_dvmg3", pts(g) * 8.0 * 16, 0, k_dvmg3, dim3(kRedBlocks), dim3(kRedThreads), 0, s, g, f, b, nwall);
  LW_LAUNCH("k_coef", kRedBlocks * 16.0, 0, k_coef, dim3(1), dim3(kRedThreads), 0, s, f, kRedBlocks, -0.5, 4);
}

LW_KERNEL void k_cvm(LW_KP(1024, 0), const Geo &g, const Fld &f, double csgsts) {
  const int i = 1 + blockIdx.x * 64 + threadIdx.x;
  const int j = 1 + blockIdx.y * 4 + threadIdx.y;
  const int k = 1 + blockIdx.z;
  if (i > g.n1m || j > g.n2m) return;
  double a[9], sr[6];
  velgrad(g, f, i, j, k, a, sr);
  const double r = vreman(a, g.f2fx[i] * g.f2fx[i], g.f2fy[j] * g.f2fy[j], g.f2fz[k] * g.f2fz[k]);
  f.nu[P(g, i, j, k)] = r == 0.0 ? 0.0 : csgsts * r;
}

void launch_sgs_cvm(hipStream_t s, const Geo &g, const Fld &f, double csgsts) {
  LW_LAUNCH("k_cvm", pts(g) * 32.0, 0, k_cvm, grid3(g.n1m, g.n2m, g.n3m), BLK3, 0, s, g, f, csgsts);
}

// ---------------------------------------------------------------------------
// eddy diffusivity, dynamic global coefficient (sgs_edm)
// buffers: a[0..5] S_ij, a[6..8] dT/dx_k, a[9..11] M_k, a[15..17] u_i,
// a[18..20] L_k, a[0..2] reused for T after the first filter pass
// ---------------------------------------------------------------------------
LW_KERNEL void k_edm1(LW_KP(1024, 0), const Geo &g, const Fld &f, const SgsBuf &b) {
  const int i = 1 + blockIdx.x * 64 + threadIdx.x;
  const int j = 1 + blockIdx.y * 4 + threadIdx.y;
  const int k = 1 + blockIdx.z;
  if (i > g.n1m || j > g.n2m) return;
  double d[9], s[6], alp[3];
  velgrad(g, f, i, j, k, d, s);
  temgrad(g, f, i, j, k, alp);
  const double del = pow(g.f2fx[i] * g.f2fy[j] * g.f2fz[k], 1. / 3.);
  double sss = s[0] * s[0] + 2 * (s[1] * s[1]) + 2 * (s[2] * s[2]) + s[3] * s[3] + 2 * (s[4] * s[4]) + s[5] * s[5];
  sss = fmax(sss, 1.0e-16);
  const double al = (del * del) * sqrt(2. * sss);
  f.al[P(g, i, j, k)] = al;
  const long q = I3(g, i, j, k);
#pragma unroll
  for (int l = 0; l < 3; ++l) b.a[9 + l][q] = al * alp[l];
#pragma unroll
  for (int l = 0; l < 6; ++l) b.a[l][q] = s[l];
#pragma unroll
  for (int l = 0; l < 3; ++l) b.a[6 + l][q] = alp[l];
}

LW_KERNEL void k_edm2(LW_KP(1024, 0), const Geo &g, const Fld &f, const SgsBuf &b) {
  const int i = 1 + blockIdx.x * 64 + threadIdx.x;
  const int j = 1 + blockIdx.y * 4 + threadIdx.y;
  const int k = 1 + blockIdx.z;
  if (i > g.n1m || j > g.n2m) return;
  const long q = I3(g, i, j, k);
  const double d1 = fwidth(g.f2fx[i - 1], g.f2fx[i], g.f2fx[i + 1], g.fixiu[i], g.fixil[i]);
  const double d2 = fwidth(g.f2fy[j - 1], g.f2fy[j], g.f2fy[j + 1], g.fixju[j], g.fixjl[j]);
  const double d3 = fwidth(g.f2fz[k - 1], g.f2fz[k], g.f2fz[k + 1], g.fixku[k], g.fixkl[k]);
  const double del = pow(d1 * d2 * d3, 1. / 3.);
  const double a1 = b.a[0][q], a2 = b.a[1][q], a3 = b.a[2][q], a4 = b.a[3][q], a5 = b.a[4][q], a6 = b.a[5][q];
  double sss = a1 * a1 + 2 * (a2 * a2) + 2 * (a3 * a3) + a4 * a4 + 2 * (a5 * a5) + a6 * a6;
  sss = fmax(sss, 1.0e-16);
  const double ami = (del * del) * sqrt(2. * sss);
#pragma unroll
  for (int l = 0; l < 3; ++l) b.a[9 + l][q] = ami * b.a[6 + l][q] - b.a[9 + l][q];
  const long c = P(g, i, j, k);
  const double u1 = 0.5 * (f.u[c] + f.u[c + 1]);
  const double u2 = 0.5 * (f.v[c] + f.v[c + g.NX]);
  const double u3 = 0.5 * (f.w[c] + f.w[c + g.NX * g.NY]);
  const double tc = f.t[c];
  b.a[15][q] = u1;
  b.a[16][q] = u2;
  b.a[17][q] = u3;
  b.a[0][q] = tc;
  b.a[1][q] = tc;
  b.a[2][q] = tc;
  b.a[18][q] = u1 * tc;
  b.a[19][q] = u2 * tc;
  b.a[20][q] = u3 * tc;
}

LW_KERNEL void k_edm3(LW_KP(kRedThreads, 0), const Geo &g, const Fld &f, const SgsBuf &b, const int64_t *nwall) {
  __shared__ double sh[kRedThreads / 64];
  const long n = (long)g.n1m * g.n2m * g.n3m;
  double lm = 0.0, mm = 0.0;
  for (long q = blockIdx.x * (long)kRedThreads + threadIdx.x; q < n; q += (long)gridDim.x * kRedThreads) {
    const int i = 1 + (int)(q % g.n1m);
    const int j = 1 + (int)((q / g.n1m) % g.n2m);
    const int k = 1 + (int)(q / ((long)g.n1m * g.n2m));
    const double volume = g.f2fx[i] * g.f2fy[j] * g.f2fz[k];
    const double l1 = b.a[18][q] - b.a[15][q] * b.a[0][q];
    const double l2 = b.a[19][q] - b.a[16][q] * b.a[1][q];
    const double l3 = b.a[20][q] - b.a[17][q] * b.a[2][q];
    const double m1 = b.a[9][q], m2 = b.a[10][q], m3 = b.a[11][q];
    const double nw = (double)nwall[q];
    lm += (l1 * m1 + l2 * m2 + l3 * m3) * nw * volume;
    mm += (m1 * m1 + m2 * m2 + m3 * m3) * nw * volume;
  }
  const double s1 = bsum<kRedThreads>(lm, sh);
  const double s2 = bsum<kRedThreads>(mm, sh);
  if (threadIdx.x == 0) {
    f.partial[2 * blockIdx.x] = s1;
    f.partial[2 * blockIdx.x + 1] = s2;
  }
}

void launch_sgs_edm(hipStream_t s, const Geo &g, const Fld &f, SgsBuf &b, int filter) {
  // u v w t in; alpha and 12 arrays out
  LW_LAUNCH("k_edm1", pts(g) * 8.0 * 17, 0, k_edm1, grid3(g.n1m, g.n2m, g.n3m), BLK3, 0, s, g, f, b);
  filter_many(s, g, b, 9, 11, filter);
  filter_many(s, g, b, 0, 8, filter);
  // 12 filtered arrays and u v w t in; 12 arrays out
  LW_LAUNCH("k_edm2", pts(g) * 8.0 * 28, 0, k_edm2, grid3(g.n1m, g.n2m, g.n3m), BLK3, 0, s, g, f, b);
  filter_many(s, g, b, 15, 17, filter);
  filter_many(s, g, b, 0, 2, filter);
  filter_many(s, g, b, 18, 20, filter);
}

void launch_sgs_edm_coef(hipStream_t s, const Geo &g, const Fld &f, const SgsBuf &b, const int64_t *nwall) {
  LW_LAUNCH("k_edm3", pts(g) * 8.0 * 13, 0, k_edm3, dim3(kRedBlocks), dim3(kRedThreads), 0, s, g, f, b, nwall);
  LW_LAUNCH("k_coef", kRedBlocks * 16.0, 0, k_coef, dim3(1), dim3(kRedThreads), 0, s, f, kRedBlocks, -1.0, 5);
}

LW_KERNEL void k_cedm(LW_KP(1024, 0), const Geo &g, const Fld &f, double csgshf) {
  const int i = 1 + blockIdx.x * 64 + threadIdx.x;
  const int j = 1 + blockIdx.y * 4 + threadIdx.y;
  const int k = 1 + blockIdx.z;
  if (i > g.n1m || j > g.n2m) return;
  double d[9], s[6];
  velgrad(g, f, i, j, k, d, s);
  const double del = pow(g.f2fx[i] * g.f2fy[j] * g.f2fz[k], 1. / 3.);
  double sss = s[0] * s[0] + 2. * (s[1] * s[1]) + 2. * (s[2] * s[2]) + s[3] * s[3] + 2. * (s[4] * s[4]) + s[5] * s[5];
  sss = fmax(sss, 1.0e-16);
  f.al[P(g, i, j, k)] = csgshf * (del * del) * sqrt(2. * sss);
}

void launch_sgs_cedm(hipStream_t s, const Geo &g, const Fld &f, double csgshf) {
  LW_LAUNCH("k_cedm", pts(g) * 32.0, 0, k_cedm, grid3(g.n1m, g.n2m, g.n3m), BLK3, 0, s, g, f, csgshf);
}

// ---------------------------------------------------------------------------
// scale by the dynamic coefficient (slot >= 0) and zero in the body and at
// the listed points (nutzero/alpzero); then z-periodic ghosts
// ---------------------------------------------------------------------------
LW_KERNEL void k_scale_zero(LW_KP(1024, 0), const Geo &g, const Fld &f, double *nu, int slot) {
  const int i = 1 + blockIdx.x * 64 + threadIdx.x;
  const int j = 1 + blockIdx.y * 4 + threadIdx.y;
  const int k = 1 + blockIdx.z;
  if (i > g.n1m || j > g.n2m) return;
  const long c = P(g, i, j, k);
  double v = nu[c];
  if (slot >= 0) v = f.red[slot] * v;
  if (f.mask[c] & (M_SOLID_C | M_NUTZERO)) v = 0.;
  nu[c] = v;
}

// listed points that fall on ghost indices (the 0-based values read as
// 1-based can be 0) are zeroed here; interior ones carry M_NUTZERO
LW_KERNEL void k_zero_list(LW_KP(1024, 0), const Geo &g, double *nu, const int *lst, int n) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  nu[P(g, lst[3 * q], lst[3 * q + 1], lst[3 * q + 2])] = 0.;
}

LW_KERNEL void k_zghost_scalar(LW_KP(1024, 0), const Geo &g, double *a) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  const int j = blockIdx.y;
  if (i > g.n1) return;
  a[P(g, i, j, 0)] = a[P(g, i, j, g.n3m)];
  a[P(g, i, j, g.n3)] = a[P(g, i, j, 1)];
}

// edge interpolation: the final value of each nu1(i,j,k,c) after the
// reference's sequence of loops (interior formula, x-boundary, y-boundary,
// z-periodic copy), last writer wins. alp = 1 selects alpinterpol's ranges.
struct EdgeCtx {
  int ib, jb, kb;
  int ybwall, ytwall;
  int xprdic;
};

__device__ inline double edge_main(const Geo &g, const double *nu, int i, int j, int k, int c) {
  const double *fx = g.f2fx, *fy = g.f2fy, *fz = g.f2fz;
#define N(a_, b_, c_) nu[P(g, a_, b_, c_)]
  if (c == 0)
    return g.c2cyi[j] * g.c2czi[k] * 0.25 *
           (fy[j - 1] * (fz[k - 1] * N(i, j, k) + fz[k] * N(i, j, k - 1)) +
            fy[j] * (fz[k - 1] * N(i, j - 1, k) + fz[k] * N(i, j - 1, k - 1)));
  if (c == 1)
    return g.c2czi[k] * g.c2cxi[i] * 0.25 *
           (fz[k - 1] * (fx[i - 1] * N(i, j, k) + fx[i] * N(i - 1, j, k)) +
            fz[k] * (fx[i - 1] * N(i, j, k - 1) + fx[i] * N(i - 1, j, k - 1)));
  return g.c2cxi[i] * g.c2cyi[j] * 0.25 *
         (fx[i - 1] * (fy[j - 1] * N(i, j, k) + fy[j] * N(i, j - 1, k)) +
          fx[i] * (fy[j - 1] * N(i - 1, j, k) + fy[j] * N(i - 1, j - 1, k)));
}

__device__ double edge_234(const Geo &g, const double *nu, const EdgeCtx &e, int i, int j, int k, int c) {
  const int n1 = g.n1, n2 = g.n2, n1m = g.n1m, n2m = g.n2m, n3m = g.n3m;
  double v = 0.0;
  // main loops
  bool inmain;
  if (c == 0) inmain = k >= e.kb && k <= n3m && j >= e.jb && j <= n2m && i >= 1 && i <= n1m;
  else if (c == 1) inmain = k >= e.kb && k <= n3m && j >= 1 && j <= n2m && i >= e.ib && i <= n1m;
  else inmain = k >= 1 && k <= n3m && j >= e.jb && j <= n2m && i >= e.ib && i <= n1m;
  if (inmain) v = edge_main(g, nu, i, j, k, c);
  const double *fx = g.f2fx, *fy = g.f2fy, *fz = g.f2fz;
  // x boundaries: the periodic copy of line 1 into n1 (sgs.f90:1032-1036),
  // or inflow/outflow (never walls)
  if (e.xprdic) {
    if (i == n1 && j >= e.jb && j <= n2m && k >= e.kb && k <= n3m) v = edge_main(g, nu, 1, j, k, c);
  } else if ((i == 1 || i == n1) && c >= 1 && j >= e.jb && j <= n2m && k >= e.kb && k <= n3m) {
    const int jm = g.jmv[j], km = g.kmv[k];
    const int is = i == 1 ? 2 : n1m;
    if (c == 1) v = g.c2czi[k] * 0.5 * (fz[k] * N(is, j, km) + fz[km] * N(is, j, k));
    else v = g.c2cyi[j] * 0.5 * (fy[j] * N(is, jm, k) + fy[jm] * N(is, j, k));
  }
  // y boundaries
  if ((j == 1 || j == n2) && k >= e.kb && k <= n3m && i >= e.ib && i <= n1m) {
    const bool wall = j == 1 ? e.ybwall : e.ytwall;
    if (wall) v = 0.;
    else if (c >= 1) {
      const int im = g.imv[i], km = g.kmv[k];
      const int js = j == 1 ? 2 : n2m;
      if (c == 1) v = g.c2czi[k] * 0.5 * (fz[k] * N(i, js, km) + fz[km] * N(i, js, k));
      else v = g.c2cxi[i] * 0.5 * (fx[i] * N(im, js, k) + fx[im] * N(i, js, k));
    }
  }
#undef N
  return v;
}

LW_KERNEL void k_edges(LW_KP(1024, 0), const Geo &g, const double *nu, double *o1, double *o2, double *o3, const EdgeCtx &e) {
  const int i = blockIdx.x * 64 + threadIdx.x;  // 0..n1
  const int j = blockIdx.y * 4 + threadIdx.y;   // 0..n2
  const int k = blockIdx.z;                     // 0..n3
  if (i > g.n1 || j > g.n2) return;
  double *out[3] = {o1, o2, o3};
  const bool zcopy = k == g.n3 && i >= e.ib && i <= g.n1m && j >= e.jb && j <= g.n2m;
  for (int c = 0; c < 3; ++c) {
    double v;
    if (zcopy) v = edge_234(g, nu, e, i, j, 1, c);
    else if (k == 0 || k == g.n3) v = 0.0;  // never written by the reference loops
    else v = edge_234(g, nu, e, i, j, k, c);
    out[c][P(g, i, j, k)] = v;
  }
}

// x-periodic ghosts of a cell array (nutinterpol, sgs.f90:960-967)
LW_KERNEL void k_xghost_scalar(LW_KP(1024, 0), const Geo &g, double *a) {
  const int j = blockIdx.x * 64 + threadIdx.x;
  const int k = blockIdx.y;
  if (j > g.n2) return;
  a[P(g, 0, j, k)] = a[P(g, g.n1m, j, k)];
  a[P(g, g.n1, j, k)] = a[P(g, 1, j, k)];
}

void launch_sgs_finish(hipStream_t s, const Geo &g, const Fld &f, double *nu, double **nu1, int alp,
                       int scale_slot, const lw_settings &st) {
  LW_LAUNCH("k_scale_zero", pts(g) * 17.0, 0, k_scale_zero, grid3(g.n1m, g.n2m, g.n3m), BLK3, 0, s, g, f, nu,
            scale_slot);
  if (g.xprdic)
    LW_LAUNCH("k_xghost_scalar", (double)(g.n2 + 1) * (g.n3 + 1) * 32.0, 0, k_xghost_scalar,
              dim3((g.n2 + 64) / 64, g.n3 + 1), dim3(64), 0, s, g, nu);
  LW_LAUNCH("k_zghost_scalar", (double)(g.n1 + 1) * (g.n2 + 1) * 32.0, 0, k_zghost_scalar,
            dim3((g.n1 + 64) / 64, g.n2 + 1), dim3(64), 0, s, g, nu);
  EdgeCtx e;
  e.ib = alp ? 1 : g.ib;
  e.jb = alp ? 1 : g.jb;
  e.kb = alp ? 1 : g.kb;
  e.ybwall = st.bc_ybtm == 0;
  e.ytwall = st.bc_ytop == 0;
  e.xprdic = g.xprdic;
  // nu in, three edge viscosities out
  LW_LAUNCH("k_edges", (double)(g.n1 + 1) * (g.n2 + 1) * (g.n3 + 1) * 32.0, 0, k_edges,
            grid3(g.n1 + 1, g.n2 + 1, g.n3 + 1), BLK3, 0, s, g, nu, nu1[0], nu1[1], nu1[2], e);
}

void launch_zero_list(hipStream_t s, const Geo &g, double *nu, const int *lst, int n) {
  if (n > 0)
    LW_LAUNCH("k_zero_list", n * 20.0, 0, k_zero_list, dim3((n + 255) / 256), dim3(256), 0, s, g, nu, lst, n);
}

}  // namespace lw
