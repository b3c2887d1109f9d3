// gfx950 kernels of the x-periodic channel (ICH != 0, XPRDIC = 1):
// the mean pressure gradient of meanpg (slv_mmtm.f90:2-177) with its
// constant-flow-rate controller, the fcmfr.dat quantities of writehistory
// (misc_hist.f90:91-137), the start-up perturbation of addperturb
// (misc_init.f90:789-882), and the x- and z-FFT Poisson solver of
// poiss_ftft.f90 (the reference Makefile's default for x-periodic grids).
//
// Channel scalars f.chan (doubles):
//   0..4  pmi(0:4)                     5  pmiavg
//   6     the controller's integral    7..9  cmfravg_int(1:3)   10 cmfravg_dur
//   11    qflux  12 ubulk  13..15 cmfravg(1:3)   (the last history row)
// Sums are fixed-order block partials (f.chanpart, kChB blocks) summed by
// one workgroup, so every total is deterministic; they differ from the
// reference's serial loops by rounding only.
#include "lw_kernels.h"

namespace lw {

namespace {

constexpr int kChB = 64;   // blocks of the partial sums
constexpr int kChT = 256;  // threads per block

__device__ inline double wsum64(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

// fixed-shape block sum (valid in thread 0)
__device__ inline double bsum(double v, double *sh) {
  v = wsum64(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double r = 0.0;
  if (w == 0) {
    r = lane < kChT / 64 ? sh[lane] : 0.0;
    r = wsum64(r);
  }
  __syncthreads();
  return r;
}

}  // namespace

// the fluid volume and the u flux over the interior u points that funcbody
// counts as fluid (at x(i), ymp(j), zmp(k)), the wall-shear terms of meanpg
// at the y walls and the volume-weighted IB forcing of u: partials
// [c * kChB + block], c = 0 flowvol, 1 qvol, 2 bottom wall, 3 top wall, 4 IB
LW_KERNEL void k_chan_sums(LW_KP(kChT, 0), const Geo &g, const Fld &f, const IBComp &ibu, double re, int ybw, int ytw) {
  __shared__ double sh[kChT / 64];
  const int n1m = g.n1m, n2m = g.n2m, n3m = g.n3m, n2 = g.n2;
  const long ncell = (long)n1m * n2m * n3m;
  const long str = (long)kChB * kChT, q0 = (long)blockIdx.x * kChT + threadIdx.x;
  double fv = 0.0, qv = 0.0, wb = 0.0, wt = 0.0, fb = 0.0;
  for (long q = q0; q < ncell; q += str) {
    const int i = 1 + (int)(q % n1m), j = 1 + (int)((q / n1m) % n2m), k = 1 + (int)(q / ((long)n1m * n2m));
    if (f.umask[q]) {
      fv = fv + g.c2cx[i] * g.f2fy[j] * g.f2fz[k];
      qv = qv + f.u[P(g, i, j, k)] * g.c2cx[i] * g.f2fy[j] * g.f2fz[k];
    }
  }
  if (ybw) {
    const double dy1 = 1.0 / g.c2cyi[1], dy2 = dy1 + 1.0 / g.c2cyi[2];
    for (long q = q0; q < (long)n1m * n3m; q += str) {
      const int i = 1 + (int)(q % n1m), k = 1 + (int)(q / n1m);
      if (f.umask[(i - 1) + (long)n1m * ((long)n2m * (k - 1))]) {
        const double gu = (f.u[P(g, i, 1, k)] * (dy2 * dy2) - f.u[P(g, i, 2, k)] * (dy1 * dy1)) / (dy1 * dy2 * (dy2 - dy1));
        wb = wb + gu * g.c2cx[i] * g.f2fz[k] / re;
      }
    }
  }
  if (ytw) {
    const double dy1 = 1.0 / g.c2cyi[n2], dy2 = dy1 + 1.0 / g.c2cyi[n2m];
    for (long q = q0; q < (long)n1m * n3m; q += str) {
      const int i = 1 + (int)(q % n1m), k = 1 + (int)(q / n1m);
      if (f.umask[(i - 1) + (long)n1m * ((n2m - 1) + (long)n2m * (k - 1))]) {
        const double gu =
            (f.u[P(g, i, n2m, k)] * (dy2 * dy2) - f.u[P(g, i, n2m - 1, k)] * (dy1 * dy1)) / (dy1 * dy2 * (dy2 - dy1));
        wt = wt + gu * g.c2cx[i] * g.f2fz[k] / re;
      }
    }
  }
  for (long n = q0; n < (long)ibu.nbody; n += str)
    fb = fb + ibu.fcv[n] * g.c2cx[ibu.ifc[n]] * g.f2fy[ibu.jfc[n]] * g.f2fz[ibu.kfc[n]];
  const double s0 = bsum(fv, sh), s1 = bsum(qv, sh), s2 = bsum(wb, sh), s3 = bsum(wt, sh), s4 = bsum(fb, sh);
  if (threadIdx.x == 0) {
    f.chanpart[0 * kChB + blockIdx.x] = s0;
    f.chanpart[1 * kChB + blockIdx.x] = s1;
    f.chanpart[2 * kChB + blockIdx.x] = s2;
    f.chanpart[3 * kChB + blockIdx.x] = s3;
    f.chanpart[4 * kChB + blockIdx.x] = s4;
  }
}

// meanpg's totals (slv_mmtm.f90:58-176): pmi(1), pmi(2) from the walls,
// pmi(0) their sum plus the IB forcing of u, all over the domain volume;
// ICH = 1: the PI controller on the bulk velocity (when ntime > 0 and
// dt > 0), ICH = 2: pmi(0) = -UDRV_I
LW_KERNEL void k_meanpg(LW_KP(kChT, 0), const Geo &g, const Fld &f, double udrv, double dtconst, int ctrl, int ybw, int ytw) {
  __shared__ double sh[kChT / 64];
  double t[5];
  for (int c = 0; c < 5; ++c) t[c] = bsum(threadIdx.x < kChB ? f.chanpart[c * kChB + threadIdx.x] : 0.0, sh);
  if (threadIdx.x != 0) return;
  const double flowvol = t[0], qvol = t[1];
  const double ubulk = flowvol > 0.0 ? qvol / flowvol : 0.0;
  double pmi[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  if (ybw) pmi[1] = pmi[1] + t[2];
  if (ytw) pmi[2] = pmi[2] + t[3];
  pmi[0] = pmi[1] + pmi[2] + pmi[3] + pmi[4];
  pmi[0] = pmi[0] + t[4];
  const double volume = g.xl * g.yl * g.zl;
  for (int q = 0; q < 5; ++q) pmi[q] = pmi[q] / volume;
  if (g.ich == 1) {
    if (ctrl) {
      const double err = udrv - ubulk;
      f.chan[6] = f.chan[6] + err * dtconst;
      pmi[0] = pmi[0] - (1.0 * err + 0.1 * f.chan[6]);
    }
  } else if (g.ich == 2) {
    pmi[0] = -udrv;
  }
  for (int q = 0; q < 5; ++q) f.chan[q] = pmi[q];
}

// substepinit's running sum pmiavg = pmiavg + pmi(0)*2*alpha
// (slv_auxi.f90:118), restarted by stepinit (:13)
LW_KERNEL void k_pmiavg(LW_KP(1024, 0), const Fld &f, double alpha, int reset) {
  const double a = reset ? 0.0 : f.chan[5];
  f.chan[5] = a + f.chan[0] * 2. * alpha;
}

// writehistory's fcmfr.dat row (misc_hist.f90:112-136): qflux, ubulk and,
// from AVG_TST on, the running averages; into hist[16..21]
LW_KERNEL void k_chan_hist(LW_KP(kChT, 0), const Geo &g, const Fld &f, double *hist, double time, double dt,
                                                  double avg_tst) {
  __shared__ double sh[kChT / 64];
  const double flowvol = bsum(threadIdx.x < kChB ? f.chanpart[threadIdx.x] : 0.0, sh);
  const double qvol = bsum(threadIdx.x < kChB ? f.chanpart[kChB + threadIdx.x] : 0.0, sh);
  if (threadIdx.x != 0) return;
  const double qflux = qvol / g.xl, ubulk = flowvol > 0.0 ? qvol / flowvol : 0.0;
  double avg[3] = {0.0, 0.0, 0.0};
  if (time >= avg_tst) {
    f.chan[7] = f.chan[7] + qflux * dt;
    f.chan[8] = f.chan[8] + ubulk * dt;
    f.chan[9] = f.chan[9] + f.chan[5] * dt;
    f.chan[10] = f.chan[10] + dt;
    if (f.chan[10] > 0.0)
      for (int q = 0; q < 3; ++q) avg[q] = f.chan[7 + q] / f.chan[10];
  }
  hist[16] = qflux;
  hist[17] = ubulk;
  hist[18] = f.chan[5];
  for (int q = 0; q < 3; ++q) hist[19 + q] = avg[q];
}

static void launch_chan_sums(hipStream_t s, const Geo &g, const Fld &f, const IBComp &ibu, const lw_settings &st) {
  LW_LAUNCH("k_chan_sums", (double)g.n1m * g.n2m * g.n3m * 9.0, 0, k_chan_sums, dim3(kChB), dim3(kChT), 0, s, g, f,
            ibu, st.re, (int)(st.bc_ybtm == 0), (int)(st.bc_ytop == 0));
}

void launch_meanpg(hipStream_t s, const Geo &g, const Fld &f, const IBComp &ibu, const lw_settings &st, double dtconst,
                   int ctrl) {
  launch_chan_sums(s, g, f, ibu, st);
  LW_LAUNCH("k_meanpg", kChB * 40.0, 0, k_meanpg, dim3(1), dim3(kChT), 0, s, g, f, st.udrv_i, dtconst, ctrl,
            (int)(st.bc_ybtm == 0), (int)(st.bc_ytop == 0));
}

void launch_pmiavg(hipStream_t s, const Fld &f, double alpha, int reset) {
  LW_LAUNCH("k_pmiavg", 24.0, 0, k_pmiavg, dim3(1), dim3(1), 0, s, f, alpha, reset);
}

void launch_chan_hist(hipStream_t s, const Geo &g, const Fld &f, const IBComp &ibu, const lw_settings &st,
                      double *hist, double time, double dt) {
  launch_chan_sums(s, g, f, ibu, st);
  LW_LAUNCH("k_chan_hist", kChB * 16.0, 0, k_chan_hist, dim3(1), dim3(kChT), 0, s, g, f, hist, time, dt, st.avg_tst);
}

// ---------------------------------------------------------------------------
// addperturb (misc_init.f90:789-882): u += ptbu(j, k) and w += ptbw(i, j) in
// the band 0.5 < ymp < 0.99 of the fluid points (funcbody > 1e-10), each
// less its mean over the fluid section; the profiles come from the host,
// computed by the reference's expressions. Fixed-order block partials.
// ---------------------------------------------------------------------------
template <int W>
LW_KERNEL void k_ptb_add(LW_KP(kChT, 0), const Geo &g, const Fld &f, const Ptb &t) {
  __shared__ double sh[kChT / 64];
  const int n1m = g.n1m, n2m = g.n2m, n3m = g.n3m;
  const long ncell = (long)n1m * n2m * n3m;
  const long str = (long)kChB * kChT;
  const uint8_t *mask = W ? t.mw : t.mu;
  double *a = W ? f.w : f.u;
  double area = 0.0, rate = 0.0;
  for (long q = (long)blockIdx.x * kChT + threadIdx.x; q < ncell; q += str) {
    if (!mask[q]) continue;
    const int i = 1 + (int)(q % n1m), j = 1 + (int)((q / n1m) % n2m), k = 1 + (int)(q / ((long)n1m * n2m));
    area = area + (W ? g.f2fx[i] * g.f2fy[j] : g.f2fy[j] * g.f2fz[k]);
    if (t.band[j]) {
      const double p = W ? t.pw[(i - 1) + (long)n1m * (j - 1)] : t.pu[(j - 1) + (long)n2m * (k - 1)];
      const long c = P(g, i, j, k);
      a[c] = a[c] + p;
      rate = rate + (W ? p * g.f2fx[i] * g.f2fy[j] : p * g.f2fy[j] * g.f2fz[k]);
    }
  }
  const double s0 = bsum(area, sh), s1 = bsum(rate, sh);
  if (threadIdx.x == 0) {
    t.part[blockIdx.x] = s0;
    t.part[kChB + blockIdx.x] = s1;
  }
}

// adj = rate / area when the section holds fluid: part[2 kChB], and the flag
LW_KERNEL void k_ptb_adj(LW_KP(kChT, 0), const Ptb &t) {
  __shared__ double sh[kChT / 64];
  const double area = bsum(threadIdx.x < kChB ? t.part[threadIdx.x] : 0.0, sh);
  const double rate = bsum(threadIdx.x < kChB ? t.part[kChB + threadIdx.x] : 0.0, sh);
  if (threadIdx.x == 0) {
    t.part[2 * kChB] = area > 0.0 ? rate / area : 0.0;
    t.part[2 * kChB + 1] = area > 0.0 ? 1.0 : 0.0;
  }
}

template <int W>
LW_KERNEL void k_ptb_sub(LW_KP(1024, 0), const Geo &g, const Fld &f, const Ptb &t) {
  const long q = (long)blockIdx.x * 256 + threadIdx.x;
  const int n1m = g.n1m, n2m = g.n2m;
  if (q >= (long)n1m * n2m * g.n3m) return;
  if (!(W ? t.mw : t.mu)[q] || t.part[2 * kChB + 1] == 0.0) return;
  const int i = 1 + (int)(q % n1m), j = 1 + (int)((q / n1m) % n2m), k = 1 + (int)(q / ((long)n1m * n2m));
  double *a = W ? f.w : f.u;
  const long c = P(g, i, j, k);
  a[c] = a[c] - t.part[2 * kChB];
}

void launch_addperturb(hipStream_t s, const Geo &g, const Fld &f, const Ptb &t) {
  const long n = (long)g.n1m * g.n2m * g.n3m;
  const unsigned nb = (unsigned)((n + 255) / 256);
  LW_LAUNCH("k_ptb_add", (double)n * 17.0, 0, k_ptb_add<0>, dim3(kChB), dim3(kChT), 0, s, g, f, t);
  LW_LAUNCH("k_ptb_adj", kChB * 16.0, 0, k_ptb_adj, dim3(1), dim3(kChT), 0, s, t);
  LW_LAUNCH("k_ptb_sub", (double)n * 17.0, 0, k_ptb_sub<0>, dim3(nb), dim3(256), 0, s, g, f, t);
  LW_LAUNCH("k_ptb_add", (double)n * 17.0, 0, k_ptb_add<1>, dim3(kChB), dim3(kChT), 0, s, g, f, t);
  LW_LAUNCH("k_ptb_adj", kChB * 16.0, 0, k_ptb_adj, dim3(1), dim3(kChT), 0, s, t);
  LW_LAUNCH("k_ptb_sub", (double)n * 17.0, 0, k_ptb_sub<1>, dim3(nb), dim3(256), 0, s, g, f, t);
}

// ---------------------------------------------------------------------------
// poiss_ftft.f90 (21-161): forward DFT in z (real, wavenumbers 1..n3/2+1)
// and in x (complex, all n1m), a y-tridiagonal solve per wavenumber pair
// (ctrdiag, 207-244) with the mean mode fixed at (i, j) = (1, n2m), and the
// inverse transforms, with FFTE's conventions (forward e^{-2 pi i k n/N}
// unnormalised, inverse e^{+} over N). The transforms are direct sums over
// a twiddle table indexed by (k n) mod N, one thread per output value, so
// any n1m and n3m work; the solve runs the reference's recurrence on one
// thread per line. Spectra are held [kz][j][kx] (kx fastest): the threads of
// a line solve walk j with consecutive kx, so their accesses coalesce.
// This alternative path is not tuned.
// ---------------------------------------------------------------------------

__device__ inline long fidx(const Geo &g, int kz, int j, int kx) {
  return (long)(kx - 1) + (long)g.n1m * ((long)(j - 1) + (long)g.n2m * (long)(kz - 1));
}

// z: X(kz) = sum_k d(k) w^{(kz-1)(k-1)}, kz = 1..n3mh, per (i, j)
LW_KERNEL void k_ftft_zf(LW_KP(1024, 0), const Geo &g, const Fld &f, const Ftft &t) {
  const int i = 1 + blockIdx.x * 64 + threadIdx.x;
  const int j = 1 + blockIdx.y;
  const int kz = 1 + blockIdx.z;
  if (i > g.n1m) return;
  double sr = 0.0, si = 0.0;
  for (int k = 1; k <= g.n3m; ++k) {
    const int q = (int)(((long)(kz - 1) * (k - 1)) % g.n3m);
    const double d = f.div[P(g, i, j, k)];
    sr = sr + d * t.twz_r[q];
    si = si + d * t.twz_i[q];
  }
  t.a[fidx(g, kz, j, i)] = make_double2(sr, si);
}

// x: Y(kx) = sum_i X(i) w^{(kx-1)(i-1)} (sg = +1), or the inverse with the
// conjugate twiddles over n1m (sg = -1)
LW_KERNEL void k_ftft_x(LW_KP(1024, 0), const Geo &g, const Ftft &t, const double2 *in, double2 *out, int sg) {
  const int kx = 1 + blockIdx.x * 64 + threadIdx.x;
  const int j = 1 + blockIdx.y;
  const int kz = 1 + blockIdx.z;
  if (kx > g.n1m) return;
  double sr = 0.0, si = 0.0;
  for (int i = 1; i <= g.n1m; ++i) {
    const int q = (int)(((long)(kx - 1) * (i - 1)) % g.n1m);
    const double wr = t.twx_r[q], wi = sg * t.twx_i[q];
    const double2 x = in[fidx(g, kz, j, i)];
    sr = sr + (x.x * wr - x.y * wi);
    si = si + (x.x * wi + x.y * wr);
  }
  if (sg < 0) {
    const double dn = 1.0 / (double)g.n1m;
    sr = sr * dn;
    si = si * dn;
  }
  out[fidx(g, kz, j, kx)] = make_double2(sr, si);
}

// the y solve of one (kz, kx) line in place in t.b (ctrdiag,
// poiss_ftft.f90:75-121, 207-244): a = ajm, b = ajc, c = ajp as the
// reference forms them; for (kz, kx) = (1, 1) row n2m is replaced by
// phi = 0 before the solve and the solution shifted after it
LW_KERNEL void k_ftft_solve(LW_KP(1024, 0), const Geo &g, const Ftft &t) {
  const int kx = 1 + blockIdx.x * 64 + threadIdx.x;
  const int kz = 1 + blockIdx.y;
  if (kx > g.n1m) return;
  const int n2m = g.n2m;
  const double ak1 = t.ak1[kx], ak3 = t.ak3[kz];
  auto coef = [&](int j, double &a, double &b, double &c) {
    const int jjp = g.jpv[j];
    const double fl = (double)g.fixjl[j], fu = (double)g.fixju[j];
    a = g.f2fyi[j] * g.c2cyi[j] * (1.0 - fl);
    c = g.f2fyi[j] * g.c2cyi[jjp] * (1.0 - fu);
    b = -((a + c + ak1 + ak3) * (1.0 - fl) * (1.0 - fu) + (g.f2fyi[1] * g.c2cyi[2] + ak1 + ak3) * fl +
          (g.f2fyi[n2m] * g.c2cyi[n2m] + ak1 + ak3) * fu);
  };
  const bool ref = kz == 1 && kx == 1;
  double2 *r = t.b;
  double a, b, c;
  coef(1, a, b, c);
  double bet = 1.0 / b;
  double2 up = r[fidx(g, kz, 1, kx)];
  up = make_double2(up.x * bet, up.y * bet);
  r[fidx(g, kz, 1, kx)] = up;
  double cprev = c, aref = 0.0, bref = 1.0;
  double2 rref = make_double2(0.0, 0.0);
  for (int j = 2; j <= n2m; ++j) {
    coef(j, a, b, c);
    double2 rv = r[fidx(g, kz, j, kx)];
    if (ref && j == n2m) {
      rref = rv;
      bref = b;
      aref = a;
      rv = make_double2(0.0, 0.0);
      b = 1.0;
      a = 0.0;
      c = 0.0;
    }
    const double gv = cprev * bet;
    bet = 1.0 / (b - a * gv);
    up = make_double2((rv.x - a * up.x) * bet, (rv.y - a * up.y) * bet);
    t.gam[fidx(g, kz, j, kx)] = gv;
    r[fidx(g, kz, j, kx)] = up;
    cprev = c;
  }
  double2 unext = up;
  for (int j = n2m - 1; j >= 1; --j) {
    const double gv = t.gam[fidx(g, kz, j + 1, kx)];
    const double2 rv = r[fidx(g, kz, j, kx)];
    unext = make_double2(rv.x - gv * unext.x, rv.y - gv * unext.y);
    r[fidx(g, kz, j, kx)] = unext;
  }
  if (ref) {
    const double2 rm = r[fidx(g, kz, n2m - 1, kx)];
    const double2 t0 = make_double2(-aref * rm.x + rref.x, -aref * rm.y + rref.y);
    const double2 ph = make_double2(t0.x / bref, t0.y / bref);
    for (int j = 1; j <= n2m; ++j) {
      const double2 v = r[fidx(g, kz, j, kx)];
      r[fidx(g, kz, j, kx)] = make_double2(v.x - ph.x, v.y - ph.y);
    }
    r[fidx(g, kz, n2m, kx)] = make_double2(0.0, 0.0);
  }
}

// inverse z of the Hermitian-extended spectrum, real part into phi (i, j, k)
LW_KERNEL void k_ftft_zi(LW_KP(1024, 0), const Geo &g, const Fld &f, const Ftft &t) {
  const int i = 1 + blockIdx.x * 64 + threadIdx.x;
  const int j = 1 + blockIdx.y;
  const int k = 1 + blockIdx.z;
  if (i > g.n1m) return;
  const int n3m = g.n3m, n3mh = t.n3mh;
  double sr = 0.0;
  for (int kz = 1; kz <= n3m; ++kz) {
    double2 x;
    if (kz <= n3mh) {
      x = t.a[fidx(g, kz, j, i)];
    } else {
      const double2 y = t.a[fidx(g, n3m + 2 - kz, j, i)];
      x = make_double2(y.x, -y.y);
    }
    const int q = (int)(((long)(kz - 1) * (k - 1)) % n3m);
    const double wr = t.twz_r[q], wi = -t.twz_i[q];
    sr = sr + (x.x * wr - x.y * wi);
  }
  f.phi[P(g, i, j, k)] = sr * (1.0 / (double)n3m);
}

void launch_poisson_ftft(hipStream_t s, const Geo &g, const Fld &f, const Ftft &t) {
  const double nc = (double)g.n1m * g.n2m;
  const unsigned bx = (unsigned)((g.n1m + 63) / 64);
  LW_LAUNCH("k_ftft_zf", nc * (8.0 * g.n3m + 16.0 * t.n3mh), 0, k_ftft_zf, dim3(bx, g.n2m, t.n3mh), dim3(64), 0, s,
            g, f, t);
  LW_LAUNCH("k_ftft_x", nc * t.n3mh * 32.0, 0, k_ftft_x, dim3(bx, g.n2m, t.n3mh), dim3(64), 0, s, g, t, t.a, t.b, 1);
  LW_LAUNCH("k_ftft_solve", nc * t.n3mh * 40.0, 0, k_ftft_solve, dim3(bx, t.n3mh), dim3(64), 0, s, g, t);
  LW_LAUNCH("k_ftft_x", nc * t.n3mh * 32.0, 0, k_ftft_x, dim3(bx, g.n2m, t.n3mh), dim3(64), 0, s, g, t, t.b, t.a, -1);
  LW_LAUNCH("k_ftft_zi", nc * (8.0 * g.n3m + 16.0 * t.n3mh), 0, k_ftft_zi, dim3(bx, g.n2m, g.n3m), dim3(64), 0, s, g,
            f, t);
}

}  // namespace lw
