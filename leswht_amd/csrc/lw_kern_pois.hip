// gfx950 kernels of the pressure Poisson solve (poiss_z_ft.f90): real-to-
// half-complex z-DFT per (i,j) column, then for every wavenumber an
// x-semicoarsened zebra-line multigrid whose lines are y-tridiagonal solves
// with the precomputed LU factors bet/gam, then the inverse z-DFT.
//
// Layout: every multigrid plane is stored [kv][i][j] with j contiguous, so a
// y-line is one contiguous row. One wavefront solves one line: each lane owns
// R consecutive rows, the forward and backward substitutions are first-order
// linear recurrences, combined across lanes with a 6-step shuffle scan of
// affine maps, and each lane then re-runs the reference recurrence on its
// chunk from the exact carry-in. All wavenumbers advance together; a
// wavenumber that met its max-norm test sets its flag on the device and every
// later kernel skips it, so each keeps the reference's own V-cycle count.
#include "lw_kernels.h"

namespace lw {

__device__ inline double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ inline double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ inline double2 rmul(double a, double2 b) { return make_double2(a * b.x, a * b.y); }
__device__ inline double2 shfl_up2(double2 v, int d) {
  return make_double2(__shfl_up(v.x, d, 64), __shfl_up(v.y, d, 64));
}
__device__ inline double2 shfl_down2(double2 v, int d) {
  return make_double2(__shfl_down(v.x, d, 64), __shfl_down(v.y, d, 64));
}

#define ROW(base, i_) ((base) + (long)(i_) * mg.LDJ)

// ---------------------------------------------------------------------------
// z-DFT with FFTE's conventions (zfft1d.f:32-85): forward e^{-2 pi i k n/N}
// unnormalised; inverse by conjugation and 1/N. Direct sums over a column
// held in registers; the column length is a template parameter.
// ---------------------------------------------------------------------------
template <int N>
__global__ void k_fft_fwd(Geo g, MG mg, const double *div) {
  const int i = 1 + blockIdx.x * 64 + threadIdx.x;
  const int j = 1 + blockIdx.y;
  if (i > g.n1m) return;
  double x[N];
#pragma unroll
  for (int n = 0; n < N; ++n) x[n] = div[P(g, i, j, n + 1)];
#pragma unroll
  for (int kv = 0; kv <= N / 2; ++kv) {
    double sr = 0.0, si = 0.0;
#pragma unroll
    for (int n = 0; n < N; ++n) {
      const int m = (kv * n) % N;
      sr += x[n] * mg.twr[m];
      si += x[n] * mg.twi[m];
    }
    mg.rhs[(long)kv * mg.PCPLANE + (long)i * mg.LDJ + j] = make_double2(sr, si);
  }
}

template <int N>
__global__ void k_fft_inv(Geo g, MG mg, double *phi) {
  const int i = 1 + blockIdx.x * 64 + threadIdx.x;
  const int j = 1 + blockIdx.y;
  if (i > g.n1m) return;
  double2 X[N / 2 + 1];
#pragma unroll
  for (int kv = 0; kv <= N / 2; ++kv) X[kv] = mg.pc[(long)kv * mg.PCPLANE + (long)i * mg.LDJ + j];
  const double dn = 1.0 / (double)N;
#pragma unroll
  for (int n = 0; n < N; ++n) {
    double s = 0.0;
#pragma unroll
    for (int kv = 0; kv < N; ++kv) {
      const int m = (kv * n) % N;
      const double c = mg.twr[m], sn = -mg.twi[m];  // e^{+i theta}
      if (kv <= N / 2) s += X[kv].x * c - X[kv].y * sn;
      else s += X[N - kv].x * c + X[N - kv].y * sn;
    }
    phi[P(g, i, j, n + 1)] = s * dn;
  }
}

// runtime-length fallback (any n3m): re-reads the column
__global__ void k_fft_fwd_any(Geo g, MG mg, const double *div, int N) {
  const int i = 1 + blockIdx.x * 64 + threadIdx.x;
  const int j = 1 + blockIdx.y;
  if (i > g.n1m) return;
  for (int kv = 0; kv <= N / 2; ++kv) {
    double sr = 0.0, si = 0.0;
    for (int n = 0; n < N; ++n) {
      const int m = (int)(((long)kv * n) % N);
      const double xv = div[P(g, i, j, n + 1)];
      sr += xv * mg.twr[m];
      si += xv * mg.twi[m];
    }
    mg.rhs[(long)kv * mg.PCPLANE + (long)i * mg.LDJ + j] = make_double2(sr, si);
  }
}

__global__ void k_fft_inv_any(Geo g, MG mg, double *phi, int N) {
  const int i = 1 + blockIdx.x * 64 + threadIdx.x;
  const int j = 1 + blockIdx.y;
  if (i > g.n1m) return;
  const double dn = 1.0 / (double)N;
  for (int n = 0; n < N; ++n) {
    double s = 0.0;
    for (int kv = 0; kv < N; ++kv) {
      const int m = (int)(((long)kv * n) % N);
      const double c = mg.twr[m], sn = -mg.twi[m];
      const int kk = kv <= N / 2 ? kv : N - kv;
      const double2 X = mg.pc[(long)kk * mg.PCPLANE + (long)i * mg.LDJ + j];
      if (kv <= N / 2) s += X.x * c - X.y * sn;
      else s += X.x * c + X.y * sn;
    }
    phi[P(g, i, j, n + 1)] = s * dn;
  }
}

void launch_fft_forward(hipStream_t s, const Geo &g, const MG &mg, const Fld &f) {
  const dim3 gr((g.n1m + 63) / 64, g.n2m), bl(64);
  switch (g.n3m) {
    case 2: hipLaunchKernelGGL(k_fft_fwd<2>, gr, bl, 0, s, g, mg, f.div); break;
    case 4: hipLaunchKernelGGL(k_fft_fwd<4>, gr, bl, 0, s, g, mg, f.div); break;
    case 8: hipLaunchKernelGGL(k_fft_fwd<8>, gr, bl, 0, s, g, mg, f.div); break;
    case 16: hipLaunchKernelGGL(k_fft_fwd<16>, gr, bl, 0, s, g, mg, f.div); break;
    case 32: hipLaunchKernelGGL(k_fft_fwd<32>, gr, bl, 0, s, g, mg, f.div); break;
    case 64: hipLaunchKernelGGL(k_fft_fwd<64>, gr, bl, 0, s, g, mg, f.div); break;
    default: hipLaunchKernelGGL(k_fft_fwd_any, gr, bl, 0, s, g, mg, f.div, g.n3m); break;
  }
}

void launch_fft_inverse(hipStream_t s, const Geo &g, const MG &mg, const Fld &f) {
  const dim3 gr((g.n1m + 63) / 64, g.n2m), bl(64);
  switch (g.n3m) {
    case 2: hipLaunchKernelGGL(k_fft_inv<2>, gr, bl, 0, s, g, mg, f.phi); break;
    case 4: hipLaunchKernelGGL(k_fft_inv<4>, gr, bl, 0, s, g, mg, f.phi); break;
    case 8: hipLaunchKernelGGL(k_fft_inv<8>, gr, bl, 0, s, g, mg, f.phi); break;
    case 16: hipLaunchKernelGGL(k_fft_inv<16>, gr, bl, 0, s, g, mg, f.phi); break;
    case 32: hipLaunchKernelGGL(k_fft_inv<32>, gr, bl, 0, s, g, mg, f.phi); break;
    case 64: hipLaunchKernelGGL(k_fft_inv<64>, gr, bl, 0, s, g, mg, f.phi); break;
    default: hipLaunchKernelGGL(k_fft_inv_any, gr, bl, 0, s, g, mg, f.phi, g.n3m); break;
  }
}

// ---------------------------------------------------------------------------
// bookkeeping kernels
// ---------------------------------------------------------------------------
__global__ void k_mg_reset(MG mg, long n) {
  const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (q < n) mg.pc[q] = make_double2(0.0, 0.0);
  if (q < mg.n3mh) {
    mg.done[q] = 0;
    mg.cyc[q] = 0;
    mg.sumres[q] = 0ull;
  }
}

void launch_mg_zero(hipStream_t s, const Geo &, const MG &mg) {
  const long n = (long)mg.n3mh * mg.PCPLANE;
  hipLaunchKernelGGL(k_mg_reset, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, mg, n);
}

// start of a V-cycle: count it, or stop a wavenumber at the MGITR cap
__global__ void k_mg_cycle_begin(MG mg) {
  const int q = threadIdx.x;
  if (q >= mg.n3mh || mg.done[q]) return;
  if (mg.cyc[q] >= mg.mgitr) {
    mg.done[q] = 2;  // iteration limit exceeded
    return;
  }
  mg.cyc[q] += 1;
  mg.sumres[q] = 0ull;
}

// after toplevel: the max-norm test (poiss_z_ft.f90:284) per wavenumber
__global__ void k_mg_check(MG mg, double test) {
  const int q = threadIdx.x;
  if (q >= mg.n3mh || mg.done[q]) return;
  const double r = __longlong_as_double((long long)mg.sumres[q]);
  if (r < test) mg.done[q] = 1;
}

// ---------------------------------------------------------------------------
// one zebra line: y-tridiagonal solve with the LU factors of line i
// (trdiag1m, poiss_z_ft.f90:383-408). MODE selects the right-hand side:
// 0 relax first half / later sweeps  resd - oldv*(aw*g(im) + ae*g(ip))
// 1 relax second half                resd - aw*g(im) - ae*g(ip)
// 2 toplevel first half              rhs - oldv*(aw*pc(im) + ae*pc(ip))
// 3 toplevel second half             rhs - aw*(g(im)-tt) - ae*(g(ip)-tt),
//   tt = g(1,n2m-1) for kv = 1 (the gauge of the singular mode), 0 otherwise
// ---------------------------------------------------------------------------
template <int R, int MODE>
__global__ __launch_bounds__(256) void k_line(Geo g, MG mg, int ibase, int nlines, double oldv, double *ttsave) {
  const int lane = threadIdx.x;
  const int t = blockIdx.x * 4 + threadIdx.y;
  const int kv = 1 + blockIdx.y;
  if (mg.done[kv - 1]) return;
  if (t >= nlines) return;
  const int n2m = g.n2m;
  const int i = ibase + 2 * t;
  const long pl = (long)(kv - 1) * mg.PLANE, ppl = (long)(kv - 1) * mg.PCPLANE;
  double2 *gg = mg.gg + pl;
  const double *bet = ROW(mg.bet + pl, i), *gam = ROW(mg.gam + pl, i);
  const int im = mg.imm[i], ip = mg.ipm[i];
  const double aw = mg.aw[i], ae = mg.ae[i];
  const double2 *src = (MODE <= 1) ? ROW(mg.resd + pl, i) : ROW(mg.rhs + ppl, i);
  const double2 *nbm = (MODE == 2) ? ROW(mg.pc + ppl, im) : ROW(gg, im);
  const double2 *nbp = (MODE == 2) ? ROW(mg.pc + ppl, ip) : ROW(gg, ip);
  double2 tt = make_double2(0.0, 0.0);
  if (MODE == 3 && kv == 1) {
    tt = ROW(gg, 1)[n2m - 1];
    if (t == 0 && lane == 0) {
      ttsave[0] = tt.x;
      ttsave[1] = tt.y;
    }
  }
  const int j0 = 1 + lane * R;
  double2 rr[R];
  double bt[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int j = j0 + r;
    if (j <= n2m) {
      const double2 s0 = src[j];
      if (MODE == 0 || MODE == 2) {
        rr[r] = csub(s0, rmul(oldv, cadd(rmul(aw, nbm[j]), rmul(ae, nbp[j]))));
      } else if (MODE == 1) {
        rr[r] = csub(csub(s0, rmul(aw, nbm[j])), rmul(ae, nbp[j]));
      } else {
        const double2 left = kv == 1 ? csub(nbm[j], tt) : nbm[j];
        rr[r] = csub(csub(s0, rmul(aw, left)), rmul(ae, csub(nbp[j], tt)));
      }
      bt[r] = bet[j];
    } else {
      rr[r] = make_double2(0.0, 0.0);
      bt[r] = 0.0;
    }
  }
  // forward substitution y_j = (rr_j - as_j y_{j-1}) bet_j as affine maps
  double M = 1.0;
  double2 Cc = make_double2(0.0, 0.0);
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int j = j0 + r;
    if (j <= n2m) {
      const double m = j == 1 ? 0.0 : -(mg.as[j] * bt[r]);
      Cc = cadd(rmul(m, Cc), rmul(bt[r], rr[r]));
      M = m * M;
    }
  }
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const double Mp = __shfl_up(M, d, 64);
    const double2 Cp = shfl_up2(Cc, d);
    if (lane >= d) {
      Cc = cadd(rmul(M, Cp), Cc);
      M = M * Mp;
    }
  }
  double2 yp = shfl_up2(Cc, 1);
  if (lane == 0) yp = make_double2(0.0, 0.0);
  double2 y[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int j = j0 + r;
    if (j <= n2m) {
      yp = j == 1 ? rmul(bt[r], rr[r]) : rmul(bt[r], csub(rr[r], rmul(mg.as[j], yp)));
      y[r] = yp;
    } else {
      y[r] = make_double2(0.0, 0.0);
    }
  }
  // backward substitution x_j = y_j - gam_{j+1} x_{j+1}, x_n2m = y_n2m
  double gm[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int j = j0 + r;
    gm[r] = j < n2m ? gam[j + 1] : 0.0;
  }
  double Nn = 1.0;
  double2 D = make_double2(0.0, 0.0);
#pragma unroll
  for (int r = R - 1; r >= 0; --r) {
    const int j = j0 + r;
    if (j <= n2m) {
      const double nj = j == n2m ? 0.0 : -gm[r];
      D = cadd(rmul(nj, D), y[r]);
      Nn = nj * Nn;
    }
  }
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const double Np = __shfl_down(Nn, d, 64);
    const double2 Dp = shfl_down2(D, d);
    if (lane + d < 64) {
      D = cadd(rmul(Nn, Dp), D);
      Nn = Nn * Np;
    }
  }
  double2 xn = shfl_down2(D, 1);
  if (lane == 63) xn = make_double2(0.0, 0.0);
  double2 *out = ROW(gg, i);
#pragma unroll
  for (int r = R - 1; r >= 0; --r) {
    const int j = j0 + r;
    if (j <= n2m) {
      xn = j == n2m ? y[r] : csub(y[r], rmul(gm[r], xn));
      out[j] = xn;
    }
  }
}

typedef void (*LineKernel)(Geo, MG, int, int, double, double *);

template <int MODE>
static LineKernel pick_line(int R) {
  switch (R) {
    case 1: return k_line<1, MODE>;
    case 2: return k_line<2, MODE>;
    case 4: return k_line<4, MODE>;
    case 8: return k_line<8, MODE>;
    default: return k_line<16, MODE>;
  }
}

static inline int lanes_r(int n2m) {
  int r = 1;
  while (r * 64 < n2m) r <<= 1;
  return r;
}

static void launch_lines(hipStream_t s, const Geo &g, const MG &mg, int mode, int ibase, int nlines, double oldv,
                         double *tts) {
  if (nlines <= 0) return;
  const int R = lanes_r(g.n2m);
  LineKernel k = mode == 0 ? pick_line<0>(R) : mode == 1 ? pick_line<1>(R) : mode == 2 ? pick_line<2>(R)
                                                                                       : pick_line<3>(R);
  hipLaunchKernelGGL(k, dim3((unsigned)((nlines + 3) / 4), (unsigned)mg.n3mh), dim3(64, 4), 0, s, g, mg, ibase,
                     nlines, oldv, tts);
}

// kv = 1: shift the odd lines by tt after the second half (toplevel 339-352)
__global__ void k_top_shift(Geo g, MG mg, const double *ttsave) {
  if (mg.done[0]) return;
  const int i = 1 + 2 * blockIdx.x;
  const double2 tt = make_double2(ttsave[0], ttsave[1]);
  double2 *row = ROW(mg.gg, i);
  for (int j = 1 + threadIdx.x; j <= g.n2m; j += blockDim.x) row[j] = csub(row[j], tt);
}

// residual on the top level, max-norm per wavenumber, pc = gg
__global__ __launch_bounds__(256) void k_top_resid(Geo g, MG mg) {
  __shared__ double sh[4];
  const int kv = 1 + blockIdx.y;
  if (mg.done[kv - 1]) return;
  const int i = 1 + blockIdx.x;
  const long pl = (long)(kv - 1) * mg.PLANE, ppl = (long)(kv - 1) * mg.PCPLANE;
  const double2 *gg = mg.gg + pl;
  const double2 *gi = ROW(gg, i), *gm = ROW(gg, mg.imm[i]), *gp = ROW(gg, mg.ipm[i]);
  const double2 *rh = ROW(mg.rhs + ppl, i);
  const double *ac = ROW(mg.ac + pl, i);
  double2 *rs = ROW(mg.resd + pl, i);
  double2 *pc = ROW(mg.pc + ppl, i);
  const double aw = mg.aw[i], ae = mg.ae[i];
  double ar = 0.0;
  for (int j = 1 + threadIdx.x; j <= g.n2m; j += 256) {
    double2 r = rh[j];
    r = csub(r, rmul(aw, gm[j]));
    r = csub(r, rmul(ae, gp[j]));
    r = csub(r, rmul(mg.as[j], gi[j - 1]));
    r = csub(r, rmul(mg.an[j], gi[j + 1]));
    r = csub(r, rmul(ac[j], gi[j]));
    rs[j] = r;
    ar = fmax(ar, hypot(r.x, r.y));
    pc[j] = gi[j];
  }
  for (int o = 32; o > 0; o >>= 1) ar = fmax(ar, __shfl_down(ar, o, 64));
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = ar;
  __syncthreads();
  if (threadIdx.x == 0) {
    const double m = fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
    atomicMax(&mg.sumres[kv - 1], (unsigned long long)__double_as_longlong(m));
  }
}

// restriction of level lf's residual into level lf-1 (toplevel 370-377,
// godown 450-458)
__global__ __launch_bounds__(256) void k_restrict(Geo g, MG mg, int lf) {
  const int kv = 1 + blockIdx.y;
  if (mg.done[kv - 1]) return;
  const int lc = lf - 1;
  const int i = mg.iimg[lc][1] + blockIdx.x;
  const int ic = mg.iimg[lf][1] + 2 * blockIdx.x;
  double2 *base = mg.resd + (long)(kv - 1) * mg.PLANE;
  double2 *rc = ROW(base, i);
  const double2 *r0 = ROW(base, ic), *r1 = ROW(base, ic + 1);
  const double c1 = mg.cor1[i], c2 = mg.cor2[i];
  for (int j = 1 + threadIdx.x; j <= g.n2m; j += 256) rc[j] = cadd(rmul(c1, r0[j]), rmul(c2, r1[j]));
}

// godown residual on level ilev, in place (poiss_z_ft.f90:438-449)
__global__ __launch_bounds__(256) void k_lev_resid(Geo g, MG mg, int ilev) {
  const int kv = 1 + blockIdx.y;
  if (mg.done[kv - 1]) return;
  const int i = mg.iimg[ilev][1] + blockIdx.x;
  const long pl = (long)(kv - 1) * mg.PLANE;
  const double2 *gg = mg.gg + pl;
  const double2 *gi = ROW(gg, i), *gm = ROW(gg, mg.imm[i]), *gp = ROW(gg, mg.ipm[i]);
  const double *ac = ROW(mg.ac + pl, i);
  double2 *rs = ROW(mg.resd + pl, i);
  const double aw = mg.aw[i], ae = mg.ae[i];
  for (int j = 1 + threadIdx.x; j <= g.n2m; j += 256) {
    double2 r = rs[j];
    r = csub(r, rmul(aw, gm[j]));
    r = csub(r, rmul(ae, gp[j]));
    r = csub(r, rmul(mg.as[j], gi[j - 1]));
    r = csub(r, rmul(mg.an[j], gi[j + 1]));
    r = csub(r, rmul(ac[j], gi[j]));
    rs[j] = r;
  }
}

// goup (poiss_z_ft.f90:411-435), or with topc the prolongation into pc of
// toplevel(1) (312-327); evaluated left to right as the reference does
__global__ __launch_bounds__(256) void k_goup(Geo g, MG mg, int ilev, int topc) {
  const int kv = 1 + blockIdx.y;
  if (mg.done[kv - 1]) return;
  const int i = mg.iimg[ilev][1] + blockIdx.x;
  const bool last = i == mg.iimg[ilev][2];
  const long pl = (long)(kv - 1) * mg.PLANE;
  const double2 *gi = ROW(mg.gg + pl, i), *gp = ROW(mg.gg + pl, mg.ipm[i]);
  double2 *dst = topc ? ROW(mg.pc + (long)(kv - 1) * mg.PCPLANE, 2 * (blockIdx.x + 1))
                      : ROW(mg.gg + pl, mg.iimg[ilev + 1][1] + 1 + 2 * (int)blockIdx.x);
  const double c1 = mg.coi1[i], c2 = mg.coi2[i];
  for (int j = 1 + threadIdx.x; j <= g.n2m; j += 256) {
    if (!last) dst[j] = cadd(cadd(dst[j], rmul(c1, gi[j])), rmul(c2, gp[j]));
    else dst[j] = cadd(dst[j], rmul(c1, gi[j]));
  }
}

static inline double *tt_slot(const MG &mg) { return (double *)(mg.sumres + mg.n3mh); }

static inline dim3 rows(int n, int n3mh) { return dim3((unsigned)n, (unsigned)n3mh); }

void launch_mg_toplevel(hipStream_t s, const Geo &g, const MG &mg, int id, double oldv, double test) {
  const int nl = mg.nlev;
  const int nc = mg.iimg[nl - 1][2] - mg.iimg[nl - 1][1] + 1;
  if (id == 1) hipLaunchKernelGGL(k_goup, rows(nc, mg.n3mh), dim3(256), 0, s, g, mg, nl - 1, 1);
  launch_lines(s, g, mg, 2, 1, (g.n1m + 1) / 2, oldv, tt_slot(mg));
  launch_lines(s, g, mg, 3, 2, g.n1m / 2, 1.0, tt_slot(mg));
  hipLaunchKernelGGL(k_top_shift, dim3((unsigned)(g.n1m / 2)), dim3(256), 0, s, g, mg, (const double *)tt_slot(mg));
  hipLaunchKernelGGL(k_top_resid, rows(g.n1m, mg.n3mh), dim3(256), 0, s, g, mg);
  hipLaunchKernelGGL(k_restrict, rows(nc, mg.n3mh), dim3(256), 0, s, g, mg, nl);
  hipLaunchKernelGGL(k_mg_check, dim3(1), dim3(128), 0, s, mg, test);
}

void launch_mg_cycle(hipStream_t s, const Geo &g, const MG &mg, double test) {
  hipLaunchKernelGGL(k_mg_cycle_begin, dim3(1), dim3(128), 0, s, mg);
  auto relax = [&](int ilev, double oldv, int iiter) {
    const int n = mg.iimg[ilev][2] - mg.iimg[ilev][1] + 1;
    for (int it = 0; it < iiter; ++it) {
      launch_lines(s, g, mg, 0, mg.iimg[ilev][1], (n + 1) / 2, it == 0 ? oldv : 1.0, nullptr);
      launch_lines(s, g, mg, 1, mg.iimg[ilev][1] + 1, n / 2, 1.0, nullptr);
    }
  };
  for (int ilev = mg.nlev - 1; ilev >= 1; --ilev) {
    relax(ilev, 0.0, 1);
    const int n = mg.iimg[ilev][2] - mg.iimg[ilev][1] + 1;
    hipLaunchKernelGGL(k_lev_resid, rows(n, mg.n3mh), dim3(256), 0, s, g, mg, ilev);
    const int nc = mg.iimg[ilev - 1][2] - mg.iimg[ilev - 1][1] + 1;
    hipLaunchKernelGGL(k_restrict, rows(nc, mg.n3mh), dim3(256), 0, s, g, mg, ilev);
  }
  relax(0, 0.0, mg.nbli);
  for (int ilev = 0; ilev <= mg.nlev - 2; ++ilev) {
    const int n = mg.iimg[ilev][2] - mg.iimg[ilev][1] + 1;
    hipLaunchKernelGGL(k_goup, rows(n, mg.n3mh), dim3(256), 0, s, g, mg, ilev, 0);
    relax(ilev + 1, 1.0, 1);
  }
  launch_mg_toplevel(s, g, mg, 1, 1.0, test);
}

}  // namespace lw
