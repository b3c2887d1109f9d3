// gfx950 kernels of the pressure Poisson solve (poiss_z_ft.f90): real-to-
// half-complex z-DFT per (i,j) column, then for every wavenumber an
// x-semicoarsened zebra-line multigrid whose lines are y-tridiagonal solves
// with the precomputed LU factors bet/gam, then the inverse z-DFT. All
// wavenumbers advance together; a wavenumber that met its max-norm test is
// masked off, so each keeps the reference's own V-cycle count.
#include "lw_kernels.h"

namespace lw {

__device__ inline bool kv_active(int kv, unsigned long long m0, unsigned long long m1) {
  return kv <= 64 ? ((m0 >> (kv - 1)) & 1ull) : ((m1 >> (kv - 65)) & 1ull);
}

__device__ inline double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ inline double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ inline double2 rmul(double a, double2 b) { return make_double2(a * b.x, a * b.y); }

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
    mg.rhs[(long)kv * mg.PCPLANE + (long)j * mg.PCLD + i] = make_double2(sr, si);
  }
}

template <int N>
__global__ void k_fft_inv(Geo g, MG mg, double *phi) {
  const int i = 1 + blockIdx.x * 64 + threadIdx.x;
  const int j = 1 + blockIdx.y;
  if (i > g.n1m) return;
  double2 X[N / 2 + 1];
#pragma unroll
  for (int kv = 0; kv <= N / 2; ++kv) X[kv] = mg.pc[(long)kv * mg.PCPLANE + (long)j * mg.PCLD + i];
  const double dn = 1.0 / (double)N;
#pragma unroll
  for (int n = 0; n < N; ++n) {
    double s = 0.0;
#pragma unroll
    for (int kv = 0; kv < N; ++kv) {
      const int m = (kv * n) % N;
      // e^{+i theta} = (twr[m], -twi[m])
      const double c = mg.twr[m], sn = -mg.twi[m];
      if (kv <= N / 2) s += X[kv].x * c - X[kv].y * sn;
      else s += X[N - kv].x * c + X[N - kv].y * sn;
    }
    phi[P(g, i, j, n + 1)] = s * dn;
  }
}

// runtime-length fallback (any n3m up to 1024): re-reads the column
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
    mg.rhs[(long)kv * mg.PCPLANE + (long)j * mg.PCLD + i] = make_double2(sr, si);
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
      const double2 X = mg.pc[(long)kk * mg.PCPLANE + (long)j * mg.PCLD + i];
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
// multigrid
// ---------------------------------------------------------------------------
#define GG(i_, j_) gg[(long)(j_) * mg.LD + (i_)]
#define RS(i_, j_) rs[(long)(j_) * mg.LD + (i_)]
#define PCX(i_, j_) pc[(long)(j_) * mg.PCLD + (i_)]
#define RH(i_, j_) rh[(long)(j_) * mg.PCLD + (i_)]
#define RE(a_, i_, j_) a_[(long)(j_) * mg.LD + (i_)]

__global__ void k_mg_zero_pc(MG mg, long n) {
  const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (q < n) mg.pc[q] = make_double2(0.0, 0.0);
}

__global__ void k_mg_zero_sumres(MG mg) {
  if (threadIdx.x < mg.n3mh) mg.sumres[threadIdx.x] = 0ull;
}

void launch_mg_zero(hipStream_t s, const Geo &g, const MG &mg) {
  const long n = (long)mg.n3mh * mg.PCPLANE;
  hipLaunchKernelGGL(k_mg_zero_pc, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, mg, n);
}

// Thomas along j with the LU factors of line i (trdiag1m, poiss_z_ft.f90:383-408);
// the right-hand side of row j comes from the functor, results into gg
template <class RHS>
__device__ inline void line_solve(const Geo &g, const MG &mg, double2 *gg, const double *bet, const double *gam,
                                  int i, RHS rhsf) {
  const int n2m = g.n2m;
  double2 up = rmul(RE(bet, i, 1), rhsf(1));
  GG(i, 1) = up;
  for (int j = 2; j <= n2m; ++j) {
    up = rmul(RE(bet, i, j), csub(rhsf(j), rmul(mg.as[j], up)));
    GG(i, j) = up;
  }
  for (int j = n2m - 1; j >= 1; --j) {
    up = csub(GG(i, j), rmul(RE(gam, i, j + 1), up));
    GG(i, j) = up;
  }
}

// toplevel, first zebra half: odd lines from the solution's neighbours
__global__ void k_top_odd(Geo g, MG mg, double oldv, unsigned long long m0, unsigned long long m1) {
  const int kv = 1 + blockIdx.y;
  const int i = 1 + 2 * (blockIdx.x * 64 + threadIdx.x);
  if (i > g.n1m || !kv_active(kv, m0, m1)) return;
  double2 *gg = mg.gg + (long)(kv - 1) * mg.PLANE;
  const double2 *pc = mg.pc + (long)(kv - 1) * mg.PCPLANE;
  const double2 *rh = mg.rhs + (long)(kv - 1) * mg.PCPLANE;
  const double *bet = mg.bet + (long)(kv - 1) * mg.PLANE, *gam = mg.gam + (long)(kv - 1) * mg.PLANE;
  const int im = mg.imm[i], ip = mg.ipm[i];
  const double aw = mg.aw[i], ae = mg.ae[i];
  line_solve(g, mg, gg, bet, gam, i, [&](int j) {
    return csub(RH(i, j), rmul(oldv, cadd(rmul(aw, PCX(im, j)), rmul(ae, PCX(ip, j)))));
  });
}

// toplevel, second half: even lines; the kv = 1 gauge value tt = gg(1,n2m-1)
// is subtracted from the odd neighbours inline (the reference shifts the odd
// lines in place first); tt is saved for k_top_shift
__global__ void k_top_even(Geo g, MG mg, unsigned long long m0, unsigned long long m1, double *ttsave) {
  const int kv = 1 + blockIdx.y;
  const int i = 2 + 2 * (blockIdx.x * 64 + threadIdx.x);
  if (i > g.n1m || !kv_active(kv, m0, m1)) return;
  double2 *gg = mg.gg + (long)(kv - 1) * mg.PLANE;
  const double2 *rh = mg.rhs + (long)(kv - 1) * mg.PCPLANE;
  const double *bet = mg.bet + (long)(kv - 1) * mg.PLANE, *gam = mg.gam + (long)(kv - 1) * mg.PLANE;
  const double2 tt = kv == 1 ? GG(1, g.n2m - 1) : make_double2(0.0, 0.0);
  if (kv == 1 && i == 2) {
    ttsave[0] = tt.x;
    ttsave[1] = tt.y;
  }
  const int im = mg.imm[i], ip = mg.ipm[i];
  const double aw = mg.aw[i], ae = mg.ae[i];
  line_solve(g, mg, gg, bet, gam, i, [&](int j) {
    const double2 left = kv == 1 ? csub(GG(im, j), tt) : GG(im, j);
    return csub(csub(RH(i, j), rmul(aw, left)), rmul(ae, csub(GG(ip, j), tt)));
  });
}

__global__ void k_top_shift(Geo g, MG mg, unsigned long long m0, const double *ttsave) {
  // kv = 1 only
  if (!(m0 & 1ull)) return;
  const int i = 1 + 2 * (blockIdx.x * 64 + threadIdx.x);
  const int j = 1 + blockIdx.y;
  if (i > g.n1m - 1) return;
  double2 *gg = mg.gg;
  GG(i, j) = csub(GG(i, j), make_double2(ttsave[0], ttsave[1]));
}

// residual on the top level, max-norm per wavenumber, pc = gg
__global__ void k_top_resid(Geo g, MG mg, unsigned long long m0, unsigned long long m1) {
  const int kv = 1 + blockIdx.z;
  const int i = 1 + blockIdx.x * 64 + threadIdx.x;
  const int j = 1 + blockIdx.y;
  double ar = 0.0;
  const bool act = kv_active(kv, m0, m1);
  if (i <= g.n1m && act) {
    double2 *gg = mg.gg + (long)(kv - 1) * mg.PLANE;
    double2 *rs = mg.resd + (long)(kv - 1) * mg.PLANE;
    double2 *pc = mg.pc + (long)(kv - 1) * mg.PCPLANE;
    const double2 *rh = mg.rhs + (long)(kv - 1) * mg.PCPLANE;
    const double *ac = mg.ac + (long)(kv - 1) * mg.PLANE;
    double2 r = RH(i, j);
    r = csub(r, rmul(mg.aw[i], GG(mg.imm[i], j)));
    r = csub(r, rmul(mg.ae[i], GG(mg.ipm[i], j)));
    r = csub(r, rmul(mg.as[j], GG(i, j - 1)));
    r = csub(r, rmul(mg.an[j], GG(i, j + 1)));
    r = csub(r, rmul(RE(ac, i, j), GG(i, j)));
    RS(i, j) = r;
    ar = hypot(r.x, r.y);
    PCX(i, j) = GG(i, j);
  }
  if (!act) return;  // uniform per block (kv is per block)
  for (int o = 32; o > 0; o >>= 1) ar = fmax(ar, __shfl_down(ar, o, 64));
  if (threadIdx.x == 0) atomicMax(&mg.sumres[kv - 1], (unsigned long long)__double_as_longlong(ar));
}

// restriction of level lf's residual into level lf-1 (toplevel 370-377,
// godown 450-458)
__global__ void k_restrict(Geo g, MG mg, int lf, unsigned long long m0, unsigned long long m1) {
  const int kv = 1 + blockIdx.z;
  const int lc = lf - 1;
  const int i = mg.iimg[lc][1] + blockIdx.x * 64 + threadIdx.x;
  const int j = 1 + blockIdx.y;
  if (i > mg.iimg[lc][2] || !kv_active(kv, m0, m1)) return;
  double2 *rs = mg.resd + (long)(kv - 1) * mg.PLANE;
  const int ic = mg.iimg[lf][1] + 2 * (i - mg.iimg[lc][1]);
  RS(i, j) = cadd(rmul(mg.cor1[i], RS(ic, j)), rmul(mg.cor2[i], RS(ic + 1, j)));
}

// relax (poiss_z_ft.f90:466-511): zebra halves on level ilev
__global__ void k_relax_odd(Geo g, MG mg, int ilev, double oldv, unsigned long long m0, unsigned long long m1) {
  const int kv = 1 + blockIdx.y;
  const int i = mg.iimg[ilev][1] + 2 * (blockIdx.x * 64 + threadIdx.x);
  if (i > mg.iimg[ilev][2] || !kv_active(kv, m0, m1)) return;
  double2 *gg = mg.gg + (long)(kv - 1) * mg.PLANE;
  const double2 *rs = mg.resd + (long)(kv - 1) * mg.PLANE;
  const double *bet = mg.bet + (long)(kv - 1) * mg.PLANE, *gam = mg.gam + (long)(kv - 1) * mg.PLANE;
  const int im = mg.imm[i], ip = mg.ipm[i];
  const double aw = mg.aw[i], ae = mg.ae[i];
  line_solve(g, mg, gg, bet, gam, i, [&](int j) {
    return csub(RS(i, j), rmul(oldv, cadd(rmul(aw, GG(im, j)), rmul(ae, GG(ip, j)))));
  });
}

__global__ void k_relax_even(Geo g, MG mg, int ilev, unsigned long long m0, unsigned long long m1) {
  const int kv = 1 + blockIdx.y;
  const int i = mg.iimg[ilev][1] + 1 + 2 * (blockIdx.x * 64 + threadIdx.x);
  if (i > mg.iimg[ilev][2] || !kv_active(kv, m0, m1)) return;
  double2 *gg = mg.gg + (long)(kv - 1) * mg.PLANE;
  const double2 *rs = mg.resd + (long)(kv - 1) * mg.PLANE;
  const double *bet = mg.bet + (long)(kv - 1) * mg.PLANE, *gam = mg.gam + (long)(kv - 1) * mg.PLANE;
  const int im = mg.imm[i], ip = mg.ipm[i];
  const double aw = mg.aw[i], ae = mg.ae[i];
  line_solve(g, mg, gg, bet, gam, i, [&](int j) {
    return csub(csub(RS(i, j), rmul(aw, GG(im, j))), rmul(ae, GG(ip, j)));
  });
}

// godown residual on level ilev, in place (poiss_z_ft.f90:438-449)
__global__ void k_lev_resid(Geo g, MG mg, int ilev, unsigned long long m0, unsigned long long m1) {
  const int kv = 1 + blockIdx.z;
  const int i = mg.iimg[ilev][1] + blockIdx.x * 64 + threadIdx.x;
  const int j = 1 + blockIdx.y;
  if (i > mg.iimg[ilev][2] || !kv_active(kv, m0, m1)) return;
  const double2 *gg = mg.gg + (long)(kv - 1) * mg.PLANE;
  double2 *rs = mg.resd + (long)(kv - 1) * mg.PLANE;
  const double *ac = mg.ac + (long)(kv - 1) * mg.PLANE;
  double2 r = RS(i, j);
  r = csub(r, rmul(mg.aw[i], GG(mg.imm[i], j)));
  r = csub(r, rmul(mg.ae[i], GG(mg.ipm[i], j)));
  r = csub(r, rmul(mg.as[j], GG(i, j - 1)));
  r = csub(r, rmul(mg.an[j], GG(i, j + 1)));
  r = csub(r, rmul(RE(ac, i, j), GG(i, j)));
  RS(i, j) = r;
}

// goup (poiss_z_ft.f90:411-435) and the prolongation into pc of toplevel(1)
__global__ void k_goup(Geo g, MG mg, int ilev, int topc, unsigned long long m0, unsigned long long m1) {
  const int kv = 1 + blockIdx.z;
  const int i = mg.iimg[ilev][1] + blockIdx.x * 64 + threadIdx.x;
  const int j = 1 + blockIdx.y;
  if (i > mg.iimg[ilev][2] || !kv_active(kv, m0, m1)) return;
  double2 *gg = mg.gg + (long)(kv - 1) * mg.PLANE;
  const int t = i - mg.iimg[ilev][1];
  double2 corr;
  if (i < mg.iimg[ilev][2]) corr = cadd(rmul(mg.coi1[i], GG(i, j)), rmul(mg.coi2[i], GG(mg.ipm[i], j)));
  else corr = rmul(mg.coi1[i], GG(i, j));
  if (topc) {
    double2 *pc = mg.pc + (long)(kv - 1) * mg.PCPLANE;
    const int ic = 2 * (t + 1);
    // pc(ic) + coi1*g + coi2*g evaluated left to right as the reference does
    if (i < mg.iimg[ilev][2])
      PCX(ic, j) = cadd(cadd(PCX(ic, j), rmul(mg.coi1[i], GG(i, j))), rmul(mg.coi2[i], GG(mg.ipm[i], j)));
    else
      PCX(ic, j) = cadd(PCX(ic, j), rmul(mg.coi1[i], GG(i, j)));
  } else {
    const int ib = mg.iimg[ilev + 1][1] + 1 + 2 * t;
    if (i < mg.iimg[ilev][2])
      GG(ib, j) = cadd(cadd(GG(ib, j), rmul(mg.coi1[i], GG(i, j))), rmul(mg.coi2[i], GG(mg.ipm[i], j)));
    else
      GG(ib, j) = cadd(GG(ib, j), rmul(mg.coi1[i], GG(i, j)));
  }
  (void)corr;
}

static inline dim3 lines_grid(int nlines, int n3mh) { return dim3((unsigned)((nlines + 63) / 64), (unsigned)n3mh); }

void launch_mg_toplevel(hipStream_t s, const Geo &g, const MG &mg, int id, double oldv, unsigned long long m0,
                        unsigned long long m1) {
  const int nl = mg.nlev;
  if (id == 1) {
    const int nc = mg.iimg[nl - 1][2] - mg.iimg[nl - 1][1] + 1;
    hipLaunchKernelGGL(k_goup, dim3((nc + 63) / 64, g.n2m, mg.n3mh), dim3(64), 0, s, g, mg, nl - 1, 1, m0, m1);
  }
  const int nodd = (g.n1m + 1) / 2, neven = g.n1m / 2;
  hipLaunchKernelGGL(k_top_odd, lines_grid(nodd, mg.n3mh), dim3(64), 0, s, g, mg, oldv, m0, m1);
  double *tts = (double *)(mg.sumres + mg.n3mh);  // two doubles after the residual slots
  hipLaunchKernelGGL(k_top_even, lines_grid(neven, mg.n3mh), dim3(64), 0, s, g, mg, m0, m1, tts);
  hipLaunchKernelGGL(k_top_shift, dim3((neven + 63) / 64, g.n2m), dim3(64), 0, s, g, mg, m0, (const double *)tts);
  hipLaunchKernelGGL(k_mg_zero_sumres, dim3(1), dim3(128), 0, s, mg);
  hipLaunchKernelGGL(k_top_resid, dim3((g.n1m + 63) / 64, g.n2m, mg.n3mh), dim3(64), 0, s, g, mg, m0, m1);
  const int nc = mg.iimg[nl - 1][2] - mg.iimg[nl - 1][1] + 1;
  hipLaunchKernelGGL(k_restrict, dim3((nc + 63) / 64, g.n2m, mg.n3mh), dim3(64), 0, s, g, mg, nl, m0, m1);
}

void launch_mg_relax(hipStream_t s, const Geo &g, const MG &mg, int ilev, double oldv, int iiter,
                     unsigned long long m0, unsigned long long m1) {
  const int n = mg.iimg[ilev][2] - mg.iimg[ilev][1] + 1;
  const int nodd = (n + 1) / 2, neven = n / 2;
  for (int it = 0; it < iiter; ++it) {
    const double ov = it == 0 ? oldv : 1.0;
    hipLaunchKernelGGL(k_relax_odd, lines_grid(nodd, mg.n3mh), dim3(64), 0, s, g, mg, ilev, ov, m0, m1);
    if (neven > 0)
      hipLaunchKernelGGL(k_relax_even, lines_grid(neven, mg.n3mh), dim3(64), 0, s, g, mg, ilev, m0, m1);
  }
}

void launch_mg_godown(hipStream_t s, const Geo &g, const MG &mg, int ilev, unsigned long long m0,
                      unsigned long long m1) {
  const int n = mg.iimg[ilev][2] - mg.iimg[ilev][1] + 1;
  hipLaunchKernelGGL(k_lev_resid, dim3((n + 63) / 64, g.n2m, mg.n3mh), dim3(64), 0, s, g, mg, ilev, m0, m1);
  const int nc = mg.iimg[ilev - 1][2] - mg.iimg[ilev - 1][1] + 1;
  hipLaunchKernelGGL(k_restrict, dim3((nc + 63) / 64, g.n2m, mg.n3mh), dim3(64), 0, s, g, mg, ilev, m0, m1);
}

void launch_mg_goup(hipStream_t s, const Geo &g, const MG &mg, int ilev, unsigned long long m0,
                    unsigned long long m1) {
  const int n = mg.iimg[ilev][2] - mg.iimg[ilev][1] + 1;
  hipLaunchKernelGGL(k_goup, dim3((n + 63) / 64, g.n2m, mg.n3mh), dim3(64), 0, s, g, mg, ilev, 0, m0, m1);
}

}  // namespace lw
