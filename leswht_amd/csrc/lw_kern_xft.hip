// gfx950 kernels of poiss_x_ft.f90: a DFT in x, then per x wavenumber a
// zebra multigrid in the (z, y) plane with tridiagonal line solves in y
// (mg2d, 235-297), or zebra line over-relaxation above IMGSOR (gsor2d,
// 485-615), and the inverse DFT. z is periodic (the only case for which the
// reference builds its z neighbour tables, coefmg 749-789).
//
// Every wavenumber iterates on its own: each kernel covers a range of
// wavenumbers and returns at once for those whose convergence flag is set,
// so the host may launch V-cycles (or sweeps) ahead of the flags and every
// wavenumber still runs exactly the reference's sequence. k_xft_check sets
// the flags after each toplevel (or resid3) from the max norm, which the
// residual kernels gather with an atomic max on the bit patterns of
// non-negative doubles (order-independent, so deterministic).
//
// Layouts (complex as double2, k fastest so that the threads of a line
// kernel, one per line, touch consecutive k):
//   rhs  [iv][j 1..n2m][k 1..n3m]         the transformed divergence
//   pc   [iv][j 0..n2m+1][k 1..n3m]       the top level's solution
//   gg   [iv][j 0..n2m+1][k 0..n3md]      ggii of every level (rows 0 and
//                                          n2m+1 stay zero)
//   resd [iv][j 1..n2m][k 1..n3md]        the levels' right-hand sides
//   ac, bet, gam as resd (real)
// This alternative Poisson path is not tuned.
#include "lw_kernels.h"

namespace lw {

namespace {

__device__ inline double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ inline double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ inline double2 crm(double a, double2 b) { return make_double2(a * b.x, a * b.y); }

struct XIdx {
  long n2m, n3m, n3md;
  __device__ long rhs(int iv, int j, int k) const { return (k - 1) + n3m * ((j - 1) + n2m * (long)(iv - 1)); }
  __device__ long pc(int iv, int j, int k) const { return (k - 1) + n3m * (j + (n2m + 2) * (long)(iv - 1)); }
  __device__ long gg(int iv, int j, int k) const { return k + (n3md + 1) * (j + (n2m + 2) * (long)(iv - 1)); }
  __device__ long lv(int iv, int j, int k) const { return (k - 1) + n3md * ((j - 1) + n2m * (long)(iv - 1)); }
};

__device__ inline XIdx xidx(const Geo &g, const Xft &x) { return XIdx{g.n2m, g.n3m, x.n3md}; }

}  // namespace

// forward DFT in x of the divergence: rhs(iv) = sum_i d(i) e^{-2 pi i (iv-1)(i-1)/n1m}
LW_KERNEL void k_xft_fwd(LW_KP(1024, 0), const Geo &g, const Fld &f, const Xft &x) {
  const int k = 1 + blockIdx.x * 64 + threadIdx.x;
  const int j = 1 + blockIdx.y;
  const int iv = 1 + blockIdx.z;
  if (k > g.n3m) return;
  double sr = 0.0, si = 0.0;
  for (int i = 1; i <= g.n1m; ++i) {
    const int q = (int)(((long)(iv - 1) * (i - 1)) % g.n1m);
    const double d = f.div[P(g, i, j, k)];
    sr = sr + d * x.twr[q];
    si = si + d * x.twi[q];
  }
  x.rhs[xidx(g, x).rhs(iv, j, k)] = make_double2(sr, si);
}

// inverse DFT of the Hermitian-extended spectrum, real part over n1m into phi
LW_KERNEL void k_xft_inv(LW_KP(1024, 0), const Geo &g, const Fld &f, const Xft &x) {
  const int i = 1 + blockIdx.x * 64 + threadIdx.x;
  const int j = 1 + blockIdx.y;
  const int k = 1 + blockIdx.z;
  if (i > g.n1m) return;
  const XIdx X = xidx(g, x);
  const int n1m = g.n1m;
  double sr = 0.0;
  for (int kx = 1; kx <= n1m; ++kx) {
    double2 c;
    if (kx <= x.n1mh) {
      c = x.pc[X.pc(kx, j, k)];
    } else {
      const double2 y = x.pc[X.pc(n1m + 2 - kx, j, k)];
      c = make_double2(y.x, -y.y);
    }
    const int q = (int)(((long)(kx - 1) * (i - 1)) % n1m);
    const double wr = x.twr[q], wi = -x.twi[q];
    sr = sr + (c.x * wr - c.y * wi);
  }
  f.phi[P(g, i, j, k)] = sr * (1.0 / (double)n1m);
}

// ICH = 1 (poiss_x_ft.f90:215-229): phiref = sum_i phi(i, n2m, 1) f2fx(i) / xl
// in the reference's order, subtracted from the interior
LW_KERNEL void k_xft_phiref(LW_KP(1024, 0), const Geo &g, const Fld &f, const Xft &x) {
  double s = 0.0;
  for (int i = 1; i <= g.n1m; ++i) s = s + f.phi[P(g, i, g.n2m, 1)] * g.f2fx[i];
  x.misc[0] = s / g.xl;
}

LW_KERNEL void k_xft_phisub(LW_KP(1024, 0), const Geo &g, const Fld &f, const Xft &x) {
  const int i = 1 + blockIdx.x * 64 + threadIdx.x;
  const int j = 1 + blockIdx.y;
  const int k = 1 + blockIdx.z;
  if (i > g.n1m) return;
  const long c = P(g, i, j, k);
  f.phi[c] = f.phi[c] - x.misc[0];
}

// one zebra half: the lines k = kb+par, kb+par+2, .. <= ke of wavenumbers
// iv0.. (one thread per line). Right-hand side R (NB: resd, or rhs when
// TOP) less the z coupling to the other half's lines, taken from gg, or
// from pc when PCN, in one of the reference's two forms:
//   formB = 0: R - w (ab n- + af n+)     (relax's first half, toplevel, gsor2d)
//   formB = 1: (R - ab n-) - af n+       (relax's second half)
// then trdiag1m (372-394) in place in gg. SOR: u = ww gg + cw u on pc.
// With gauge, line 1 stores its value at j = n2m (the mean mode's shift).
template <int TOP, int PCN, int SOR>
LW_KERNEL void k_xft_line(LW_KP(64, 0), const Geo &g, const Xft &x, int iv0, int kb, int ke, int par, double w,
                                                 int formB, double ww, double cw, int gauge) {
  const int k = kb + par + 2 * (int)(blockIdx.x * 64 + threadIdx.x);
  const int iv = iv0 + blockIdx.y;
  if (k > ke) return;
  if (x.done[iv - 1]) return;
  const XIdx X = xidx(g, x);
  const int n2m = g.n2m;
  const double ab = x.ab[k], af = x.af[k];
  const int km = x.kmm[k], kp = x.kpm[k];
  double2 prev = make_double2(0.0, 0.0);
  for (int j = 1; j <= n2m; ++j) {
    const double2 R = TOP ? x.rhs[X.rhs(iv, j, k)] : x.resd[X.lv(iv, j, k)];
    const double2 a = PCN ? x.pc[X.pc(iv, j, km)] : x.gg[X.gg(iv, j, km)];
    const double2 b = PCN ? x.pc[X.pc(iv, j, kp)] : x.gg[X.gg(iv, j, kp)];
    double2 r;
    if (formB) r = csub(csub(R, crm(ab, a)), crm(af, b));
    else r = csub(R, crm(w, cadd(crm(ab, a), crm(af, b))));
    const double bet = x.bet[X.lv(iv, j, k)];
    const double2 u = j == 1 ? crm(bet, r) : crm(bet, csub(r, crm(x.as[j], prev)));
    x.gg[X.gg(iv, j, k)] = u;
    prev = u;
  }
  double2 nx = prev;
  for (int j = n2m - 1; j >= 1; --j) {
    const long q = X.gg(iv, j, k);
    const double2 v = csub(x.gg[q], crm(x.gam[X.lv(iv, j + 1, k)], nx));
    x.gg[q] = v;
    nx = v;
  }
  if (SOR) {
    for (int j = 1; j <= n2m; ++j) {
      const long q = X.pc(iv, j, k);
      x.pc[q] = cadd(crm(ww, x.gg[X.gg(iv, j, k)]), crm(cw, x.pc[q]));
    }
  }
  if (gauge && k == 1) x.tt[iv - 1] = SOR ? x.pc[X.pc(iv, n2m, 1)] : x.gg[X.gg(iv, n2m, 1)];
}

// the mean mode (wavenumber 1): the odd lines less the stored shift
// (toplevel 326-335 on gg, gsor2d 520-527 on pc)
template <int SOR>
LW_KERNEL void k_xft_gauge(LW_KP(1024, 0), const Geo &g, const Xft &x) {
  const int k = 1 + 2 * (int)(blockIdx.x * 64 + threadIdx.x);
  const int j = 1 + blockIdx.y;
  if (k > g.n3m) return;
  if (x.done[0]) return;
  const XIdx X = xidx(g, x);
  const double2 tt = x.tt[0];
  const long q = SOR ? X.pc(1, j, k) : X.gg(1, j, k);
  if (SOR) x.pc[q] = csub(x.pc[q], tt);
  else x.gg[q] = csub(x.gg[q], tt);
}

// the top level's residual (toplevel 344-354; resid3 618-638 when SOR on
// pc): its max norm per wavenumber and, for the multigrid, resd and pc = gg
template <int SOR>
LW_KERNEL void k_xft_res(LW_KP(1024, 0), const Geo &g, const Xft &x, int iv0) {
  const int k = 1 + blockIdx.x * 64 + threadIdx.x;
  const int j = 1 + blockIdx.y;
  const int iv = iv0 + blockIdx.z;
  if (k > g.n3m) return;
  if (x.done[iv - 1]) return;
  const XIdx X = xidx(g, x);
  auto S = [&](int kk, int jj) { return SOR ? x.pc[X.pc(iv, jj, kk)] : x.gg[X.gg(iv, jj, kk)]; };
  const double2 c = S(k, j);
  double2 r = x.rhs[X.rhs(iv, j, k)];
  r = csub(r, crm(x.ab[k], S(x.kmm[k], j)));
  r = csub(r, crm(x.af[k], S(x.kpm[k], j)));
  r = csub(r, crm(x.as[j], S(k, j - 1)));
  r = csub(r, crm(x.an[j], S(k, j + 1)));
  r = csub(r, crm(x.ac[X.lv(iv, j, k)], c));
  if (!SOR) {
    x.resd[X.lv(iv, j, k)] = r;
    x.pc[X.pc(iv, j, k)] = c;
  }
  const double a = hypot(r.x, r.y);
  atomicMax(&x.sres[iv - 1], (unsigned long long)__double_as_longlong(a));
}

// toplevel's restriction to level nlev-1 (359-367)
LW_KERNEL void k_xft_restrict_top(LW_KP(1024, 0), const Geo &g, const Xft &x, int iv0) {
  const int lv = x.nlev - 1;
  const int k = x.kb[lv] + blockIdx.x * 64 + threadIdx.x;
  const int j = 1 + blockIdx.y;
  const int iv = iv0 + blockIdx.z;
  if (k > x.ke[lv]) return;
  if (x.done[iv - 1]) return;
  const XIdx X = xidx(g, x);
  const int kc = 2 * (k - x.kb[lv]) + 1;
  x.resd[X.lv(iv, j, k)] =
      cadd(crm(x.cor1[k], x.resd[X.lv(iv, j, kc)]), crm(x.cor2[k], x.resd[X.lv(iv, j, kc + 1)]));
}

// toplevel's interpolation from level nlev-1 to the even top lines (305-313)
LW_KERNEL void k_xft_interp_top(LW_KP(1024, 0), const Geo &g, const Xft &x, int iv0) {
  const int lv = x.nlev - 1;
  const int k = x.kb[lv] + blockIdx.x * 64 + threadIdx.x;
  const int j = 1 + blockIdx.y;
  const int iv = iv0 + blockIdx.z;
  if (k > x.ke[lv]) return;
  if (x.done[iv - 1]) return;
  const XIdx X = xidx(g, x);
  const int kc = 2 * (k - x.kb[lv] + 1);
  const long q = X.pc(iv, j, kc);
  x.pc[q] = cadd(cadd(x.pc[q], crm(x.coi1[k], x.gg[X.gg(iv, j, k)])), crm(x.coi2[k], x.gg[X.gg(iv, j, x.kpm[k])]));
}

// godown (415-437): the residual of level l's two cells under each cell of
// level l-1, then their restriction
LW_KERNEL void k_xft_godown(LW_KP(1024, 0), const Geo &g, const Xft &x, int iv0, int l) {
  const int k = x.kb[l - 1] + blockIdx.x * 64 + threadIdx.x;
  const int j = 1 + blockIdx.y;
  const int iv = iv0 + blockIdx.z;
  if (k > x.ke[l - 1]) return;
  if (x.done[iv - 1]) return;
  const XIdx X = xidx(g, x);
  const int kf = x.kb[l] + 2 * (k - x.kb[l - 1]);
  double2 rf[2];
  for (int q = 0; q < 2; ++q) {
    const int kk = kf + q;
    double2 r = x.resd[X.lv(iv, j, kk)];
    r = csub(r, crm(x.ab[kk], x.gg[X.gg(iv, j, x.kmm[kk])]));
    r = csub(r, crm(x.af[kk], x.gg[X.gg(iv, j, x.kpm[kk])]));
    r = csub(r, crm(x.as[j], x.gg[X.gg(iv, j - 1, kk)]));
    r = csub(r, crm(x.an[j], x.gg[X.gg(iv, j + 1, kk)]));
    r = csub(r, crm(x.ac[X.lv(iv, j, kk)], x.gg[X.gg(iv, j, kk)]));
    x.resd[X.lv(iv, j, kk)] = r;
    rf[q] = r;
  }
  x.resd[X.lv(iv, j, k)] = cadd(crm(x.cor1[k], rf[0]), crm(x.cor2[k], rf[1]));
}

// goup (397-412): level l's correction added to the even cells of level l+1
LW_KERNEL void k_xft_goup(LW_KP(1024, 0), const Geo &g, const Xft &x, int iv0, int l) {
  const int k = x.kb[l] + blockIdx.x * 64 + threadIdx.x;
  const int j = 1 + blockIdx.y;
  const int iv = iv0 + blockIdx.z;
  if (k > x.ke[l]) return;
  if (x.done[iv - 1]) return;
  const XIdx X = xidx(g, x);
  const int kh = x.kb[l + 1] + 1 + 2 * (k - x.kb[l]);
  const long q = X.gg(iv, j, kh);
  x.gg[q] = cadd(cadd(x.gg[q], crm(x.coi1[k], x.gg[X.gg(iv, j, k)])), crm(x.coi2[k], x.gg[X.gg(iv, j, x.kpm[k])]));
}

// the convergence test after a toplevel (or resid3) of wavenumbers
// iv0..iv0+nv-1: the max norm against `test`, the cycle count of those that
// pass, the norm kept, the gathering slot cleared; nact[slot] = the number
// still iterating
LW_KERNEL void k_xft_check(LW_KP(256, 0), const Xft &x, int iv0, int nv, int cyc, double test, int slot) {
  __shared__ int sh[4];
  int act = 0;
  for (int q = threadIdx.x; q < nv; q += 256) {
    const int iv = iv0 + q;
    if (x.done[iv - 1]) continue;
    const double s = __longlong_as_double((long long)x.sres[iv - 1]);
    x.sfin[iv - 1] = s;
    x.sres[iv - 1] = 0ull;
    if (s < test) {
      x.done[iv - 1] = 1;
      x.ii[iv - 1] = cyc;
    } else {
      act += 1;
    }
  }
  for (int o = 32; o > 0; o >>= 1) act += __shfl_down(act, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = act;
  __syncthreads();
  if (threadIdx.x == 0) x.nact[slot] = sh[0] + sh[1] + sh[2] + sh[3];
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static unsigned nb64(int n) { return (unsigned)((n + 63) / 64); }

void launch_xft_forward(hipStream_t s, const Geo &g, const Fld &f, const Xft &x) {
  LW_LAUNCH("k_xft_fwd", (double)g.n1m * g.n2m * g.n3m * (8.0 + 16.0 * x.n1mh / g.n1m), 0, k_xft_fwd,
            dim3(nb64(g.n3m), g.n2m, x.n1mh), dim3(64), 0, s, g, f, x);
}

void launch_xft_inverse(hipStream_t s, const Geo &g, const Fld &f, const Xft &x, int ich) {
  LW_LAUNCH("k_xft_inv", (double)g.n1m * g.n2m * g.n3m * 24.0, 0, k_xft_inv, dim3(nb64(g.n1m), g.n2m, g.n3m),
            dim3(64), 0, s, g, f, x);
  if (ich == 1) {
    LW_LAUNCH("k_xft_phiref", (double)g.n1m * 16.0, 0, k_xft_phiref, dim3(1), dim3(1), 0, s, g, f, x);
    LW_LAUNCH("k_xft_phisub", (double)g.n1m * g.n2m * g.n3m * 16.0, 0, k_xft_phisub,
              dim3(nb64(g.n1m), g.n2m, g.n3m), dim3(64), 0, s, g, f, x);
  }
}

// one zebra half over wavenumbers iv0 .. iv0+nv-1
static void half(hipStream_t s, const Geo &g, const Xft &x, int iv0, int nv, int kb, int ke, int par, double w,
                 int formB, int top, int pcn, int sor, double ww, double cw, int gauge) {
  if (ke - kb - par < 0) return;  // a one-line level has no second half
  const int nl = (ke - kb - par) / 2 + 1;
  const dim3 gr(nb64(nl), nv), bl(64);
  const double bytes = (double)nl * nv * g.n2m * 80.0;
  if (top && pcn && sor)
    LW_LAUNCH("k_xft_line", bytes, 0, (k_xft_line<1, 1, 1>), gr, bl, 0, s, g, x, iv0, kb, ke, par, w, formB, ww, cw,
              gauge);
  else if (top && pcn)
    LW_LAUNCH("k_xft_line", bytes, 0, (k_xft_line<1, 1, 0>), gr, bl, 0, s, g, x, iv0, kb, ke, par, w, formB, ww, cw,
              gauge);
  else if (top)
    LW_LAUNCH("k_xft_line", bytes, 0, (k_xft_line<1, 0, 0>), gr, bl, 0, s, g, x, iv0, kb, ke, par, w, formB, ww, cw,
              gauge);
  else
    LW_LAUNCH("k_xft_line", bytes, 0, (k_xft_line<0, 0, 0>), gr, bl, 0, s, g, x, iv0, kb, ke, par, w, formB, ww, cw,
              gauge);
}

// relax (440-482) of level l
static void relax(hipStream_t s, const Geo &g, const Xft &x, int iv0, int nv, int l, double oldv, int iiter) {
  for (int pass = 0; pass < iiter; ++pass) {
    half(s, g, x, iv0, nv, x.kb[l], x.ke[l], 0, pass == 0 ? oldv : 1.0, 0, 0, 0, 0, 0.0, 0.0, 0);
    half(s, g, x, iv0, nv, x.kb[l], x.ke[l], 1, 1.0, 1, 0, 0, 0, 0.0, 0.0, 0);
  }
}

static void check(hipStream_t s, const Xft &x, int iv0, int nv, int cyc, double test, int slot) {
  LW_LAUNCH("k_xft_check", (double)nv * 32.0, 0, k_xft_check, dim3(1), dim3(256), 0, s, x, iv0, nv, cyc, test, slot);
}

// toplevel (300-370) with id 0 (oldv = 0, no interpolation) or 1, then the
// convergence test with cycle count cyc
void launch_xft_toplevel(hipStream_t s, const Geo &g, const Xft &x, int iv0, int nv, int id, double test, int cyc) {
  const int l = x.nlev - 1;
  const unsigned nlc = nb64(x.ke[l] - x.kb[l] + 1);
  if (id == 1)
    LW_LAUNCH("k_xft_interp_top", (double)nv * g.n2m * (x.ke[l] - x.kb[l] + 1) * 64.0, 0, k_xft_interp_top,
              dim3(nlc, g.n2m, nv), dim3(64), 0, s, g, x, iv0);
  const int mean = iv0 == 1;
  half(s, g, x, iv0, nv, 1, g.n3m, 0, id == 1 ? 1.0 : 0.0, 0, 1, 1, 0, 0.0, 0.0, mean);
  if (mean)
    LW_LAUNCH("k_xft_gauge", (double)g.n2m * g.n3m * 16.0, 0, k_xft_gauge<0>, dim3(nb64(g.n3m / 2 + 1), g.n2m),
              dim3(64), 0, s, g, x);
  half(s, g, x, iv0, nv, 1, g.n3m, 1, 1.0, 0, 1, 0, 0, 0.0, 0.0, 0);
  LW_LAUNCH("k_xft_res", (double)nv * g.n2m * g.n3m * 120.0, 0, k_xft_res<0>, dim3(nb64(g.n3m), g.n2m, nv),
            dim3(64), 0, s, g, x, iv0);
  check(s, x, iv0, nv, cyc, test, 0);
  LW_LAUNCH("k_xft_restrict_top", (double)nv * g.n2m * (x.ke[l] - x.kb[l] + 1) * 48.0, 0, k_xft_restrict_top,
            dim3(nlc, g.n2m, nv), dim3(64), 0, s, g, x, iv0);
}

// the V-cycle of mg2d's loop (253-266) before its toplevel
void launch_xft_vcycle(hipStream_t s, const Geo &g, const Xft &x, int iv0, int nv) {
  for (int l = x.nlev - 1; l >= 1; --l) {
    relax(s, g, x, iv0, nv, l, 0.0, 1);
    const int nc = x.ke[l - 1] - x.kb[l - 1] + 1;
    LW_LAUNCH("k_xft_godown", (double)nv * g.n2m * nc * 240.0, 0, k_xft_godown, dim3(nb64(nc), g.n2m, nv), dim3(64),
              0, s, g, x, iv0, l);
  }
  relax(s, g, x, iv0, nv, 0, 0.0, x.nbli);
  for (int l = 0; l <= x.nlev - 2; ++l) {
    const int nc = x.ke[l] - x.kb[l] + 1;
    LW_LAUNCH("k_xft_goup", (double)nv * g.n2m * nc * 64.0, 0, k_xft_goup, dim3(nb64(nc), g.n2m, nv), dim3(64), 0, s,
              g, x, iv0, l);
    relax(s, g, x, iv0, nv, l + 1, 1.0, 1);
  }
}

// one gsor2d sweep (wgt 0 for the first, 1 after) with its resid3 and test
void launch_xft_sor(hipStream_t s, const Geo &g, const Xft &x, int iv0, int nv, double wgt, double ww, double test,
                    int cyc) {
  const double cw = wgt * (1.0 - ww);
  const int mean = iv0 == 1;
  half(s, g, x, iv0, nv, 1, g.n3m, 0, wgt, 0, 1, 1, 1, ww, cw, mean);
  if (mean)
    LW_LAUNCH("k_xft_gauge", (double)g.n2m * g.n3m * 16.0, 0, k_xft_gauge<1>, dim3(nb64(g.n3m / 2 + 1), g.n2m),
              dim3(64), 0, s, g, x);
  half(s, g, x, iv0, nv, 1, g.n3m, 1, 1.0, 0, 1, 1, 1, ww, cw, 0);
  LW_LAUNCH("k_xft_res", (double)nv * g.n2m * g.n3m * 96.0, 0, k_xft_res<1>, dim3(nb64(g.n3m), g.n2m, nv), dim3(64),
            0, s, g, x, iv0);
  check(s, x, iv0, nv, cyc, test, 1);
}

}  // namespace lw
