// gfx950 kernels of the alternative Poisson solver, poiss_mgrd.f90:
// BiCGSTAB (25-224) preconditioned by one 3-D multigrid V-cycle (mg3d,
// 227-268) that coarsens x and z and relaxes y-lines with a TDMA in a zebra
// (checkerboard in i and k) order. Every expression keeps the reference's
// operation order (built with -ffp-contract=off); the BiCGSTAB dot products
// are fixed-order block sums instead of the reference's serial loops, so the
// scalars, and with them the fields, agree with the reference to rounding.
//
// Device scalars M3::sc: 0 rho, 1 rv, 2 uppp, 3 bttt, 4 alp, 5 w, 6 bet,
// 7 rho_ol, 8 the iteration's max |res|, 9 its cell (last in k, j, i order,
// as (i-1) + n1m (j-1) + n1m n2m (k-1)), 10 test1 (the stop bound, set by
// the host each solve), 11 phi(1, n2m, 1) at the end of the solve.
#include "lw_kernels.h"

namespace lw {

enum { S_RHO = 0, S_RV, S_UP, S_BT, S_ALP, S_W, S_BET, S_RHO_OL, S_RESM, S_LOC, S_TEST, S_REF };

namespace {

constexpr int kT = 256;  // threads of the point kernels

__device__ inline double wsum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  return v;
}

// fixed-shape block sum of kT threads (valid in thread 0)
__device__ inline double bsum(double v, double *sh) {
  v = wsum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double r = 0.0;
  if (w == 0) {
    r = lane < kT / 64 ? sh[lane] : 0.0;
    r = wsum(r);
  }
  __syncthreads();
  return r;
}

__device__ inline long G(const M3 &m, int i, int j, int k) { return (long)i + m.GX * ((long)j + m.GY * (long)k); }
__device__ inline long R(const M3 &m, const Geo &g, int i, int j, int k) {
  return (long)(i - 1) + (long)m.n1md * ((long)(j - 1) + (long)g.n2m * (long)(k - 1));
}
__device__ inline double acc3(const M3 &m, int i, int j, int k) {
  return -1.0 * (m.aw[i] + m.ae[i] + m.as[j] + m.an[j] + m.ab[k] + m.af[k]);
}

// one zebra half of toplevel/relax (poiss_mgrd.f90:298-344, 500-591) on
// level l: the lines (i, k) with i = ibg + mod(kpm(k), 2) + 2q (half 0) or
// ibg + mod(k, 2) + 2q (half 1) take the right-hand side rhs (the finest
// level's padded array) or resid, minus wgt times the x and z neighbours,
// then the y TDMA with the stored pivots (gam(j) = an(j-1) bet(j-1)). One
// thread per line; the lines of one half read only lines of the other half,
// or their own old values where a level has one z line. out (the last
// toplevel of a V-cycle): the lines' final values also go to out, the
// BiCGSTAB vector in the fields' padded layout.
LW_KERNEL void k3_half(LW_KP(64, 0), const Geo &g, const M3 &m, int l, int half, double wgt, const double *rhs,
                                              double *out) {
  const int k = m.kkmg[l][0] + blockIdx.y;
  const int is = m.iimg[l][0] + (half == 0 ? m.kpm[k] % 2 : k % 2);
  const int i = is + 2 * (int)(blockIdx.x * 64 + threadIdx.x);
  if (i > m.iimg[l][1]) return;
  const int n2m = g.n2m, im = m.imm[i], ip = m.ipm[i], km = m.kmm[k], kp = m.kpm[k];
  const double aw = m.aw[i], ae = m.ae[i], ab = m.ab[k], af = m.af[k];
  double *gi = m.gi;
  for (int j = 1; j <= n2m; ++j) {
    const double r = rhs ? rhs[P(g, i, j, k)] : m.resid[R(m, g, i, j, k)];
    const double nb = aw * gi[G(m, im, j, k)] + ae * gi[G(m, ip, j, k)] + ab * gi[G(m, i, j, km)] +
                      af * gi[G(m, i, j, kp)];
    gi[G(m, i, j, k)] = r - wgt * nb;
  }
  double prev = gi[G(m, i, 1, k)] * m.bet[R(m, g, i, 1, k)];
  gi[G(m, i, 1, k)] = prev;
  for (int j = 2; j <= n2m; ++j) {
    prev = (gi[G(m, i, j, k)] - m.as[j] * prev) * m.bet[R(m, g, i, j, k)];
    gi[G(m, i, j, k)] = prev;
  }
  if (out) out[P(g, i, n2m, k)] = prev;
  double next = prev;
  for (int j = n2m - 1; j >= 1; --j) {
    const double gam = m.an[j] * m.bet[R(m, g, i, j, k)];
    next = gi[G(m, i, j, k)] - gam * next;
    gi[G(m, i, j, k)] = next;
    if (out) out[P(g, i, j, k)] = next;
  }
}

// the residual of level l (toplevel 349-360, rescal 407-432): resid =
// rhs - A gi (finest level, rhs given) or resid - A gi
LW_KERNEL void k3_resid(LW_KP(kT, 0), const Geo &g, const M3 &m, int l, const double *rhs) {
  const int ni = m.iimg[l][2];
  const int q = blockIdx.x * kT + threadIdx.x;
  if (q >= ni) return;
  const int i = m.iimg[l][0] + q, j = 1 + blockIdx.y, k = m.kkmg[l][0] + blockIdx.z;
  const double *gi = m.gi;
  const double acc = acc3(m, i, j, k);
  const long c = R(m, g, i, j, k);
  const double r = rhs ? rhs[P(g, i, j, k)] : m.resid[c];
  m.resid[c] = r - acc * gi[G(m, i, j, k)] - m.aw[i] * gi[G(m, m.imm[i], j, k)] -
               m.ae[i] * gi[G(m, m.ipm[i], j, k)] - m.as[j] * gi[G(m, i, j - 1, k)] -
               m.an[j] * gi[G(m, i, j + 1, k)] - m.ab[k] * gi[G(m, i, j, m.kmm[k])] -
               m.af[k] * gi[G(m, i, j, m.kpm[k])];
}

// godown (435-458), and the restriction at the end of toplevel (364-375):
// level l from level l + 1
LW_KERNEL void k3_godown(LW_KP(kT, 0), const Geo &g, const M3 &m, int l) {
  const int ni = m.iimg[l][2];
  const int q = blockIdx.x * kT + threadIdx.x;
  if (q >= ni) return;
  const int i = m.iimg[l][0] + q, j = 1 + blockIdx.y, k = m.kkmg[l][0] + blockIdx.z;
  const double *r = m.resid;
  const double fi = m.fidw[i], fk = m.fkdw[k];
  m.resid[R(m, g, i, j, k)] = (1.0 - fi) * (1.0 - fk) * r[R(m, g, m.ih1[i], j, m.kh1[k])] +
                              fi * (1.0 - fk) * r[R(m, g, m.ih2[i], j, m.kh1[k])] +
                              (1.0 - fi) * fk * r[R(m, g, m.ih1[i], j, m.kh2[k])] +
                              fi * fk * r[R(m, g, m.ih2[i], j, m.kh2[k])];
}

// goup (461-486): the lines i = ibg + mod(k, 2) + 2q of level l add the
// bilinear (x, z) interpolation of level l - 1
LW_KERNEL void k3_goup(LW_KP(kT, 0), const Geo &g, const M3 &m, int l) {
  const int k = m.kkmg[l][0] + blockIdx.z, j = 1 + blockIdx.y;
  const int i = m.iimg[l][0] + k % 2 + 2 * (int)(blockIdx.x * kT + threadIdx.x);
  if (i > m.iimg[l][1]) return;
  double *gi = m.gi;
  const double fi = m.fiup[i], fk = m.fkup[k];
  const long c = G(m, i, j, k);
  gi[c] = gi[c] + (1.0 - fi) * (1.0 - fk) * gi[G(m, m.il1[i], j, m.kl1[k])] +
          fi * (1.0 - fk) * gi[G(m, m.il2[i], j, m.kl1[k])] + (1.0 - fi) * fk * gi[G(m, m.il1[i], j, m.kl2[k])] +
          fi * fk * gi[G(m, m.il2[i], j, m.kl2[k])];
}

// ---- BiCGSTAB (25-224) ----------------------------------------------------

// grid-stride walk over the interior in the reference's k, j, i order
#define INTERIOR_LOOP(g, body)                                                     \
  {                                                                                \
    const long ni_ = (long)(g).n1m, nij_ = ni_ * (g).n2m, n_ = nij_ * (g).n3m;     \
    for (long q_ = blockIdx.x * (long)kT + threadIdx.x; q_ < n_; q_ += (long)gridDim.x * kT) { \
      const int k = 1 + (int)(q_ / nij_);                                          \
      const int j = 1 + (int)((q_ % nij_) / ni_);                                  \
      const int i = 1 + (int)(q_ % ni_);                                           \
      const long c = P(g, i, j, k);                                                \
      body                                                                         \
    }                                                                              \
  }

// start of a solve (IOLDV = 0): phi = 0 (the whole array, solver.f90:117),
// res = divgsum, and the scalars rho_ol = alp = w = 1
LW_KERNEL void k3_cg_init(LW_KP(kT, 0), const Geo &g, const M3 &m, const Fld &f, double test1) {
  const long n = (long)g.NX * g.NY * g.NZ;
  for (long q = blockIdx.x * (long)kT + threadIdx.x; q < n; q += (long)gridDim.x * kT) f.phi[q] = 0.0;
  INTERIOR_LOOP(g, m.res[c] = f.div[c];)
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    m.sc[S_RHO_OL] = 1.0;
    m.sc[S_ALP] = 1.0;
    m.sc[S_W] = 1.0;
    m.sc[S_TEST] = test1;
  }
}

// block partial sums of up to two products over the interior
template <int WHICH>
LW_KERNEL void k3_dot(LW_KP(kT, 0), const Geo &g, const M3 &m, const Fld &f) {
  __shared__ double sh[kT / 64];
  double a = 0.0, b = 0.0;
  if constexpr (WHICH == 0) {  // rho = sum divgsum * res (84-91)
    INTERIOR_LOOP(g, a = a + f.div[c] * m.res[c];)
  } else if constexpr (WHICH == 1) {  // rv = sum divgsum * v (120-127)
    INTERIOR_LOOP(g, a = a + f.div[c] * m.v[c];)
  } else {  // uppp = sum t s, bttt = sum t^2 (157-165)
    INTERIOR_LOOP(g, {
      const double t = m.t[c];
      a = a + t * m.s[c];
      b = b + t * t;
    })
  }
  const double sa = bsum(a, sh);
  const double sb = WHICH == 2 ? bsum(b, sh) : 0.0;
  if (threadIdx.x == 0) {
    m.part[blockIdx.x] = sa;
    m.part[kRedBlocks + blockIdx.x] = sb;
  }
}

// one workgroup: the fixed-order sum of the partials and the scalar update
// that follows it: 0 rho and bet (93), 1 rv and alp (129), 2 uppp, bttt, w (166)
template <int WHICH>
LW_KERNEL void k3_scal(LW_KP(kT, 0), const M3 &m) {
  __shared__ double sh[kT / 64];
  double a = 0.0, b = 0.0;
  for (int q = threadIdx.x; q < kRedBlocks; q += kT) {
    a += m.part[q];
    b += m.part[kRedBlocks + q];
  }
  const double sa = bsum(a, sh);
  const double sb = WHICH == 2 ? bsum(b, sh) : 0.0;
  if (threadIdx.x != 0) return;
  double *sc = m.sc;
  if constexpr (WHICH == 0) {
    sc[S_RHO] = sa;
    sc[S_BET] = sa / sc[S_RHO_OL] * sc[S_ALP] / sc[S_W];
  } else if constexpr (WHICH == 1) {
    sc[S_RV] = sa;
    sc[S_ALP] = sc[S_RHO] / sa;
  } else {
    sc[S_UP] = sa;
    sc[S_BT] = sb;
    sc[S_W] = sa / sb;
  }
}

// p = res + bet (p - w v) (96-103); the first iteration reads p = v = 0
LW_KERNEL void k3_cg_p(LW_KP(kT, 0), const Geo &g, const M3 &m, int first) {
  const double bet = m.sc[S_BET], w = m.sc[S_W];
  if (first) {
    INTERIOR_LOOP(g, m.p[c] = m.res[c] + bet * (0.0 - w * 0.0);)
  } else {
    INTERIOR_LOOP(g, m.p[c] = m.res[c] + bet * (m.p[c] - w * m.v[c]);)
  }
}

// v (or t) = A y (or A z) on the finest level (107-118, 143-154); y's
// ghost cells stay 0 as the reference's do
LW_KERNEL void k3_apply(LW_KP(kT, 0), const Geo &g, const M3 &m, const double *y, double *v) {
  INTERIOR_LOOP(g, {
    const double acc = acc3(m, i, j, k);
    v[c] = m.aw[i] * y[P(g, m.imm[i], j, k)] + m.ae[i] * y[P(g, m.ipm[i], j, k)] + m.as[j] * y[c - g.NX] +
           m.an[j] * y[c + g.NX] + m.ab[k] * y[P(g, i, j, m.kmm[k])] + m.af[k] * y[P(g, i, j, m.kpm[k])] +
           acc * y[c];
  })
}

// s = res - alp v (131-138)
LW_KERNEL void k3_cg_s(LW_KP(kT, 0), const Geo &g, const M3 &m) {
  const double alp = m.sc[S_ALP];
  INTERIOR_LOOP(g, m.s[c] = m.res[c] - alp * m.v[c];)
}

// phi += alp y + w z; res = s - w t with its max norm and the last cell in
// k, j, i order that holds it (168-193): block (max, cell) pairs
LW_KERNEL void k3_cg_update(LW_KP(kT, 0), const Geo &g, const M3 &m, const Fld &f) {
  __shared__ double shm[kT];
  __shared__ long shi[kT];
  const double alp = m.sc[S_ALP], w = m.sc[S_W];
  double mx = -1.0;
  long at = -1;
  const long ni = (long)g.n1m, nij = ni * g.n2m;
  INTERIOR_LOOP(g, {
    f.phi[c] = f.phi[c] + alp * m.y[c] + w * m.z[c];
    const double r = m.s[c] - w * m.t[c];
    m.res[c] = r;
    const double a = fabs(r);
    if (a >= mx) {  // later cells of a thread come later in k, j, i order
      mx = a;
      at = (long)(i - 1) + ni * (long)(j - 1) + nij * (long)(k - 1);
    }
  })
  shm[threadIdx.x] = mx;
  shi[threadIdx.x] = at;
  __syncthreads();
  for (int o = kT / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      const double b = shm[threadIdx.x + o];
      const long bi = shi[threadIdx.x + o];
      if (b > shm[threadIdx.x] || (b == shm[threadIdx.x] && bi > shi[threadIdx.x])) {
        shm[threadIdx.x] = b;
        shi[threadIdx.x] = bi;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    m.part[2 * kRedBlocks + blockIdx.x] = shm[0];
    m.part[3 * kRedBlocks + blockIdx.x] = (double)shi[0];
  }
}

// the iteration's max norm and its cell from the block pairs; rho_ol = rho
LW_KERNEL void k3_cg_resm(LW_KP(kT, 0), const M3 &m) {
  __shared__ double shm[kT];
  __shared__ double shi[kT];
  double mx = -1.0, at = -1.0;
  for (int q = threadIdx.x; q < kRedBlocks; q += kT) {
    const double b = m.part[2 * kRedBlocks + q], bi = m.part[3 * kRedBlocks + q];
    if (b > mx || (b == mx && bi > at)) {
      mx = b;
      at = bi;
    }
  }
  shm[threadIdx.x] = mx;
  shi[threadIdx.x] = at;
  __syncthreads();
  for (int o = kT / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      const double b = shm[threadIdx.x + o], bi = shi[threadIdx.x + o];
      if (b > shm[threadIdx.x] || (b == shm[threadIdx.x] && bi > shi[threadIdx.x])) {
        shm[threadIdx.x] = b;
        shi[threadIdx.x] = bi;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    m.sc[S_RESM] = shm[0];
    m.sc[S_LOC] = shi[0];
    m.sc[S_RHO_OL] = m.sc[S_RHO];
  }
}

// phi -= phi(1, n2m, 1) over the interior (210-218): the reference value is
// taken by one thread first, since the subtraction overwrites it
LW_KERNEL void k3_phiref_take(LW_KP(1024, 0), const Geo &g, const M3 &m, const Fld &f) { m.sc[S_REF] = f.phi[P(g, 1, g.n2m, 1)]; }
LW_KERNEL void k3_phiref(LW_KP(kT, 0), const Geo &g, const M3 &m, const Fld &f) {
  const double ref = m.sc[S_REF];
  INTERIOR_LOOP(g, f.phi[c] = f.phi[c] - ref;)
}

}  // namespace

static inline dim3 lines_grid(const M3 &m, int l) {
  const int nk = m.kkmg[l][2];
  const int ni = m.iimg[l][2];
  const int pairs = (ni + 1) / 2;  // lines of one half on a z line
  return dim3((unsigned)((pairs + 63) / 64), (unsigned)nk, 1);
}
static inline dim3 level_grid(const Geo &g, const M3 &m, int l) {
  return dim3((unsigned)((m.iimg[l][2] + kT - 1) / kT), (unsigned)g.n2m, (unsigned)m.kkmg[l][2]);
}
static inline double level_pts(const Geo &g, const M3 &m, int l) {
  return (double)m.iimg[l][2] * g.n2m * m.kkmg[l][2];
}

static void half(hipStream_t s, const Geo &g, const M3 &m, int l, int h, double wgt, const double *rhs,
                 double *out = nullptr) {
  // per line point: rhs or resid, 4 neighbours (two of them shared with the
  // adjacent line), bet, and gi written and read back: about 48 B
  LW_LAUNCH("k3_half", level_pts(g, m, l) * 0.5 * 48.0, 0, k3_half, lines_grid(m, l), dim3(64), 0, s, g, m, l, h,
            wgt, rhs, out);
}

// mg3d (227-268) with rhs on the finest level; the solution goes to sol
static void mg3d(hipStream_t s, const Geo &g, const M3 &m, const double *rhs, double *sol) {
  const int L = m.nlev;
  auto resid = [&](int l, const double *r) {
    LW_LAUNCH("k3_resid", level_pts(g, m, l) * 8.0 * 4, 0, k3_resid, level_grid(g, m, l), dim3(kT), 0, s, g, m, l,
              r);
  };
  auto godown = [&](int l) {
    LW_LAUNCH("k3_godown", level_pts(g, m, l) * 8.0 * 5, 0, k3_godown, level_grid(g, m, l), dim3(kT), 0, s, g, m,
              l);
  };
  auto goup = [&](int l) {
    dim3 gr = level_grid(g, m, l);
    gr.x = (unsigned)(((m.iimg[l][2] + 1) / 2 + kT - 1) / kT);
    LW_LAUNCH("k3_goup", level_pts(g, m, l) * 0.5 * 8.0 * 6, 0, k3_goup, gr, dim3(kT), 0, s, g, m, l);
  };
  auto relax = [&](int l, int iter, int iold) {
    half(s, g, m, l, 0, (double)iold, nullptr);
    half(s, g, m, l, 1, 1.0, nullptr);
    for (int q = 1; q <= iter - 1; ++q) {
      half(s, g, m, l, 0, 1.0, nullptr);
      half(s, g, m, l, 1, 1.0, nullptr);
    }
  };
  // toplevel(iold = 0): both halves, the residual, its restriction
  half(s, g, m, L, 0, 0.0, rhs);
  half(s, g, m, L, 1, 1.0, rhs);
  resid(L, rhs);
  if (L != 0) {
    godown(L - 1);
    for (int l = L - 1; l >= 1; --l) {
      relax(l, 1, 0);
      resid(l, nullptr);
      godown(l - 1);
    }
    relax(0, m.nbli, 0);
    for (int l = 1; l <= L - 1; ++l) {
      goup(l);
      relax(l, 1, 1);
    }
    goup(L);
  }
  // toplevel(iold = 1); its residual and restriction are not read
  half(s, g, m, L, 0, 1.0, rhs, sol);
  half(s, g, m, L, 1, 1.0, rhs, sol);
}

void launch_mgrd_init(hipStream_t s, const Geo &g, const M3 &m, const Fld &f, double test1) {
  LW_LAUNCH("k3_cg_init", (double)g.NX * g.NY * g.NZ * 8.0 + (double)g.n1m * g.n2m * g.n3m * 16.0, 0, k3_cg_init,
            dim3(kRedBlocks), dim3(kT), 0, s, g, m, f, test1);
}

// one BiCGSTAB iteration (81-199); the host reads sc[S_RESM] afterwards
void launch_mgrd_iteration(hipStream_t s, const Geo &g, const M3 &m, const Fld &f, int first) {
  const double n = (double)g.n1m * g.n2m * g.n3m;
  LW_LAUNCH("k3_dot", n * 16.0, 0, k3_dot<0>, dim3(kRedBlocks), dim3(kT), 0, s, g, m, f);
  LW_LAUNCH("k3_scal", 0, 0, k3_scal<0>, dim3(1), dim3(kT), 0, s, m);
  LW_LAUNCH("k3_cg_p", n * (first ? 16.0 : 32.0), 0, k3_cg_p, dim3(kRedBlocks), dim3(kT), 0, s, g, m, first);
  mg3d(s, g, m, m.p, m.y);
  LW_LAUNCH("k3_apply", n * 16.0, 0, k3_apply, dim3(kRedBlocks), dim3(kT), 0, s, g, m, m.y, m.v);
  LW_LAUNCH("k3_dot", n * 16.0, 0, k3_dot<1>, dim3(kRedBlocks), dim3(kT), 0, s, g, m, f);
  LW_LAUNCH("k3_scal", 0, 0, k3_scal<1>, dim3(1), dim3(kT), 0, s, m);
  LW_LAUNCH("k3_cg_s", n * 24.0, 0, k3_cg_s, dim3(kRedBlocks), dim3(kT), 0, s, g, m);
  mg3d(s, g, m, m.s, m.z);
  LW_LAUNCH("k3_apply", n * 16.0, 0, k3_apply, dim3(kRedBlocks), dim3(kT), 0, s, g, m, m.z, m.t);
  LW_LAUNCH("k3_dot", n * 16.0, 0, k3_dot<2>, dim3(kRedBlocks), dim3(kT), 0, s, g, m, f);
  LW_LAUNCH("k3_scal", 0, 0, k3_scal<2>, dim3(1), dim3(kT), 0, s, m);
  LW_LAUNCH("k3_cg_update", n * 56.0, 0, k3_cg_update, dim3(kRedBlocks), dim3(kT), 0, s, g, m, f);
  LW_LAUNCH("k3_cg_resm", 0, 0, k3_cg_resm, dim3(1), dim3(kT), 0, s, m);
}

void launch_mgrd_finish(hipStream_t s, const Geo &g, const M3 &m, const Fld &f) {
  LW_LAUNCH("k3_phiref_take", 8.0, 0, k3_phiref_take, dim3(1), dim3(1), 0, s, g, m, f);
  LW_LAUNCH("k3_phiref", (double)g.n1m * g.n2m * g.n3m * 16.0, 0, k3_phiref, dim3(kRedBlocks), dim3(kT), 0, s, g,
            m, f);
}

}  // namespace lw
