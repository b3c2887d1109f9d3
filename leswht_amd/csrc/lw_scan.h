// Wavefront-wide scans for line solves on gfx950: DPP moves of doubles and
// tridiagonal solves with one line per wavefront or half-wavefront. The
// scans reassociate the serial recurrences anyway, so their arithmetic uses
// fused multiply-adds (one rounding, half the dependent latency of a
// multiply followed by an add).
#pragma once
#include <hip/hip_runtime.h>

namespace lw {

// DPP moves of a double (both halves with the same control; lanes whose
// source is outside their 16-lane row, or rows outside ROWS, receive `old`)
template <int CTRL, int ROWS = 0xF>
__device__ inline double dpp_d(double v, double old) {
  const long long vb = __double_as_longlong(v), ob = __double_as_longlong(old);
  const int lo = __builtin_amdgcn_update_dpp((int)ob, (int)vb, CTRL, ROWS, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp((int)(ob >> 32), (int)(vb >> 32), CTRL, ROWS, 0xF, false);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
// DPP move of a double with zero fill: lanes without a source (outside the
// 16-lane row, or past the wavefront's end) read 0 (bound_ctrl), so no
// identity value has to be placed in the destination first
template <int CTRL>
__device__ inline double dpp_z(double v) {
  const long long vb = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)vb, CTRL, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(vb >> 32), CTRL, 0xF, 0xF, true);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
__device__ inline double rdlane(double v, int l) {
  const long long vb = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)vb, l), hi = __builtin_amdgcn_readlane((int)(vb >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
constexpr int kShr1 = 0x111, kShr2 = 0x112, kShr4 = 0x114, kShr8 = 0x118;  // row_shr:n
constexpr int kShl1 = 0x101, kShl2 = 0x102, kShl4 = 0x104, kShl8 = 0x108;  // row_shl:n
constexpr int kBcast15 = 0x142, kBcast31 = 0x143;  // row_bcast:15, row_bcast:31
constexpr int kWshl1 = 0x130, kWshr1 = 0x138;      // wave_shl:1, wave_shr:1

// 2x2 matrices of the pivot recurrence, row-major (p, q; r, s)
struct M22 {
  double p, q, r, s;
};
__device__ inline M22 m22_mul(const M22 &A, const M22 &B) {  // A*B
  return M22{fma(A.p, B.p, A.q * B.r), fma(A.p, B.q, A.q * B.s), fma(A.r, B.p, A.s * B.r), fma(A.r, B.q, A.s * B.s)};
}
// scale by the power of two that brings the largest entry near 1 (exact)
__device__ inline M22 m22_norm(const M22 &A) {
  const double m = fmax(fmax(fabs(A.p), fabs(A.q)), fmax(fabs(A.r), fabs(A.s)));
  if (m == 0.0) return A;
  const int e = -__builtin_amdgcn_frexp_exp(m);
  return M22{ldexp(A.p, e), ldexp(A.q, e), ldexp(A.r, e), ldexp(A.s, e)};
}
template <int CTRL, int ROWS = 0xF>
__device__ inline M22 m22_dpp(const M22 &A) {  // identity where nothing moves in
  if constexpr (ROWS == 0xF)
    return M22{dpp_d<CTRL>(A.p, 1.0), dpp_z<CTRL>(A.q), dpp_z<CTRL>(A.r), dpp_d<CTRL>(A.s, 1.0)};
  else
    return M22{dpp_d<CTRL, ROWS>(A.p, 1.0), dpp_d<CTRL, ROWS>(A.q, 0.0), dpp_d<CTRL, ROWS>(A.r, 0.0),
               dpp_d<CTRL, ROWS>(A.s, 1.0)};
}

// Solves a_t x_{t-1} + b_t x_t + c_t x_{t+1} = f_t, t = 1..N (a_1 and c_N
// unused) with one line per wavefront: lane l owns rows t = 1 + l*R .. l*R+R
// (N <= 64*R; FULL: N == 64*R). The arithmetic of a row is Thomas's
// (trdiag1/trdiag2, slv_mmtm.f90:1365-1455): gam_t = c_{t-1} bet_{t-1},
// bet_t = 1/(b_t - a_t gam_t), y_t = (f_t - a_t y_{t-1}) bet_t,
// x_t = y_t - gam_{t+1} x_{t+1}. What a lane needs from the rows before it
// comes from scans: the pivot d = 1/bet of its first row from the products
// of the 2x2 maps (p_t, p_{t-1}) -> (b_t p_t - a_t c_{t-1} p_{t-1}, p_t),
// d_t = p_t / p_{t-1} (normalised by exact powers of two), and y, x from the
// affine maps of the two substitutions.
// LANES = 32: each half of the wavefront solves its own line (lanes 0-31,
// 32-63), with half the scan steps per line.
template <int R, bool FULL, int LANES = 64>
__device__ inline void tri_line(const double (&a)[R], const double (&b)[R], const double (&c)[R],
                                const double (&f)[R], int N, int lane, double (&x)[R]) {
  const int ll = LANES == 64 ? lane : (lane & 31);  // lane within the line
  const int t0 = 1 + ll * R;
  auto ok = [&](int r) { return FULL || t0 + r <= N; };
  // c of the row before this lane's first (unused on a line's first lane)
  const double cl = dpp_z<kWshr1>(c[R - 1]);
  // pivot maps: this lane's product, then the exclusive scan over the lanes
  M22 L{1.0, 0.0, 0.0, 1.0};
#pragma unroll
  for (int r = 0; r < R; ++r) {
    if (ok(r)) {
      const double cp = r == 0 ? cl : c[r - 1];
      const double q = t0 + r == 1 ? 0.0 : -(a[r] * cp);
      L = M22{fma(b[r], L.p, q * L.r), fma(b[r], L.q, q * L.s), L.p, L.q};
    }
  }
  L = m22_norm(L);
#define PIV_STEP(CT)                       \
  {                                        \
    const M22 B_ = m22_dpp<CT>(L);         \
    L = m22_norm(m22_mul(L, B_));          \
  }
  PIV_STEP(kShr1)
  PIV_STEP(kShr2)
  PIV_STEP(kShr4)
  PIV_STEP(kShr8)
#undef PIV_STEP
  if (LANES == 64) {
    L = m22_norm(m22_mul(L, m22_dpp<kBcast15, 0x2>(L)));
    L = m22_norm(m22_mul(L, m22_dpp<kBcast31, 0x4>(L)));
    L = m22_norm(m22_mul(L, m22_dpp<kBcast15, 0x8>(L)));
  } else {
    L = m22_norm(m22_mul(L, m22_dpp<kBcast15, 0xA>(L)));  // rows 1 and 3 of each line
  }
  const M22 E = m22_dpp<kWshr1>(L);  // the rows before this lane (identity for its first lane)
  // bet of the row before this lane's first: p_{t0-1} / p_{t0-2}, inverted
  double bet = ll == 0 ? 0.0 : 1.0 / (E.p / E.r);
  double bt[R], gm[R];
  double cprev = cl;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int t = t0 + r;
    if (ok(r)) {
      if (t == 1) {
        bet = 1.0 / b[r];
      } else {
        const double gam = cprev * bet;
        bet = 1.0 / fma(-a[r], gam, b[r]);
      }
      bt[r] = bet;
      gm[r] = t == N ? 0.0 : c[r] * bet;  // gam_{t+1}
      cprev = c[r];
    } else {
      bt[r] = 0.0;
      gm[r] = 0.0;
    }
  }
  // forward substitution y_t = (f_t - a_t y_{t-1}) bet_t as affine maps
  double M = 1.0, Cc = 0.0;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    if (ok(r)) {
      const double m = t0 + r == 1 ? 0.0 : -(a[r] * bt[r]);
      Cc = fma(m, Cc, bt[r] * f[r]);
      M = m * M;
    }
  }
#define FWD_R(CT)                       \
  {                                     \
    const double Mp = dpp_d<CT>(M, 1.0); \
    const double Cp = dpp_z<CT>(Cc); \
    Cc = fma(M, Cp, Cc);                \
    M = M * Mp;                         \
  }
  FWD_R(kShr1)
  FWD_R(kShr2)
  FWD_R(kShr4)
  FWD_R(kShr8)
#undef FWD_R
  if (LANES == 64) {
    Cc = fma(M, dpp_d<kBcast15, 0x2>(Cc, 0.0), Cc);
    Cc = fma(M, dpp_d<kBcast31, 0x4>(Cc, 0.0), Cc);
    Cc = fma(M, dpp_d<kBcast15, 0x8>(Cc, 0.0), Cc);
  } else {
    Cc = fma(M, dpp_d<kBcast15, 0xA>(Cc, 0.0), Cc);
  }
  double yp = dpp_z<kWshr1>(Cc);
  if (LANES == 32 && ll == 0) yp = 0.0;
  double y[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    if (ok(r)) {
      yp = t0 + r == 1 ? f[r] * bt[r] : fma(-a[r], yp, f[r]) * bt[r];
      y[r] = yp;
    } else {
      y[r] = 0.0;
    }
  }
  // back substitution x_t = y_t - gam_{t+1} x_{t+1}
  double Nn = 1.0, D = 0.0;
#pragma unroll
  for (int r = R - 1; r >= 0; --r) {
    if (ok(r)) {
      const double nj = t0 + r == N ? 0.0 : -gm[r];
      D = fma(nj, D, y[r]);
      Nn = nj * Nn;
    }
  }
#define BWD_R(CT)                        \
  {                                      \
    const double Np = dpp_d<CT>(Nn, 1.0); \
    const double Dp = dpp_z<CT>(D);  \
    D = fma(Nn, Dp, D);                  \
    Nn = Nn * Np;                        \
  }
  BWD_R(kShl1)
  BWD_R(kShl2)
  BWD_R(kShl4)
  BWD_R(kShl8)
#undef BWD_R
  if (LANES == 64) {
    const double n1 = rdlane(Nn, 16), n2 = rdlane(Nn, 32);
    const double d1 = rdlane(D, 16), d2 = rdlane(D, 32), d3 = rdlane(D, 48);
    const double s2 = fma(n2, d3, d2);  // start of row 2
    const double s1 = fma(n1, s2, d1);  // start of row 1
    const int row = lane >> 4;
    const double next = row == 3 ? 0.0 : row == 2 ? d3 : row == 1 ? s2 : s1;
    D = fma(Nn, next, D);
  } else {
    const double d1 = rdlane(D, 16), d3 = rdlane(D, 48);
    const int row = lane >> 4;
    const double next = row == 0 ? d1 : row == 2 ? d3 : 0.0;
    D = fma(Nn, next, D);
  }
  double xn = dpp_z<kWshl1>(D);
  if (LANES == 32 && ll == 31) xn = 0.0;
#pragma unroll
  for (int r = R - 1; r >= 0; --r) {
    if (ok(r)) {
      xn = t0 + r == N ? y[r] : fma(-gm[r], xn, y[r]);
      x[r] = xn;
    } else {
      x[r] = 0.0;
    }
  }
}

// Periodic tridiagonal solve (trdiag3p, slv_mmtm.f90:1572-1641) of
// a_t x_{t-1} + b_t x_t + c_t x_{t+1} = f_t, t = 1..N, x_0 = x_N and
// x_{N+1} = x_1, on lines of 32 lanes (each half of the wavefront solves its
// own line) with N = 32*R: lane l owns rows 1 + l*R .. l*R + R. The rows
// follow the reference's recurrences: pn_t = 1/(b_t + a_t q_{t-1}),
// q_t = -c_t pn_t, s_t = -a_t s_{t-1} pn_t, f_t = (f_t - a_t f_{t-1}) pn_t;
// then from s_N = 1, qe_N = 0 down: s_t += q_t s_{t+1},
// qe_t = f_t + q_t qe_{t+1}; x_N from rows 1, N-1 and N, and
// x_t = x_N s_t + qe_t. Across lanes: the pivots from tri_line's scan of 2x2
// maps, s and f from one affine scan (s is its product part, multiplier
// -a_t pn_t), the backward pair from one scan with the shared multiplier q_t.
template <int R>
__device__ inline void tri_per_line32(const double (&a)[R], const double (&b)[R], const double (&c)[R],
                                      const double (&f)[R], int lane, double (&x)[R]) {
  constexpr int N = 32 * R;
  const int ll = lane & 31, hb = lane & 32;
  const int t0 = 1 + ll * R;
  const double cl = dpp_z<kWshr1>(c[R - 1]);  // c of the row before this lane's first
  M22 L{1.0, 0.0, 0.0, 1.0};
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const double cp = r == 0 ? cl : c[r - 1];
    const double q = t0 + r == 1 ? 0.0 : -(a[r] * cp);
    L = M22{fma(b[r], L.p, q * L.r), fma(b[r], L.q, q * L.s), L.p, L.q};
  }
  L = m22_norm(L);
  L = m22_norm(m22_mul(L, m22_dpp<kShr1>(L)));
  L = m22_norm(m22_mul(L, m22_dpp<kShr2>(L)));
  L = m22_norm(m22_mul(L, m22_dpp<kShr4>(L)));
  L = m22_norm(m22_mul(L, m22_dpp<kShr8>(L)));
  L = m22_norm(m22_mul(L, m22_dpp<kBcast15, 0xA>(L)));
  const M22 E = m22_dpp<kWshr1>(L);
  // q of the row before this lane's first (unused on a line's first lane)
  double qprev = ll == 0 ? 0.0 : -(cl * (1.0 / (E.p / E.r)));
  double pn[R], qv[R], mv[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    pn[r] = t0 + r == 1 ? 1.0 / b[r] : 1.0 / fma(a[r], qprev, b[r]);
    qv[r] = -c[r] * pn[r];
    mv[r] = -a[r] * pn[r];
    qprev = qv[r];
  }
  // forward pair: y_t = m_t y_{t-1} + pn_t f_t, s_t = m_t s_{t-1} (s_0 = 1)
  double M = 1.0, Cc = 0.0;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    Cc = fma(mv[r], Cc, pn[r] * f[r]);
    M = mv[r] * M;
  }
#define PFWD(CT, ROWS)                           \
  {                                              \
    const double Mp = dpp_d<CT, ROWS>(M, 1.0);   \
    const double Cp = ROWS == 0xF ? dpp_z<CT>(Cc) : dpp_d<CT, ROWS>(Cc, 0.0); \
    Cc = fma(M, Cp, Cc);                         \
    M = M * Mp;                                  \
  }
  PFWD(kShr1, 0xF)
  PFWD(kShr2, 0xF)
  PFWD(kShr4, 0xF)
  PFWD(kShr8, 0xF)
  PFWD(kBcast15, 0xA)
#undef PFWD
  double sp = dpp_d<kWshr1>(M, 1.0), yp = dpp_z<kWshr1>(Cc);
  if (ll == 0) {  // s_0 = 1, f_0 = 0: row 1 becomes s_1 = -a_1 pn_1, f_1 = f_1 pn_1
    sp = 1.0;
    yp = 0.0;
  }
  double sv[R], yv[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    sv[r] = -(a[r] * sp) * pn[r];
    yv[r] = fma(-a[r], yp, f[r]) * pn[r];
    sp = sv[r];
    yp = yv[r];
  }
  // backward pair: X_t = q_t X_{t+1} + (s_t | f_t) for t < N, X_N = (1 | 0)
  double Nn = 1.0, Ds = 0.0, Dq = 0.0;
#pragma unroll
  for (int r = R - 1; r >= 0; --r) {
    const bool last = t0 + r == N;
    const double nj = last ? 0.0 : qv[r];
    Ds = fma(nj, Ds, last ? 1.0 : sv[r]);
    Dq = fma(nj, Dq, last ? 0.0 : yv[r]);
    Nn = nj * Nn;
  }
#define PBWD(CT)                                \
  {                                             \
    const double Np = dpp_d<CT>(Nn, 1.0);       \
    const double Sp = dpp_z<CT>(Ds);       \
    const double Qp = dpp_z<CT>(Dq);       \
    Ds = fma(Nn, Sp, Ds);                       \
    Dq = fma(Nn, Qp, Dq);                       \
    Nn = Nn * Np;                               \
  }
  PBWD(kShl1)
  PBWD(kShl2)
  PBWD(kShl4)
  PBWD(kShl8)
#undef PBWD
  {
    const double s1 = rdlane(Ds, 16), s3 = rdlane(Ds, 48), q1 = rdlane(Dq, 16), q3 = rdlane(Dq, 48);
    const int row = lane >> 4;
    const double ns = row == 0 ? s1 : row == 2 ? s3 : 0.0, nq = row == 0 ? q1 : row == 2 ? q3 : 0.0;
    Ds = fma(Nn, ns, Ds);
    Dq = fma(Nn, nq, Dq);
  }
  double xs = dpp_z<kWshl1>(Ds), xq = dpp_z<kWshl1>(Dq);
  if (ll == 31) {  // holds row N, which does not read its successor
    xs = 0.0;
    xq = 0.0;
  }
  double sf[R], qf[R];
#pragma unroll
  for (int r = R - 1; r >= 0; --r) {
    if (t0 + r == N) {
      xs = 1.0;
      xq = 0.0;
    } else {
      xs = fma(qv[r], xs, sv[r]);
      xq = fma(qv[r], xq, yv[r]);
    }
    sf[r] = xs;
    qf[r] = xq;
  }
  // x_N = (f_N - c_N qe_1 - a_N qe_{N-1}) / (c_N s_1 + a_N s_{N-1} + b_N)
  const int l1 = hb, lN = hb | 31;
  const double s1 = __shfl(sf[0], l1, 64), qe1 = __shfl(qf[0], l1, 64);
  const double sN1 = R >= 2 ? __shfl(sf[R >= 2 ? R - 2 : 0], lN, 64) : __shfl(sf[0], lN - 1, 64);
  const double qN1 = R >= 2 ? __shfl(qf[R >= 2 ? R - 2 : 0], lN, 64) : __shfl(qf[0], lN - 1, 64);
  const double aN = __shfl(a[R - 1], lN, 64), bN = __shfl(b[R - 1], lN, 64), cN = __shfl(c[R - 1], lN, 64);
  const double fN = __shfl(f[R - 1], lN, 64);
  const double xN = (fN - cN * qe1 - aN * qN1) / (cN * s1 + aN * sN1 + bN);
#pragma unroll
  for (int r = 0; r < R; ++r) x[r] = t0 + r == N ? xN : fma(xN, sf[r], qf[r]);
}

}  // namespace lw
