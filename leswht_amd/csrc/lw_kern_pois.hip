// gfx950 kernels of the pressure Poisson solve (poiss_z_ft.f90): real-to-
// half-complex z-DFT per (i,j) column, then for every wavenumber an
// x-semicoarsened zebra-line multigrid whose lines are y-tridiagonal solves
// with the precomputed LU factors bet/gam, then the inverse z-DFT.
//
// The x neighbours of row i are rows i-1 and i+1 on every level (imm, ipm of
// coefmg, poiss_z_ft.f90:733-741; x-periodic grids, which rewire them, are
// outside this path).
//
// Layout: every multigrid plane is stored [kv][i][j] with j contiguous, so a
// y-line is one contiguous row. One wavefront solves one line: each lane owns
// R consecutive rows, the forward and backward substitutions are first-order
// linear recurrences, combined across lanes with a scan of affine maps (DPP
// within 16-lane rows, then across rows from the row totals), and each lane
// then re-runs the reference recurrence on its chunk from the carry-in.
// All wavenumbers advance together; a wavenumber that met its max-norm test
// sets its flag on the device and every later kernel skips it, so each keeps
// the reference's own V-cycle count.
#include "lw_kernels.h"
#include "lw_scan.h"

namespace lw {

__device__ inline double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ inline double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ inline double2 rmul(double a, double2 b) { return make_double2(a * b.x, a * b.y); }
// a*b + c with one rounding per component: the scans' combination steps,
// which have no counterpart in the reference's serial sweep, use it to
// halve the dependent latency of each step
__device__ inline double2 cfma(double a, double2 b, double2 c) { return make_double2(fma(a, b.x, c.x), fma(a, b.y, c.y)); }
// diagonal of the wavenumber-kv operator, ac(i,j,kv) (pmat and coefmg,
// poiss_z_ft.f90:78-88, 696-703): the same additions in the same order as the
// stored table, so the same bits (ak3(1) = 0 and x - 0 = x)
__device__ inline double ac_of(double awe, double as, double an, double ak) { return -((awe + as) + an) - ak; }

// The level planes below the top (gg, resd, gd, rsu) are stored as FP64
// (double2) or, in the mixed-precision form, as FP32 (float2): values are
// rounded to FP32 when stored and widened when loaded, and every operation
// stays FP64. The top level (rhs, pc) and the LU pivots stay FP64.
__device__ inline double2 ld2(const double2 *p) { return *p; }
__device__ inline double2 ld2(const float2 *p) {
  const float2 v = *p;
  return make_double2((double)v.x, (double)v.y);
}
__device__ inline void st2(double2 *p, double2 v) { *p = v; }
__device__ inline void st2(float2 *p, double2 v) { *p = make_float2((float)v.x, (float)v.y); }

// an LDS read of a row that may not exist: it reads one that does and
// applies the condition to the value, so every read of a stage is one
// unconditional 16-byte access and they issue together (a read under a
// branch is split into two 8-byte reads, each waited for on its own)
__device__ inline double2 lds_if(bool on, const double2 *p_on, const double2 *p_off) {
  const double2 v = *(on ? p_on : p_off);
  return on ? v : make_double2(0.0, 0.0);
}

#define ROW(base, i_) ((base) + (long)(i_) * mg.LDJ)

// prefix step of the affine maps (M, C) of the forward substitution: this
// lane's map after the map of the lane D below it (identity across rows)
#define FWD_STEP(CT)                                                             \
  {                                                                              \
    const double Mp = dpp_d<CT>(M, 1.0);                                         \
    const double2 Cp = make_double2(dpp_z<CT>(Cc.x), dpp_z<CT>(Cc.y)); \
    Cc = cfma(M, Cp, Cc);                                                        \
    M = M * Mp;                                                                  \
  }
// carry from another row into the rows of mask RM: C = M*C(source) + C
#define CROSS_STEP(CT, RM)                                                                   \
  {                                                                                          \
    const double2 Cp = make_double2(dpp_d<CT, RM>(Cc.x, 0.0), dpp_d<CT, RM>(Cc.y, 0.0));   \
    Cc = cfma(M, Cp, Cc);                                                                    \
  }
// suffix step of the back substitution's maps (N, D)
#define BWD_STEP(CT)                                                           \
  {                                                                            \
    const double Np = dpp_d<CT>(Nn, 1.0);                                      \
    const double2 Dp = make_double2(dpp_z<CT>(D.x), dpp_z<CT>(D.y)); \
    D = cfma(Nn, Dp, D);                                                       \
    Nn = Nn * Np;                                                              \
  }



// ---------------------------------------------------------------------------
// z-DFT with FFTE's conventions (zfft1d.f:32-85): forward e^{-2 pi i k n/N}
// unnormalised; inverse by conjugation and 1/N. Direct sums over a column
// held in registers; the column length is a template parameter.
// ---------------------------------------------------------------------------
// Tiles of 16 x 16 (i, j) columns per workgroup. A thread holds one column:
// the real data are read along i (contiguous in the field), and the spectrum
// leaves through LDS so that it is written along j (contiguous in the
// multigrid planes), kFftKb wavenumbers at a time.
constexpr int kFftT = 16;
constexpr int kFftKb = 8;
constexpr int kFftLd = kFftT + 1;  // padded LDS row

// Radix-2 decimation-in-frequency FFT of length M held in registers:
// natural order in, bit-reversed order out. w(q) = e^{-2 pi i q / N} from the
// table (uniform index: scalar loads), M divides N, INV conjugates it. With
// every loop unrolled the twiddle indices are compile-time constants and the
// multiplications by w(0) = 1 drop out.
template <int M, int N, bool INV>
__device__ inline void fft_dif(double2 (&z)[M], const MG &mg) {
#pragma unroll
  for (int len = M; len >= 2; len >>= 1) {
    const int half = len / 2;
#pragma unroll
    for (int st = 0; st < M; st += len) {
#pragma unroll
      for (int k = 0; k < half; ++k) {
        const double2 a = z[st + k], b = z[st + k + half];
        z[st + k] = cadd(a, b);
        const double2 d = csub(a, b);
        const int q = k * (M / len) * (N / M);
        if (q == 0) {
          z[st + k + half] = d;
        } else {
          const double wr = mg.twr[q], wi = INV ? -mg.twi[q] : mg.twi[q];
          z[st + k + half] = make_double2(d.x * wr - d.y * wi, d.x * wi + d.y * wr);
        }
      }
    }
  }
}

__host__ __device__ constexpr int bitrev(int k, int M) {
  int r = 0;
  for (int b = 1; b < M; b <<= 1) r = (r << 1) | ((k & b) ? 1 : 0);
  return r;
}

// Forward z-DFT of real columns with FFTE's conventions (zfft1d.f:32-85:
// e^{-2 pi i k n/N}, unnormalised), kv = 0..N/2: the even and odd samples
// form one complex sequence of length M = N/2 whose FFT is split into the
// spectra of the two halves and recombined, X_k = E_k + w^k O_k. Tiles of
// 16 x 16 (i, j) columns per workgroup, a column per thread: the real data
// are read along i (contiguous in the field) and the spectrum leaves through
// LDS so that it is written along j (contiguous in the multigrid planes),
// kFftKb wavenumbers at a time.
template <int N>
LW_KERNEL void k_fft_fwd(LW_KP(256, 0), const Geo &g, const MG &mg, const double *div) {
  constexpr int M = N / 2, NH = N / 2 + 1;
  __shared__ double2 S[kFftKb * kFftT * kFftLd];
  const int t = threadIdx.x, ti = t & (kFftT - 1), tj = t >> 4;
  const int i0 = 1 + blockIdx.x * kFftT, j0 = 1 + blockIdx.y * kFftT;
  const int i = i0 + ti, j = j0 + tj;
  const bool in = i <= g.n1m && j <= g.n2m;
  const int oi = i0 + tj, oj = j0 + ti;  // writer of row oi, column oj
  const bool oin = oi <= g.n1m && oj <= g.n2m;
  double2 z[M];
#pragma unroll
  for (int m = 0; m < M; ++m)
    z[m] = in ? make_double2(div[P(g, i, j, 2 * m + 1)], div[P(g, i, j, 2 * m + 2)]) : make_double2(0.0, 0.0);
  fft_dif<M, N, false>(z, mg);
  // X_k for k = 0..M from Z_k and Z_{M-k} (Z_M = Z_0), in place by pairs
  // (k, M-k): X_k goes where Z_k was (slot bitrev(k)), X_M to its own slot
  double2 XM;
  {
    const double2 z0 = z[0];
    z[0] = make_double2(z0.x + z0.y, 0.0);
    XM = make_double2(z0.x - z0.y, 0.0);
  }
#pragma unroll
  for (int k = 1; 2 * k <= M; ++k) {
    const double2 zk = z[bitrev(k, M)], zc = z[bitrev(M - k, M)];
    {
      const double2 e = make_double2(0.5 * (zk.x + zc.x), 0.5 * (zk.y - zc.y));
      const double2 o = make_double2(0.5 * (zk.y + zc.y), -0.5 * (zk.x - zc.x));
      const double wr = mg.twr[k], wi = mg.twi[k];
      z[bitrev(k, M)] = make_double2(e.x + (o.x * wr - o.y * wi), e.y + (o.x * wi + o.y * wr));
    }
    if (2 * k < M) {
      const double2 e = make_double2(0.5 * (zc.x + zk.x), 0.5 * (zc.y - zk.y));
      const double2 o = make_double2(0.5 * (zc.y + zk.y), -0.5 * (zc.x - zk.x));
      const double wr = mg.twr[M - k], wi = mg.twi[M - k];
      z[bitrev(M - k, M)] = make_double2(e.x + (o.x * wr - o.y * wi), e.y + (o.x * wi + o.y * wr));
    }
  }
  auto spec = [&](int k) -> double2 { return k == M ? XM : z[bitrev(k < M ? k : 0, M)]; };
#pragma unroll
  for (int kb = 0; kb < NH; kb += kFftKb) {
#pragma unroll
    for (int u = 0; u < kFftKb; ++u)
      if (kb + u < NH) S[(u * kFftT + tj) * kFftLd + ti] = spec(kb + u);
    __syncthreads();
#pragma unroll
    for (int u = 0; u < kFftKb; ++u) {
      const int kv = kb + u;
      if (kv < NH && oin) mg.rhs[(long)kv * mg.PCPLANE + (long)oi * mg.LDJ + oj] = S[(u * kFftT + ti) * kFftLd + tj];
    }
    __syncthreads();
  }
}

// Inverse z-DFT (conjugate twiddles, 1/N): from X_0..X_M the spectra of the
// even and odd samples, E_k = (X_k + conj X_{M-k})/2 and
// O_k = (X_k - conj X_{M-k}) conj(w^k)/2, give Z_k = E_k + i O_k, whose
// inverse FFT of length M (scaled 1/M) holds x_{2m} + i x_{2m+1}.
template <int N>
LW_KERNEL void k_fft_inv(LW_KP(256, 0), const Geo &g, const MG &mg, double *phi) {
  constexpr int M = N / 2, NH = N / 2 + 1;
  __shared__ double2 S[kFftKb * kFftT * kFftLd];
  const int t = threadIdx.x, ti = t & (kFftT - 1), tj = t >> 4;
  const int i0 = 1 + blockIdx.x * kFftT, j0 = 1 + blockIdx.y * kFftT;
  const int i = i0 + ti, j = j0 + tj;
  const bool in = i <= g.n1m && j <= g.n2m;
  const int oi = i0 + tj, oj = j0 + ti;  // reader of row oi, column oj
  const bool oin = oi <= g.n1m && oj <= g.n2m;
  // X_0..X_{M-1} in z, X_M apart; then Z_k in place by pairs (k, M-k)
  double2 z[M], XM = make_double2(0.0, 0.0);
#pragma unroll
  for (int kb = 0; kb < NH; kb += kFftKb) {
#pragma unroll
    for (int u = 0; u < kFftKb; ++u) {
      const int kv = kb + u;
      if (kv < NH)
        S[(u * kFftT + ti) * kFftLd + tj] =
            oin ? (mg.pcsel[kv] ? mg.pc2 : mg.pc)[(long)kv * mg.PCPLANE + (long)oi * mg.LDJ + oj]
                : make_double2(0.0, 0.0);
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < kFftKb; ++u) {
      const int kv = kb + u;
      if (kv < M) z[kv] = S[(u * kFftT + tj) * kFftLd + ti];
      else if (kv == M) XM = S[(u * kFftT + tj) * kFftLd + ti];
    }
    __syncthreads();
  }
  if (!in) return;
  auto zk_of = [&](const double2 xk, const double2 xc, int k) -> double2 {
    const double2 e = make_double2(0.5 * (xk.x + xc.x), 0.5 * (xk.y - xc.y));
    const double2 d = make_double2(0.5 * (xk.x - xc.x), 0.5 * (xk.y + xc.y));
    double2 o = d;
    if (k != 0) {
      const double wr = mg.twr[k], wi = -mg.twi[k];
      o = make_double2(d.x * wr - d.y * wi, d.x * wi + d.y * wr);
    }
    return make_double2(e.x - o.y, e.y + o.x);
  };
  z[0] = zk_of(z[0], XM, 0);
#pragma unroll
  for (int k = 1; 2 * k <= M; ++k) {
    const double2 xk = z[k], xc = z[M - k];
    z[k] = zk_of(xk, xc, k);
    if (2 * k < M) z[M - k] = zk_of(xc, xk, M - k);
  }
  fft_dif<M, N, true>(z, mg);
  const double dm = 1.0 / (double)M;
#pragma unroll
  for (int m = 0; m < M; ++m) {
    const double2 v = z[bitrev(m, M)];
    phi[P(g, i, j, 2 * m + 1)] = v.x * dm;
    phi[P(g, i, j, 2 * m + 2)] = v.y * dm;
  }
}

// runtime-length fallback (any n3m): re-reads the column
LW_KERNEL void k_fft_fwd_any(LW_KP(1024, 0), const Geo &g, const MG &mg, const double *div, int N) {
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

LW_KERNEL void k_fft_inv_any(LW_KP(1024, 0), const Geo &g, const MG &mg, double *phi, int N) {
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
      const double2 X = (mg.pcsel[kk] ? mg.pc2 : mg.pc)[(long)kk * mg.PCPLANE + (long)i * mg.LDJ + j];
      if (kv <= N / 2) s += X.x * c - X.y * sn;
      else s += X.x * c + X.y * sn;
    }
    phi[P(g, i, j, n + 1)] = s * dn;
  }
}

// algorithmic bytes of the z-DFTs: the real column in, the half spectrum out
static double fft_bytes(const Geo &g, const MG &mg) {
  return (double)g.n1m * g.n2m * (8.0 * g.n3m + 16.0 * mg.n3mh);
}

void launch_fft_forward(hipStream_t s, const Geo &g, const MG &mg, const Fld &f) {
  const dim3 gr((g.n1m + 63) / 64, g.n2m), bl(64);
  const dim3 gt((g.n1m + kFftT - 1) / kFftT, (g.n2m + kFftT - 1) / kFftT), bt(256);
  ::lw::KFn<Geo, MG, const double *> k;
  switch (g.n3m) {
    case 2: k = LW_KFN(k_fft_fwd<2>); break;
    case 4: k = LW_KFN(k_fft_fwd<4>); break;
    case 8: k = LW_KFN(k_fft_fwd<8>); break;
    case 16: k = LW_KFN(k_fft_fwd<16>); break;
    case 32: k = LW_KFN(k_fft_fwd<32>); break;
    case 64: k = LW_KFN(k_fft_fwd<64>); break;
    default: break;
  }
  if (k) LW_LAUNCH_K("k_fft_fwd", fft_bytes(g, mg), 0, k, gt, bt, 0, s, g, mg, f.div);
  else LW_LAUNCH("k_fft_fwd_any", fft_bytes(g, mg), 0, k_fft_fwd_any, gr, bl, 0, s, g, mg, f.div, g.n3m);
}

void launch_fft_inverse(hipStream_t s, const Geo &g, const MG &mg, const Fld &f) {
  const dim3 gr((g.n1m + 63) / 64, g.n2m), bl(64);
  const dim3 gt((g.n1m + kFftT - 1) / kFftT, (g.n2m + kFftT - 1) / kFftT), bt(256);
  ::lw::KFn<Geo, MG, double *> k;
  switch (g.n3m) {
    case 2: k = LW_KFN(k_fft_inv<2>); break;
    case 4: k = LW_KFN(k_fft_inv<4>); break;
    case 8: k = LW_KFN(k_fft_inv<8>); break;
    case 16: k = LW_KFN(k_fft_inv<16>); break;
    case 32: k = LW_KFN(k_fft_inv<32>); break;
    case 64: k = LW_KFN(k_fft_inv<64>); break;
    default: break;
  }
  if (k) LW_LAUNCH_K("k_fft_inv", fft_bytes(g, mg), 0, k, gt, bt, 0, s, g, mg, f.phi);
  else LW_LAUNCH("k_fft_inv_any", fft_bytes(g, mg), 0, k_fft_inv_any, gr, bl, 0, s, g, mg, f.phi, g.n3m);
}

// ---------------------------------------------------------------------------
// multigrid (poiss_z_ft.f90:258-511)
//
// Launches per V-cycle: on each level going down the two zebra halves and one
// residual+restriction kernel; NBLI pairs of halves on the coarsest level;
// going up the prolongation (goup) is folded into the first half of the finer
// level, which is the only reader of the prolonged rows before the second
// half overwrites them; the top level ends with one kernel that forms the
// residual, restricts it, applies the kv = 1 gauge shift and, in its last
// workgroup, runs the convergence test of every wavenumber.
// ---------------------------------------------------------------------------
LW_KERNEL void k_mg_reset(LW_KP(1024, 0), const MG &mg, long n) {
  const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (q < n) mg.pc[q] = make_double2(0.0, 0.0);
  if (q < mg.n3mh) {
    mg.done[q] = 0;
    mg.cyc[q] = 0;
    mg.pcsel[q] = 0;
    mg.sumres[q] = 0ull;
  }
}

void launch_mg_zero(hipStream_t s, const Geo &, const MG &mg) {
  const long n = (long)mg.n3mh * mg.PCPLANE;
  LW_LAUNCH("k_mg_reset", 16.0 * n, 0, k_mg_reset, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, mg, n);
}

// goup correction of a destination row as the reference forms it:
// (dst + coi1*g(ic)) + coi2*g(ic+1), or dst + coi1*g(ic) on the last row
struct Prolong {
  const double2 *g0, *g1;
  double c1, c2;
  bool on, last;
  __device__ inline double2 apply(double2 v, int j) const {
    if (!on) return v;
    return last ? cadd(v, rmul(c1, g0[j])) : cadd(cadd(v, rmul(c1, g0[j])), rmul(c2, g1[j]));
  }
};

__device__ inline Prolong make_prolong(const MG &mg, const double2 *gg, int ic, bool on, int cend) {
  Prolong p;
  p.on = on;
  p.last = false;
  p.g0 = p.g1 = nullptr;
  p.c1 = p.c2 = 0.0;
  if (on) {
    p.g0 = ROW(gg, ic);
    p.g1 = ROW(gg, ic + 1);
    p.c1 = mg.coi1[ic];
    p.c2 = mg.coi2[ic];
    p.last = ic == cend;
  }
  return p;
}

// forward and back substitution of one line whose right-hand side and LU
// factors sit in registers: lane `lane` owns rows j0..j0+R-1 (j0 = 1 + lane*R);
// rows past n2m carry zero factors and return x = 0
template <int R, bool FULL>
__device__ inline void tri_regs(const double2 (&rr)[R], const double (&bt)[R], const double (&asj)[R],
                                const double (&gm)[R], int j0, int n2m, int lane, double2 (&x)[R]) {
  // forward substitution y_j = (rr_j - as_j y_{j-1}) bet_j as affine maps
  // y -> M y + C composed over this lane's rows
  double M = 1.0;
  double2 Cc = make_double2(0.0, 0.0);
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int j = j0 + r;
    if ((FULL || j <= n2m)) {
      const double m = j == 1 ? 0.0 : -(asj[r] * bt[r]);
      Cc = cfma(m, Cc, rmul(bt[r], rr[r]));
      M = m * M;
    }
  }
  // wavefront scan: within 16-lane rows by DPP, then across the rows one
  // row at a time: row 1 takes the total of row 0 (row_bcast:15 into row 1
  // only), row 2 that of rows 0-1 (row_bcast:31 into row 2), row 3 that of
  // rows 0-2 (row_bcast:15 into row 3). Rows outside a step's mask read the
  // identity 0. yp = the inclusive value of the lane below (wave_shr:1).
  FWD_STEP(kShr1)
  FWD_STEP(kShr2)
  FWD_STEP(kShr4)
  FWD_STEP(kShr8)
  CROSS_STEP(kBcast15, 0x2)
  CROSS_STEP(kBcast31, 0x4)
  CROSS_STEP(kBcast15, 0x8)
  double2 yp = make_double2(dpp_z<kWshr1>(Cc.x), dpp_z<kWshr1>(Cc.y));
  double2 y[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int j = j0 + r;
    if ((FULL || j <= n2m)) {
      yp = j == 1 ? rmul(bt[r], rr[r]) : rmul(bt[r], cfma(-asj[r], yp, rr[r]));
      y[r] = yp;
    } else {
      y[r] = make_double2(0.0, 0.0);
    }
  }
  // backward substitution x_j = y_j - gam_{j+1} x_{j+1}, x_n2m = y_n2m
  double Nn = 1.0;
  double2 D = make_double2(0.0, 0.0);
#pragma unroll
  for (int r = R - 1; r >= 0; --r) {
    const int j = j0 + r;
    if ((FULL || j <= n2m)) {
      const double nj = j == n2m ? 0.0 : -gm[r];
      D = cfma(nj, D, y[r]);
      Nn = nj * Nn;
    }
  }
  BWD_STEP(kShl1)
  BWD_STEP(kShl2)
  BWD_STEP(kShl4)
  BWD_STEP(kShl8)
  double2 xn;
  {
    const double n1 = rdlane(Nn, 16), n2 = rdlane(Nn, 32);
    const double2 d1 = make_double2(rdlane(D.x, 16), rdlane(D.y, 16));
    const double2 d2 = make_double2(rdlane(D.x, 32), rdlane(D.y, 32));
    const double2 d3 = make_double2(rdlane(D.x, 48), rdlane(D.y, 48));
    const double2 s2 = cfma(n2, d3, d2);  // start of row 2
    const double2 s1 = cfma(n1, s2, d1);  // start of row 1
    const int row = lane >> 4;
    const double2 next = row == 3 ? make_double2(0.0, 0.0) : row == 2 ? d3 : row == 1 ? s2 : s1;
    D = cfma(Nn, next, D);  // inclusive suffix over the wavefront
    xn = make_double2(dpp_z<kWshl1>(D.x), dpp_z<kWshl1>(D.y));
  }
#pragma unroll
  for (int r = R - 1; r >= 0; --r) {
    const int j = j0 + r;
    if ((FULL || j <= n2m)) {
      xn = j == n2m ? y[r] : cfma(-gm[r], xn, y[r]);
      x[r] = xn;
    } else {
      x[r] = make_double2(0.0, 0.0);
    }
  }
}

// ---------------------------------------------------------------------------
// one zebra line: y-tridiagonal solve with the LU factors of line i
// (trdiag1m, poiss_z_ft.f90:383-408). MODE selects the right-hand side:
// 0 relax first half        resd - oldv*(aw*g(im) + ae*g(ip))
// 1 relax second half       resd - aw*g(im) - ae*g(ip)
// 2 toplevel first half     rhs - oldv*(aw*pc(im) + ae*pc(ip))
// 3 toplevel second half    rhs - aw*(g(im)-tt) - ae*(g(ip)-tt),
//   tt = g(1,n2m-1) for kv = 1 (the gauge of the singular mode), 0 otherwise
// 4 as 0, the neighbours prolonged from level clev first (goup)
// 5 as 2, the neighbours prolonged from level clev first (toplevel(1))
// The neighbour rows of line t that receive a prolongation are the
// destinations of coarse rows t-1 (when t > 0) and t of level clev.
// ---------------------------------------------------------------------------
template <int R, int MODE, bool FULL>
__device__ inline void line_solve(const Geo &g, const MG &mg, int kv, int ibase, int t, double oldv, int clev,
                                  int lane) {
  const int n2m = FULL ? 64 * R : g.n2m;
  const int i = ibase + 2 * t;
  const long pl = (long)(kv - 1) * mg.PLANE, ppl = (long)(kv - 1) * mg.PCPLANE;
  double2 *gg = mg.gg + pl;
  const double *bet = ROW(mg.bet + pl, i);
  const int im = i - 1, ip = i + 1;
  const double aw = mg.aw[i], ae = mg.ae[i];
  constexpr bool TOP = MODE == 2 || MODE == 3 || MODE == 5;
  constexpr bool PCN = MODE == 2 || MODE == 5;
  constexpr bool PRO = MODE == 4 || MODE == 5;
  const double2 *src = TOP ? ROW(mg.rhs + ppl, i) : ROW(mg.resd + pl, i);
  const double2 *nbm = PCN ? ROW(mg.pc + ppl, im) : ROW(gg, im);
  const double2 *nbp = PCN ? ROW(mg.pc + ppl, ip) : ROW(gg, ip);
  Prolong pm, pp;
  if (PRO) {
    const int cb = mg.iimg[clev][1], ce = mg.iimg[clev][2];
    pm = make_prolong(mg, gg, cb + t - 1, t > 0, ce);
    pp = make_prolong(mg, gg, cb + t, true, ce);
  }
  double2 tt = make_double2(0.0, 0.0);
  if (MODE == 3 && kv == 1) {
    tt = ROW(gg, 1)[n2m - 1];
    if (t == 0 && lane == 0) {
      double *ts = (double *)(mg.sumres + mg.n3mh);
      ts[0] = tt.x;
      ts[1] = tt.y;
    }
  }
  const int j0 = 1 + lane * R;
  // every load is issued unconditionally (rows past n2m read row n2m), so a
  // line waits for memory once; values past n2m are masked out afterwards
  double2 rr[R];
  double bt[R], asj[R], gm[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int j = j0 + r;
    const int jc = (FULL || j <= n2m) ? j : n2m;
    const double2 s0 = src[jc];
    double2 v;
    if (MODE == 0 || MODE == 2) {
      v = csub(s0, rmul(oldv, cadd(rmul(aw, nbm[jc]), rmul(ae, nbp[jc]))));
    } else if (PRO) {
      const double2 vm = pm.apply(nbm[jc], jc), vp = pp.apply(nbp[jc], jc);
      v = csub(s0, rmul(oldv, cadd(rmul(aw, vm), rmul(ae, vp))));
    } else if (MODE == 1) {
      v = csub(csub(s0, rmul(aw, nbm[jc])), rmul(ae, nbp[jc]));
    } else {
      const double2 left = kv == 1 ? csub(nbm[jc], tt) : nbm[jc];
      v = csub(csub(s0, rmul(aw, left)), rmul(ae, csub(nbp[jc], tt)));
    }
    const double b = bet[jc];
    const double a = mg.as[jc];
    const double gnext = mg.an[jc] * b;  // gam(j+1)
    // masks as factors (exact for finite values): a select would let the
    // compiler move the loads into a branch per row again
    const double okf = (FULL || j <= n2m) ? 1.0 : 0.0, okg = j < n2m ? 1.0 : 0.0;
    rr[r] = rmul(okf, v);
    bt[r] = okf * b;
    asj[r] = a;
    gm[r] = okg * gnext;
  }
  double2 x[R];
  tri_regs<R, FULL>(rr, bt, asj, gm, j0, n2m, lane, x);
  double2 *out = ROW(gg, i);
#pragma unroll
  for (int r = 0; r < R; ++r)
    if ((FULL || j0 + r <= n2m)) out[j0 + r] = x[r];
}

// one wavefront per line, four lines per workgroup, all wavenumbers
template <int R, int MODE, bool FULL>
LW_KERNEL void k_line(LW_KP(256, 0), const Geo &g, const MG &mg, int ibase, int nlines, double oldv, int clev) {
  const int t = blockIdx.x * 4 + threadIdx.y;
  const int kv = mg.kv0 + blockIdx.y;
  if (mg.done[kv - 1]) return;
  if (t >= nlines) return;
  line_solve<R, MODE, FULL>(g, mg, kv, ibase, t, oldv, clev, threadIdx.x);
}

typedef ::lw::KFn<Geo, MG, int, int, double, int> LineKernel;

template <int MODE, bool F>
static LineKernel pick_line_rf(int R) {
  switch (R) {
    case 1: return LW_KFN((k_line<1, MODE, F>));
    case 2: return LW_KFN((k_line<2, MODE, F>));
    case 4: return LW_KFN((k_line<4, MODE, F>));
    case 8: return LW_KFN((k_line<8, MODE, F>));
    default: return LW_KFN((k_line<16, MODE, F>));
  }
}

// FULL: n2m == 64*R, every lane's rows exist and the row guards vanish
template <int MODE>
static LineKernel pick_line(int R, bool full) {
  return full ? pick_line_rf<MODE, true>(R) : pick_line_rf<MODE, false>(R);
}

static inline int lanes_r(int n2m) {
  int r = 1;
  while (r * 64 < n2m) r <<= 1;
  return r;
}

// every lane's R rows exist (n2m == 64*R): the kernels drop the row guards
static inline bool full_rows(const Geo &g) { return g.n2m == 64 * lanes_r(g.n2m); }

static void launch_lines(hipStream_t s, const Geo &g, const MG &mg, int mode, int ibase, int nlines, double oldv,
                         int clev) {
  if (nlines <= 0) return;
  const int R = lanes_r(g.n2m);
  const bool full = full_rows(g);
  LineKernel k;
  switch (mode) {
    case 0: k = pick_line<0>(R, full); break;
    case 1: k = pick_line<1>(R, full); break;
    case 2: k = pick_line<2>(R, full); break;
    case 3: k = pick_line<3>(R, full); break;
    case 4: k = pick_line<4>(R, full); break;
    default: k = pick_line<5>(R, full); break;
  }
  static const char *names[6] = {"k_line_relax1", "k_line_relax2", "k_line_top1",
                                 "k_line_top2",   "k_line_relax1_up", "k_line_top1_up"};
  // per row: own right-hand side 16, one neighbour row 16 (each shared by two
  // lines), bet 8, solution 16; the prolonging modes also read one coarse row
  // per line (16)
  const double per_row = mode >= 4 ? 72.0 : 56.0;
  LW_LAUNCH_K(names[mode], 0, per_row * g.n2m * nlines, k, dim3((unsigned)((nlines + 3) / 4), (unsigned)mg.nkv),
            dim3(64, 4), 0, s, g, mg, ibase, nlines, oldv, clev);
}

// residual of row i from the rows around it (poiss_z_ft.f90:360-369, 438-449)
#define RESID_ROW(r0_, gi_, gm_, gp_, awe_, ak_, j_)                          \
  ({                                                                         \
    double2 r_ = (r0_);                                                      \
    r_ = csub(r_, rmul(aw, (gm_)));                                          \
    r_ = csub(r_, rmul(ae, (gp_)));                                          \
    r_ = csub(r_, rmul(mg.as[j_], gi_[(j_) - 1]));                           \
    r_ = csub(r_, rmul(mg.an[j_], gi_[(j_) + 1]));                           \
    r_ = csub(r_, rmul(ac_of(awe_, mg.as[j_], mg.an[j_], ak_), gi_[j_]));    \
    r_;                                                                      \
  })

// going down (godown, poiss_z_ft.f90:438-458): residual of the fine rows
// i0 = start + 2t and i0 + 1 in place, then the coarse row t from them
__device__ inline void lev_rr(const Geo &g, const MG &mg, int kv, int ilev, int t, int jfirst, int jstep) {
  const long pl = (long)(kv - 1) * mg.PLANE;
  const double2 *gg = mg.gg + pl;
  double2 *rsb = mg.resd + pl;
  const int ia = mg.iimg[ilev][1] + 2 * t, ib = ia + 1;
  const int ic = mg.iimg[ilev - 1][1] + t;
  const double2 *ga = ROW(gg, ia), *gb = ROW(gg, ib), *gam_ = ROW(gg, ia - 1), *gbp = ROW(gg, ib + 1);
  const double awea = mg.awe[ia], aweb = mg.awe[ib], ak = mg.ak3[kv];
  double2 *ra = ROW(rsb, ia), *rb = ROW(rsb, ib), *rc = ROW(rsb, ic);
  const double c1 = mg.cor1[ic], c2 = mg.cor2[ic];
  const double awa = mg.aw[ia], aea = mg.ae[ia], awb = mg.aw[ib], aeb = mg.ae[ib];
  for (int j = jfirst; j <= g.n2m; j += jstep) {
    double aw = awa, ae = aea;
    const double2 x = RESID_ROW(ra[j], ga, gam_[j], gb[j], awea, ak, j);
    aw = awb;
    ae = aeb;
    const double2 y = RESID_ROW(rb[j], gb, ga[j], gbp[j], aweb, ak, j);
    ra[j] = x;
    rb[j] = y;
    rc[j] = cadd(rmul(c1, x), rmul(c2, y));
  }
}

LW_KERNEL void k_lev_rr(LW_KP(256, 0), const Geo &g, const MG &mg, int ilev) {
  const int kv = mg.kv0 + blockIdx.y;
  if (mg.done[kv - 1]) return;
  lev_rr(g, mg, kv, ilev, blockIdx.x, 1 + threadIdx.x, 256);
}

// ---------------------------------------------------------------------------
// the coarse levels 0..lc of a V-cycle in one launch: one workgroup of
// kCoarseWaves wavefronts per wavenumber runs, in the reference's order, the
// down-sweep from level lc (whose right-hand side the previous launch
// restricted), the NBLI sweeps of level 0 and the up-sweep back to level lc.
// Each stage ends at a workgroup barrier; the levels are small enough that
// the barrier is far cheaper than a kernel boundary.
// ---------------------------------------------------------------------------
constexpr int kCoarseWaves = 8;

template <int R, int MODE, bool FULL>
__device__ inline void wg_sweep(const Geo &g, const MG &mg, int kv, int ibase, int nlines, double oldv, int clev) {
  for (int t = threadIdx.y; t < nlines; t += kCoarseWaves)
    line_solve<R, MODE, FULL>(g, mg, kv, ibase, t, oldv, clev, threadIdx.x);
  __syncthreads();
}

template <int R, bool FULL>
LW_KERNEL void k_mg_coarse(LW_KP(64 * kCoarseWaves, 0), const Geo &g, const MG &mg, int lc) {
  const int kv = mg.kv0 + blockIdx.x;
  if (mg.done[kv - 1]) return;
  if (threadIdx.x == 0 && threadIdx.y == 0) mg.sumres[kv - 1] = 0ull;  // for the next toplevel's max norm
  const int tid = threadIdx.y * 64 + threadIdx.x;
  for (int ilev = lc; ilev >= 1; --ilev) {
    const int n = mg.iimg[ilev][2] - mg.iimg[ilev][1] + 1, b = mg.iimg[ilev][1];
    wg_sweep<R, 0, FULL>(g, mg, kv, b, (n + 1) / 2, 0.0, 0);
    wg_sweep<R, 1, FULL>(g, mg, kv, b + 1, n / 2, 1.0, 0);
    for (int q = tid; q < (n / 2) * g.n2m; q += 64 * kCoarseWaves)
      lev_rr(g, mg, kv, ilev, q / g.n2m, 1 + q % g.n2m, g.n2m);
    __syncthreads();
  }
  {
    const int n = mg.iimg[0][2] - mg.iimg[0][1] + 1, b = mg.iimg[0][1];
    for (int it = 0; it < mg.nbli; ++it) {
      wg_sweep<R, 0, FULL>(g, mg, kv, b, (n + 1) / 2, it == 0 ? 0.0 : 1.0, 0);
      wg_sweep<R, 1, FULL>(g, mg, kv, b + 1, n / 2, 1.0, 0);
    }
  }
  for (int ilev = 0; ilev < lc; ++ilev) {
    const int n = mg.iimg[ilev + 1][2] - mg.iimg[ilev + 1][1] + 1, b = mg.iimg[ilev + 1][1];
    wg_sweep<R, 4, FULL>(g, mg, kv, b, (n + 1) / 2, 1.0, ilev);
    wg_sweep<R, 1, FULL>(g, mg, kv, b + 1, n / 2, 1.0, 0);
  }
}

// ---------------------------------------------------------------------------
// levels 1 and 0 of a V-cycle with the solution planes of both levels in
// LDS. Wavefront w owns the level-1 lines 2w, 2w+1 and the level-0 line w
// (level 1 has twice as many lines as level 0): their right-hand sides and
// residuals stay in its registers from the first stage to the last, and only
// the solutions, which the neighbouring lines read, pass through LDS. The
// stages are those of k_mg_coarse with lc = 1, in the reference's order:
//   relax(1, oldv 0): the first half has no neighbour term
//   godown(1): residual of both lines, restriction into line w of level 0
//   relax(0, NBLI sweeps, oldv 0 on the first)
//   goup(0) folded into relax(1, oldv 1), as in k_line mode 4
// Only the new level-1 solution leaves the kernel (the next finer level's
// prolongation reads it); every other value is rebuilt by the next cycle.
// LDS slot of row j0 + r of local line q for lane l: (q*R + r)*64 + l, so a
// wavefront reads or writes one row index of a line as one contiguous block.
// ---------------------------------------------------------------------------
constexpr int kCoarseLdsWaves = 12;
constexpr int kCoarseLdsBytes = 160 * 1024;

template <int R>
struct LineLU {
  double bt[R], gm[R];
};

template <int R, bool FULL>
__device__ inline void load_lu(const MG &mg, long pl, int i, int j0, int n2m, LineLU<R> &c) {
  const double *bet = ROW(mg.bet + pl, i);
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int j = j0 + r;
    const int jc = (FULL || j <= n2m) ? j : n2m;
    const double okf = (FULL || j <= n2m) ? 1.0 : 0.0, okg = j < n2m ? 1.0 : 0.0;
    const double b = bet[jc];
    c.bt[r] = okf * b;
    c.gm[r] = okg * (mg.an[jc] * b);  // gam(j+1)
  }
}

template <int R, bool FULL, typename S>
__device__ __forceinline__ void mg_coarse_lds_body(const Geo &g, const MG &mg, int kvi) {
  extern __shared__ double2 lds[];
  const int kv = mg.kv0 + kvi;
  if (mg.done[kv - 1]) return;
  if (threadIdx.x == 0 && threadIdx.y == 0) mg.sumres[kv - 1] = 0ull;  // for the next toplevel's max norm
  // wavefront y takes level-0 line w (and level-1 lines 2w, 2w+1): the even
  // lines on wavefronts 0.., the odd ones after them, so that the level-0
  // half-sweeps, which run on every other line, spread their wavefronts
  // over the four SIMDs instead of two (consecutive wavefronts of a
  // workgroup go to consecutive SIMDs)
  const int half = (blockDim.y + 1) / 2;
  const int wy = __builtin_amdgcn_readfirstlane(threadIdx.y);  // uniform: scalar loads of line coefficients
  const int lane = threadIdx.x, w = wy < half ? 2 * wy : 2 * (wy - half) + 1;
  const int n2m = FULL ? 64 * R : g.n2m;
  const int b1 = mg.iimg[1][1], n1 = mg.iimg[1][2] - b1 + 1;
  const int b0 = mg.iimg[0][1], n0 = mg.iimg[0][2] - b0 + 1;
  double2 *L1 = lds, *L0 = lds + (long)n1 * R * 64;
  auto slot = [&](int q, int r) { return (q * R + r) * 64 + lane; };
  const long pl = (long)(kv - 1) * mg.PLANE;
  S *gg = (S *)mg.gg + pl;
  S *resd = (S *)mg.resd + pl;
  const int la = 2 * w, lb = la + 1;  // local level-1 lines
  const int ia = b1 + la, ib = ia + 1, ic = b0 + w;
  const int j0 = 1 + lane * R;
  const double2 z2 = make_double2(0.0, 0.0);
  // every stage's line coefficients requested up front (uniform: scalar loads)
  const double awA = mg.aw[ia], aeA = mg.ae[ia], weA = mg.awe[ia];
  const double awB = mg.aw[ib], aeB = mg.ae[ib], weB = mg.awe[ib];
  const double aw0 = mg.aw[ic], ae0 = mg.ae[ic], cr1 = mg.cor1[ic], cr2 = mg.cor2[ic], akv = mg.ak3[kv];
  const double ci1m = w > 0 ? mg.coi1[ic - 1] : 0.0, ci2m = w > 0 ? mg.coi2[ic - 1] : 0.0;
  const double ci1 = mg.coi1[ic], ci2 = mg.coi2[ic];

  // registers held through the kernel: as, the level-0 line's right-hand
  // side and LU factors. The level-1 residuals go to resd, as godown leaves
  // them, and come back with the level-1 factors for the up-sweep (the
  // register file holds twelve wavefronts)
  double asj[R];
  double2 resA[R], resB[R], res0[R];
  LineLU<R> cA, cB, c0;
  {
    const S *ra = ROW(resd, ia), *rb = ROW(resd, ib);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int j = j0 + r;
      const int jc = (FULL || j <= n2m) ? j : n2m;
      const double okf = (FULL || j <= n2m) ? 1.0 : 0.0;
      asj[r] = mg.as[jc];
      resA[r] = rmul(okf, ld2(ra + jc));
      resB[r] = rmul(okf, ld2(rb + jc));
    }
    load_lu<R, FULL>(mg, pl, ia, j0, n2m, cA);
    load_lu<R, FULL>(mg, pl, ib, j0, n2m, cB);
    load_lu<R, FULL>(mg, pl, ic, j0, n2m, c0);
  }
  double2 rr[R], x[R];

  // relax(1), first half, oldv = 0: the right-hand side is resd itself
#pragma unroll
  for (int r = 0; r < R; ++r) rr[r] = resA[r];
  tri_regs<R, FULL>(rr, cA.bt, asj, cA.gm, j0, n2m, lane, x);
#pragma unroll
  for (int r = 0; r < R; ++r) L1[slot(la, r)] = x[r];
  __syncthreads();

  // relax(1), second half: resd - aw*g(i-1) - ae*g(i+1)
  {
    const double aw = awB, ae = aeB;
    const bool hp = lb + 1 < n1;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const double2 gmv = L1[slot(la, r)];
      const double2 gpv = lds_if(hp, &L1[slot(lb + 1, r)], &L1[slot(lb, r)]);
      rr[r] = rmul((FULL || j0 + r <= n2m) ? 1.0 : 0.0, csub(csub(resB[r], rmul(aw, gmv)), rmul(ae, gpv)));
    }
    tri_regs<R, FULL>(rr, cB.bt, asj, cB.gm, j0, n2m, lane, x);
#pragma unroll
    for (int r = 0; r < R; ++r) L1[slot(lb, r)] = x[r];
  }
  __syncthreads();

  // godown(1): residuals of lines 2w and 2w+1, written to resd as godown
  // leaves them, and restricted into line w of level 0 (res0). One line at a
  // time: the scheduling barrier keeps the second line's loads from being
  // hoisted over the first line's arithmetic.
  {
    const double c1 = cr1, c2 = cr2, ak = akv;
    double anj[R];
#pragma unroll
    for (int r = 0; r < R; ++r) anj[r] = mg.an[(FULL || j0 + r <= n2m) ? j0 + r : n2m];
    {
      const double aw = awA, ae = aeA, awe = weA;
      S *out = ROW(resd, ia);
      const double2 lo = lds_if(lane > 0, &L1[(la * R + R - 1) * 64 + lane - 1], &L1[(la * R + R - 1) * 64 + lane]);
      const double2 hi = lds_if(lane < 63, &L1[(la * R) * 64 + lane + 1], &L1[(la * R) * 64 + lane]);
      double2 gi[R];
#pragma unroll
      for (int r = 0; r < R; ++r) gi[r] = L1[slot(la, r)];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const double2 gmv = lds_if(la > 0, &L1[slot(la - 1, r)], &L1[slot(la, r)]);
        const double2 gpv = L1[slot(lb, r)];
        const double2 dn = r > 0 ? gi[r > 0 ? r - 1 : 0] : lo, up = r < R - 1 ? gi[r < R - 1 ? r + 1 : 0] : hi;
        double2 u = resA[r];
        u = csub(u, rmul(aw, gmv));
        u = csub(u, rmul(ae, gpv));
        u = csub(u, rmul(asj[r], dn));
        u = csub(u, rmul(anj[r], up));
        u = csub(u, rmul(ac_of(awe, asj[r], anj[r], ak), gi[r]));
        resA[r] = u;
        if ((FULL || j0 + r <= n2m)) st2(out + j0 + r, u);
      }
    }
    __builtin_amdgcn_sched_barrier(0);
    {
      const double aw = awB, ae = aeB, awe = weB;
      S *out = ROW(resd, ib);
      const bool hp = lb + 1 < n1;
      const double2 lo = lds_if(lane > 0, &L1[(lb * R + R - 1) * 64 + lane - 1], &L1[(lb * R + R - 1) * 64 + lane]);
      const double2 hi = lds_if(lane < 63, &L1[(lb * R) * 64 + lane + 1], &L1[(lb * R) * 64 + lane]);
      double2 gi[R];
#pragma unroll
      for (int r = 0; r < R; ++r) gi[r] = L1[slot(lb, r)];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const double2 gmv = L1[slot(la, r)];
        const double2 gpv = lds_if(hp, &L1[slot(lb + 1, r)], &L1[slot(lb, r)]);
        const double2 dn = r > 0 ? gi[r > 0 ? r - 1 : 0] : lo, up = r < R - 1 ? gi[r < R - 1 ? r + 1 : 0] : hi;
        double2 v = resB[r];
        v = csub(v, rmul(aw, gmv));
        v = csub(v, rmul(ae, gpv));
        v = csub(v, rmul(asj[r], dn));
        v = csub(v, rmul(anj[r], up));
        v = csub(v, rmul(ac_of(awe, asj[r], anj[r], ak), gi[r]));
        if ((FULL || j0 + r <= n2m)) st2(out + j0 + r, v);
        res0[r] = rmul((FULL || j0 + r <= n2m) ? 1.0 : 0.0, cadd(rmul(c1, resA[r]), rmul(c2, v)));
      }
    }
  }

  // relax(0, NBLI): first half on even lines, second half on odd lines
  {
    const double aw = aw0, ae = ae0;
    const bool hm = w > 0, hp = w + 1 < n0;
    for (int it = 0; it < mg.nbli; ++it) {
      if ((w & 1) == 0) {
        if (it == 0) {
#pragma unroll
          for (int r = 0; r < R; ++r) rr[r] = res0[r];
        } else {
#pragma unroll
          for (int r = 0; r < R; ++r) {
            const double2 gmv = lds_if(hm, &L0[slot(w - 1, r)], &L0[slot(w, r)]);
            const double2 gpv = lds_if(hp, &L0[slot(w + 1, r)], &L0[slot(w, r)]);
            rr[r] = rmul((FULL || j0 + r <= n2m) ? 1.0 : 0.0, csub(res0[r], rmul(1.0, cadd(rmul(aw, gmv), rmul(ae, gpv)))));
          }
        }
        tri_regs<R, FULL>(rr, c0.bt, asj, c0.gm, j0, n2m, lane, x);
#pragma unroll
        for (int r = 0; r < R; ++r) L0[slot(w, r)] = x[r];
      }
      __syncthreads();
      if (w & 1) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const double2 gmv = L0[slot(w - 1, r)];
          const double2 gpv = lds_if(hp, &L0[slot(w + 1, r)], &L0[slot(w, r)]);
          rr[r] = rmul((FULL || j0 + r <= n2m) ? 1.0 : 0.0, csub(csub(res0[r], rmul(aw, gmv)), rmul(ae, gpv)));
        }
        tri_regs<R, FULL>(rr, c0.bt, asj, c0.gm, j0, n2m, lane, x);
#pragma unroll
        for (int r = 0; r < R; ++r) L0[slot(w, r)] = x[r];
      }
      __syncthreads();
    }
  }

  // goup(0) + relax(1), first half: the neighbours 2w-1 and 2w+1 of line 2w
  // are the destinations of level-0 rows w-1 and w, prolonged as they are read
  // (cA and the level-1 residual resA are still in registers)
  {
    const double aw = awA, ae = aeA;
    const bool hm = w > 0, plast = w + 1 == n0;
    const double pm1 = ci1m, pm2 = ci2m;
    const double pp1 = ci1, pp2 = ci2;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      double2 vm = z2;
      const double2 l1m = lds_if(hm, &L1[slot(la - 1, r)], &L1[slot(la, r)]);
      const double2 l0m = lds_if(hm, &L0[slot(w - 1, r)], &L0[slot(w, r)]), l0 = L0[slot(w, r)];
      const double2 l0p = lds_if(!plast, &L0[slot(w + 1, r)], &L0[slot(w, r)]);
      if (hm) vm = cadd(cadd(l1m, rmul(pm1, l0m)), rmul(pm2, l0));
      double2 vp = cadd(L1[slot(lb, r)], rmul(pp1, L0[slot(w, r)]));
      if (!plast) vp = cadd(vp, rmul(pp2, l0p));
      rr[r] = rmul((FULL || j0 + r <= n2m) ? 1.0 : 0.0, csub(resA[r], rmul(1.0, cadd(rmul(aw, vm), rmul(ae, vp)))));
      __builtin_amdgcn_sched_barrier(0);  // five rows of reads per row: one row at a time
    }
  }
  // the new even rows are written where no wavefront of this stage reads
  tri_regs<R, FULL>(rr, cA.bt, asj, cA.gm, j0, n2m, lane, x);
  {
    S *out = ROW(gg, ia);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      L1[slot(la, r)] = x[r];
      if ((FULL || j0 + r <= n2m)) st2(out + j0 + r, x[r]);
    }
  }
  __syncthreads();

  // relax(1), second half
  load_lu<R, FULL>(mg, pl, ib, j0, n2m, cB);
  {
    const S *rb = ROW(resd, ib);
#pragma unroll
    for (int r = 0; r < R; ++r) resB[r] = ld2(rb + ((FULL || j0 + r <= n2m) ? j0 + r : n2m));
  }
  {
    const double aw = awB, ae = aeB;
    const bool hp = lb + 1 < n1;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const double2 gmv = L1[slot(la, r)];
      const double2 gpv = lds_if(hp, &L1[slot(lb + 1, r)], &L1[slot(lb, r)]);
      rr[r] = rmul((FULL || j0 + r <= n2m) ? 1.0 : 0.0, csub(csub(resB[r], rmul(aw, gmv)), rmul(ae, gpv)));
    }
    tri_regs<R, FULL>(rr, cB.bt, asj, cB.gm, j0, n2m, lane, x);
    S *out = ROW(gg, ib);
#pragma unroll
    for (int r = 0; r < R; ++r)
      if ((FULL || j0 + r <= n2m)) st2(out + j0 + r, x[r]);
  }
}

template <int R, bool FULL, typename S>
LW_KERNEL void k_mg_coarse_lds(LW_KP(64 * kCoarseLdsWaves, 0), const Geo &g, const MG &mg) {
  mg_coarse_lds_body<R, FULL, S>(g, mg, blockIdx.x);
}

typedef ::lw::KFn<Geo, MG> CoarseLdsKernel;

template <bool F, typename S>
static CoarseLdsKernel pick_coarse_lds_rf(int R) {
  switch (R) {
    case 1: return LW_KFN((k_mg_coarse_lds<1, F, S>));
    case 2: return LW_KFN((k_mg_coarse_lds<2, F, S>));
    case 4: return LW_KFN((k_mg_coarse_lds<4, F, S>));
    default: return {};
  }
}

static CoarseLdsKernel pick_coarse_lds(int R, bool full, bool f32 = false) {
  if (f32) return full ? pick_coarse_lds_rf<true, float2>(R) : pick_coarse_lds_rf<false, float2>(R);
  return full ? pick_coarse_lds_rf<true, double2>(R) : pick_coarse_lds_rf<false, double2>(R);
}

// LDS bytes of k_mg_coarse_lds, or 0 when the levels do not fit its shape
// (lc = 1, at most kCoarseLdsWaves level-0 lines, R <= 4, planes in LDS)
static long coarse_lds_bytes(const Geo &g, const MG &mg) {
  if (mg.lc != 1) return 0;
  const int R = lanes_r(g.n2m);
  if (R > 4) return 0;
  const int n1 = mg.iimg[1][2] - mg.iimg[1][1] + 1, n0 = mg.iimg[0][2] - mg.iimg[0][1] + 1;
  if (n1 != 2 * n0 || n0 < 1 || n0 > kCoarseLdsWaves) return 0;
  const long bytes = (long)(n1 + n0) * R * 64 * (long)sizeof(double2);
  return bytes <= kCoarseLdsBytes ? bytes : 0;
}

typedef ::lw::KFn<Geo, MG, int> CoarseKernel;

template <bool F>
static CoarseKernel pick_coarse_rf(int R) {
  switch (R) {
    case 1: return LW_KFN((k_mg_coarse<1, F>));
    case 2: return LW_KFN((k_mg_coarse<2, F>));
    case 4: return LW_KFN((k_mg_coarse<4, F>));
    case 8: return LW_KFN((k_mg_coarse<8, F>));
    default: return LW_KFN((k_mg_coarse<16, F>));
  }
}

static CoarseKernel pick_coarse(int R, bool full) {
  return full ? pick_coarse_rf<true>(R) : pick_coarse_rf<false>(R);
}

static void launch_coarse(hipStream_t s, const Geo &g, const MG &mg) {
  // unique bytes per wavenumber: the restricted rows of level lc in (16)
  // and its solution out (16) per point; bet (8) on every level; the
  // levels below lc never need to leave the chip
  double per_kv = 0.0;
  for (int l = 0; l <= mg.lc; ++l) {
    const double pts = (double)(mg.iimg[l][2] - mg.iimg[l][1] + 1) * g.n2m;
    per_kv += pts * (8.0 + (l == mg.lc ? 32.0 : 0.0));
  }
  const long lds = mg.coarse_lds ? coarse_lds_bytes(g, mg) : 0;
  if (lds > 0) {
    const int n0 = mg.iimg[0][2] - mg.iimg[0][1] + 1;
    LW_LAUNCH_K("k_mg_coarse", 0, per_kv, pick_coarse_lds(lanes_r(g.n2m), full_rows(g), mg.f32), dim3((unsigned)mg.nkv), dim3(64, n0),
              (unsigned)lds, s, g, mg);
    return;
  }
  LW_LAUNCH_K("k_mg_coarse", 0, per_kv, pick_coarse(lanes_r(g.n2m), full_rows(g)), dim3((unsigned)mg.nkv), dim3(64, kCoarseWaves),
            0, s, g, mg, mg.lc);
}

// top level (toplevel, poiss_z_ft.f90:339-380): rows ia = 2t+1 (odd) and
// ib = 2t+2. For kv = 1 the reference shifts every odd row of the top level
// by tt before the residual; here the shift is applied as the rows are read
// (the shifted odd rows reach pc; gg's odd rows are overwritten by the next
// toplevel's first half before anything reads them). Residual into resd,
// pc = solution, coarse row t, per-wavenumber max norm.
LW_KERNEL void k_top_rr(LW_KP(256, 0), const Geo &g, const MG &mg) {
  __shared__ double sh[4];
  const int kv = mg.kv0 + blockIdx.y;
  if (!mg.done[kv - 1]) {
    const int t = blockIdx.x;
    const long pl = (long)(kv - 1) * mg.PLANE, ppl = (long)(kv - 1) * mg.PCPLANE;
    const double2 *gg = mg.gg + pl;
    const int ia = 1 + 2 * t, ib = ia + 1;
    const int nl = mg.nlev;
    const int ic = mg.iimg[nl - 1][1] + t;
    double2 tt = make_double2(0.0, 0.0);
    const bool sh1 = kv == 1;
    if (sh1) {
      const double *ts = (const double *)(mg.sumres + mg.n3mh);
      tt = make_double2(ts[0], ts[1]);
    }
    // neighbours: imm(ia) = 2t (even), ipm(ib) = 2t+3 (odd, shifted unless
    // beyond the top level)
    const int ibp = ib + 1;
    const bool sbp = sh1 && ibp <= g.n1m;
    const double2 *ga = ROW(gg, ia);
    const double2 *gb = ROW(gg, ib), *gam_ = ROW(gg, ia - 1), *gbp = ROW(gg, ibp);
    const double awea = mg.awe[ia], aweb = mg.awe[ib], ak = mg.ak3[kv];
    const double2 *rha = ROW(mg.rhs + ppl, ia), *rhb = ROW(mg.rhs + ppl, ib);
    double2 *ra = ROW(mg.resd + pl, ia), *rb = ROW(mg.resd + pl, ib), *rc = ROW(mg.resd + pl, ic);
    double2 *pa = ROW(mg.pc + ppl, ia), *pb = ROW(mg.pc + ppl, ib);
    const double c1 = mg.cor1[ic], c2 = mg.cor2[ic];
    const double awa = mg.aw[ia], aea = mg.ae[ia], awb = mg.aw[ib], aeb = mg.ae[ib];
    double ar = 0.0;
    // the shift covers j = 1..n2m only; the ghost rows j = 0, n2 stay zero
    auto GA = [&](int j) -> double2 { return (sh1 && j >= 1 && j <= g.n2m) ? csub(ga[j], tt) : ga[j]; };
    for (int j = 1 + threadIdx.x; j <= g.n2m; j += 256) {
      double aw = awa, ae = aea;
      const double2 gaj = GA(j);
      double2 x = rha[j];
      x = csub(x, rmul(aw, gam_[j]));
      x = csub(x, rmul(ae, gb[j]));
      x = csub(x, rmul(mg.as[j], GA(j - 1)));
      x = csub(x, rmul(mg.an[j], GA(j + 1)));
      x = csub(x, rmul(ac_of(awea, mg.as[j], mg.an[j], ak), gaj));
      aw = awb;
      ae = aeb;
      const double2 gbpj = sbp ? csub(gbp[j], tt) : gbp[j];
      double2 y = rhb[j];
      y = csub(y, rmul(aw, gaj));
      y = csub(y, rmul(ae, gbpj));
      y = csub(y, rmul(mg.as[j], gb[j - 1]));
      y = csub(y, rmul(mg.an[j], gb[j + 1]));
      y = csub(y, rmul(ac_of(aweb, mg.as[j], mg.an[j], ak), gb[j]));
      ra[j] = x;
      rb[j] = y;
      rc[j] = cadd(rmul(c1, x), rmul(c2, y));
      pa[j] = gaj;
      pb[j] = gb[j];
      ar = fmax(ar, fmax(hypot(x.x, x.y), hypot(y.x, y.y)));
    }
    for (int o = 32; o > 0; o >>= 1) ar = fmax(ar, __shfl_down(ar, o, 64));
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = ar;
    __syncthreads();
    if (threadIdx.x == 0) {
      const double m = fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
      atomicMax(&mg.sumres[kv - 1], (unsigned long long)__double_as_longlong(m));
    }
  }
}

// the convergence test after a toplevel (poiss_z_ft.f90:284-296): id = 1
// counts a completed V-cycle; a wavenumber stops when its max norm is below
// `test`, or when it has used MGITR cycles. Clears the norms for the next
// toplevel.
LW_KERNEL void k_mg_check(LW_KP(1024, 0), const MG &mg, double test, int id, int flip) {
  for (int q = mg.kv0 - 1 + threadIdx.x; q < mg.kv0 - 1 + mg.nkv; q += blockDim.x) {
    if (!mg.done[q]) {
      const int c = mg.cyc[q] + id;
      mg.cyc[q] = c;
      if (flip) mg.pcsel[q] ^= 1;  // k_mg_top wrote the new pc to the other buffer
      const double r = __longlong_as_double((long long)mg.sumres[q]);
      if (r < test) mg.done[q] = 1;
      else if (c >= mg.mgitr) mg.done[q] = 2;
      if (mg.done[q]) mg.rfin[q] = r;
    }
    mg.sumres[q] = 0ull;
  }
}

// ---------------------------------------------------------------------------
// fused level kernels: one launch per level and direction instead of one per
// zebra half and one per residual. A workgroup owns a tile of T consecutive
// lines of its level (T even, dividing the level) and recomputes the few
// halo lines its last stage needs, so no workgroup waits for another; the
// halo solves are the same computation as the owner's and give the same
// values. Stages inside the workgroup meet at barriers, solutions pass
// through LDS in the lane-major slot layout of k_mg_coarse_lds.
//   k_mg_down(L): relax(L, oldv 0) + godown(L) (poiss_z_ft.f90:272-275)
//   k_mg_up(L):   goup(L-1) + relax(L, oldv 1) (poiss_z_ft.f90:279-282)
//   k_mg_top:     toplevel: both halves, residual, restriction, pc
//                 (poiss_z_ft.f90:302-380)
// The down-sweep's second-half lines, the rows goup adds to, go to mg.gd
// rather than mg.gg, so the up kernel of a tile reads its neighbours' down
// values while those neighbours write their up values to mg.gg.
// ---------------------------------------------------------------------------
template <int R, bool FULL, typename T>
__device__ inline void load_rows(const T *row, int j0, int n2m, double2 (&v)[R]) {
#pragma unroll
  for (int r = 0; r < R; ++r) v[r] = ld2(row + ((FULL || j0 + r <= n2m) ? j0 + r : n2m));
}

template <int R, bool FULL, typename T>
__device__ inline void store_rows(T *row, int j0, int n2m, const double2 (&v)[R]) {
#pragma unroll
  for (int r = 0; r < R; ++r)
    if ((FULL || j0 + r <= n2m)) st2(row + j0 + r, v[r]);
}

// XCD-aware placement of the fused kernels' workgroups. gfx950 deals
// consecutive workgroups of a launch to its 8 XCDs in turn, and each XCD has
// its own L2; neighbouring tiles of a wavenumber read each other's halo rows
// and coarse rows. Each wavenumber's tiles are cut into 8 runs of
// c = ceil(ntile / 8) consecutive tiles, one run per XCD: workgroup b runs on
// XCD x = b % 8 and takes tile x*c + (b/8) % c of wavenumber (b/8) / c
// (tiles past the end exit). A wavenumber still spreads over all XCDs, which
// matters when only one or two are left iterating.
constexpr int kXcds = 8;
__device__ inline void xcd_tile(int b, int ntile, int &kvi, int &tile) {
  const int c = (ntile + kXcds - 1) / kXcds, slot = b / kXcds;
  kvi = slot / c;
  tile = (b % kXcds) * c + slot % c;
}
static inline dim3 xcd_grid(int ntile, int nkv) {
  return dim3((unsigned)(kXcds * ((ntile + kXcds - 1) / kXcds) * nkv));
}

// tile size of the fused kernels on a level of n lines
constexpr int kTileMax = 8;
// most wavefronts a workgroup of a multigrid kernel has (k_mg_kv1: the
// level-0 lines of k_mg_coarse_lds)
constexpr int kWaveMax = 16;
__host__ __device__ inline int fused_tile(int n, int tmax) {
  for (int t = tmax; t > 2; t >>= 1)
    if (n % t == 0) return t;
  return 2;
}

// the residual at rows j0+r of the two lines whose LDS slots are qa, qa+1
// (neighbours qa-1, qa+2), in the reference's order (poiss_z_ft.f90:440-443)
#define TILE_AT(q_, r_) lds[((q_) * R + (r_)) * 64 + lane]
#define TILE_DN(q_, r_) \
  ((r_) > 0 ? lds[((q_) * R + (r_) - 1) * 64 + lane] : lane > 0 ? lds[((q_) * R + R - 1) * 64 + lane - 1] : z2)
#define TILE_UP(q_, r_) \
  ((r_) < R - 1 ? lds[((q_) * R + (r_) + 1) * 64 + lane] : lane < 63 ? lds[((q_) * R) * 64 + lane + 1] : z2)

// godown (poiss_z_ft.f90:431-458) on a tile of level `ilev` >= 1: the relax
// stages write the tile's lines f0-2 .. f0+T to LDS slots 0 .. T+2 (lines
// outside the level as zeros); then residual and restriction per pair.
template <int R, bool FULL, typename S>
__device__ __forceinline__ void mg_down_body(const Geo &g, const MG &mg, int ilev, int T, double test, int cyc, int bid) {
  extern __shared__ double2 lds[];
  const int ntile = (mg.iimg[ilev][2] - mg.iimg[ilev][1] + 1) / T;
  int kvi, tile;
  xcd_tile(bid, ntile, kvi, tile);
  if (tile >= ntile) return;
  const int kv = mg.kv0 + kvi;
  if (mg.done[kv - 1]) return;
  if (cyc >= 0) {
    // the convergence test of the toplevel just done (k_mg_check's), folded
    // into the first launch of the next cycle: every workgroup decides from
    // the max norm, which no kernel of this launch writes; workgroup (0, kv)
    // records it. An active wavenumber has completed `cyc` V-cycles.
    const double r = __longlong_as_double((long long)mg.sumres[kv - 1]);
    const int dn = r < test ? 1 : cyc >= mg.mgitr ? 2 : 0;
    if (tile == 0 && threadIdx.x == 0 && threadIdx.y == 0) {
      mg.cyc[kv - 1] = cyc;
      mg.pcsel[kv - 1] ^= 1;  // k_mg_top wrote the new pc to the other buffer
      if (dn) {
        mg.done[kv - 1] = dn;
        mg.rfin[kv - 1] = r;
      }
    }
    if (dn) return;
  }
  // the wavefront index as a uniform value: line coefficients come in by
  // scalar loads and row addresses stay in scalar registers
  const int lane = threadIdx.x, w = __builtin_amdgcn_readfirstlane(threadIdx.y);
  const int n2m = FULL ? 64 * R : g.n2m, j0 = 1 + lane * R;
  const int bL = mg.iimg[ilev][1], nL = mg.iimg[ilev][2] - bL + 1, bC = mg.iimg[ilev - 1][1];
  const int f0 = tile * T;
  const long pl = (long)(kv - 1) * mg.PLANE;
  S *resdS = (S *)mg.resd + pl, *gdS = (S *)mg.gd + pl, *rsuS = (S *)mg.rsu + pl;
  const double2 z2 = make_double2(0.0, 0.0);
  double asj[R];
#pragma unroll
  for (int r = 0; r < R; ++r) asj[r] = mg.as[(FULL || j0 + r <= n2m) ? j0 + r : n2m];
  double2 x[R];

  // wavefront w solves line e = f0-2+2w of the first half (w <= T/2+1) and
  // line o = f0-1+2w of the second (w <= T/2); for 1 <= w <= T/2 these are
  // the tile's pair w-1, whose residual it then forms from the right-hand
  // sides still in its registers
  const int e = f0 - 2 + 2 * w;
  const bool eok = w <= T / 2 + 1 && e >= 0 && e < nL;
  const int o = f0 - 1 + 2 * w;
  const bool ook = w <= T / 2 && o >= 0;
  const bool pair = w >= 1 && w <= T / 2;
  // an(j) for the residual stage, through LDS (row j at j-1; rows past n2m
  // read n2m), and the stages' line coefficients, all requested up front
  double *san = (double *)(lds + (T + 3) * R * 64);
  for (int q = lane + 64 * w; q < 64 * R; q += 64 * (int)blockDim.y) san[q] = mg.an[q + 1 <= n2m ? q + 1 : n2m];
  const int iaq = pair ? bL + e : bL, ibq = ook ? bL + o : bL, icq = pair ? bC + f0 / 2 + w - 1 : bC;
  const double aw2 = mg.aw[ibq], ae2 = mg.ae[ibq];
  const double awa_ = mg.aw[iaq], aea_ = mg.ae[iaq], wea = mg.awe[iaq], web = mg.awe[ibq], ak = mg.ak3[kv];
  const double c1 = mg.cor1[icq], c2 = mg.cor2[icq];
  double2 so[R], se[R];
  LineLU<R> co;
  if (ook) {
    load_rows<R, FULL>(ROW(resdS, bL + o), j0, n2m, so);
    load_lu<R, FULL>(mg, pl, bL + o, j0, n2m, co);
  }
  // first half (oldv 0): the right-hand side is resd itself
  if (eok) {
    LineLU<R> ce;
    load_rows<R, FULL>(ROW(resdS, bL + e), j0, n2m, se);
    load_lu<R, FULL>(mg, pl, bL + e, j0, n2m, ce);
    tri_regs<R, FULL>(se, ce.bt, asj, ce.gm, j0, n2m, lane, x);
  } else {
#pragma unroll
    for (int r = 0; r < R; ++r) x[r] = z2;
  }
  if (w <= T / 2 + 1) {
#pragma unroll
    for (int r = 0; r < R; ++r) TILE_AT(2 * w, r) = x[r];
  }
  __syncthreads();

  // second half: resd - aw*g(i-1) - ae*g(i+1)
  if (ook) {
    double2 rr[R];
#pragma unroll
    for (int r = 0; r < R; ++r) rr[r] = csub(csub(so[r], rmul(aw2, TILE_AT(2 * w, r))), rmul(ae2, TILE_AT(2 * w + 2, r)));
    tri_regs<R, FULL>(rr, co.bt, asj, co.gm, j0, n2m, lane, x);
    if (w >= 1) store_rows<R, FULL>(ROW(gdS, bL + o), j0, n2m, x);
  } else {
#pragma unroll
    for (int r = 0; r < R; ++r) x[r] = z2;
  }
  if (w <= T / 2) {
#pragma unroll
    for (int r = 0; r < R; ++r) TILE_AT(2 * w + 1, r) = x[r];
  }
  __syncthreads();

  // residual of the pair a = f0+2p, b = a+1 (p = w-1), restricted into the
  // coarse row bC + f0/2 + p. The residual goes to rsu: neighbouring tiles
  // still read resd rows of this level as their halo lines' right-hand sides.
  if (pair) {
    const int qa = 2 * w, ia = bL + e, ib = ia + 1, ic = bC + f0 / 2 + w - 1;
    const double awa = awa_, aea = aea_, awb = aw2, aeb = ae2;
    double acA[R], acB[R], anj[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      anj[r] = san[j0 - 1 + r];
      acA[r] = ac_of(wea, asj[r], anj[r], ak);
      acB[r] = ac_of(web, asj[r], anj[r], ak);
    }
    S *ra = ROW(rsuS, ia), *rb = ROW(rsuS, ib), *rc = ROW(resdS, ic);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int j = j0 + r;
      if (!FULL && j > n2m) continue;
      const double2 ga = TILE_AT(qa, r), gb = TILE_AT(qa + 1, r);
      double2 xa = se[r];
      xa = csub(xa, rmul(awa, TILE_AT(qa - 1, r)));
      xa = csub(xa, rmul(aea, gb));
      xa = csub(xa, rmul(asj[r], TILE_DN(qa, r)));
      xa = csub(xa, rmul(anj[r], TILE_UP(qa, r)));
      xa = csub(xa, rmul(acA[r], ga));
      double2 xb = so[r];
      xb = csub(xb, rmul(awb, ga));
      xb = csub(xb, rmul(aeb, TILE_AT(qa + 2, r)));
      xb = csub(xb, rmul(asj[r], TILE_DN(qa + 1, r)));
      xb = csub(xb, rmul(anj[r], TILE_UP(qa + 1, r)));
      xb = csub(xb, rmul(acB[r], gb));
      st2(ra + j, xa);
      st2(rb + j, xb);
      st2(rc + j, cadd(rmul(c1, xa), rmul(c2, xb)));
    }
  }
}

template <int R, bool FULL, typename S>
LW_KERNEL void k_mg_down(LW_KP(64 * (kTileMax / 2 + 2), 4), const Geo &g, const MG &mg, int ilev, int T, double test,
                                                                      int cyc) {
  mg_down_body<R, FULL, S>(g, mg, ilev, T, test, cyc, blockIdx.x);
}

// goup(ilev-1) + relax(ilev, oldv 1) on a tile: first half on lines f0+2w
// (w <= T/2, the last one a halo) with its neighbours prolonged as read,
// then the second half on the tile's odd lines
template <int R, bool FULL, typename S>
__device__ __forceinline__ void mg_up_body(const Geo &g, const MG &mg, int ilev, int T, int bid) {
  extern __shared__ double2 lds[];
  const int ntile = (mg.iimg[ilev][2] - mg.iimg[ilev][1] + 1) / T;
  int kvi, tile;
  xcd_tile(bid, ntile, kvi, tile);
  if (tile >= ntile) return;
  const int kv = mg.kv0 + kvi;
  if (mg.done[kv - 1]) return;
  // the wavefront index as a uniform value: line coefficients come in by
  // scalar loads and row addresses stay in scalar registers
  const int lane = threadIdx.x, w = __builtin_amdgcn_readfirstlane(threadIdx.y);
  const int n2m = FULL ? 64 * R : g.n2m, j0 = 1 + lane * R;
  const int bL = mg.iimg[ilev][1], nL = mg.iimg[ilev][2] - bL + 1;
  const int cb = mg.iimg[ilev - 1][1], ce = mg.iimg[ilev - 1][2];
  const int f0 = tile * T;
  const long pl = (long)(kv - 1) * mg.PLANE;
  const double2 z2 = make_double2(0.0, 0.0);
  const S *gg = (const S *)mg.gg + pl, *gd = (const S *)mg.gd + pl;
  S *ggW = (S *)mg.gg + pl;
  const S *rsuS = (const S *)mg.rsu + pl;
  double asj[R];
#pragma unroll
  for (int r = 0; r < R; ++r) asj[r] = mg.as[(FULL || j0 + r <= n2m) ? j0 + r : n2m];
  double2 x[R];

  const int o = f0 + 1 + 2 * w;
  const bool ook = w < T / 2;
  double2 so[R];
  LineLU<R> co;
  if (ook) {
    load_rows<R, FULL>(ROW(rsuS, bL + o), j0, n2m, so);
    load_lu<R, FULL>(mg, pl, bL + o, j0, n2m, co);
  }
  // wavefronts past T/2 (a launch with more of them) take no line
  const int e = f0 + 2 * w;
  if (w <= T / 2 && e < nL) {
    const int i = bL + e;
    const double aw = mg.aw[i], ae = mg.ae[i];
    double2 rr[R];
    LineLU<R> cl;
    load_rows<R, FULL>(ROW(rsuS, i), j0, n2m, rr);
    load_lu<R, FULL>(mg, pl, i, j0, n2m, cl);
    // line e-1 is the destination of coarse row c-1, line e+1 of row c
    const int c = cb + e / 2;
    const S *dm = ROW(gd, i - 1), *dp = ROW(gd, i + 1);
    const S *cm = ROW(gg, c - 1), *c0 = ROW(gg, c), *c1 = ROW(gg, c < ce ? c + 1 : c);
    const bool hm = e > 0, last = c == ce;
    const double pm1 = hm ? mg.coi1[c - 1] : 0.0, pm2 = hm ? mg.coi2[c - 1] : 0.0;
    const double pp1 = mg.coi1[c], pp2 = mg.coi2[c];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int jc = (FULL || j0 + r <= n2m) ? j0 + r : n2m;
      const double2 vm = hm ? cadd(cadd(ld2(dm + jc), rmul(pm1, ld2(cm + jc))), rmul(pm2, ld2(c0 + jc))) : z2;
      double2 vp = cadd(ld2(dp + jc), rmul(pp1, ld2(c0 + jc)));
      if (!last) vp = cadd(vp, rmul(pp2, ld2(c1 + jc)));
      rr[r] = csub(rr[r], rmul(1.0, cadd(rmul(aw, vm), rmul(ae, vp))));
    }
    tri_regs<R, FULL>(rr, cl.bt, asj, cl.gm, j0, n2m, lane, x);
    if (w < T / 2) store_rows<R, FULL>(ROW(ggW, i), j0, n2m, x);
  } else {
#pragma unroll
    for (int r = 0; r < R; ++r) x[r] = z2;
  }
  if (w <= T / 2) {
#pragma unroll
    for (int r = 0; r < R; ++r) TILE_AT(w, r) = x[r];
  }
  __syncthreads();

  if (ook) {
    const double aw = mg.aw[bL + o], ae = mg.ae[bL + o];
    double2 rr[R];
#pragma unroll
    for (int r = 0; r < R; ++r) rr[r] = csub(csub(so[r], rmul(aw, TILE_AT(w, r))), rmul(ae, TILE_AT(w + 1, r)));
    tri_regs<R, FULL>(rr, co.bt, asj, co.gm, j0, n2m, lane, x);
    store_rows<R, FULL>(ROW(ggW, bL + o), j0, n2m, x);
  }
}

template <int R, bool FULL, typename S>
LW_KERNEL void k_mg_up(LW_KP(64 * (kTileMax / 2 + 1), 0), const Geo &g, const MG &mg, int ilev, int T) {
  mg_up_body<R, FULL, S>(g, mg, ilev, T, blockIdx.x);
}

// first half of toplevel for line e (0-based; row 1 + e) reading the pc
// planes `pcr`: rhs - oldv*(aw*P(e-1) + ae*P(e+1)), P = pc plus, for id = 1,
// the prolongation of level nlev-1 added as it is read; the line's rhs rows
// stay in `rh` for the residual
template <int R, bool FULL, typename S>
__device__ inline void top_first(const Geo &g, const MG &mg, long pl, long ppl, const double2 *pcr, int e,
                                 double oldv, int id, int j0, int n2m, int lane, const double (&asj)[R],
                                 double2 (&rh)[R], double2 (&x)[R]) {
  const int i = 1 + e;
  const double2 z2 = make_double2(0.0, 0.0);
  const S *gg = (const S *)mg.gg + pl;
  const int cb = mg.iimg[mg.nlev - 1][1], ce = mg.iimg[mg.nlev - 1][2];
  const int c = cb + e / 2;  // line e-1 <- coarse row c-1, line e+1 <- row c
  const bool hm = e > 0, pro = id == 1, last = c == ce;
  const double aw = mg.aw[i], ae = mg.ae[i];
  const double pm1 = hm ? mg.coi1[c - 1] : 0.0, pm2 = hm ? mg.coi2[c - 1] : 0.0;
  const double pp1 = mg.coi1[c], pp2 = mg.coi2[c];
  const double2 *pm = ROW(pcr, i - 1), *pp = ROW(pcr, i + 1);
  const S *cm = ROW(gg, c - 1), *c0 = ROW(gg, c), *c1 = ROW(gg, c < ce ? c + 1 : c);
  LineLU<R> cl;
  load_rows<R, FULL>(ROW(mg.rhs + ppl, i), j0, n2m, rh);
  load_lu<R, FULL>(mg, pl, i, j0, n2m, cl);
  double2 rr[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int jc = (FULL || j0 + r <= n2m) ? j0 + r : n2m;
    double2 vm = hm ? pm[jc] : z2, vp = pp[jc];
    if (pro) {
      if (hm) vm = cadd(cadd(vm, rmul(pm1, ld2(cm + jc))), rmul(pm2, ld2(c0 + jc)));
      vp = cadd(vp, rmul(pp1, ld2(c0 + jc)));
      if (!last) vp = cadd(vp, rmul(pp2, ld2(c1 + jc)));
    }
    rr[r] = csub(rh[r], rmul(oldv, cadd(rmul(aw, vm), rmul(ae, vp))));
  }
  tri_regs<R, FULL>(rr, cl.bt, asj, cl.gm, j0, n2m, lane, x);
}

// toplevel (poiss_z_ft.f90:302-380) in one launch per cycle: on a tile of
// top lines, wavefront w solves first-half line f0-2+2w (w <= T/2+1) and
// second-half line f0-1+2w (w <= T/2); for 1 <= w <= T/2 these are pair w-1,
// whose residual, restriction, pc and max norm it then forms. For kv = 1 the
// gauge tt = first-half line 0 at row n2m-1 shifts the first-half lines;
// wavefront T/2+2 recomputes line 0 so every tile has it. The tiles read the
// old pc and write the new one into the other buffer (mg.pcsel per
// wavenumber says which holds the current pc; k_mg_check flips it), because
// a tile's halo lines read pc rows its neighbours write.
template <int R, bool FULL, typename S>
__device__ __forceinline__ void mg_top_body(const Geo &g, const MG &mg, int T, double oldv, int id, int bid) {
  extern __shared__ double2 lds[];
  __shared__ double wmax[kWaveMax];
  __shared__ double2 tts;
  const int ntile = g.n1m / T;
  int kvi, tile;
  xcd_tile(bid, ntile, kvi, tile);
  if (tile >= ntile) return;
  const int kv = mg.kv0 + kvi;
  if (mg.done[kv - 1]) return;
  const int lane = threadIdx.x, w = __builtin_amdgcn_readfirstlane(threadIdx.y), W = blockDim.y;
  const int n2m = FULL ? 64 * R : g.n2m, n1m = g.n1m, j0 = 1 + lane * R;
  const int f0 = tile * T;
  const long pl = (long)(kv - 1) * mg.PLANE, ppl = (long)(kv - 1) * mg.PCPLANE;
  const int sel = mg.pcsel[kv - 1];
  const double2 *pcr = (sel ? mg.pc2 : mg.pc) + ppl;
  double2 *pcw = (sel ? mg.pc : mg.pc2) + ppl;
  const double2 z2 = make_double2(0.0, 0.0);
  double asj[R];
#pragma unroll
  for (int r = 0; r < R; ++r) asj[r] = mg.as[(FULL || j0 + r <= n2m) ? j0 + r : n2m];

  const bool pair = w >= 1 && w <= T / 2;
  const int e = f0 - 2 + 2 * w, o = e + 1;
  const bool eok = w <= T / 2 + 1 && e >= 0 && e < n1m;
  const bool ook = w <= T / 2 && o >= 0;
  // an(j) for the residual stage, through LDS (row j at j-1; rows past n2m
  // read n2m), and the stages' line coefficients, all requested up front
  double *san = (double *)(lds + (T + 3) * R * 64);
  for (int q = lane + 64 * w; q < 64 * R; q += 64 * W) san[q] = mg.an[q + 1 <= n2m ? q + 1 : n2m];
  const int ia = 1 + e, ib = ia + 1, ic = mg.iimg[mg.nlev - 1][1] + f0 / 2 + w - 1;
  const int icq = pair ? ic : mg.iimg[mg.nlev - 1][1], iaq = pair ? ia : 1, ibq = ook ? ib : 1;
  const double aw2 = mg.aw[ibq], ae2 = mg.ae[ibq];
  const double awa = mg.aw[iaq], aea = mg.ae[iaq], wea = mg.awe[iaq], web = mg.awe[ibq], ak = mg.ak3[kv];
  const double c1 = mg.cor1[icq], c2 = mg.cor2[icq];
  double2 sa[R], sb[R], x[R];
  LineLU<R> co;
  if (ook) {
    load_rows<R, FULL>(ROW(mg.rhs + ppl, 1 + o), j0, n2m, sb);
    load_lu<R, FULL>(mg, pl, 1 + o, j0, n2m, co);
  }
  if (eok) {
    top_first<R, FULL, S>(g, mg, pl, ppl, pcr, e, oldv, id, j0, n2m, lane, asj, sa, x);
  } else if (w == T / 2 + 2 && kv == 1) {
    double2 rh[R];
    top_first<R, FULL, S>(g, mg, pl, ppl, pcr, 0, oldv, id, j0, n2m, lane, asj, rh, x);
#pragma unroll
    for (int r = 0; r < R; ++r)
      if (j0 + r == n2m - 1) tts = x[r];
  } else {
#pragma unroll
    for (int r = 0; r < R; ++r) x[r] = z2;
  }
  __syncthreads();
  // the first-half lines as the second half and the residual see them:
  // shifted by tt on rows 1..n2m, the ghost rows past n2m zero
  const double2 tt = kv == 1 ? tts : z2;
  if (w <= T / 2 + 1) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      double2 v = eok && kv == 1 ? csub(x[r], tt) : x[r];
      if (!FULL && j0 + r > n2m) v = z2;
      TILE_AT(2 * w, r) = v;
    }
  }
  __syncthreads();

  // second half: rhs - aw*(g(i-1)-tt) - ae*(g(i+1)-tt)
  if (ook) {
    double2 rr[R];
#pragma unroll
    for (int r = 0; r < R; ++r) rr[r] = csub(csub(sb[r], rmul(aw2, TILE_AT(2 * w, r))), rmul(ae2, TILE_AT(2 * w + 2, r)));
    tri_regs<R, FULL>(rr, co.bt, asj, co.gm, j0, n2m, lane, x);
  } else {
#pragma unroll
    for (int r = 0; r < R; ++r) x[r] = z2;
  }
  if (w <= T / 2) {
#pragma unroll
    for (int r = 0; r < R; ++r) TILE_AT(2 * w + 1, r) = x[r];
  }
  __syncthreads();

  // residual of the pair (rows ia = 1+e, ib = ia+1), restriction, pc, norm
  double ar = 0.0;
  if (pair) {
    const int qa = 2 * w;
    const double awb = aw2, aeb = ae2;
    double acA[R], acB[R], anj[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      anj[r] = san[j0 - 1 + r];
      acA[r] = ac_of(wea, asj[r], anj[r], ak);
      acB[r] = ac_of(web, asj[r], anj[r], ak);
    }
    S *rc = ROW((S *)mg.resd + pl, ic);
    double2 *pa = ROW(pcw, ia), *pb = ROW(pcw, ib);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int j = j0 + r;
      if (!FULL && j > n2m) continue;
      const double2 ga = TILE_AT(qa, r), gb = TILE_AT(qa + 1, r);
      double2 xa = sa[r];
      xa = csub(xa, rmul(awa, TILE_AT(qa - 1, r)));
      xa = csub(xa, rmul(aea, gb));
      xa = csub(xa, rmul(asj[r], TILE_DN(qa, r)));
      xa = csub(xa, rmul(anj[r], TILE_UP(qa, r)));
      xa = csub(xa, rmul(acA[r], ga));
      double2 xb = sb[r];
      xb = csub(xb, rmul(awb, ga));
      xb = csub(xb, rmul(aeb, TILE_AT(qa + 2, r)));
      xb = csub(xb, rmul(asj[r], TILE_DN(qa + 1, r)));
      xb = csub(xb, rmul(anj[r], TILE_UP(qa + 1, r)));
      xb = csub(xb, rmul(acB[r], gb));
      st2(rc + j, cadd(rmul(c1, xa), rmul(c2, xb)));
      pa[j] = ga;
      pb[j] = gb;
      ar = fmax(ar, fmax(hypot(xa.x, xa.y), hypot(xb.x, xb.y)));
    }
  }
  for (int s = 32; s > 0; s >>= 1) ar = fmax(ar, __shfl_down(ar, s, 64));
  if (lane == 0) wmax[w] = ar;
  __syncthreads();
  if (w == 0 && lane == 0) {
    double m = 0.0;
    for (int q = 0; q < W; ++q) m = fmax(m, wmax[q]);
    atomicMax(&mg.sumres[kv - 1], (unsigned long long)__double_as_longlong(m));
  }
}

template <int R, bool FULL, typename S>
LW_KERNEL void k_mg_top(LW_KP(64 * (kTileMax / 2 + 3), 0), const Geo &g, const MG &mg, int T, double oldv, int id) {
  mg_top_body<R, FULL, S>(g, mg, T, oldv, id, blockIdx.x);
}

// ---------------------------------------------------------------------------
// gsor2d (poiss_z_ft.f90:513-587) for the wavenumbers above IMGSOR: zebra
// y-line over-relaxation of the top level alone, one launch per sweep, each
// sweep ending with resid3's max norm (:590-616). The tiles are k_mg_top's:
// wavefront w solves first-half line f0-2+2w (w <= T/2+1) and second-half
// line f0-1+2w (w <= T/2), and for 1 <= w <= T/2 forms the residual of the
// pair w-1 from the relaxed values in LDS. A line's relaxed value is
// u = ww*x + (wgt*ww2)*u_old, u_old from the current pc buffer; the new
// values go to the other buffer (the check flips pcsel). wgt is gsor2d's
// oldv on the sweep before its loop (0: u starts at zero) and 1 in it; it
// also scales the first half's x coupling, which the second half's never is:
// rhs - wgt*(aw*u(i-1) + ae*u(i+1)) against rhs - (aw*u(i-1) + ae*u(i+1)).
// ---------------------------------------------------------------------------
template <int R, bool FULL>
LW_KERNEL void k_sor(LW_KP(64 * (kTileMax / 2 + 2), 0), const Geo &g, const MG &mg, int T, double wgt, double ww) {
  extern __shared__ double2 lds[];
  __shared__ double wmax[kWaveMax];
  const int ntile = g.n1m / T;
  int kvi, tile;
  xcd_tile(blockIdx.x, ntile, kvi, tile);
  if (tile >= ntile) return;
  const int kv = mg.kv0 + kvi;
  if (mg.done[kv - 1]) return;
  const int lane = threadIdx.x, w = __builtin_amdgcn_readfirstlane(threadIdx.y), W = blockDim.y;
  const int n2m = FULL ? 64 * R : g.n2m, n1m = g.n1m, j0 = 1 + lane * R;
  const int f0 = tile * T;
  const long pl = (long)(kv - 1) * mg.PLANE, ppl = (long)(kv - 1) * mg.PCPLANE;
  const int sel = mg.pcsel[kv - 1];
  const double2 *pcr = (sel ? mg.pc2 : mg.pc) + ppl;
  double2 *pcw = (sel ? mg.pc : mg.pc2) + ppl;
  const double2 z2 = make_double2(0.0, 0.0);
  const double cw = wgt * (1.0 - ww);
  double asj[R];
#pragma unroll
  for (int r = 0; r < R; ++r) asj[r] = mg.as[(FULL || j0 + r <= n2m) ? j0 + r : n2m];

  const bool pair = w >= 1 && w <= T / 2;
  const int e = f0 - 2 + 2 * w, o = e + 1;
  const bool eok = w <= T / 2 + 1 && e >= 0 && e < n1m;
  const bool ook = w <= T / 2 && o >= 0;
  double *san = (double *)(lds + (T + 3) * R * 64);
  for (int q = lane + 64 * w; q < 64 * R; q += 64 * W) san[q] = mg.an[q + 1 <= n2m ? q + 1 : n2m];
  const int ia = 1 + e, ib = ia + 1;
  const int iaq = pair ? ia : 1, ibq = ook ? ib : 1;
  const double aw2 = mg.aw[ibq], ae2 = mg.ae[ibq];
  const double awa = mg.aw[iaq], aea = mg.ae[iaq], wea = mg.awe[iaq], web = mg.awe[ibq], ak = mg.ak3[kv];
  double2 sa[R], sb[R], ub[R], x[R];
  LineLU<R> co;
  if (ook) {
    load_rows<R, FULL>(ROW(mg.rhs + ppl, 1 + o), j0, n2m, sb);
    load_rows<R, FULL>(ROW(pcr, 1 + o), j0, n2m, ub);
    load_lu<R, FULL>(mg, pl, 1 + o, j0, n2m, co);
  }
  // first half: line i = 1 + e
  if (eok) {
    const int i = ia;
    const double aw = mg.aw[i], ae = mg.ae[i];
    const double2 *pm = ROW(pcr, i - 1), *pp = ROW(pcr, i + 1), *p0 = ROW(pcr, i);
    LineLU<R> cl;
    load_rows<R, FULL>(ROW(mg.rhs + ppl, i), j0, n2m, sa);
    load_lu<R, FULL>(mg, pl, i, j0, n2m, cl);
    double2 rr[R], u0[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int jc = (FULL || j0 + r <= n2m) ? j0 + r : n2m;
      const double2 vm = e > 0 ? pm[jc] : z2, vp = pp[jc];
      u0[r] = p0[jc];
      rr[r] = csub(sa[r], rmul(wgt, cadd(rmul(aw, vm), rmul(ae, vp))));
    }
    tri_regs<R, FULL>(rr, cl.bt, asj, cl.gm, j0, n2m, lane, x);
#pragma unroll
    for (int r = 0; r < R; ++r) x[r] = cadd(rmul(ww, x[r]), rmul(cw, u0[r]));
  } else {
#pragma unroll
    for (int r = 0; r < R; ++r) x[r] = z2;
  }
  if (w <= T / 2 + 1) {
#pragma unroll
    for (int r = 0; r < R; ++r) TILE_AT(2 * w, r) = (!FULL && j0 + r > n2m) ? z2 : x[r];
  }
  __syncthreads();

  // second half: line i = 1 + o, its neighbours relaxed above
  if (ook) {
    double2 rr[R];
#pragma unroll
    for (int r = 0; r < R; ++r) rr[r] = csub(sb[r], cadd(rmul(aw2, TILE_AT(2 * w, r)), rmul(ae2, TILE_AT(2 * w + 2, r))));
    tri_regs<R, FULL>(rr, co.bt, asj, co.gm, j0, n2m, lane, x);
#pragma unroll
    for (int r = 0; r < R; ++r) x[r] = cadd(rmul(ww, x[r]), rmul(cw, ub[r]));
  } else {
#pragma unroll
    for (int r = 0; r < R; ++r) x[r] = z2;
  }
  if (w <= T / 2) {
#pragma unroll
    for (int r = 0; r < R; ++r) TILE_AT(2 * w + 1, r) = (!FULL && j0 + r > n2m) ? z2 : x[r];
  }
  __syncthreads();

  // resid3 of the pair (rows ia, ib), the new u and the max norm
  double ar = 0.0;
  if (pair) {
    const int qa = 2 * w;
    double acA[R], acB[R], anj[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      anj[r] = san[j0 - 1 + r];
      acA[r] = ac_of(wea, asj[r], anj[r], ak);
      acB[r] = ac_of(web, asj[r], anj[r], ak);
    }
    double2 *pa = ROW(pcw, ia), *pb = ROW(pcw, ib);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int j = j0 + r;
      if (!FULL && j > n2m) continue;
      const double2 ga = TILE_AT(qa, r), gb = TILE_AT(qa + 1, r);
      double2 xa = sa[r];
      xa = csub(xa, rmul(awa, TILE_AT(qa - 1, r)));
      xa = csub(xa, rmul(aea, gb));
      xa = csub(xa, rmul(asj[r], TILE_DN(qa, r)));
      xa = csub(xa, rmul(anj[r], TILE_UP(qa, r)));
      xa = csub(xa, rmul(acA[r], ga));
      double2 xb = sb[r];
      xb = csub(xb, rmul(aw2, ga));
      xb = csub(xb, rmul(ae2, TILE_AT(qa + 2, r)));
      xb = csub(xb, rmul(asj[r], TILE_DN(qa + 1, r)));
      xb = csub(xb, rmul(anj[r], TILE_UP(qa + 1, r)));
      xb = csub(xb, rmul(acB[r], gb));
      pa[j] = ga;
      pb[j] = gb;
      ar = fmax(ar, fmax(hypot(xa.x, xa.y), hypot(xb.x, xb.y)));
    }
  }
  for (int s = 32; s > 0; s >>= 1) ar = fmax(ar, __shfl_down(ar, s, 64));
  if (lane == 0) wmax[w] = ar;
  __syncthreads();
  if (w == 0 && lane == 0) {
    double m = 0.0;
    for (int q = 0; q < W; ++q) m = fmax(m, wmax[q]);
    atomicMax(&mg.sumres[kv - 1], (unsigned long long)__double_as_longlong(m));
  }
}

// the test after a sweep of gsor2d (poiss_z_ft.f90:551-552, 585-587): as
// k_mg_check, and for the sweeps of the loop (id = 1) the max norm of sweep
// c of wavenumber q is kept at log[(c-1)*ld + q] for the residual rows
LW_KERNEL void k_sor_check(LW_KP(1024, 0), const MG &mg, double test, int id, double *log, int ld) {
  for (int q = mg.kv0 - 1 + threadIdx.x; q < mg.kv0 - 1 + mg.nkv; q += blockDim.x) {
    if (!mg.done[q]) {
      const int c = mg.cyc[q] + id;
      mg.cyc[q] = c;
      mg.pcsel[q] ^= 1;
      const double r = __longlong_as_double((long long)mg.sumres[q]);
      if (id == 1) log[(long)(c - 1) * ld + q] = r;
      if (r < test) mg.done[q] = 1;
      else if (c >= mg.mgitr) mg.done[q] = 2;
      if (mg.done[q]) mg.rfin[q] = r;
    }
    mg.sumres[q] = 0ull;
  }
}

#undef TILE_AT
#undef TILE_DN
#undef TILE_UP

// ---------------------------------------------------------------------------
// Every V-cycle of kv = 1 in one launch (mg.kv0 = 1, mg.nkv = 1). One
// wavenumber cannot fill the chip, and as one launch per level its cycle
// time is set by the gaps between launches rather than by its work. Here a
// grid of co-resident workgroups runs the stages of each cycle in the order
// of launch_mg_cycle (down-sweeps, the LDS coarse levels, up-sweeps,
// toplevel), every stage the device code of its own kernel (so the same
// values), with a grid-wide barrier in place of each kernel boundary. The
// loop ends when the convergence test folded into the first down-stage has
// set the wavenumber's flag (every workgroup reads the flag after that
// stage's barrier, so all leave together), or after ncyc cycles, when
// workgroup 0 runs k_mg_check's test on the last toplevel. Every barrier
// wait is bounded: a workgroup that waits longer than tlim ticks of the
// constant-rate clock, or sees another's failure, sets *mg.kerr and leaves,
// so the launch always ends; the host turns the flag into an error.
// ---------------------------------------------------------------------------
__device__ inline bool grid_barrier(const MG &mg, unsigned target, long long tlim) {
  __shared__ int ok;
  // the workgroup barrier does not wait for global stores and atomics:
  // every wavefront waits for its own first (vmcnt(0))
  __builtin_amdgcn_s_waitcnt(0x0F70);
  __syncthreads();
  if (threadIdx.x == 0 && threadIdx.y == 0) {
    // release at device scope: the workgroup's writes are visible to every
    // XCD before its arrival counts
    __hip_atomic_fetch_add(mg.kbar, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    int good = 1;
    const long long t0 = wall_clock64();
    while (__hip_atomic_load(mg.kbar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
      if (__hip_atomic_load(mg.kerr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 || wall_clock64() - t0 > tlim) {
        good = 0;
        __hip_atomic_store(mg.kerr, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    // acquire at device scope: the other workgroups' writes become visible
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    ok = good;
  }
  __syncthreads();
  return ok != 0;
}

__device__ inline int nrow_of(const MG &mg, int l) { return mg.iimg[l][2] - mg.iimg[l][1] + 1; }
__device__ inline int xcd_blocks(int ntile) { return kXcds * ((ntile + kXcds - 1) / kXcds); }

template <int R, bool FULL, typename S>
LW_KERNEL void k_mg_kv1(LW_KP(64 * kCoarseLdsWaves, 0), const Geo &g, const MG &mg, double test, int ncyc, long long tlim) {
  const int nb = gridDim.x;
  unsigned target = 0;
  bool alive = true, conv = false;
  for (int c = 0; c < ncyc && alive && !conv; ++c) {
    // down-sweeps; the first runs the test of the toplevel before it
    for (int ilev = mg.nlev - 1; ilev > mg.lc && alive; --ilev) {
      const int T = fused_tile(nrow_of(mg, ilev), mg.tile[0]), nblk = xcd_blocks(nrow_of(mg, ilev) / T);
      for (int b = blockIdx.x; b < nblk; b += nb) {
        __syncthreads();  // the previous tile's last LDS reads are done
        mg_down_body<R, FULL, S>(g, mg, ilev, T, test, ilev == mg.nlev - 1 ? c : -1, b);
      }
      alive = grid_barrier(mg, (++target) * nb, tlim);
      if (alive && ilev == mg.nlev - 1 && __hip_atomic_load(&mg.done[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
        conv = true;
        break;
      }
    }
    if (!alive || conv) break;
    if (blockIdx.x == 0) mg_coarse_lds_body<R, FULL, S>(g, mg, 0);
    alive = grid_barrier(mg, (++target) * nb, tlim);
    for (int ilev = mg.lc + 1; ilev <= mg.nlev - 1 && alive; ++ilev) {
      const int T = fused_tile(nrow_of(mg, ilev), mg.tile[1]), nblk = xcd_blocks(nrow_of(mg, ilev) / T);
      for (int b = blockIdx.x; b < nblk; b += nb) {
        __syncthreads();
        mg_up_body<R, FULL, S>(g, mg, ilev, T, b);
      }
      alive = grid_barrier(mg, (++target) * nb, tlim);
    }
    if (!alive) break;
    {
      const int T = fused_tile(g.n1m, mg.tile[2]), nblk = xcd_blocks(g.n1m / T);
      for (int b = blockIdx.x; b < nblk; b += nb) {
        __syncthreads();
        mg_top_body<R, FULL, S>(g, mg, T, 1.0, 1, b);
      }
      alive = grid_barrier(mg, (++target) * nb, tlim);
    }
  }
  // ncyc cycles without the flag: the test of the last toplevel (k_mg_check
  // with id 1 and the pc buffers flipped)
  if (alive && !conv && blockIdx.x == 0 && threadIdx.x == 0 && threadIdx.y == 0) {
    if (!mg.done[0]) {
      const int c = mg.cyc[0] + 1;
      mg.cyc[0] = c;
      mg.pcsel[0] ^= 1;
      const double r = __longlong_as_double((long long)mg.sumres[0]);
      if (r < test) mg.done[0] = 1;
      else if (c >= mg.mgitr) mg.done[0] = 2;
      if (mg.done[0]) mg.rfin[0] = r;
    }
    mg.sumres[0] = 0ull;
  }
}

LW_KERNEL void k_mg_kv1_reset(LW_KP(1024, 0), const MG &mg) {
  *mg.kbar = 0u;
  *mg.kerr = 0;
}

typedef ::lw::KFn<Geo, MG, int, int> FusedLevelKernel;
typedef ::lw::KFn<Geo, MG, int, int, double, int> FusedDownKernel;

template <bool F, typename S>
static FusedDownKernel pick_down_rf(int R) {
  switch (R) {
    case 1: return LW_KFN((k_mg_down<1, F, S>));
    case 2: return LW_KFN((k_mg_down<2, F, S>));
    case 4: return LW_KFN((k_mg_down<4, F, S>));
    case 8: return LW_KFN((k_mg_down<8, F, S>));
    default: return LW_KFN((k_mg_down<16, F, S>));
  }
}

static FusedDownKernel pick_down(int R, bool full, bool f32 = false) {
  if (f32) return full ? pick_down_rf<true, float2>(R) : pick_down_rf<false, float2>(R);
  return full ? pick_down_rf<true, double2>(R) : pick_down_rf<false, double2>(R);
}
template <bool F, typename S>
static FusedLevelKernel pick_up_rf(int R) {
  switch (R) {
    case 1: return LW_KFN((k_mg_up<1, F, S>));
    case 2: return LW_KFN((k_mg_up<2, F, S>));
    case 4: return LW_KFN((k_mg_up<4, F, S>));
    case 8: return LW_KFN((k_mg_up<8, F, S>));
    default: return LW_KFN((k_mg_up<16, F, S>));
  }
}

static FusedLevelKernel pick_up(int R, bool full, bool f32 = false) {
  if (f32) return full ? pick_up_rf<true, float2>(R) : pick_up_rf<false, float2>(R);
  return full ? pick_up_rf<true, double2>(R) : pick_up_rf<false, double2>(R);
}

typedef ::lw::KFn<Geo, MG, int, double, int> FusedTopKernel;
template <bool F, typename S>
static FusedTopKernel pick_top_rf(int R) {
  switch (R) {
    case 1: return LW_KFN((k_mg_top<1, F, S>));
    case 2: return LW_KFN((k_mg_top<2, F, S>));
    case 4: return LW_KFN((k_mg_top<4, F, S>));
    case 8: return LW_KFN((k_mg_top<8, F, S>));
    default: return LW_KFN((k_mg_top<16, F, S>));
  }
}
static FusedTopKernel pick_top(int R, bool full, bool f32 = false) {
  if (f32) return full ? pick_top_rf<true, float2>(R) : pick_top_rf<false, float2>(R);
  return full ? pick_top_rf<true, double2>(R) : pick_top_rf<false, double2>(R);
}

// LDS of the fused kernels for tile T: T+3 line slots (down, top2) or
// T/2+1 (up), 64*R rows each
// dynamic LDS of the fused kernels: the tile's line slots and, for the down
// and top kernels, the an(j) table their residual stage reads
static inline unsigned fused_lds(int R, int T, bool up) {
  return (unsigned)(up ? (T / 2 + 1) * R * 64 * (int)sizeof(double2)
                       : (T + 3) * R * 64 * (int)sizeof(double2) + 64 * R * (int)sizeof(double));
}

static void launch_down(hipStream_t s, const Geo &g, const MG &mg, int ilev, double test, int cyc) {
  const int n = mg.iimg[ilev][2] - mg.iimg[ilev][1] + 1, T = fused_tile(n, mg.tile[0]), R = lanes_r(g.n2m);
  // unique bytes per point of the level: reads resd 16, bet 8; writes the
  // second-half solution 16 on half the lines (8), the residual 16 and the
  // restricted row 16 per pair (8)
  LW_LAUNCH_K("k_mg_down", 0, (mg.f32 ? 32.0 : 56.0) * g.n2m * n, pick_down(R, full_rows(g), mg.f32),
            xcd_grid(n / T, mg.nkv), dim3(64, T / 2 + 2), fused_lds(R, T, false), s, g, mg, ilev, T, test, cyc);
}

static void launch_up(hipStream_t s, const Geo &g, const MG &mg, int ilev) {
  const int n = mg.iimg[ilev][2] - mg.iimg[ilev][1] + 1, T = fused_tile(n, mg.tile[1]), R = lanes_r(g.n2m);
  // unique bytes per point of the level: reads residual 16, bet 8, the
  // down-sweep row on half the lines (8) and the coarse row per pair (8);
  // writes the solution 16
  LW_LAUNCH_K("k_mg_up", 0, (mg.f32 ? 32.0 : 56.0) * g.n2m * n, pick_up(R, full_rows(g), mg.f32), xcd_grid(n / T, mg.nkv), dim3(64, T / 2 + 1),
            fused_lds(R, T, true), s, g, mg, ilev, T);
}

static void launch_top(hipStream_t s, const Geo &g, const MG &mg, int id, double oldv) {
  const int n = g.n1m, T = fused_tile(n, mg.tile[2]), R = lanes_r(g.n2m);
  // unique bytes per top point: reads rhs 16, bet 8, the old pc on half the
  // lines (8) and for id = 1 the coarse row per pair (8); writes the new pc
  // 16 and the restricted row per pair (8)
  // (the coarse rows take 4 and 4 in the mixed-precision form)
  const double per_pt = (id == 1 ? 64.0 : 56.0) - (mg.f32 ? (id == 1 ? 8.0 : 4.0) : 0.0);
  LW_LAUNCH_K(id == 1 ? "k_mg_top_up" : "k_mg_top", 0, per_pt * g.n2m * n, pick_top(R, full_rows(g), mg.f32),
            xcd_grid(n / T, mg.nkv), dim3(64, T / 2 + 3), fused_lds(R, T, false), s, g, mg, T, oldv, id);
}

static inline dim3 rows(int n, int nkv) { return dim3((unsigned)n, (unsigned)nkv); }

void launch_mg_toplevel(hipStream_t s, const Geo &g, const MG &mg, int id, double oldv, double test, bool check) {
  const int nl = mg.nlev;
  if (mg.fused) {
    launch_top(s, g, mg, id, oldv);
  } else {
    launch_lines(s, g, mg, id == 1 ? 5 : 2, 1, (g.n1m + 1) / 2, oldv, nl - 1);
    launch_lines(s, g, mg, 3, 2, g.n1m / 2, 1.0, 0);
    // rhs 16, solution 16, residual 16, pc 16 per point; coarse row 16 per pair
    LW_LAUNCH("k_top_rr", 0, 72.0 * g.n2m * g.n1m, k_top_rr, rows(g.n1m / 2, mg.nkv), dim3(256), 0, s, g, mg);
  }
  if (check) LW_LAUNCH("k_mg_check", 0, 0, k_mg_check, dim3(1), dim3(128), 0, s, mg, test, id, mg.fused);
}

// whether the convergence test after a toplevel may move into the next
// cycle's first launch: the fused path with a k_mg_down level above a coarse
// kernel (which clears the max norms for the next toplevel)
bool mg_folds_check(const MG &mg) { return mg.fused && mg.lc >= 0 && mg.nlev - 1 > mg.lc; }

typedef ::lw::KFn<Geo, MG, double, int, long long> Kv1Kernel;
template <bool F, typename S>
static Kv1Kernel pick_kv1_rf(int R) {
  switch (R) {
    case 1: return LW_KFN((k_mg_kv1<1, F, S>));
    case 2: return LW_KFN((k_mg_kv1<2, F, S>));
    default: return LW_KFN((k_mg_kv1<4, F, S>));  // the LDS coarse levels need R <= 4
  }
}
static Kv1Kernel pick_kv1(int R, bool full, bool f32) {
  if (f32) return full ? pick_kv1_rf<true, float2>(R) : pick_kv1_rf<false, float2>(R);
  return full ? pick_kv1_rf<true, double2>(R) : pick_kv1_rf<false, double2>(R);
}

// k_mg_kv1's workgroup is the level-0 lines' wavefronts of the LDS coarse
// kernel; every tile kernel's stages must fit in that many
static int kv1_waves(const MG &a) { return a.iimg[0][2] - a.iimg[0][1] + 1; }

bool mg_kv1_supported(const Geo &g, const MG &a) {
  if (!a.fused || !a.coarse_lds || a.lc != 1 || a.kv0 != 1 || a.nkv != 1 || !mg_folds_check(a)) return false;
  if (lanes_r(g.n2m) > 4 || coarse_lds_bytes(g, a) <= 0) return false;
  int need = fused_tile(g.n1m, a.tile[2]) / 2 + 3;
  for (int l = a.lc + 1; l <= a.nlev - 1; ++l) {
    const int n = a.iimg[l][2] - a.iimg[l][1] + 1;
    need = std::max(need, std::max(fused_tile(n, a.tile[0]) / 2 + 2, fused_tile(n, a.tile[1]) / 2 + 1));
  }
  const int nw = kv1_waves(a);
  return nw >= need && nw <= kCoarseLdsWaves;
}

typedef ::lw::KFn<Geo, MG, int, double, double> SorKernel;
template <bool F>
static SorKernel pick_sor_rf(int R) {
  switch (R) {
    case 1: return LW_KFN((k_sor<1, F>));
    case 2: return LW_KFN((k_sor<2, F>));
    case 4: return LW_KFN((k_sor<4, F>));
    case 8: return LW_KFN((k_sor<8, F>));
    default: return LW_KFN((k_sor<16, F>));
  }
}
static SorKernel pick_sor(int R, bool full) { return full ? pick_sor_rf<true>(R) : pick_sor_rf<false>(R); }

void launch_sor_sweep(hipStream_t s, const Geo &g, const MG &m, double wgt, double ww, double test, int id, double *log,
                      int ld) {
  const int n = g.n1m, T = fused_tile(n, m.tile[2]), R = lanes_r(g.n2m);
  // per top point: rhs 16, bet 8, the old u on every line 16 (and its
  // neighbours on the first half, from L2) read; the new u 16 written
  LW_LAUNCH_K("k_sor", 0, 56.0 * g.n2m * n, pick_sor(R, full_rows(g)), xcd_grid(n / T, m.nkv), dim3(64, T / 2 + 2),
            fused_lds(R, T, false), s, g, m, T, wgt, ww);
  LW_LAUNCH("k_sor_check", 0, 0, k_sor_check, dim3(1), dim3(128), 0, s, m, test, id, log, ld);
}

void launch_mg_kv1(hipStream_t s, const Geo &g, const MG &a, double test, int ncyc, long long tlim) {
  const int R = lanes_r(g.n2m);
  // dynamic LDS: the largest stage's; grid: the most tiles a stage has
  unsigned lds = (unsigned)coarse_lds_bytes(g, a);
  int tiles = g.n1m / fused_tile(g.n1m, a.tile[2]);
  lds = std::max(lds, fused_lds(R, fused_tile(g.n1m, a.tile[2]), false));
  for (int l = a.lc + 1; l <= a.nlev - 1; ++l) {
    const int n = a.iimg[l][2] - a.iimg[l][1] + 1, td = fused_tile(n, a.tile[0]), tu = fused_tile(n, a.tile[1]);
    lds = std::max(lds, std::max(fused_lds(R, td, false), fused_lds(R, tu, true)));
    tiles = std::max(tiles, std::max(n / td, n / tu));
  }
  const unsigned grid = (unsigned)(kXcds * ((tiles + kXcds - 1) / kXcds));
  LW_LAUNCH("k_mg_kv1_reset", 0, 0, k_mg_kv1_reset, dim3(1), dim3(1), 0, s, a);
  LW_LAUNCH_K("k_mg_kv1", 0, 0, pick_kv1(R, full_rows(g), a.f32 != 0), dim3(grid), dim3(64, kv1_waves(a)), lds, s, g, a,
            test, ncyc, tlim);
}

// one V-cycle. cyc >= 0: the first launch runs the convergence test of the
// preceding toplevel, after which an active wavenumber has completed cyc
// cycles (mg_folds_check); cyc < 0: a k_mg_check already ran it. check: end
// the cycle with k_mg_check.
void launch_mg_cycle(hipStream_t s, const Geo &g, const MG &mg, double test, int cyc, bool check) {
  auto nrow = [&](int l) { return mg.iimg[l][2] - mg.iimg[l][1] + 1; };
  // levels above lc: one launch per zebra half, one per residual+restriction
  for (int ilev = mg.nlev - 1; ilev >= 1 && ilev > mg.lc; --ilev) {
    const int n = nrow(ilev);
    if (mg.fused) {
      launch_down(s, g, mg, ilev, test, ilev == mg.nlev - 1 ? cyc : -1);
      continue;
    }
    launch_lines(s, g, mg, 0, mg.iimg[ilev][1], (n + 1) / 2, 0.0, 0);
    launch_lines(s, g, mg, 1, mg.iimg[ilev][1] + 1, n / 2, 1.0, 0);
    // residual in/out 32, solution 16 per fine point; coarse row 8 per fine point
    LW_LAUNCH("k_lev_rr", 0, 56.0 * g.n2m * n, k_lev_rr, rows(n / 2, mg.nkv), dim3(256), 0, s, g, mg, ilev);
  }
  if (mg.lc >= 0) {
    launch_coarse(s, g, mg);
  } else {
    const int n = nrow(0);
    for (int it = 0; it < mg.nbli; ++it) {
      launch_lines(s, g, mg, 0, mg.iimg[0][1], (n + 1) / 2, it == 0 ? 0.0 : 1.0, 0);
      launch_lines(s, g, mg, 1, mg.iimg[0][1] + 1, n / 2, 1.0, 0);
    }
  }
  for (int ilev = mg.lc < 0 ? 0 : mg.lc; ilev <= mg.nlev - 2; ++ilev) {
    const int n = nrow(ilev + 1);
    if (mg.fused) {
      launch_up(s, g, mg, ilev + 1);
      continue;
    }
    launch_lines(s, g, mg, 4, mg.iimg[ilev + 1][1], (n + 1) / 2, 1.0, ilev);
    launch_lines(s, g, mg, 1, mg.iimg[ilev + 1][1] + 1, n / 2, 1.0, 0);
  }
  launch_mg_toplevel(s, g, mg, 1, 1.0, test, check);
}

// once per handle, before the first cycle: dynamic LDS beyond 64 KiB must be
// allowed explicitly
// the fused kernels need R <= 8 (n2m <= 512) to keep their tiles in LDS
hipError_t prepare_mg_kernels(const Geo &g, MG &mg) {
  const int R = lanes_r(g.n2m);
  // tile sizes (lines per workgroup) of the down, up and top kernels: 2, 2, 4
  // (LW_MG_TILES="d,u,t" overrides them, each 2, 4 or 8); tile1 for the
  // kv = 1 side of the two-stream solve (LW_MG_TILES1, default the same)
  auto parse = [](const char *var, int (&t)[3]) {
    if (const char *e = getenv(var)) {
      int v[3] = {0, 0, 0};
      if (sscanf(e, "%d,%d,%d", &v[0], &v[1], &v[2]) == 3)
        for (int q = 0; q < 3; ++q)
          if (v[q] == 2 || v[q] == 4 || v[q] == 8) t[q] = v[q];
    }
  };
  mg.tile[0] = 2;
  mg.tile[1] = 2;
  mg.tile[2] = 4;
  parse("LW_MG_TILES", mg.tile);
  for (int q = 0; q < 3; ++q) mg.tile1[q] = mg.tile[q];
  parse("LW_MG_TILES1", mg.tile1);
  if (R > 8) mg.fused = 0;
  if (R <= 8) {
    // the line-SOR sweeps take the top kernel's tiles
    hipError_t e = ::lw::set_max_lds(pick_sor(R, full_rows(g)),
                                       (int)fused_lds(R, kTileMax, false));
    if (e != hipSuccess) return e;
  }
  if (mg.fused) {
    const int big = (int)fused_lds(R, kTileMax, false);
    hipError_t e = hipSuccess;
    for (int f32 = 0; f32 <= 1 && e == hipSuccess; ++f32) {
      e = ::lw::set_max_lds(pick_down(R, full_rows(g), f32),
                              big);
      if (e == hipSuccess)
        e = ::lw::set_max_lds(pick_up(R, full_rows(g), f32),
                                big);
      if (e == hipSuccess)
        e = ::lw::set_max_lds(pick_top(R, full_rows(g), f32),
                                big);
    }
    if (e != hipSuccess) return e;
  }
  if (!mg.coarse_lds || coarse_lds_bytes(g, mg) == 0) return hipSuccess;
  hipError_t e = ::lw::set_max_lds(pick_coarse_lds(R, full_rows(g), false), kCoarseLdsBytes);
  if (e == hipSuccess)
    e = ::lw::set_max_lds(pick_coarse_lds(R, full_rows(g), true), kCoarseLdsBytes);
  if (R > 4) return e;
  // k_mg_kv1 also holds a few static LDS words (the top stage's norms, the
  // barrier's flag): the dynamic part may take the rest
  for (int f32 = 0; f32 < 2 && e == hipSuccess; ++f32) {
    const void *fn = pick_kv1(R, full_rows(g), f32 != 0).one_;
    hipFuncAttributes at{};
    e = hipFuncGetAttributes(&at, fn);
    if (e == hipSuccess)
      e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, kCoarseLdsBytes - (int)at.sharedSizeBytes);
  }
  return e;
}

// the mixed-precision form needs every kernel that touches the level planes
// to read them as FP32: the fused level kernels and the LDS-resident coarse
// kernel (or no coarse kernel)
bool mg_mixed_supported(const Geo &g, const MG &mg) {
  return mg.fused && (mg.lc < 0 || (mg.coarse_lds && coarse_lds_bytes(g, mg) > 0));
}

}  // namespace lw
