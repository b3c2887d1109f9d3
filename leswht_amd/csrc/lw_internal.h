// Internal definitions shared by the gfx950 kernels and the host driver of
// the LESwHT time step. Fields live in HBM in the reference's padded layout
// (0:n1, 0:n2, 0:n3) with i fastest, so every stencil reads coalesced along
// i and the C ABI copies arrays without reshaping.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/leswht.h"

namespace lw {

// 1-D metrics and index maps on the device (Fortran index = array index)
struct Geo {
  int n1, n2, n3, n1m, n2m, n3m;
  long NX, NY, NZ;  // n1+1, n2+1, n3+1
  double xl, yl, zl;
  const double *x, *y, *z;
  const double *c2cx, *c2cy, *c2cz, *c2cxi, *c2cyi, *c2czi;
  const double *f2fx, *f2fy, *f2fz, *f2fxi, *f2fyi, *f2fzi;
  const double *xmp, *ymp, *zmp;
  const int *ipv, *jpv, *kpv, *imv, *jmv, *kmv;
  const int *fixil, *fixjl, *fixkl, *fixiu, *fixju, *fixku;
  // ADI coefficients (slv_mmtm.f90:180-249)
  const double *aiu, *ciu, *aivw, *civw, *ajv, *cjv, *ajuw, *cjuw, *akw, *ckw, *akuv, *ckuv;
  // SGS test-filter weights cf?1(n,-1:1) stored [3*i + d + 1]
  const double *cfx1, *cfy1, *cfz1;
  int ib, jb, kb;  // i_bgpx, j_bgpy, k_bgpz
  int xprdic;      // 1: x periodic (the channel)
  int ich;         // channel driving: 0 none, 1 constant flow rate, 2 constant pressure gradient
  int adi_wave;    // sweeps with whole lines in registers: bit 1 y, bit 2 x, bit 4 z
};

__host__ __device__ inline long P(const Geo &g, int i, int j, int k) {
  return (long)i + g.NX * ((long)j + g.NY * (long)k);
}

// per-substep scalars (substepinit, slv_auxi.f90:105-123)
struct Sub {
  double dt, gam, ro, alpha, dtconst, dtconsti, acoef, acoefi, re, pr;
  int msub;
};

// bits of the per-cell mask array
enum : uint8_t {
  M_BODY_U = 1,   // u point is an IB forcing point
  M_BODY_V = 2,
  M_BODY_W = 4,
  M_BODY_T = 8,
  M_SOLID_C = 16,  // cell centre has funcbody < 1e-10 (nutzero/alpzero)
  M_NUTZERO = 32,  // listed in ibmpre_nutzero.bin (shifted by the D7 rule)
};

struct IBComp {
  int nintp = 0, nbody = 0;
  int *ifc = nullptr, *jfc = nullptr, *kfc = nullptr;  // nbody
  int *indx = nullptr;                                   // 3*nintp
  double *geomfac = nullptr;                             // 27*nintp
  double *targ = nullptr;                                // nintp
  double *fcv = nullptr, *fcvavg = nullptr, *dudtr = nullptr;
};

struct MG {
  int nlev, n1md, n3mh, nbli, mgitr;
  int kv0, nkv;  // the wavenumbers kv0 .. kv0+nkv-1 (1-based) a launch covers
  int lc;        // levels 0..lc run in one launch per V-cycle (-1: none)
  int coarse_lds;  // 1: levels 1 and 0 with LDS-resident planes when they fit
  long LDJ;      // row length of every MG plane: n2 + 1 rounded up to 8
  long PLANE;    // (n1md + 1) * LDJ, one wavenumber of the level stack
  long PCPLANE;  // (n1 + 1) * LDJ, one wavenumber of the top level only
  int iimg[21][4];
  // coefficients
  double *aw, *ae, *as, *an, *coi1, *coi2, *cor1, *cor2;
  double *awe;             // aw + ae per line: ac = -((awe + as) + an) - ak3
  double *ak3;             // modified z wavenumber squared per kv (ak3[1] = 0)
  double *bet;             // [kv][i][j], real; gam(j+1) = an(j) bet(j)
  double2 *gg, *resd;      // [kv][i][j], all levels stacked along i
  double2 *gd, *rsu;       // fused kernels: down-sweep second halves, residuals
  int fused;               // 1: one kernel per level and direction
  int f32;                 // 1: the level planes (gg, resd, gd, rsu) hold FP32 values (mixed precision)
  int tile[3], tile1[3];   // lines per workgroup of the fused down, up, top kernels (tile1: kv = 1 side)
  double2 *pc, *rhs;       // [kv][i][j], top level
  double2 *pc2;            // fused path: the other pc buffer (pcsel picks)
  unsigned long long *sumres;  // n3mh max-norm bit patterns, then tt (2 doubles)
  int *done;               // n3mh: 0 iterating, 1 converged, 2 MGITR reached
  int *cyc;                // n3mh: V-cycles completed
  int *pcsel;              // n3mh: 1 when pc2 holds the current pc
  double *rfin;            // n3mh: the max norm of the toplevel that ended each wavenumber's solve
  unsigned *kbar;          // k_mg_kv1: the grid barrier's arrival counter
  int *kerr;               // k_mg_kv1: set when a barrier wait ran out
  double *twr, *twi;       // z-DFT twiddles cos/sin(-2 pi m / n3m), m = 0..n3m-1
};

// The 3-D multigrid-preconditioned BiCGSTAB Poisson solver (poiss_mgrd.f90).
// Levels are stacked along i and k as in the reference: level l holds the
// 1-based stacked lines iimg[l][0]..iimg[l][1] (iimg[l][2] of them) and
// kkmg[l][0]..[1]; the finest level is nlev (1..n1m, 1..n3m).
struct M3 {
  int nlev, n1md, n3md, nbli;
  int iimg[21][3], kkmg[21][3];
  long GX, GY;  // gi(i, j, k) at i + GX (j + GY k): GX = n1md + 1, GY = n2 + 1
  // coefficients with the reference's 1-based index (aw(i) at aw[i])
  const double *aw, *ae, *as, *an, *ab, *af, *fidw, *fkdw, *fiup, *fkup;
  const int *ih1, *ih2, *kh1, *kh2, *il1, *il2, *kl1, *kl2, *ipm, *imm, *kpm, *kmm;
  const double *bet;    // (n1md, n2m, n3md), i fastest; gam(j) = an(j-1) bet(j-1)
  double *gi, *resid;   // (0:n1md, 0:n2, 0:n3md) and (n1md, n2m, n3md)
  double *p, *s, *y, *z, *v, *t, *res;  // BiCGSTAB vectors, the fields' padded layout
  double *sc;           // scalars, see lw_kern_mgrd.hip
  double *part;         // block partials: 4 * kRedBlocks sums, then kRedBlocks (max, index) pairs
};

}  // namespace lw
