// Kernel launchers of the LESwHT time step (gfx950). Each launcher enqueues
// on the given stream; all arrays are device pointers.
#pragma once
#include <hip/hip_ext.h>

#include "lw_launch.h"

#include "lw_internal.h"

namespace lw {

// ---- per-launch event profiler (lw_solver.hip) ----
// While a profile is active on the calling thread, every launch made through
// LW_LAUNCH of a selected kernel is launched with hipExtLaunchKernelGGL, whose
// start and stop events take the dispatch's own begin and end timestamps on
// its stream, and is recorded with its algorithmic HBM bytes: BYTES, plus
// BYTES_KV times the number of wavenumbers still iterating in the multigrid
// cycle it belongs to. BYTES and BYTES_KV are evaluated only while profiling.
struct Prof;
extern thread_local Prof *g_prof;
bool prof_take(const char *name, hipEvent_t *a, hipEvent_t *b);
void prof_record(const char *name, hipEvent_t a, hipEvent_t b, double bytes, double bytes_kv);

// KERNEL names a kernel body (lw_launch.h); LW_LAUNCH_K takes a KFn value
// chosen at run time. While the calling thread records (batched stepping),
// the launch is appended to its recorder instead.
#define LW_LAUNCH(NAME, BYTES, BYTES_KV, KERNEL, GRID, BLOCK, SHM, STREAM, ...) \
  LW_LAUNCH_K(NAME, BYTES, BYTES_KV, LW_KFN(KERNEL), GRID, BLOCK, SHM, STREAM, __VA_ARGS__)

#define LW_LAUNCH_K(NAME, BYTES, BYTES_KV, KFN, GRID, BLOCK, SHM, STREAM, ...)                      \
  do {                                                                                               \
    const auto lw_k_ = (KFN);                                                                        \
    if (::lw::g_rec != nullptr) {                                                                    \
      ::lw::rec_launch(lw_k_, NAME, GRID, BLOCK, SHM, __VA_ARGS__);                                  \
      break;                                                                                         \
    }                                                                                                \
    hipEvent_t lw_a_ = nullptr, lw_b_ = nullptr;                                                     \
    if (::lw::g_prof != nullptr && ::lw::prof_take(NAME, &lw_a_, &lw_b_)) {                          \
      hipExtLaunchKernelGGL(lw_k_.one(), GRID, BLOCK, SHM, STREAM, lw_a_, lw_b_, 0u, __VA_ARGS__);   \
      ::lw::prof_record(NAME, lw_a_, lw_b_, (double)(BYTES), (double)(BYTES_KV));                    \
    } else {                                                                                         \
      hipLaunchKernelGGL(lw_k_.one(), GRID, BLOCK, SHM, STREAM, __VA_ARGS__);                        \
    }                                                                                                \
  } while (0)

struct Fld {
  double *u, *v, *w, *p, *t;
  double *rhs[4];
  double *rk3o[4], *rk3oo[4];
  double *nu, *nu1[3], *al, *al1[3];
  double *qmass, *div, *phi;
  double *sa[3], *sb[3];  // ADI scratch per momentum component (gam, or q and s of the periodic solve)
  uint8_t *mask;
  // outlet planes (0:n2, 0:n3), index j + NY*k
  double *uout, *vout, *wout, *tout, *duout, *dvout, *dwout, *dtout;
  // reduction scratch: red[0..63] doubles, partials[...] per block
  double *red, *partial;  // partial: 4 * kRedBlocks
  double *obc;            // outlet-plane reduction partials (8 * 64)
  double *dpart;          // per-step diagnostic partials (D_NSUM * 512)
  unsigned long long *redmax;  // [0] cfl, [1] dvmax, [2] qmmax
  // channel (ICH != 0): scalars (see lw_kern_chan.hip), block partials, and
  // 1 per interior u point that funcbody counts as fluid
  double *chan, *chanpart;
  uint8_t *umask;
};

struct Body {
  int kind;
  double p0, p1, p2, p3;
};

// block-partial sum buffer size (blocks of the two-level reductions)
constexpr int kRedBlocks = 512;
constexpr int kRedThreads = 256;

// ---- momentum / continuity / projection / energy (lw_kern_flow.hip) ----
void launch_rhs_mom(hipStream_t s, const Geo &g, const Sub &sb, Fld &f, int les);
void launch_ibm(hipStream_t s, const Geo &g, const Sub &sb, const Fld &f, IBComp *ib, int comp);
void launch_ibm_mom(hipStream_t s, const Geo &g, const Sub &sb, const Fld &f, IBComp *ib);
void launch_convbc_mom(hipStream_t s, const Geo &g, const Sub &sb, const Fld &f, const Body &b, int les);
void launch_adi_mom(hipStream_t s, const Geo &g, const Sub &sb, const Fld &f, int les);
hipError_t prepare_adi_kernels(const Geo &g, int *use_wave);
void launch_retrv_mom(hipStream_t s, const Geo &g, const Fld &f);
void launch_divgs(hipStream_t s, const Geo &g, const Sub &sb, const Fld &f, int masson);
void launch_projection(hipStream_t s, const Geo &g, const Sub &sb, const Fld &f);
void launch_rhs_t(hipStream_t s, const Geo &g, const Sub &sb, Fld &f, int les);
void launch_convbc_t(hipStream_t s, const Geo &g, const Sub &sb, const Fld &f, const Body &b,
                     int les, const lw_settings &st);
void launch_adi_t(hipStream_t s, const Geo &g, const Sub &sb, const Fld &f, int les);
void launch_energy_bcs(hipStream_t s, const Geo &g, const Fld &f, const Body &b, const lw_settings &st);
void launch_cfl(hipStream_t s, const Geo &g, const Fld &f, double dt);
void launch_diag(hipStream_t s, const Geo &g, const Fld &f, int masson);
void launch_lagforce(hipStream_t s, const Geo &g, const Sub &sb, const Fld &f, IBComp *ib);

// per-step diagnostics: sums s of k_step_diag, results in red[kDiagRed + s]
enum {
  D_IBU = 0, D_IBV, D_IBW,                                  // IB forcing integrals (draglift)
  D_YB_U, D_YB_W, D_YT_U, D_YT_W, D_ZB_U, D_ZB_V, D_ZT_U, D_ZT_V,  // wall shear (draglift)
  D_HXM, D_HXP, D_HYM, D_HYP, D_HZM, D_HZP,                // boundary heat fluxes (heatflux)
  D_TFV, D_TFI, D_TSV, D_TSI,                              // bulk temperatures (writehistory)
  D_NSUM
};
constexpr int kDiagRed = 32;
// fluid tests of the boundary planes (1 = funcbody >= 1e-10): walls indexed
// i + (n1+1)*k or i + (n1+1)*j, x faces j + (n2+1)*k
struct DiagPlanes {
  const uint8_t *wyb, *wyt, *wzb, *wzt, *hxm, *hxp, *hym, *hyp, *hzm, *hzp;
};
void launch_step_diag(hipStream_t s, const Geo &g, const Fld &f, IBComp *ib, const DiagPlanes &dp, unsigned active,
                      double re);
void launch_hist_update(hipStream_t s, const Fld &f, double *hist, unsigned active, int ibmon, int heat, int hstep,
                        double time, double dt, double avg_tst, double vol);

// ---- subgrid scale (lw_kern_sgs.hip) ----
struct SgsBuf {
  double *a[24];  // interior arrays (n1m*n2m*n3m each)
  double *tmp;
};
void launch_sgs_dvmg(hipStream_t s, const Geo &g, const Fld &f, SgsBuf &b, int filter);
void launch_sgs_cvm(hipStream_t s, const Geo &g, const Fld &f, double csgsts);
// the dynamic models in one pass each, filtering on the fly (no SgsBuf);
// coefficient into red[4] (Vreman) or red[5] (eddy diffusivity).
// LW_SGS_FUSED=0 selects the separate kernels and filter launches instead.
bool sgs_fused_enabled();
// fp32: the test filter's arithmetic in FP32 (the mixed-precision form)
void launch_sgs_dvmg_fused(hipStream_t s, const Geo &g, const Fld &f, const int64_t *nwall, int filter, int fp32);
void launch_sgs_edm_fused(hipStream_t s, const Geo &g, const Fld &f, const int64_t *nwall, int filter, int fp32);
void launch_sgs_edm(hipStream_t s, const Geo &g, const Fld &f, SgsBuf &b, int filter);
void launch_sgs_cedm(hipStream_t s, const Geo &g, const Fld &f, double csgshf);
void launch_sgs_finish(hipStream_t s, const Geo &g, const Fld &f, double *nu, double **nu1,
                       int alp, int scale_slot, const lw_settings &st);

// ---- time-averaged statistics (lw_kern_stats.hip) ----
// running sums of field_avg (misc_hist.f90:287-472), interior arrays:
// 0-2 u, v, w; 3-8 uiuj (11, 12, 13, 22, 23, 33); 9 p; 10 p^2; 11-13 vorticity;
// 14-19 its products (xx, xy, xz, yy, yz, zz); 20 S:S; 21 T; 22 T^2
constexpr int kAvgArrays = 23;
struct AvgBuf {
  double *a[kAvgArrays];
};
void launch_field_avg(hipStream_t s, const Geo &g, const Fld &f, const AvgBuf &b, double dt, int heat);
// tracer: u, v, w at the centre and p of n <= 64 cells (pts: i, j, k each)
void launch_trace(hipStream_t s, const Geo &g, const Fld &f, const int *pts, int n, double *out);

// ---- Poisson (lw_kern_pois.hip) ----
void launch_fft_forward(hipStream_t s, const Geo &g, const MG &mg, const Fld &f);
void launch_fft_inverse(hipStream_t s, const Geo &g, const MG &mg, const Fld &f);
hipError_t prepare_mg_kernels(const Geo &g, MG &mg);
bool mg_mixed_supported(const Geo &g, const MG &mg);
void launch_mg_zero(hipStream_t s, const Geo &g, const MG &mg);
// toplevel (poiss_z_ft.f90:302-380), id 0 or 1, ending with the per-wavenumber
// convergence test against `test`
void launch_mg_toplevel(hipStream_t s, const Geo &g, const MG &mg, int id, double oldv, double test, bool check);
// the x-periodic channel (lw_kern_chan.hip): meanpg, the substep sum of
// pmi, the fcmfr.dat quantities, and poiss_ftft
void launch_meanpg(hipStream_t s, const Geo &g, const Fld &f, const IBComp &ibu, const lw_settings &st, double dtconst,
                   int ctrl);
void launch_pmiavg(hipStream_t s, const Fld &f, double alpha, int reset);
void launch_chan_hist(hipStream_t s, const Geo &g, const Fld &f, const IBComp &ibu, const lw_settings &st,
                      double *hist, double time, double dt);
// addperturb's tables (lw_kern_chan.hip): fluid masks of the interior u
// and w points (funcbody > 1e-10, i fastest), the profiles ptbu(j, k) and
// ptbw(i, j), the band 0.5 < ymp(j) < 0.99, and partial sums (2 * 64 + 2)
struct Ptb {
  const uint8_t *mu, *mw, *band;
  const double *pu, *pw;
  double *part;
};
void launch_addperturb(hipStream_t s, const Geo &g, const Fld &f, const Ptb &t);
// prdic_adj_uvw(0) (slv_mmtm.f90:1746-1857): the x (x-periodic grids), then
// z periodic ghost planes of u, v, w
void launch_prdic_uvw(hipStream_t s, const Geo &g, const Fld &f);
struct Ftft {
  int n3mh;
  const double *twx_r, *twx_i;  // e^{-2 pi i q / n1m}, q = 0..n1m-1
  const double *twz_r, *twz_i;  // e^{-2 pi i q / n3m}
  const double *ak1, *ak3;      // modified wavenumbers (1-based, poisinit of poiss_ftft.f90:164-204)
  double2 *a, *b;               // spectra [kz][j][kx], n3mh * n2m * n1m
  double *gam;                  // the line solve's gam, same shape
};
void launch_poisson_ftft(hipStream_t s, const Geo &g, const Fld &f, const Ftft &t);
// poiss_x_ft (lw_kern_xft.hip): x DFT, per x wavenumber a periodic-z zebra
// multigrid (mg2d) or line-SOR (gsor2d), inverse DFT
struct Xft {
  int n1mh, n3md, nlev, nbli;
  int kb[20], ke[20];  // the levels' first and last k (kkmg(l, 1:2))
  const double *ab, *af, *as, *an, *coi1, *coi2, *cor1, *cor2;  // 1-based (pmat, coefmg)
  const int *kpm, *kmm;
  const double *ac, *bet, *gam;  // [iv][j][k], k fastest over n3md
  const double *twr, *twi;       // e^{-2 pi i q / n1m}
  double2 *rhs, *pc, *gg, *resd, *tt;
  int *done, *ii, *nact;
  unsigned long long *sres;  // max-norm gathering slot per wavenumber
  double *sfin, *misc;       // the last max norm per wavenumber; misc[0] phiref
};
void launch_xft_forward(hipStream_t s, const Geo &g, const Fld &f, const Xft &x);
void launch_xft_inverse(hipStream_t s, const Geo &g, const Fld &f, const Xft &x, int ich);
void launch_xft_toplevel(hipStream_t s, const Geo &g, const Xft &x, int iv0, int nv, int id, double test, int cyc);
void launch_xft_vcycle(hipStream_t s, const Geo &g, const Xft &x, int iv0, int nv);
void launch_xft_sor(hipStream_t s, const Geo &g, const Xft &x, int iv0, int nv, double wgt, double ww, double test,
                    int cyc);
bool mg_folds_check(const MG &mg);
// kv = 1 of the two-stream solve in one launch (k_mg_kv1): whether the
// tiles and levels of `mg` (kv0 = 1, nkv = 1) allow it, and the launch of up
// to ncyc V-cycles after the first toplevel, each barrier wait bounded by
// tlim ticks of the constant-rate clock
bool mg_kv1_supported(const Geo &g, const MG &mg);
void launch_mg_kv1(hipStream_t s, const Geo &g, const MG &mg, double test, int ncyc, long long tlim);
// one sweep of gsor2d over the wavenumbers of `mg` (kv0.., nkv) with its
// test: wgt 0 for the sweep before the loop (id 0), 1 in the loop (id 1,
// which keeps each active wavenumber's max norm at log[(sweep-1)*ld + q])
void launch_sor_sweep(hipStream_t s, const Geo &g, const MG &mg, double wgt, double ww, double test, int id, double *log,
                      int ld);
// one V-cycle of every wavenumber still iterating (poiss_z_ft.f90:258-299)
void launch_mg_cycle(hipStream_t s, const Geo &g, const MG &mg, double test, int cyc, bool check);

// ---- the 3-D MG-preconditioned BiCGSTAB Poisson solver (lw_kern_mgrd.hip) ----
// scalars of M3::sc the host reads after an iteration
constexpr int kM3Resm = 8, kM3Loc = 9;
void launch_mgrd_init(hipStream_t s, const Geo &g, const M3 &m, const Fld &f, double test1);
void launch_mgrd_iteration(hipStream_t s, const Geo &g, const M3 &m, const Fld &f, int first);
void launch_mgrd_finish(hipStream_t s, const Geo &g, const M3 &m, const Fld &f);

}  // namespace lw

// Kernel bodies take their workgroup index and grid size as parameters
// (lw_launch.h): in the kernel sources these names refer to them.
#define blockIdx lw_bid_
#define gridDim lw_gdim_
