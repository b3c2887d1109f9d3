// Kernel launchers of the LESwHT time step (gfx950). Each launcher enqueues
// on the given stream; all arrays are device pointers.
#pragma once
#include "lw_internal.h"

namespace lw {

// ---- per-launch event profiler (lw_solver.hip) ----
// While a profile is active on the calling thread, every launch made through
// LW_LAUNCH of a selected kernel is bracketed by two events on its stream and
// recorded with its algorithmic HBM bytes: BYTES, plus BYTES_KV times the
// number of wavenumbers still iterating in the multigrid cycle it belongs to.
// BYTES and BYTES_KV are evaluated only while profiling.
struct Prof;
extern thread_local Prof *g_prof;
bool prof_pre(const char *name, hipStream_t s);
void prof_post(const char *name, hipStream_t s, double bytes, double bytes_kv);

#define LW_LAUNCH(NAME, BYTES, BYTES_KV, KERNEL, GRID, BLOCK, SHM, STREAM, ...)       \
  do {                                                                                \
    const bool lw_p_ = ::lw::g_prof != nullptr && ::lw::prof_pre(NAME, STREAM);       \
    hipLaunchKernelGGL(KERNEL, GRID, BLOCK, SHM, STREAM, __VA_ARGS__);                \
    if (lw_p_) ::lw::prof_post(NAME, STREAM, (double)(BYTES), (double)(BYTES_KV));    \
  } while (0)

struct Fld {
  double *u, *v, *w, *p, *t;
  double *rhs[4];
  double *rk3o[4], *rk3oo[4];
  double *nu, *nu1[3], *al, *al1[3];
  double *qmass, *div, *phi;
  double *s1, *s2;  // ADI scratch (gam, or q and s of the periodic solve)
  uint8_t *mask;
  // outlet planes (0:n2, 0:n3), index j + NY*k
  double *uout, *vout, *wout, *tout, *duout, *dvout, *dwout, *dtout;
  // reduction scratch: red[0..63] doubles, partials[...] per block
  double *red, *partial;  // partial: 4 * kRedBlocks
  double *obc;            // outlet-plane reduction partials (8 * 64)
  unsigned long long *redmax;  // [0] cfl, [1] dvmax, [2] qmmax
};

struct Body {
  int kind;
  double p0, p1, p2, p3;
};

// block-partial sum buffer size (blocks of the two-level reductions)
constexpr int kRedBlocks = 512;
constexpr int kRedThreads = 256;

// ---- momentum / continuity / projection / energy (lw_kern_flow.hip) ----
void launch_rhs_mom(hipStream_t s, const Geo &g, const Sub &sb, const Fld &f, int les);
void launch_ibm(hipStream_t s, const Geo &g, const Sub &sb, const Fld &f, IBComp *ib, int comp);
void launch_convbc_mom(hipStream_t s, const Geo &g, const Sub &sb, const Fld &f, const Body &b, int les);
void launch_adi_mom(hipStream_t s, const Geo &g, const Sub &sb, const Fld &f, int comp, int les);
void launch_retrv_mom(hipStream_t s, const Geo &g, const Fld &f);
void launch_divgs(hipStream_t s, const Geo &g, const Sub &sb, const Fld &f, int masson);
void launch_projection(hipStream_t s, const Geo &g, const Sub &sb, const Fld &f);
void launch_rhs_t(hipStream_t s, const Geo &g, const Sub &sb, const Fld &f, int les);
void launch_convbc_t(hipStream_t s, const Geo &g, const Sub &sb, const Fld &f, const Body &b,
                     int les, const lw_settings &st);
void launch_adi_t(hipStream_t s, const Geo &g, const Sub &sb, const Fld &f, int les);
void launch_energy_bcs(hipStream_t s, const Geo &g, const Fld &f, const Body &b, const lw_settings &st);
void launch_cfl(hipStream_t s, const Geo &g, const Fld &f, double dt);
void launch_diag(hipStream_t s, const Geo &g, const Fld &f, int masson);
void launch_lagforce(hipStream_t s, const Geo &g, const Sub &sb, const Fld &f, IBComp *ib);

// ---- subgrid scale (lw_kern_sgs.hip) ----
struct SgsBuf {
  double *a[24];  // interior arrays (n1m*n2m*n3m each)
  double *tmp;
};
void launch_sgs_dvmg(hipStream_t s, const Geo &g, const Fld &f, SgsBuf &b, int filter);
void launch_sgs_cvm(hipStream_t s, const Geo &g, const Fld &f, double csgsts);
void launch_sgs_edm(hipStream_t s, const Geo &g, const Fld &f, SgsBuf &b, int filter);
void launch_sgs_cedm(hipStream_t s, const Geo &g, const Fld &f, double csgshf);
void launch_sgs_finish(hipStream_t s, const Geo &g, const Fld &f, double *nu, double **nu1,
                       int alp, int scale_slot, const lw_settings &st);

// ---- Poisson (lw_kern_pois.hip) ----
void launch_fft_forward(hipStream_t s, const Geo &g, const MG &mg, const Fld &f);
void launch_fft_inverse(hipStream_t s, const Geo &g, const MG &mg, const Fld &f);
void launch_mg_zero(hipStream_t s, const Geo &g, const MG &mg);
// toplevel (poiss_z_ft.f90:302-380), id 0 or 1, ending with the per-wavenumber
// convergence test against `test`
void launch_mg_toplevel(hipStream_t s, const Geo &g, const MG &mg, int id, double oldv, double test);
// one V-cycle of every wavenumber still iterating (poiss_z_ft.f90:258-299)
void launch_mg_cycle(hipStream_t s, const Geo &g, const MG &mg, double test);

}  // namespace lw
