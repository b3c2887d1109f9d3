// Host driver and C ABI of the LESwHT time step on MI355X (gfx950).
//
// One handle owns one case's state in HBM and enqueues the step on its own
// HIP stream in the reference's order (solver.f90:85-181). Set-up work that
// the reference does once (lhsinit, poisinit/coefmg, sgsfilterinit, the IB
// masks derived from funcbody) is computed here on the host in FP64 with the
// reference's expressions and uploaded.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <filesystem>
#include <mutex>
#include <thread>
#include <string>
#include <unordered_map>
#include <vector>

#include "lw_kernels.h"

namespace lw {
size_t static_area_bytes();
hipError_t static_area_acquire(int *dev, void **ptr, size_t *bytes);
void static_area_release(int dev);
void launch_sgs_dvmg_coef(hipStream_t s, const Geo &g, const Fld &f, const SgsBuf &b, const int64_t *nwall);
void launch_sgs_edm_coef(hipStream_t s, const Geo &g, const Fld &f, const SgsBuf &b, const int64_t *nwall);
}  // namespace lw

using namespace lw;

static thread_local std::string g_err;
static int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(x)                                                                   \
  do {                                                                              \
    hipError_t e_ = (x);                                                            \
    if (e_ != hipSuccess) {                                                         \
      g_err = std::string("HIP error: ") + hipGetErrorString(e_) + " at " #x;       \
      throw LW_ERR_DEVICE;                                                          \
    }                                                                               \
  } while (0)

// A launch the device refuses (resources, LDS size) reports only through
// the calling thread's last-error state. The step checks it after every
// phase, so a kernel that did not run ends the call with LW_ERR_DEVICE
// instead of leaving wrong fields. hipErrorNotReady is what the host's
// polling of a busy stream leaves there, not a failure.
static void check_launches(const char *where) {
  if (g_rec) return;  // a recording thread launches nothing itself
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess && e != hipErrorNotReady) {
    g_err = std::string("HIP error: ") + hipGetErrorString(e) + " in " + where;
    throw LW_ERR_DEVICE;
  }
}

// Stream operations of the step. While the calling thread records for a
// batch (g_rec set), they join its recording, and a host synchronisation is
// the batch's rendezvous: the recorded work of every case runs, and the
// read-backs have arrived when it returns.
struct lw_handle;
static void st_zero(lw_handle *h, void *p, size_t bytes);
static void st_read(lw_handle *h, void *dst, const void *src, size_t bytes);
static void st_sync(lw_handle *h);
static void st_event(lw_handle *h, hipEvent_t e);
// a batch's gate (lw_batch.hip): the cases of a batch wait for each other
// here, so that they record the same phase of the step together
static inline void st_gate() {
  if (g_rec) batch_gate();
}

// ---------------------------------------------------------------------------
// per-launch event profiler
// ---------------------------------------------------------------------------
namespace lw {
struct ProfRec {
  int cls, cycle, kv0, nkv;
  double bytes, bytes_kv;
  hipEvent_t a, b;
};
struct Prof {
  bool on = false, all = true;
  std::vector<std::string> sel;
  std::vector<std::string> names;
  std::unordered_map<std::string, int> idx;
  std::vector<int64_t> launches;
  std::vector<double> ms, bytes;
  std::vector<ProfRec> recs;
  std::vector<hipEvent_t> pool;
  size_t used = 0;
  int cycle = -1;      // multigrid cycle being enqueued (0: the first toplevel)
  int kv0 = 1, nkv = 0;  // its wavenumbers kv0 .. kv0+nkv-1 (nkv 0: all)
  size_t mg_first = 0; // first record of the Poisson solve being enqueued

  hipEvent_t take() {
    if (used == pool.size()) {
      hipEvent_t e = nullptr;
      HIPCHK(hipEventCreate(&e));
      pool.push_back(e);
    }
    return pool[used++];
  }
  int cls(const char *name) {
    auto it = idx.find(name);
    if (it != idx.end()) return it->second;
    const int c = (int)names.size();
    names.emplace_back(name);
    idx.emplace(name, c);
    launches.push_back(0);
    ms.push_back(0.0);
    bytes.push_back(0.0);
    return c;
  }
  bool selected(const char *name) const {
    if (all) return true;
    for (const auto &n : sel)
      if (n == name) return true;
    return false;
  }
  // bytes of the per-wavenumber launches of the Poisson solve just enqueued;
  // cyc[q] = V-cycles wavenumber q started
  void resolve_mg(const int *cyc, int n3mh) {
    for (size_t r = mg_first; r < recs.size(); ++r) {
      ProfRec &p = recs[r];
      if (p.bytes_kv == 0.0) continue;
      int act = 0;
      const int q0 = p.nkv > 0 ? p.kv0 - 1 : 0, q1 = p.nkv > 0 ? p.kv0 - 1 + p.nkv : n3mh;
      for (int q = q0; q < q1; ++q) act += (p.cycle == 0 || cyc[q] >= p.cycle) ? 1 : 0;
      p.bytes += p.bytes_kv * act;
      p.bytes_kv = 0.0;
    }
    cycle = -1;
    kv0 = 1;
    nkv = 0;
  }
  // after the stream drained: fold the records into the per-kernel totals
  void flush() {
    for (const ProfRec &p : recs) {
      float t = 0.f;
      HIPCHK(hipEventElapsedTime(&t, p.a, p.b));
      launches[p.cls] += 1;
      ms[p.cls] += (double)t;
      bytes[p.cls] += p.bytes;
    }
    recs.clear();
    used = 0;
    mg_first = 0;
  }
  ~Prof() {
    for (hipEvent_t e : pool) hipEventDestroy(e);
  }
};

thread_local Prof *g_prof = nullptr;

bool prof_take(const char *name, hipEvent_t *a, hipEvent_t *b) {
  if (!g_prof->selected(name)) return false;
  *a = g_prof->take();
  *b = g_prof->take();
  return true;
}

void prof_record(const char *name, hipEvent_t a, hipEvent_t b, double bytes, double bytes_kv) {
  g_prof->recs.push_back(ProfRec{g_prof->cls(name), g_prof->cycle, g_prof->kv0, g_prof->nkv, bytes, bytes_kv, a, b});
}
}  // namespace lw

// activates a handle's profile on this thread for the duration of one call
struct ProfScope {
  explicit ProfScope(Prof *p) { g_prof = (p && p->on) ? p : nullptr; }
  ~ProfScope() { g_prof = nullptr; }
};

struct lw_handle {
  hipStream_t s = nullptr;
  hipStream_t s2 = nullptr;  // the Poisson solve of kv = 1, beside the other wavenumbers
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  lw_settings st;
  lw_body body;
  int les = 0, heat = 0, ncomp = 3;
  Geo g;
  Fld f;
  MG mg;
  IBComp ib[4];
  SgsBuf sb;
  int64_t *nwall = nullptr;
  std::vector<void *> allocs;
  int64_t bytes = 0;
  // host copies of metrics (Fortran indices)
  std::vector<double> xmp, ymp, zmp, x;
  // scalars
  int64_t ntime = 0, ihist = 0, m = 0;
  double dt = 0, time = 0, gamma[4] = {0}, ro[4] = {0}, cflmax = 0, dvmax = 0, qmmax = 0;
  double csgsts = 0, csgshf = 0;
  int64_t cyc_max = 0, cyc_sum = 0;
  Sub sub;
  hipEvent_t ev[8];
  unsigned long long *h_sumres = nullptr;  // pinned
  int *h_flags = nullptr;                  // pinned copy of mg.done / mg.cyc
  int *h_flags2 = nullptr;                 // the same, read on s2
  int mg_hint = 4;                         // V-cycles of the previous Poisson solve
  int mg_hint1 = 4;                        // of its kv = 1 alone (two-stream solve)
  // kv = 1 of the two-stream solve as one launch of k_mg_kv1 (LW_MG_KV1=1
  // or option "mg_kv1" 1; off by default: on MI355X its stages and grid
  // barriers take longer than the launches per level they replace), its
  // tiles (LW_MG_KV1_TILES) and the bound on each barrier wait in clock ticks
  int mg_kv1 = 0;
  int kv1_tiles[3] = {8, 8, 8};
  long long kv1_tlim = 0;
  int mg_split = 1;                        // kv = 1 on its own stream (LW_MG_SPLIT=0: one stream)
  int sgs_fused = 1;                       // dynamic SGS in one pass (LW_SGS_FUSED=0: separate kernels)
  int sgs_fp32 = 0;                        // the fused models' test filter in FP32 (option "sgs_fp32")
  DiagPlanes dplanes{};
  double *hist = nullptr;  // device: the step's history values and their running averages
  double h_hist[32] = {0};
  // mgftresiduemax.dat rows of the last step (poiss_z_ft.f90:293-297): per
  // substep, kv = 1..min(3, n3mh) with its cycle count and final max norm
  int64_t mgres_n = 0;
  int64_t mgres_kv[9] = {0}, mgres_kk[9] = {0};
  double mgres_rm[9] = {0};
  // every row of mgftresiduemax.dat of the last step in file order: mg2d's
  // (kind 0, RM) and gsor2d's per sweep (kind 1, RG; poiss_z_ft.f90:584)
  std::vector<int64_t> rrow_kind, rrow_kv, rrow_kk;
  std::vector<double> rrow_val;
  // gsor2d for kv > IMGSOR: the max norm of every sweep of its loop on the
  // device, [sweep][kv]; the host copy; the sweeps of the last solve
  double *sorlog = nullptr;
  std::vector<double> h_sorlog;
  int sor_hint = 1;
  double *h_rfin = nullptr;  // pinned: rfin of kv = 1..3 after a solve
  double hs_rfin[4] = {0};
  // poiss_mgrd: the solver's state and the log of the last step
  M3 m3{};
  // poiss_ftft: tables and spectra
  Ftft ftft{};
  // poiss_x_ft: tables, level planes and per-wavenumber state
  Xft xft{};
  // addperturb (IREAD = 0, EPS_PTR != 0): tables; pending before the first
  // step when PTB_TST <= 0; applied once
  Ptb ptb{};
  int ptb_armed = 0, ptb_pending = 0, ptb_applied = 0;
  std::vector<int64_t> bicg_sub, bicg_icg, bicg_loc, solve_icg;
  std::vector<double> bicg_rm, solve_rm;
  // tracer (misc_hist.f90:184-209): cells whose centre values are traced
  int64_t ntrace = 0, ntr = 1;
  int64_t trpts[64 * 3] = {0};
  int *d_trpts = nullptr;
  double *d_trace = nullptr;
  int hist_written = 0;    // the last step was a history step (mod(ntime, npin) = 0)
  hipEvent_t tev[3][8] = {};  // phase boundaries per substep: sgs, rhsnlhs, divgs+poisson, energy
  double s_phase[4] = {0, 0, 0, 0};
  bool tev_valid = false;  // tev hold a whole step recorded by lw_step
  Prof *prof = nullptr;
  struct lw_batch *batch = nullptr;        // the batch the handle steps in, if any
  double *h_red = nullptr;                 // pinned
  // lw_create_static: every device array comes from the device's static
  // area (arena, a device address) instead of hipMalloc, and the host copies
  // live in the handle itself instead of pinned memory
  unsigned char *arena = nullptr;
  size_t arena_cap = 0, arena_used = 0;
  int arena_dev = -1;
  unsigned long long hs_sumres[256] = {0};
  int hs_flags[256] = {0}, hs_flags2[512] = {0};
  double hs_red[64] = {0};
  // field_avg state (misc_hist.f90:287-527; mod_common avg_started,
  // npriavg_count, timeinit, ihistinit); avg_dir empty: no fav files
  AvgBuf avg{};
  int navg = 0;
  bool avg_started = false;
  int64_t avg_count = 0, avg_ihistinit = 0;
  double avg_timeinit = 0;
  std::string avg_dir;

  template <class T>
  T *dalloc(size_t n) {
    void *p = nullptr;
    size_t b = sizeof(T) * (n ? n : 1);
    if (arena) {
      const size_t off = (arena_used + 255) & ~(size_t)255;
      if (off + b > arena_cap) {
        g_err = "lw_create_static: the case needs more than the " + std::to_string(arena_cap >> 20) +
                " MiB static device area (use lw_create)";
        throw LW_ERR_ARG;
      }
      p = arena + off;
      arena_used = off + b;
      HIPCHK(hipMemsetAsync(p, 0, b, s));
      bytes += (int64_t)b;
      return (T *)p;
    }
    HIPCHK(hipMalloc(&p, b));
    HIPCHK(hipMemsetAsync(p, 0, b, s));
    allocs.push_back(p);
    bytes += (int64_t)b;
    return (T *)p;
  }
  template <class T>
  T *upload(const std::vector<T> &v) {
    T *p = dalloc<T>(v.size());
    if (!v.empty()) HIPCHK(hipMemcpyAsync(p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, s));
    return p;
  }
};

// The body's level set, > 0 in fluid, evaluated on the host at set-up (solid
// masks, outlet and inlet cells, zero-viscosity cells). Kind 0 is the
// reference's cylinder (geometry/funcbody_example/funcbody_cylinder.f90:13);
// kinds 1 (square: half side p[0], angle p[1]) and 2 (ellipse: semi-axes
// p[0], p[1], angle p[2]) are funcbody examples of the same form, evaluated
// with the case builder's operations so the masks agree with its IB tables.
static double funcbody(const lw_body &b, double x, double y, double z) {
  (void)z;
  if (b.kind == 1 || b.kind == 2) {
    const double ang = b.kind == 1 ? b.p[1] : b.p[2];
    const double c = cos(ang), s = sin(ang);
    const double xr = c * x + s * y, yr = (-s) * x + c * y;
    if (b.kind == 1) return fmax(fabs(xr), fabs(yr)) - b.p[0];
    const double ta = xr / b.p[0], tb = yr / b.p[1];
    return (sqrt(ta * ta + tb * tb) - 1.0) * (b.p[0] < b.p[1] ? b.p[0] : b.p[1]);
  }
  return sqrt(x * x + y * y) - b.p[0];
}

static std::vector<double> vd(const double *src, int64_t lo, int64_t hi) {
  std::vector<double> v((size_t)(hi + 2), 0.0);
  for (int64_t q = lo; q <= hi; ++q) v[q] = src[q];
  return v;
}
static std::vector<int> vi(const int64_t *src, int64_t lo, int64_t hi) {
  std::vector<int> v((size_t)(hi + 2), 0);
  for (int64_t q = lo; q <= hi; ++q) v[q] = (int)src[q];
  return v;
}

// ---------------------------------------------------------------------------
// set-up: lhsinit, poisinit/pmat/mwavenumber/coefmg, sgsfilterinit
// ---------------------------------------------------------------------------
// poisinit, mgcoef and tridcoef of poiss_mgrd.f90 (678-993): the 1-D
// coefficients of every level (levels stacked along i and k), the
// restriction and prolongation weights and indices, and the TDMA pivots of
// every y-line of every level, in the reference's expressions and order.
// At a level's first and last line the reference multiplies a quotient by
// 0 (iwest, ieast, kback, kforw) whose divisor holds a centre of another
// level or one not yet set; that coefficient is 0 here.
static void setup_ftft(lw_handle *h, const lw_grid *gr);
static void setup_xft(lw_handle *h, const lw_grid *gr);
static void setup_mgrd(lw_handle *h, const lw_grid *gr) {
  const lw_settings &st = h->st;
  const Geo &g = h->g;
  M3 &m = h->m3;
  const int n1 = g.n1, n2 = g.n2, n3 = g.n3, n1m = g.n1m, n2m = g.n2m, n3m = g.n3m;
  const int nlev = (int)st.nlev, n1md = 2 * n1m, n3md = 2 * n3m;
  m.nlev = nlev;
  m.n1md = n1md;
  m.n3md = n3md;
  m.nbli = (int)st.nbli;
  m.GX = n1md + 1;
  m.GY = n2 + 1;
  auto c2cxi = vd(gr->c2cxi, 0, n1), c2cyi = vd(gr->c2cyi, 0, n2), c2czi = vd(gr->c2czi, 0, n3);
  auto f2fxi = vd(gr->f2fxi, 0, n1), f2fyi = vd(gr->f2fyi, 0, n2), f2fzi = vd(gr->f2fzi, 0, n3);
  auto zmp = vd(gr->zmp, 0, n3);
  auto x = vd(gr->x, 1, n1), z = vd(gr->z, 1, n3);
  auto ipv = vi(gr->ipv, 1, n1m), jpv = vi(gr->jpv, 1, n2m), kpv = vi(gr->kpv, 1, n3m);
  auto fixil = vi(gr->fixil, 1, n1m), fixjl = vi(gr->fixjl, 1, n2m), fixkl = vi(gr->fixkl, 1, n3m);
  auto fixiu = vi(gr->fixiu, 1, n1m), fixju = vi(gr->fixju, 1, n2m), fixku = vi(gr->fixku, 1, n3m);
  std::vector<double> aw(n1md + 2, 0.0), ae(n1md + 2, 0.0), as(n2m + 2, 0.0), an(n2m + 2, 0.0), ab(n3md + 2, 0.0),
      af(n3md + 2, 0.0), fidw(n1md + 2, 0.0), fiup(n1md + 2, 0.0), fkdw(n3md + 2, 0.0), fkup(n3md + 2, 0.0);
  std::vector<int> ih1(n1md + 2, 0), ih2(n1md + 2, 0), il1(n1md + 2, 0), il2(n1md + 2, 0), ipm(n1md + 2, 0),
      imm(n1md + 2, 0), kh1(n3md + 2, 0), kh2(n3md + 2, 0), kl1(n3md + 2, 0), kl2(n3md + 2, 0), kpm(n3md + 2, 0),
      kmm(n3md + 2, 0);
  for (int i = 1; i <= n1m; ++i) {
    const int ip = ipv[i];
    aw[i] = (1.0 - fixil[i]) * c2cxi[i] * f2fxi[i];
    ae[i] = (1.0 - fixiu[i]) * c2cxi[ip] * f2fxi[i];
  }
  for (int j = 1; j <= n2m; ++j) {
    const int jp = jpv[j];
    as[j] = (1.0 - fixjl[j]) * c2cyi[j] * f2fyi[j];
    an[j] = (1.0 - fixju[j]) * c2cyi[jp] * f2fyi[j];
  }
  for (int k = 1; k <= n3m; ++k) {
    const int kp = kpv[k];
    ab[k] = (1.0 - fixkl[k]) * c2czi[k] * f2fzi[k];
    af[k] = (1.0 - fixku[k]) * c2czi[kp] * f2fzi[k];
  }
  // level ranges: [l][0] first, [1] last, [2] count (1-based stacked lines)
  int(*iimg)[3] = m.iimg, (*kkmg)[3] = m.kkmg;
  iimg[nlev][0] = 1; iimg[nlev][1] = n1m; iimg[nlev][2] = n1m;
  kkmg[nlev][0] = 1; kkmg[nlev][1] = n3m; kkmg[nlev][2] = n3m;
  for (int l = nlev - 1; l >= 0; --l) {
    iimg[l][0] = iimg[l + 1][0] + iimg[l + 1][2];
    iimg[l][2] = (n1m / (1 << nlev)) * (1 << l);
    iimg[l][1] = iimg[l][0] + iimg[l][2] - 1;
    kkmg[l][0] = kkmg[l + 1][0] + kkmg[l + 1][2];
    kkmg[l][2] = (n3m / (1 << nlev)) * (1 << l);
    kkmg[l][1] = kkmg[l][0] + kkmg[l][2] - 1;
  }
  std::vector<double> xmpm(n1md + 2, 0.0), zmpm(n3md + 2, 0.0);
  std::vector<int> iwest(n1md + 2, 0), ieast(n1md + 2, 0), kback(n3md + 2, 0), kforw(n3md + 2, 0);
  for (int i = 1; i <= n1m; ++i) {
    xmpm[i] = h->xmp[i];
    iwest[i] = 1 - 1 / i;
    ieast[i] = 1 - i / n1m;
  }
  for (int k = 1; k <= n3m; ++k) {
    zmpm[k] = zmp[k];
    kback[k] = 1 - 1 / k;
    kforw[k] = 1 - k / n3m;
  }
  for (int l = nlev - 1; l >= 0; --l) {
    const int ibg = iimg[l][0], iend = iimg[l][1], isp = 1 << (nlev - l);
    const int kbg = kkmg[l][0], kend = kkmg[l][1], ksp = 1 << (nlev - l);
    for (int i = ibg, ic = 1; i <= iend; ++i, ++ic) {
      xmpm[i] = 0.5 * (x[(ic - 1) * isp + 1] + x[ic * isp + 1]);
      iwest[i] = 1 - ibg / i;
      ieast[i] = 1 - i / iend;
    }
    for (int k = kbg, kc = 1; k <= kend; ++k, ++kc) {
      zmpm[k] = 0.5 * (z[(kc - 1) * ksp + 1] + z[kc * ksp + 1]);
      kback[k] = 1 - kbg / k;
      kforw[k] = 1 - k / kend;
    }
    for (int i = ibg, ic = 1; i <= iend; ++i, ++ic) {
      const double w = x[ic * isp + 1] - x[(ic - 1) * isp + 1];
      aw[i] = iwest[i] ? 1.0 / ((xmpm[i] - xmpm[i - 1]) * w) : 0.0;
      ae[i] = ieast[i] ? 1.0 / ((xmpm[i + 1] - xmpm[i]) * w) : 0.0;
    }
    for (int k = kbg, kc = 1; k <= kend; ++k, ++kc) {
      const double w = z[kc * ksp + 1] - z[(kc - 1) * ksp + 1];
      ab[k] = kback[k] ? 1.0 / ((zmpm[k] - zmpm[k - 1]) * w) : 0.0;
      af[k] = kforw[k] ? 1.0 / ((zmpm[k + 1] - zmpm[k]) * w) : 0.0;
    }
  }
  for (int l = 0; l <= nlev - 1; ++l) {
    int ibgh = iimg[l + 1][0];
    for (int i = iimg[l][0]; i <= iimg[l][1]; ++i, ibgh += 2) {
      fidw[i] = (xmpm[i] - xmpm[ibgh]) / (xmpm[ibgh + 1] - xmpm[ibgh]);
      ih1[i] = ibgh;
      ih2[i] = ibgh + 1;
    }
    int kbgh = kkmg[l + 1][0];
    for (int k = kkmg[l][0]; k <= kkmg[l][1]; ++k, kbgh += 2) {
      fkdw[k] = (zmpm[k] - zmpm[kbgh]) / (zmpm[kbgh + 1] - zmpm[kbgh]);
      kh1[k] = kbgh;
      kh2[k] = kbgh + 1;
    }
  }
  for (int l = 1; l <= nlev; ++l) {
    int ibgl = iimg[l - 1][0];
    for (int i = iimg[l][0]; i <= iimg[l][1]; i += 2, ibgl += 1) {
      fiup[i] = (xmpm[i] - xmpm[ibgl - 1]) / (xmpm[ibgl] - xmpm[ibgl - 1]);
      fiup[i + 1] = (xmpm[i + 1] - xmpm[ibgl]) / (xmpm[ibgl + 1] - xmpm[ibgl]);
      il1[i] = ibgl - 1;
      il2[i] = ibgl;
      il1[i + 1] = ibgl;
      il2[i + 1] = ibgl + 1;
    }
    fiup[iimg[l][0]] = 1.0;
    fiup[iimg[l][1]] = 0.0;
    int kbgl = kkmg[l - 1][0];
    for (int k = kkmg[l][0]; k <= kkmg[l][1]; k += 2, kbgl += 1) {
      fkup[k] = (zmpm[k] - zmpm[kbgl - 1]) / (zmpm[kbgl] - zmpm[kbgl - 1]);
      fkup[k + 1] = (zmpm[k + 1] - zmpm[kbgl]) / (zmpm[kbgl + 1] - zmpm[kbgl]);
      kl1[k] = kbgl - 1;
      kl2[k] = kbgl;
      kl1[k + 1] = kbgl;
      kl2[k + 1] = kbgl + 1;
    }
    fkup[kkmg[l][0]] = 1.0;
    fkup[kkmg[l][1]] = 0.0;
  }
  for (int l = nlev; l >= 0; --l) {
    for (int k = kkmg[l][0]; k <= kkmg[l][1]; ++k) { kpm[k] = k + 1; kmm[k] = k - 1; }
    for (int i = iimg[l][0]; i <= iimg[l][1]; ++i) { ipm[i] = i + 1; imm[i] = i - 1; }
  }
  if (st.zprdic == 1) {
    for (int l = nlev; l >= 0; --l) {
      kpm[kkmg[l][1]] = kkmg[l][0];
      kmm[kkmg[l][0]] = kkmg[l][1];
    }
    for (int l = nlev - 1; l >= 0; --l) {
      const int kbg = kkmg[l][0], kend = kkmg[l][1], ksp = 1 << (nlev - l);
      const double vdz = zmpm[kbg] - z[1] + z[n3] - zmpm[kend];
      const double sdz_kbg = z[1 + ksp] - z[1], sdz_kend = z[n3] - z[n3 - ksp];
      ab[kbg] = 1.0 / (vdz * sdz_kbg);
      af[kend] = 1.0 / (vdz * sdz_kend);
    }
    for (int l = 1; l <= nlev; ++l) {
      const int kbg = kkmg[l][0], kend = kkmg[l][1], kbgl = kkmg[l - 1][0], kendl = kkmg[l - 1][1];
      const double dz = zmpm[kbgl] - z[1] + z[n3] - zmpm[kendl];
      const double dz1 = zmpm[kend] - zmpm[kendl];
      const double dz2 = z[n3] - zmpm[kendl] + zmpm[kbg] - z[1];
      fkup[kbg] = dz2 / dz;
      fkup[kend] = dz1 / dz;
      kl1[kbg] = kendl;
      kl2[kbg] = kbgl;
      kl1[kend] = kendl;
      kl2[kend] = kbgl;
    }
  }
  // tridcoef: pivots of every line; gam(j) = an(j-1) bet(j-1) is formed
  // again by the kernels with the same product
  std::vector<double> bet((size_t)n1md * n2m * n3md, 0.0);
  auto B = [&](int i, int j, int k) -> double & {
    return bet[(size_t)(i - 1) + (size_t)n1md * ((size_t)(j - 1) + (size_t)n2m * (size_t)(k - 1))];
  };
  for (int l = 0; l <= nlev; ++l)
    for (int k = kkmg[l][0]; k <= kkmg[l][1]; ++k)
      for (int i = iimg[l][0]; i <= iimg[l][1]; ++i) {
        const double acc1 = -1.0 * (aw[i] + ae[i] + as[1] + an[1] + ab[k] + af[k]);
        B(i, 1, k) = 1.0 / acc1;
        for (int j = 2; j <= n2m; ++j) {
          const double acc = -1.0 * (aw[i] + ae[i] + as[j] + an[j] + ab[k] + af[k]);
          const double gam = an[j - 1] * B(i, j - 1, k);
          B(i, j, k) = 1.0 / (acc - as[j] * gam);
        }
      }
  m.aw = h->upload(aw); m.ae = h->upload(ae); m.as = h->upload(as); m.an = h->upload(an);
  m.ab = h->upload(ab); m.af = h->upload(af);
  m.fidw = h->upload(fidw); m.fkdw = h->upload(fkdw); m.fiup = h->upload(fiup); m.fkup = h->upload(fkup);
  m.ih1 = h->upload(ih1); m.ih2 = h->upload(ih2); m.kh1 = h->upload(kh1); m.kh2 = h->upload(kh2);
  m.il1 = h->upload(il1); m.il2 = h->upload(il2); m.kl1 = h->upload(kl1); m.kl2 = h->upload(kl2);
  m.ipm = h->upload(ipm); m.imm = h->upload(imm); m.kpm = h->upload(kpm); m.kmm = h->upload(kmm);
  m.bet = h->upload(bet);
  // gi and resid start at 0 and every value read from them later is finite
  m.gi = h->dalloc<double>((size_t)m.GX * m.GY * (n3md + 1));
  m.resid = h->dalloc<double>((size_t)n1md * n2m * n3md);
  const size_t np = (size_t)g.NX * g.NY * g.NZ;
  m.p = h->dalloc<double>(np); m.s = h->dalloc<double>(np); m.y = h->dalloc<double>(np);
  m.z = h->dalloc<double>(np); m.v = h->dalloc<double>(np); m.t = h->dalloc<double>(np);
  m.res = h->dalloc<double>(np);
  m.sc = h->dalloc<double>(16);
  m.part = h->dalloc<double>(4 * kRedBlocks);
}

static void set_mg_fp32(lw_handle *h, bool on);

static void setup(lw_handle *h, const lw_grid *gr, const lw_tables *tb) {
  const lw_settings &st = h->st;
  Geo &g = h->g;
  g.n1 = (int)gr->n1; g.n2 = (int)gr->n2; g.n3 = (int)gr->n3;
  g.n1m = (int)gr->n1m; g.n2m = (int)gr->n2m; g.n3m = (int)gr->n3m;
  g.NX = g.n1 + 1; g.NY = g.n2 + 1; g.NZ = g.n3 + 1;
  g.xl = gr->xl; g.yl = gr->yl; g.zl = gr->zl;
  const int n1 = g.n1, n2 = g.n2, n3 = g.n3, n1m = g.n1m, n2m = g.n2m, n3m = g.n3m;
  auto c2cx = vd(gr->c2cx, 0, n1), c2cy = vd(gr->c2cy, 0, n2), c2cz = vd(gr->c2cz, 0, n3);
  auto c2cxi = vd(gr->c2cxi, 0, n1), c2cyi = vd(gr->c2cyi, 0, n2), c2czi = vd(gr->c2czi, 0, n3);
  auto f2fx = vd(gr->f2fx, 0, n1), f2fy = vd(gr->f2fy, 0, n2), f2fz = vd(gr->f2fz, 0, n3);
  auto f2fxi = vd(gr->f2fxi, 0, n1), f2fyi = vd(gr->f2fyi, 0, n2), f2fzi = vd(gr->f2fzi, 0, n3);
  h->xmp = vd(gr->xmp, 0, n1); h->ymp = vd(gr->ymp, 0, n2); h->zmp = vd(gr->zmp, 0, n3);
  h->x = vd(gr->x, 1, n1);
  auto y = vd(gr->y, 1, n2), z = vd(gr->z, 1, n3);
  auto ipv = vi(gr->ipv, 1, n1m), jpv = vi(gr->jpv, 1, n2m), kpv = vi(gr->kpv, 1, n3m);
  auto imv = vi(gr->imv, 1, n1m), jmv = vi(gr->jmv, 1, n2m), kmv = vi(gr->kmv, 1, n3m);
  auto fixil = vi(gr->fixil, 1, n1m), fixjl = vi(gr->fixjl, 1, n2m), fixkl = vi(gr->fixkl, 1, n3m);
  auto fixiu = vi(gr->fixiu, 1, n1m), fixju = vi(gr->fixju, 1, n2m), fixku = vi(gr->fixku, 1, n3m);
  g.x = h->upload(h->x); g.y = h->upload(y); g.z = h->upload(z);
  g.c2cx = h->upload(c2cx); g.c2cy = h->upload(c2cy); g.c2cz = h->upload(c2cz);
  g.c2cxi = h->upload(c2cxi); g.c2cyi = h->upload(c2cyi); g.c2czi = h->upload(c2czi);
  g.f2fx = h->upload(f2fx); g.f2fy = h->upload(f2fy); g.f2fz = h->upload(f2fz);
  g.f2fxi = h->upload(f2fxi); g.f2fyi = h->upload(f2fyi); g.f2fzi = h->upload(f2fzi);
  g.xmp = h->upload(h->xmp); g.ymp = h->upload(h->ymp); g.zmp = h->upload(h->zmp);
  g.ipv = h->upload(ipv); g.jpv = h->upload(jpv); g.kpv = h->upload(kpv);
  g.imv = h->upload(imv); g.jmv = h->upload(jmv); g.kmv = h->upload(kmv);
  g.fixil = h->upload(fixil); g.fixjl = h->upload(fixjl); g.fixkl = h->upload(fixkl);
  g.fixiu = h->upload(fixiu); g.fixju = h->upload(fixju); g.fixku = h->upload(fixku);
  // lhsinit (slv_mmtm.f90:180-249)
  g.ib = st.xprdic == 1 ? 1 : 2;
  g.xprdic = st.xprdic == 1 ? 1 : 0;
  g.ich = (int)st.ich;
  g.jb = st.yprdic == 1 ? 1 : 2;
  g.kb = st.zprdic == 1 ? 1 : 2;
  // 32 zero entries past the end: the chunked line sweeps read up to two
  // chunks beyond a line's last point (values never used)
  const int pad = 32;
  std::vector<double> aiu(n1 + 2 + pad), ciu(n1 + 2 + pad), aivw(n1 + 2 + pad), civw(n1 + 2 + pad),
      ajv(n2 + 2 + pad), cjv(n2 + 2 + pad), ajuw(n2 + 2 + pad), cjuw(n2 + 2 + pad), akw(n3 + 2 + pad),
      ckw(n3 + 2 + pad), akuv(n3 + 2 + pad), ckuv(n3 + 2 + pad);
  for (int ic = g.ib; ic <= n1m; ++ic) {
    const int im = imv[ic];
    aiu[ic] = -c2cxi[ic] * f2fxi[im];
    ciu[ic] = -c2cxi[ic] * f2fxi[ic];
  }
  for (int ic = 1; ic <= n1m; ++ic) {
    const int ip = ipv[ic];
    aivw[ic] = -c2cxi[ic] * f2fxi[ic];
    civw[ic] = -c2cxi[ip] * f2fxi[ic];
  }
  for (int jc = g.jb; jc <= n2m; ++jc) {
    const int jm = jmv[jc];
    ajv[jc] = -c2cyi[jc] * f2fyi[jm];
    cjv[jc] = -c2cyi[jc] * f2fyi[jc];
  }
  for (int jc = 1; jc <= n2m; ++jc) {
    const int jp = jpv[jc];
    ajuw[jc] = -c2cyi[jc] * f2fyi[jc];
    cjuw[jc] = -c2cyi[jp] * f2fyi[jc];
  }
  for (int kc = g.kb; kc <= n3m; ++kc) {
    const int km = kmv[kc];
    akw[kc] = -c2czi[kc] * f2fzi[km];
    ckw[kc] = -c2czi[kc] * f2fzi[kc];
  }
  for (int kc = 1; kc <= n3m; ++kc) {
    const int kp = kpv[kc];
    akuv[kc] = -c2czi[kc] * f2fzi[kc];
    ckuv[kc] = -c2czi[kp] * f2fzi[kc];
  }
  g.aiu = h->upload(aiu); g.ciu = h->upload(ciu); g.aivw = h->upload(aivw); g.civw = h->upload(civw);
  g.ajv = h->upload(ajv); g.cjv = h->upload(cjv); g.ajuw = h->upload(ajuw); g.cjuw = h->upload(cjuw);
  g.akw = h->upload(akw); g.ckw = h->upload(ckw); g.akuv = h->upload(akuv); g.ckuv = h->upload(ckuv);
  // sgsfilterinit (sgs.f90:2-285), cf?1 only (the filter uses no cf?2)
  auto filt = [](const std::vector<double> &c2c, int nm) {
    std::vector<double> cf(3 * (nm + 2), 0.0);
    if (nm != 1)
      for (int i = 1; i <= nm; ++i) {
        const double h1 = c2c[i], h2 = c2c[i + 1];
        cf[3 * i + 0] = (2. * h1 - h2) / 6. / h1;
        cf[3 * i + 1] = ((h1 + h2) * (h1 + h2)) / 6. / h1 / h2;
        cf[3 * i + 2] = (2. * h2 - h1) / 6. / h2;
      }
    return cf;
  };
  g.cfx1 = h->upload(filt(c2cx, n1m));
  g.cfy1 = h->upload(filt(c2cy, n2m));
  g.cfz1 = h->upload(filt(c2cz, n3m));

  // ---- fields ----
  const size_t np = (size_t)g.NX * g.NY * g.NZ, ni = (size_t)n1m * n2m * n3m, nb = (size_t)g.NY * g.NZ;
  Fld &f = h->f;
  f.u = h->dalloc<double>(np); f.v = h->dalloc<double>(np); f.w = h->dalloc<double>(np);
  f.p = h->dalloc<double>(np); f.t = h->dalloc<double>(np);
  for (int c = 0; c < 4; ++c) {
    f.rhs[c] = (c < 3 || h->heat) ? h->dalloc<double>(np) : nullptr;
    f.rk3o[c] = (c < 3 || h->heat) ? h->dalloc<double>(np) : nullptr;
    f.rk3oo[c] = (c < 3 || h->heat) ? h->dalloc<double>(np) : nullptr;
  }
  f.nu = h->dalloc<double>(np);
  f.al = h->dalloc<double>(np);
  for (int c = 0; c < 3; ++c) {
    f.nu1[c] = h->dalloc<double>(np);
    f.al1[c] = h->dalloc<double>(np);
  }
  f.qmass = h->dalloc<double>(np);
  f.div = h->dalloc<double>(np);
  f.phi = h->dalloc<double>(np);
  for (int c = 0; c < 3; ++c) {
    f.sa[c] = h->dalloc<double>(np);
    f.sb[c] = h->dalloc<double>(np);
  }
  f.uout = h->dalloc<double>(nb); f.vout = h->dalloc<double>(nb); f.wout = h->dalloc<double>(nb);
  f.tout = h->dalloc<double>(nb); f.duout = h->dalloc<double>(nb); f.dvout = h->dalloc<double>(nb);
  f.dwout = h->dalloc<double>(nb); f.dtout = h->dalloc<double>(nb);
  f.red = h->dalloc<double>(64);
  f.partial = h->dalloc<double>(4 * kRedBlocks);
  f.obc = h->dalloc<double>(8 * 64);
  f.dpart = h->dalloc<double>(D_NSUM * 512);
  h->hist = h->dalloc<double>(32);
  h->d_trpts = h->dalloc<int>(64 * 3);
  h->d_trace = h->dalloc<double>(64 * 4);
  f.redmax = h->dalloc<unsigned long long>(8);
  if (h->st.iavg == 1) {
    h->navg = h->heat ? kAvgArrays : kAvgArrays - 2;
    for (int l = 0; l < h->navg; ++l) h->avg.a[l] = h->dalloc<double>(ni);
  }
  if (h->les) {
    if (!h->sgs_fused) {  // the separate-kernel path keeps 24 interior arrays and a spare
      for (int l = 0; l < 24; ++l) h->sb.a[l] = h->dalloc<double>(ni);
      h->sb.tmp = h->dalloc<double>(ni);
    }
    std::vector<int64_t> nw(tb->nwall_dvm, tb->nwall_dvm + ni);
    h->nwall = h->upload(nw);
  }

  // ---- IB tables and masks ----
  std::vector<uint8_t> mask(np, 0);
  auto PIh = [&](int i, int j, int k) { return (size_t)i + (size_t)g.NX * ((size_t)j + (size_t)g.NY * k); };
  const uint8_t bodybit[4] = {M_BODY_U, M_BODY_V, M_BODY_W, M_BODY_T};
  for (int c = 0; c < h->ncomp; ++c) {
    IBComp &ib = h->ib[c];
    ib.nintp = (int)tb->nintp[c];
    ib.nbody = (int)tb->nbody[c];
    std::vector<int> ifc(ib.nbody), jfc(ib.nbody), kfc(ib.nbody), ix(3 * (size_t)ib.nintp);
    for (int n = 0; n < ib.nbody; ++n) {
      ifc[n] = (int)tb->ifc[c][n];
      jfc[n] = (int)tb->jfc[c][n];
      kfc[n] = (int)tb->kfc[c][n];
      if (ifc[n] < 0 || ifc[n] > n1 || jfc[n] < 0 || jfc[n] > n2 || kfc[n] < 0 || kfc[n] > n3) {
        g_err = "IB forcing point outside the grid";
        throw LW_ERR_ARG;
      }
      mask[PIh(ifc[n], jfc[n], kfc[n])] |= bodybit[c];
    }
    for (size_t q = 0; q < ix.size(); ++q) ix[q] = (int)tb->intpindx[c][q];
    ib.ifc = h->upload(ifc); ib.jfc = h->upload(jfc); ib.kfc = h->upload(kfc); ib.indx = h->upload(ix);
    std::vector<double> gf(tb->geomfac[c], tb->geomfac[c] + 27 * (size_t)ib.nintp);
    ib.geomfac = h->upload(gf);
    ib.targ = h->dalloc<double>(ib.nintp);
    ib.fcv = h->dalloc<double>(ib.nbody);
    ib.fcvavg = h->dalloc<double>(ib.nbody);
    ib.dudtr = c < 3 ? h->dalloc<double>(ib.nbody) : nullptr;
  }
  for (int k = 1; k <= n3m; ++k)
    for (int j = 1; j <= n2m; ++j)
      for (int i = 1; i <= n1m; ++i)
        if (!(funcbody(h->body, h->xmp[i], h->ymp[j], h->zmp[k]) >= 1.e-10)) mask[PIh(i, j, k)] |= M_SOLID_C;
  if (h->les)
    for (int64_t l = 0; l < tb->nzero; ++l) {
      const int64_t i = tb->inz[l], j = tb->jnz[l], k = tb->knz[l];
      if (i >= 1 && i <= n1m && j >= 1 && j <= n2m && k >= 1 && k <= n3m) mask[PIh(i, j, k)] |= M_NUTZERO;
    }
  for (int k = 1; k <= n3m; ++k)
    for (int j = 1; j <= n2m; ++j)
      if (funcbody(h->body, h->xmp[n1m], h->ymp[j], h->zmp[k]) <= 1.e-10) mask[PIh(n1m, j, k)] |= 64;
  for (int k = 0; k <= n3; ++k)
    for (int j = 0; j <= n2; ++j)
      if (funcbody(h->body, h->xmp[0], h->ymp[j], h->zmp[k]) <= 1.e-10) mask[PIh(0, j, k)] |= 128;
  f.mask = h->upload(mask);
  if (st.iread == 0 && st.eps_ptr != 0.0) {
    // addperturb's profiles by the reference's expressions
    // (misc_init.f90:805-866) and its fluid tests (funcbody > 1e-10)
    h->ptb_armed = 1;
    const double ret = 180.0, pi = acos(-1.0), yr = y[n2 - 2];
    std::vector<uint8_t> mu((size_t)n1m * n2m * n3m, 0), mw((size_t)n1m * n2m * n3m, 0), band(n2 + 1, 0);
    std::vector<double> pu((size_t)n2m * n3m, 0.0), pw((size_t)n1m * n2m, 0.0);
    for (int j = 1; j <= n2m; ++j) band[j] = h->ymp[j] > .5 && h->ymp[j] < .99;
    for (int k = 1; k <= n3m; ++k)
      for (int j = 1; j <= n2m; ++j) {
        const double d = fabs(yr - h->ymp[j]);
        pu[(size_t)(j - 1) + (size_t)n2m * (k - 1)] =
            st.eps_ptr * st.udrv_i * cos(h->zmp[k] / gr->zl * pi) * d * ret * exp(-.01 * pow(d * ret, 2.0) + .5);
      }
    for (int j = 1; j <= n2m; ++j)
      for (int i = 1; i <= n1m; ++i) {
        const double d = fabs(yr - h->ymp[j]);
        pw[(size_t)(i - 1) + (size_t)n1m * (j - 1)] =
            st.eps_ptr * sin(h->xmp[i] / gr->xl * pi) * d * ret * exp(-.01 * pow(d * ret, 2.0));
      }
    for (int k = 1; k <= n3m; ++k)
      for (int j = 1; j <= n2m; ++j)
        for (int i = 1; i <= n1m; ++i) {
          const size_t q = (size_t)(i - 1) + (size_t)n1m * ((j - 1) + (size_t)n2m * (k - 1));
          mu[q] = funcbody(h->body, h->x[i], h->ymp[j], h->zmp[k]) > 1.e-10;
          mw[q] = funcbody(h->body, h->xmp[i], h->ymp[j], z[k]) > 1.e-10;
        }
    h->ptb.mu = h->upload(mu);
    h->ptb.mw = h->upload(mw);
    h->ptb.band = h->upload(band);
    h->ptb.pu = h->upload(pu);
    h->ptb.pw = h->upload(pw);
    h->ptb.part = h->dalloc<double>(2 * 64 + 2);
  }
  if (st.ich != 0) {
    // meanpg's and writehistory's fluid test at the u points (x(i), ymp(j),
    // zmp(k)) of the interior, i fastest
    std::vector<uint8_t> um((size_t)n1m * n2m * n3m, 0);
    for (int k = 1; k <= n3m; ++k)
      for (int j = 1; j <= n2m; ++j)
        for (int i = 1; i <= n1m; ++i)
          um[(size_t)(i - 1) + (size_t)n1m * ((j - 1) + (size_t)n2m * (k - 1))] =
              funcbody(h->body, h->x[i], h->ymp[j], h->zmp[k]) >= 1.e-10 ? 1 : 0;
    f.umask = h->upload(um);
    f.chan = h->dalloc<double>(32);
    f.chanpart = h->dalloc<double>(5 * 64);
    HIPCHK(hipMemsetAsync(f.chan, 0, 32 * sizeof(double), h->s));
  }
  // fluid tests of the boundary planes used by draglift's wall terms and by
  // heatflux (slv_auxi.f90:590-640, 733-829): funcbody >= 1e-10
  {
    auto fl = [&](double xx, double yy, double zz) -> uint8_t {
      return funcbody(h->body, xx, yy, zz) >= 1.e-10 ? 1 : 0;
    };
    const std::vector<double> &X = h->x;
    std::vector<double> Y = vd(gr->y, 1, n2), Z = vd(gr->z, 1, n3);
    std::vector<uint8_t> wyb((size_t)(n1 + 1) * (n3 + 1)), wyt(wyb.size()), hym(wyb.size()), hyp(wyb.size());
    std::vector<uint8_t> wzb((size_t)(n1 + 1) * (n2 + 1)), wzt(wzb.size()), hzm(wzb.size()), hzp(wzb.size());
    std::vector<uint8_t> hxm((size_t)(n2 + 1) * (n3 + 1)), hxp(hxm.size());
    for (int k = 1; k <= n3m; ++k)
      for (int i = 1; i <= n1m; ++i) {
        const size_t q = i + (size_t)(n1 + 1) * k;
        wyb[q] = fl(X[i], h->ymp[1], h->zmp[k]);
        wyt[q] = fl(X[i], h->ymp[n2m], h->zmp[k]);
        hym[q] = fl(h->xmp[i], Y[1], h->zmp[k]);
        hyp[q] = fl(h->xmp[i], Y[n2], h->zmp[k]);
      }
    for (int j = 1; j <= n2m; ++j)
      for (int i = 1; i <= n1m; ++i) {
        const size_t q = i + (size_t)(n1 + 1) * j;
        wzb[q] = fl(X[i], h->ymp[j], h->zmp[1]);
        wzt[q] = fl(X[i], h->ymp[j], h->zmp[n3m]);
        hzm[q] = fl(h->xmp[i], h->ymp[j], Z[1]);
        hzp[q] = fl(h->xmp[i], h->ymp[j], Z[n3]);
      }
    for (int k = 1; k <= n3m; ++k)
      for (int j = 1; j <= n2m; ++j) {
        const size_t q = j + (size_t)(n2 + 1) * k;
        hxm[q] = fl(X[1], h->ymp[j], h->zmp[k]);
        hxp[q] = fl(X[n1], h->ymp[j], h->zmp[k]);
      }
    DiagPlanes &dp = h->dplanes;
    dp.wyb = h->upload(wyb); dp.wyt = h->upload(wyt); dp.wzb = h->upload(wzb); dp.wzt = h->upload(wzt);
    dp.hxm = h->upload(hxm); dp.hxp = h->upload(hxp); dp.hym = h->upload(hym); dp.hyp = h->upload(hyp);
    dp.hzm = h->upload(hzm); dp.hzp = h->upload(hzp);
  }

  if (st.poisson == 1) {
    setup_mgrd(h, gr);
    HIPCHK(lw::prepare_adi_kernels(g, &g.adi_wave));
    return;
  }
  if (st.poisson == 2) {
    setup_ftft(h, gr);
    HIPCHK(lw::prepare_adi_kernels(g, &g.adi_wave));
    return;
  }
  if (st.poisson == 3) {
    setup_xft(h, gr);
    HIPCHK(lw::prepare_adi_kernels(g, &g.adi_wave));
    return;
  }
  // ---- Poisson (poisinit, poiss_z_ft.f90:18-113 and coefmg 643-798) ----
  MG &mg = h->mg;
  const int nlev = (int)st.nlev;
  mg.nlev = nlev;
  mg.n1md = 2 * n1m;
  mg.n3mh = n3m / 2 + 1;
  mg.kv0 = 1;
  // the multigrid takes kv = 1..IMGSOR, line-SOR (gsor2d) the rest
  // (poiss_z_ft.f90:186-190)
  mg.nkv = (int)std::min<int64_t>(st.imgsor, mg.n3mh);
  mg.nbli = (int)st.nbli;
  mg.mgitr = (int)st.mgitr;
  mg.LDJ = ((long)(n2 + 1) + 7) / 8 * 8;
  mg.PLANE = (long)(mg.n1md + 1) * mg.LDJ;
  mg.PCPLANE = (long)(n1 + 1) * mg.LDJ;
  const int n1md = mg.n1md, n3mh = mg.n3mh;
  std::vector<double> aw(n1md + 2), ae(n1md + 2), as(n2 + 2), an(n2 + 2), coi1(n1md + 2), coi2(n1md + 2),
      cor1(n1md + 2), cor2(n1md + 2), ak3(n3mh + 2);
  std::vector<double> ac((size_t)n3mh * mg.PLANE, 0.0), gam(ac.size(), 0.0), bet(ac.size(), 0.0);
  auto AC = [&](int i, int j, int k) -> double & { return ac[(size_t)(k - 1) * mg.PLANE + (size_t)i * mg.LDJ + j]; };
  auto GAMh = [&](int i, int j, int k) -> double & { return gam[(size_t)(k - 1) * mg.PLANE + (size_t)i * mg.LDJ + j]; };
  auto BETh = [&](int i, int j, int k) -> double & { return bet[(size_t)(k - 1) * mg.PLANE + (size_t)i * mg.LDJ + j]; };
  for (int i = 1; i <= n1m; ++i) {
    const int ip = ipv[i];
    aw[i] = (1.0 - fixil[i]) * c2cxi[i] * f2fxi[i];
    ae[i] = (1.0 - fixiu[i]) * c2cxi[ip] * f2fxi[i];
  }
  for (int j = 1; j <= n2m; ++j) {
    const int jp = jpv[j];
    as[j] = (1.0 - fixjl[j]) * c2cyi[j] * f2fyi[j];
    an[j] = (1.0 - fixju[j]) * c2cyi[jp] * f2fyi[j];
  }
  for (int j = 1; j <= n2m; ++j)
    for (int i = 1; i <= n1m; ++i) AC(i, j, 1) = -1.0 * (aw[i] + ae[i] + as[j] + an[j]);
  if (n3m > 1) {
    const double pi = acos(-1.0);
    for (int k = 1; k <= n3mh; ++k)
      ak3[k] = 2.0 * (1.0 - cos(2.0 * (double)(k - 1) * pi / (double)n3m)) * f2fzi[1] * f2fzi[1];
  }
  for (int k = 2; k <= n3mh; ++k)
    for (int j = 1; j <= n2m; ++j)
      for (int i = 1; i <= n1m; ++i) AC(i, j, k) = AC(i, j, 1) - ak3[k];
  // coefmg
  {
    const int minrow = n1m / (1 << nlev);
    std::vector<double> xmpm(n1md + 2, 0.0), iwz(n1md + 2, 0.0), iez(n1md + 2, 0.0);
    auto &iimg = mg.iimg;
    iimg[nlev][1] = 1;
    iimg[nlev][2] = n1m;
    iimg[nlev][3] = n1m;
    for (int i = 1; i <= n1m; ++i) {
      xmpm[i] = h->xmp[i];
      iwz[i] = (double)(1 - 1 / i);
      iez[i] = (double)(1 - i / n1m);
    }
    const std::vector<double> &X = h->x;
    for (int ilev = nlev - 1; ilev >= 0; --ilev) {
      iimg[ilev][1] = iimg[ilev + 1][1] + iimg[ilev + 1][3];
      iimg[ilev][3] = minrow * (1 << ilev);
      iimg[ilev][2] = iimg[ilev][1] + iimg[ilev][3] - 1;
      const int ibg = iimg[ilev][1], iend = iimg[ilev][2];
      const int isp = 1 << (nlev - ilev);
      int ic = 0;
      for (int i = ibg; i <= iend; ++i) {
        ic += 1;
        xmpm[i] = 0.5 * (X[ic * isp + 1] + X[(ic - 1) * isp + 1]);
        iwz[i] = (double)(1 - ibg / i);
        iez[i] = (double)(1 - i / iend);
      }
      ic = 0;
      for (int i = ibg; i <= iend; ++i) {
        ic += 1;
        const double dxc = X[ic * isp + 1] - X[(ic - 1) * isp + 1];
        aw[i] = iwz[i] / ((xmpm[i] - xmpm[i - 1]) * dxc);
        ae[i] = iez[i] / ((xmpm[i + 1] - xmpm[i]) * dxc);
      }
      for (int j = 1; j <= n2m; ++j)
        for (int i = ibg; i <= iend; ++i) AC(i, j, 1) = -1.0 * (aw[i] + ae[i] + as[j] + an[j]);
      for (int k = 2; k <= n3mh; ++k)
        for (int j = 1; j <= n2m; ++j)
          for (int i = ibg; i <= iend; ++i) AC(i, j, k) = AC(i, j, 1) - ak3[k];
    }
    for (int ilev = nlev; ilev >= 1; --ilev) {
      int ibgh = iimg[ilev][1];
      for (int i = iimg[ilev - 1][1]; i <= iimg[ilev - 1][2]; ++i) {
        cor1[i] = (xmpm[ibgh + 1] - xmpm[i]) / (xmpm[ibgh + 1] - xmpm[ibgh]);
        cor2[i] = 1.0 - cor1[i];
        ibgh += 2;
      }
    }
    for (int ilev = 0; ilev <= nlev - 1; ++ilev) {
      int ibgh = iimg[ilev + 1][1] + 1;
      for (int i = iimg[ilev][1]; i <= iimg[ilev][2] - 1; ++i) {
        coi1[i] = (xmpm[i + 1] - xmpm[ibgh]) / (xmpm[i + 1] - xmpm[i]);
        coi2[i] = 1.0 - coi1[i];
        ibgh += 2;
      }
      coi1[iimg[ilev][2]] = 1.0;
    }
    // coarse levels handled by one workgroup per wavenumber: up to 24 rows
    // (LW_MG_COARSE_ROWS overrides the limit; 0 launches every level apart)
    int coarse_rows = 24;
    if (const char *e = getenv("LW_MG_COARSE_ROWS")) coarse_rows = atoi(e);
    mg.lc = -1;
    for (int ilev = 0; ilev <= nlev - 1; ++ilev)
      if (iimg[ilev][3] <= coarse_rows) mg.lc = ilev;
    // LW_MG_COARSE_LDS=0 keeps the coarse levels in global memory
    mg.coarse_lds = 1;
    if (const char *e = getenv("LW_MG_COARSE_LDS")) mg.coarse_lds = atoi(e) != 0;
    for (int ilev = 0; ilev <= nlev; ++ilev)
      for (int k = 1; k <= n3mh; ++k) {
        for (int i = iimg[ilev][1]; i <= iimg[ilev][2]; ++i) BETh(i, 1, k) = 1.0 / AC(i, 1, k);
        for (int j = 2; j <= n2m; ++j)
          for (int i = iimg[ilev][1]; i <= iimg[ilev][2]; ++i) {
            GAMh(i, j, k) = an[j - 1] * BETh(i, j - 1, k);
            BETh(i, j, k) = 1.0 / (AC(i, j, k) - as[j] * GAMh(i, j, k));
          }
      }
  }
  // The kernels form ac(i,j,kv) = -(((aw+ae)+as)+an) - ak3(kv) and
  // gam(i,j+1,kv) = an(j) bet(i,j,kv) as they go, with the additions and the
  // product the tables above were built from, instead of reading two more
  // arrays of n3mh planes (ak3(1) = 0 and x - 0 = x). Only bet is stored.
  std::vector<double> awe(n1md + 2, 0.0);
  for (int i = 0; i <= n1md + 1; ++i) awe[i] = aw[i] + ae[i];
  for (int ilev = 0; ilev <= nlev; ++ilev)
    for (int k = 1; k <= n3mh; ++k)
      for (int j = 1; j <= n2m; ++j)
        for (int i = mg.iimg[ilev][1]; i <= mg.iimg[ilev][2]; ++i) {
          const double a = -((awe[i] + as[j]) + an[j]) - ak3[k];
          if (memcmp(&a, &AC(i, j, k), sizeof(double)) != 0) {
            g_err = "Poisson diagonal differs from -(aw+ae+as+an) - ak3";
            throw LW_ERR_ARG;
          }
        }
  mg.aw = h->upload(aw); mg.ae = h->upload(ae); mg.as = h->upload(as); mg.an = h->upload(an);
  mg.awe = h->upload(awe); mg.ak3 = h->upload(ak3);
  mg.coi1 = h->upload(coi1); mg.coi2 = h->upload(coi2); mg.cor1 = h->upload(cor1); mg.cor2 = h->upload(cor2);
  mg.bet = h->upload(bet);
  std::vector<double>().swap(ac);
  std::vector<double>().swap(gam);
  mg.gg = h->dalloc<double2>((size_t)n3mh * mg.PLANE);
  mg.resd = h->dalloc<double2>((size_t)n3mh * mg.PLANE);
  // LW_MG_FUSED=0 runs every zebra half and residual as its own launch
  mg.fused = 1;
  if (const char *e = getenv("LW_MG_FUSED")) mg.fused = atoi(e) != 0;
  HIPCHK(lw::prepare_mg_kernels(g, mg));  // clears fused where the tiles do not fit LDS
  HIPCHK(lw::prepare_adi_kernels(g, &g.adi_wave));
  mg.pc = h->dalloc<double2>((size_t)n3mh * mg.PCPLANE);
  mg.rhs = h->dalloc<double2>((size_t)n3mh * mg.PCPLANE);
  if (mg.fused) {
    mg.gd = h->dalloc<double2>((size_t)n3mh * mg.PLANE);
    mg.rsu = h->dalloc<double2>((size_t)n3mh * mg.PLANE);
  }
  // the fused kernels and the line-SOR sweeps read one pc buffer and write
  // the other (their tiles read rows their neighbours write)
  const bool sor = mg.nkv < n3mh;
  mg.pc2 = mg.fused || sor ? h->dalloc<double2>((size_t)n3mh * mg.PCPLANE) : mg.pc;
  if (sor) h->sorlog = h->dalloc<double>((size_t)std::max(mg.mgitr, 1) * n3mh);
  mg.sumres = h->dalloc<unsigned long long>(n3mh + 4);
  mg.done = h->dalloc<int>(392);
  mg.cyc = mg.done + 128;
  mg.pcsel = mg.done + 256;
  mg.kbar = (unsigned *)(mg.done + 384);
  mg.kerr = mg.done + 385;
  mg.rfin = h->dalloc<double>(128);
  std::vector<double> twr(n3m), twi(n3m);
  const double pi2 = 8.0 * atan(1.0);
  for (int q = 0; q < n3m; ++q) {
    const double ang = -pi2 * (double)q / (double)n3m;
    twr[q] = cos(ang);
    twi[q] = sin(ang);
  }
  mg.twr = h->upload(twr);
  mg.twi = h->upload(twi);
}

// poisinit of poiss_ftft.f90 (164-204): the modified wavenumbers by the
// reference's expressions, the twiddle tables e^{-2 pi i q/N} of both
// transforms, and the spectra of the solve
static void setup_ftft(lw_handle *h, const lw_grid *gr) {
  Ftft &t = h->ftft;
  const int n1m = (int)h->g.n1m, n2m = (int)h->g.n2m, n3m = (int)h->g.n3m;
  const int n3mh = n3m / 2 + 1, n1mh = n1m / 2 + 1;
  t.n3mh = n3mh;
  const double pi = 2.0 * asin(1.0);
  std::vector<double> ak3(n3mh + 2, 0.0), ak1(n1m + 2, 0.0);
  const double sdzis = gr->f2fzi[1], sdxis = gr->f2fxi[1];
  for (int kk = 1; kk <= n3mh; ++kk) {
    const double ai = kk == 1 ? 0.0 : (double)(kk - 1) * 2.0 * pi;
    ak3[kk] = 2.0 * (1.0 - cos(ai / (double)n3m)) * sdzis * sdzis;
  }
  for (int kk = 1; kk <= n1mh; ++kk) {
    const double ai = kk == 1 ? 0.0 : (double)(kk - 1) * 2.0 * pi;
    ak1[kk] = 2.0 * (1.0 - cos(ai / (double)n1m)) * sdxis * sdxis;
  }
  for (int kk = n1m; kk >= n1mh + 1; --kk) ak1[kk] = ak1[n1m + 2 - kk];
  auto table = [](int n, std::vector<double> &re, std::vector<double> &im) {
    re.resize(n);
    im.resize(n);
    const double pi2 = 8.0 * atan(1.0);
    for (int q = 0; q < n; ++q) {
      const double ang = -pi2 * (double)q / (double)n;
      re[q] = cos(ang);
      im[q] = sin(ang);
    }
  };
  std::vector<double> xr, xi, zr, zi;
  table(n1m, xr, xi);
  table(n3m, zr, zi);
  t.twx_r = h->upload(xr);
  t.twx_i = h->upload(xi);
  t.twz_r = h->upload(zr);
  t.twz_i = h->upload(zi);
  t.ak1 = h->upload(ak1);
  t.ak3 = h->upload(ak3);
  const size_t ns = (size_t)n3mh * n2m * n1m;
  t.a = h->dalloc<double2>(ns);
  t.b = h->dalloc<double2>(ns);
  t.gam = h->dalloc<double>(ns);
}

// poisinit of poiss_x_ft.f90 (18-52) with pmat (53-97), mwavenumber
// (99-114) and coefmg (640-792), by the reference's expressions: the z
// levels stacked after the top level, their couplings with the periodic
// wrap of the end cells, the transfer weights, and the LU pivots of the y
// lines per (k, j, x wavenumber)
static void setup_xft(lw_handle *h, const lw_grid *gr) {
  Xft &x = h->xft;
  const lw_settings &st = h->st;
  const int n1m = (int)gr->n1m, n2m = (int)gr->n2m, n3m = (int)gr->n3m, n3 = (int)gr->n3, nlev = (int)st.nlev;
  const int n1mh = n1m / 2 + 1, n3md = 2 * n3m;
  x.n1mh = n1mh;
  x.n3md = n3md;
  x.nlev = nlev;
  x.nbli = (int)st.nbli;
  auto z = vd(gr->z, 1, n3), zmp = vd(gr->zmp, 0, n3);
  auto c2czi = vd(gr->c2czi, 0, n3), f2fzi = vd(gr->f2fzi, 0, n3), c2cyi = vd(gr->c2cyi, 0, gr->n2),
       f2fyi = vd(gr->f2fyi, 0, gr->n2), f2fxi = vd(gr->f2fxi, 0, gr->n1);
  auto kpv = vi(gr->kpv, 1, n3m), jpv = vi(gr->jpv, 1, n2m);
  auto fixkl = vi(gr->fixkl, 1, n3m), fixku = vi(gr->fixku, 1, n3m), fixjl = vi(gr->fixjl, 1, n2m),
       fixju = vi(gr->fixju, 1, n2m);
  std::vector<double> ab(n3md + 2, 0.0), af(n3md + 2, 0.0), as(n2m + 2, 0.0), an(n2m + 2, 0.0);
  std::vector<double> coi1(n3md + 2, 0.0), coi2(n3md + 2, 0.0), cor1(n3md + 2, 0.0), cor2(n3md + 2, 0.0);
  std::vector<double> ai3(n1mh + 1, 0.0);
  std::vector<int> kpm(n3md + 2, 0), kmm(n3md + 2, 0);
  const size_t na = (size_t)n3md * n2m * n1mh;
  std::vector<double> ac(na, 0.0), bet(na, 0.0), gam(na, 0.0);
  auto L = [&](int k, int j, int i) { return (size_t)(k - 1) + (size_t)n3md * ((size_t)(j - 1) + (size_t)n2m * (i - 1)); };
  for (int k = 1; k <= n3m; ++k) {
    const int kp = kpv[k];
    ab[k] = (1.0 - (double)fixkl[k]) * c2czi[k] * f2fzi[k];
    af[k] = (1.0 - (double)fixku[k]) * c2czi[kp] * f2fzi[k];
  }
  for (int j = 1; j <= n2m; ++j) {
    const int jp = jpv[j];
    as[j] = (1.0 - (double)fixjl[j]) * c2cyi[j] * f2fyi[j];
    an[j] = (1.0 - (double)fixju[j]) * c2cyi[jp] * f2fyi[j];
  }
  const double pi = acos(-1.0);
  for (int i = 1; i <= n1mh; ++i)
    ai3[i] = 2.0 * (1.0 - cos(2.0 * pi * (double)(i - 1) / (double)n1m)) * f2fxi[1] * f2fxi[1];
  auto diag = [&](int kb, int ke) {
    for (int j = 1; j <= n2m; ++j)
      for (int k = kb; k <= ke; ++k) ac[L(k, j, 1)] = -1.0 * (ab[k] + af[k] + as[j] + an[j]);
    for (int i = 2; i <= n1mh; ++i)
      for (int j = 1; j <= n2m; ++j)
        for (int k = kb; k <= ke; ++k) ac[L(k, j, i)] = ac[L(k, j, 1)] - ai3[i];
  };
  diag(1, n3m);
  int kk[20][4] = {{0}};
  const int minrow = n3m >> nlev;
  std::vector<double> zmpm(n3md + 2, 0.0), kbz(n3md + 2, 0.0), kfz(n3md + 2, 0.0);
  kk[nlev][1] = 1; kk[nlev][2] = n3m; kk[nlev][3] = n3m;
  for (int k = 1; k <= n3m; ++k) zmpm[k] = zmp[k];
  for (int l = nlev - 1; l >= 0; --l) {
    kk[l][1] = kk[l + 1][1] + kk[l + 1][3];
    kk[l][3] = minrow << l;
    kk[l][2] = kk[l][1] + kk[l][3] - 1;
    const int kbg = kk[l][1], kend = kk[l][2], ksp = 1 << (nlev - l);
    int kc = 0;
    for (int k = kbg; k <= kend; ++k) {
      kc += 1;
      zmpm[k] = 0.5 * (z[kc * ksp + 1] + z[(kc - 1) * ksp + 1]);
      kbz[k] = (double)(1 - kbg / k);
      kfz[k] = (double)(1 - k / kend);
    }
    kc = 0;
    for (int k = kbg; k <= kend; ++k) {
      kc += 1;
      const double dzc = z[kc * ksp + 1] - z[(kc - 1) * ksp + 1];
      ab[k] = kbz[k] / ((zmpm[k] - zmpm[k - 1]) * dzc);
      af[k] = kfz[k] / ((zmpm[k + 1] - zmpm[k]) * dzc);
    }
    diag(kbg, kend);
  }
  for (int l = nlev; l >= 1; --l) {
    int kbgh = kk[l][1];
    for (int k = kk[l - 1][1]; k <= kk[l - 1][2]; ++k) {
      cor1[k] = (zmpm[kbgh + 1] - zmpm[k]) / (zmpm[kbgh + 1] - zmpm[kbgh]);
      cor2[k] = 1.0 - cor1[k];
      kbgh += 2;
    }
  }
  for (int l = 0; l <= nlev - 1; ++l) {
    int kbgh = kk[l + 1][1] + 1;
    for (int k = kk[l][1]; k <= kk[l][2] - 1; ++k) {
      coi1[k] = (zmpm[k + 1] - zmpm[kbgh]) / (zmpm[k + 1] - zmpm[k]);
      coi2[k] = 1.0 - coi1[k];
      kbgh += 2;
    }
    coi1[kk[l][2]] = 1.0;
  }
  // z periodicity (749-789)
  for (int l = nlev; l >= 0; --l) {
    const int kbg = kk[l][1], kend = kk[l][2];
    for (int k = kbg; k <= kend; ++k) {
      kpm[k] = k + 1;
      kmm[k] = k - 1;
    }
    kpm[kend] = kbg;
    kmm[kbg] = kend;
  }
  for (int l = nlev - 1; l >= 0; --l) {
    const int kbg = kk[l][1], kend = kk[l][2], ksp = 1 << (nlev - l);
    const double vdz = zmpm[kbg] - z[1] + z[n3] - zmpm[kend];
    ab[kbg] = 1.0 / (vdz * (z[1 + ksp] - z[1]));
    af[kend] = 1.0 / (vdz * (z[n3] - z[n3 - ksp]));
  }
  for (int l = nlev - 1; l >= 0; --l) diag(kk[l][1], kk[l][2]);
  for (int l = 0; l <= nlev - 1; ++l) {
    const int kbg = kk[l][1], kend = kk[l][2];
    const double vdz = zmpm[kbg] - z[1] + z[n3] - zmpm[kend];
    coi2[kend] = (zmpm[kk[l + 1][2]] - zmpm[kend]) / vdz;
    coi1[kend] = 1.0 - coi2[kend];
  }
  for (int l = 0; l <= nlev; ++l)
    for (int i = 1; i <= n1mh; ++i) {
      for (int k = kk[l][1]; k <= kk[l][2]; ++k) bet[L(k, 1, i)] = 1.0 / ac[L(k, 1, i)];
      for (int j = 2; j <= n2m; ++j)
        for (int k = kk[l][1]; k <= kk[l][2]; ++k) {
          gam[L(k, j, i)] = an[j - 1] * bet[L(k, j - 1, i)];
          bet[L(k, j, i)] = 1.0 / (ac[L(k, j, i)] - as[j] * gam[L(k, j, i)]);
        }
    }
  for (int l = 0; l <= nlev; ++l) {
    x.kb[l] = kk[l][1];
    x.ke[l] = kk[l][2];
  }
  std::vector<double> twr(n1m), twi(n1m);
  const double pi2 = 8.0 * atan(1.0);
  for (int q = 0; q < n1m; ++q) {
    const double ang = -pi2 * (double)q / (double)n1m;
    twr[q] = cos(ang);
    twi[q] = sin(ang);
  }
  x.ab = h->upload(ab); x.af = h->upload(af); x.as = h->upload(as); x.an = h->upload(an);
  x.coi1 = h->upload(coi1); x.coi2 = h->upload(coi2); x.cor1 = h->upload(cor1); x.cor2 = h->upload(cor2);
  x.kpm = h->upload(kpm); x.kmm = h->upload(kmm);
  x.ac = h->upload(ac); x.bet = h->upload(bet); x.gam = h->upload(gam);
  x.twr = h->upload(twr); x.twi = h->upload(twi);
  x.rhs = h->dalloc<double2>((size_t)n1mh * n2m * n3m);
  x.pc = h->dalloc<double2>((size_t)n1mh * (n2m + 2) * n3m);
  x.gg = h->dalloc<double2>((size_t)n1mh * (n2m + 2) * (n3md + 1));
  x.resd = h->dalloc<double2>((size_t)n1mh * n2m * n3md);
  x.tt = h->dalloc<double2>((size_t)n1mh);
  x.done = h->dalloc<int>((size_t)n1mh);
  x.ii = h->dalloc<int>((size_t)n1mh);
  x.nact = h->dalloc<int>(2);
  x.sres = h->dalloc<unsigned long long>((size_t)n1mh);
  x.sfin = h->dalloc<double>((size_t)n1mh);
  x.misc = h->dalloc<double>(2);
}

static void check_scope(const lw_settings *s, const lw_grid *g) {
  auto bad = [](const char *m) {
    g_err = m;
    throw LW_ERR_ARG;
  };
  if (s->yprdic) bad("y-periodic grids are not supported");
  // the x-periodic channel takes poiss_ftft (the reference Makefile's
  // choice for it); inflow/outflow grids the other two solvers
  if (s->xprdic && s->poisson != 2 && s->poisson != 3)
    bad("an x-periodic grid needs poiss_ftft (poisson = 2) or poiss_x_ft (poisson = 3)");
  if (!s->xprdic && (s->poisson == 2 || s->poisson == 3)) bad("poiss_ftft and poiss_x_ft need an x-periodic grid");
  if (s->xprdic && s->ihtrans == 1) bad("heat transfer on an x-periodic grid is not supported");
  if (s->ich < 0 || s->ich > 2) bad("ICH must be 0, 1 or 2");
  if (s->ich != 0 && !s->xprdic) bad("ICH != 0 (channel driving) needs an x-periodic grid");
  if (s->zprdic != 1) bad("the z direction must be periodic (the step's z sweeps and poiss_z_ft)");
  if (s->iconjg) bad("ICONJG = 1 (conjugate heat transfer) is not supported");
  if (s->imovingon) bad("IMOVING = 1 is not supported");
  if (s->idecomp) bad("IDECOMP = 1 is not supported");
  if (s->gr != 0.0) bad("GR != 0 (buoyancy) is not supported");
  if (s->iles == 1 && s->insmdl != 1) bad("only the Vreman model (INSMDL = 1) exists in the reference");
  if (s->iles == 1 && s->ihtrans == 1 && s->itemdl != 0) bad("ITEMDL must be 0");
  if (s->iles == 1 && (s->filter < 1 || s->filter > 3)) bad("FILTER must be 1, 2 or 3");
  if (s->nlev < 1 || s->nlev > 19) bad("NLEV must be in 1..19");
  if (s->poisson <= 1 && g->n1m % (1 << s->nlev) != 0) bad("n1m must be divisible by 2**NLEV");
  if (g->n1m % 2 != 0) bad("n1m must be even");
  if (g->n3m < 2) bad("n3m must be at least 2");
  if (s->nbli < 1) bad("NBLI must be >= 1");
  if (s->poisson < 0 || s->poisson > 3)
    bad("poisson must be 0 (poiss_z_ft), 1 (poiss_mgrd), 2 (poiss_ftft) or 3 (poiss_x_ft)");
  if (s->poisson == 3) {
    // the z multigrid coarsens n3m into n3m / 2**NLEV lines; a level with
    // an odd number of lines above 1 would put two coupled lines in one
    // zebra half, whose result then depends on the order they are solved in
    if (g->n3m % (1 << s->nlev) != 0) bad("poiss_x_ft: n3m must be divisible by 2**NLEV");
    const int64_t nz0 = g->n3m >> s->nlev;
    if (nz0 > 1 && nz0 % 2 == 1) bad("poiss_x_ft: n3m / 2**NLEV must be 1 or even");
    if (s->nlev > 19) bad("poiss_x_ft: NLEV must be at most 19");
  }
  if (s->poisson == 1) {
    // poiss_mgrd coarsens z as well; with IOLDV = 1 its V-cycle starts from
    // an uninitialised local array (poiss_mgrd.f90:236, 239)
    if (s->ioldv != 0) bad("poiss_mgrd: IOLDV must be 0");
    if (g->n3m % (1 << s->nlev) != 0) bad("poiss_mgrd: n3m must be divisible by 2**NLEV");
    const int64_t nz0 = g->n3m >> s->nlev;
    // an odd number of z lines above 1 on the coarsest level puts two
    // neighbouring lines in one zebra half; the reference's result then
    // depends on the order its loop visits them
    if (nz0 > 1 && nz0 % 2 == 1) bad("poiss_mgrd: n3m / 2**NLEV must be 1 or even");
  } else if (s->poisson == 0) {
    // kv = 1, the singular mean mode, needs the multigrid's gauge shift
    if (s->imgsor < 1) bad("IMGSOR must be at least 1");
    if (s->imgsor < g->n3m / 2 + 1 && (g->n2m + 63) / 64 > 8)
      bad("line-SOR wavenumbers (IMGSOR < n3m/2+1) need n2m <= 512");
    if (g->n3m / 2 + 1 > 128) bad("n3m must be at most 254");
  }
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

const char *lw_last_error(void) { return g_err.c_str(); }
}  // extern "C"

void lw_set_last_error(const std::string &msg) { g_err = msg; }

extern "C" {

static int create_handle(const lw_settings *s, const lw_grid *gr, const lw_tables *t, const lw_body *body,
                         lw_handle **out, bool in_static_area) {
  if (!s || !gr || !t || !out) return fail(LW_ERR_ARG, "null argument");
  *out = nullptr;
  lw_handle *h = nullptr;
  try {
    check_scope(s, gr);
    if (body && (body->kind < 0 || body->kind > 2 || !(body->p[0] > 0.0) || (body->kind == 2 && !(body->p[1] > 0.0))))
      return fail(LW_ERR_ARG, "body: kind must be 0 (cylinder), 1 (square) or 2 (ellipse) with positive sizes");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(LW_ERR_DEVICE, "no HIP device");
    h = new lw_handle();
    if (in_static_area) {
      void *a = nullptr;
      size_t cap = 0;
      int dev = -1;
      const hipError_t e = static_area_acquire(&dev, &a, &cap);
      if (e == hipErrorAlreadyAcquired) {
        delete h;
        return fail(LW_ERR_ARG, "lw_create_static: another handle holds this device's static area");
      }
      if (e != hipSuccess) {
        delete h;
        return fail(LW_ERR_DEVICE, std::string("lw_create_static: static area: ") + hipGetErrorString(e));
      }
      h->arena = (unsigned char *)a;
      h->arena_cap = cap;
      h->arena_dev = dev;
    }
    h->st = *s;
    if (body) h->body = *body;
    else {
      h->body.kind = 0;
      h->body.p[0] = 0.5;
    }
    h->les = s->iles == 1;
    h->heat = s->ihtrans == 1;
    h->ncomp = h->heat ? 4 : 3;
    h->csgsts = s->csgsts;
    h->csgshf = s->csgshf;
    HIPCHK(hipStreamCreateWithFlags(&h->s, hipStreamNonBlocking));
    HIPCHK(hipStreamCreateWithFlags(&h->s2, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&h->ev_fork, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&h->ev_join, hipEventDisableTiming));
    // kv = 1 on its own stream pays on large planes (384 x 256: 4% per step);
    // on small ones one stream is as fast alone and leaves the hardware
    // queues to other cases (LW_MG_SPLIT or lw_set_option override)
    h->mg_split = (int64_t)gr->n1m * gr->n2m >= 65536;
    if (const char *e = getenv("LW_MG_SPLIT")) h->mg_split = atoi(e) != 0;
    // the two-stream solve polls copies into pinned memory; a static handle
    // has none, so it solves on one stream (the same kernels and results)
    if (h->arena) h->mg_split = 0;
    if (const char *e = getenv("LW_MG_KV1")) h->mg_kv1 = atoi(e) != 0;
    if (const char *e = getenv("LW_MG_KV1_TILES")) {
      int v[3] = {0, 0, 0};
      if (sscanf(e, "%d,%d,%d", &v[0], &v[1], &v[2]) == 3)
        for (int q = 0; q < 3; ++q)
          if (v[q] == 2 || v[q] == 4 || v[q] == 8) h->kv1_tiles[q] = v[q];
    }
    {
      // 2 s of the constant-rate clock (its rate in kHz)
      int dev = 0, khz = 0;
      HIPCHK(hipGetDevice(&dev));
      HIPCHK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
      h->kv1_tlim = 2000LL * (long long)(khz > 0 ? khz : 100000);
    }
    h->sgs_fused = sgs_fused_enabled() ? 1 : 0;
    for (int e = 0; e < 8; ++e) HIPCHK(hipEventCreate(&h->ev[e]));
    for (int q = 0; q < 3; ++q)
      for (int e = 0; e < 8; ++e) HIPCHK(hipEventCreate(&h->tev[q][e]));
    if (h->arena) {
      h->h_sumres = h->hs_sumres;
      h->h_flags = h->hs_flags;
      h->h_flags2 = h->hs_flags2;
      h->h_red = h->hs_red;
      h->h_rfin = h->hs_rfin;
    } else {
      HIPCHK(hipHostMalloc((void **)&h->h_rfin, 4 * sizeof(double)));
      HIPCHK(hipHostMalloc((void **)&h->h_sumres, 256 * sizeof(unsigned long long)));
      HIPCHK(hipHostMalloc((void **)&h->h_flags, 256 * sizeof(int)));
      HIPCHK(hipHostMalloc((void **)&h->h_flags2, 512 * sizeof(int)));
      HIPCHK(hipHostMalloc((void **)&h->h_red, 64 * sizeof(double)));
    }
    setup(h, gr, t);
    // dttimeinit / rk3coefinit (misc_init.f90:1001-1044)
    h->gamma[1] = 8. / 15.; h->gamma[2] = 5. / 12.; h->gamma[3] = 3. / 4.;
    h->ro[1] = 0.; h->ro[2] = -17. / 60.; h->ro[3] = -5. / 12.;
    h->dt = s->dt_size;
    // LW_MG_FP32=1 asks for the mixed-precision multigrid wherever the case
    // allows it (the option "mg_fp32" insists and reports when it does not)
    if (const char *e = getenv("LW_MG_FP32"))
      if (atoi(e) != 0 && h->st.poisson == 0 && mg_mixed_supported(h->g, h->mg)) set_mg_fp32(h, true);
    // LW_SGS_FP32=1: the FP32 test filter wherever the dynamic models run fused
    if (const char *e = getenv("LW_SGS_FP32"))
      if (atoi(e) != 0 && h->sgs_fused) h->sgs_fp32 = 1;
    HIPCHK(hipStreamSynchronize(h->s));
    *out = h;
    return LW_OK;
  } catch (int code) {
    if (h) lw_destroy(h);
    return code;
  }
}

int lw_create(const lw_settings *s, const lw_grid *gr, const lw_tables *t, const lw_body *body, lw_handle **out) {
  return create_handle(s, gr, t, body, out, false);
}

int lw_create_static(const lw_settings *s, const lw_grid *gr, const lw_tables *t, const lw_body *body,
                     lw_handle **out) {
  return create_handle(s, gr, t, body, out, true);
}

int64_t lw_static_area_bytes(void) { return (int64_t)static_area_bytes(); }

int lw_destroy(lw_handle *h) {
  if (!h) return LW_OK;
  if (h->batch) return fail(LW_ERR_ARG, "lw_destroy: the handle belongs to a batch (lw_batch_destroy first)");
  if (h->s) hipStreamSynchronize(h->s);
  if (h->s2) hipStreamSynchronize(h->s2);
  for (void *p : h->allocs) hipFree(p);
  for (int e = 0; e < 8; ++e)
    if (h->ev[e]) hipEventDestroy(h->ev[e]);
  for (int q = 0; q < 3; ++q)
    for (int e = 0; e < 8; ++e)
      if (h->tev[q][e]) hipEventDestroy(h->tev[q][e]);
  if (!h->arena) {
    if (h->h_sumres) hipHostFree(h->h_sumres);
    if (h->h_flags) hipHostFree(h->h_flags);
    if (h->h_flags2) hipHostFree(h->h_flags2);
    if (h->h_red) hipHostFree(h->h_red);
    if (h->h_rfin) hipHostFree(h->h_rfin);
  }
  delete h->prof;
  if (h->ev_fork) hipEventDestroy(h->ev_fork);
  if (h->ev_join) hipEventDestroy(h->ev_join);
  if (h->s2) hipStreamDestroy(h->s2);
  if (h->s) hipStreamDestroy(h->s);
  if (h->arena) static_area_release(h->arena_dev);
  delete h;
  return LW_OK;
}

int lw_get_dims(const lw_handle *h, int64_t n[3]) {
  if (!h || !n) return fail(LW_ERR_ARG, "null argument");
  n[0] = h->g.n1; n[1] = h->g.n2; n[2] = h->g.n3;
  return LW_OK;
}

int64_t lw_device_bytes(const lw_handle *h) { return h ? h->bytes : 0; }

double lw_body_eval(const lw_body *b, double x, double y, double z) {
  lw_body c{};
  c.p[0] = 0.5;
  return funcbody(b ? *b : c, x, y, z);
}

// Upper bound of what lw_create allocates for these sizes, from the same
// array counts as the allocations above: 34 padded FP64 arrays (30 without T),
// 25 interior arrays with the dynamic SGS model, the multigrid plane stacks
// (gg, resd, gd, rsu over all levels, pc, rhs, pc2 on the top level, bet) and
// a margin for the masks, IB tables, 1-D metrics and reduction buffers.
int64_t lw_estimate_device_bytes(const lw_settings *s, int64_t n1m, int64_t n2m, int64_t n3m) {
  if (!s || n1m < 2 || n2m < 2 || n3m < 2) return fail(LW_ERR_ARG, "lw_estimate_device_bytes: bad sizes");
  const int64_t np = (n1m + 2) * (n2m + 2) * (n3m + 2), ni = n1m * n2m * n3m;
  const int64_t heat = s->ihtrans == 1 ? 1 : 0;
  int64_t b = (30 + 4 * heat) * np * 8 + 2 * np;  // fields, mask and a byte of slack per cell
  if (s->iles == 1) b += (sgs_fused_enabled() ? 0 : 25 * ni * 8) + ni * 8;  // SGS work arrays, nwall
  if (s->iavg == 1) b += (heat ? 23 : 21) * ni * 8;  // field_avg running sums
  if (s->poisson == 1) {
    // poiss_mgrd: gi, resid and the pivots over the stacked levels, and the
    // seven BiCGSTAB vectors
    b += (2 * n1m + 1) * (n2m + 2) * (2 * n3m + 1) * 8 + 2 * (2 * n1m) * n2m * (2 * n3m) * 8 + 7 * np * 8;
  } else if (s->poisson == 2) {
    // poiss_ftft: two complex spectra and the solve's gam
    b += (n3m / 2 + 1) * n2m * n1m * (2 * 16 + 8) + ni;
  } else if (s->poisson == 3) {
    // poiss_x_ft: rhs, pc, gg and resd per x wavenumber, ac, bet and gam
    const int64_t n1mh = n1m / 2 + 1;
    b += n1mh * (n2m + 2) * (n3m + (2 * n3m + 1)) * 16 + n1mh * n2m * (n3m + 2 * n3m) * 16 +
         3 * n1mh * n2m * 2 * n3m * 8 + ni;
  } else {
    const int64_t ldj = (n2m + 2 + 7) / 8 * 8, plane = (2 * n1m + 1) * ldj, pcplane = (n1m + 2) * ldj;
    const int64_t n3mh = n3m / 2 + 1;
    b += n3mh * (4 * plane + 3 * pcplane) * 16 + n3mh * plane * 8;
  }
  b += 8 * (n2m + 2) * (n3m + 2) * 8;                // outlet planes
  // IB tables: at most every cell of a component a forcing point, but bodies
  // cover a small share; 64 MB covers tables, metrics and partial sums
  return b + b / 32 + ((int64_t)64 << 20);
}

int lw_set_fields(lw_handle *h, const double *u, const double *v, const double *w, const double *p, const double *t,
                  int64_t ihist, int64_t m, double time, double dt) {
  if (!h || !u || !v || !w || !p) return fail(LW_ERR_ARG, "null argument");
  if (h->heat && !t) return fail(LW_ERR_ARG, "temperature field required (IHTRANS = 1)");
  try {
    const size_t b = sizeof(double) * (size_t)h->g.NX * h->g.NY * h->g.NZ;
    HIPCHK(hipMemcpyAsync(h->f.u, u, b, hipMemcpyHostToDevice, h->s));
    HIPCHK(hipMemcpyAsync(h->f.v, v, b, hipMemcpyHostToDevice, h->s));
    HIPCHK(hipMemcpyAsync(h->f.w, w, b, hipMemcpyHostToDevice, h->s));
    HIPCHK(hipMemcpyAsync(h->f.p, p, b, hipMemcpyHostToDevice, h->s));
    if (t) HIPCHK(hipMemcpyAsync(h->f.t, t, b, hipMemcpyHostToDevice, h->s));
    HIPCHK(hipStreamSynchronize(h->s));
    h->ihist = ihist;
    h->m = m;
    h->time = time;
    h->dt = dt;
    // dttimeinit (misc_init.f90:1001-1028)
    if (h->st.idtopt == 0) h->dt = h->st.dt_size;
    if (h->st.idtopt != 0 && h->st.iread == 0) h->dt = h->st.dt_size;
    if (h->st.ireset == 1) {
      h->ihist = 0;
      h->time = 0.;
    }
    h->ntime = 0;
    // addperturb before the first step when PTB_TST <= 0 (solver.f90:76-80):
    // after the initial field was taken (meanpg below) and could be written
    h->ptb_applied = 0;
    h->ptb_pending = h->ptb_armed && h->st.ptb_tst <= 0.0;
    // solver.f90:54: the channel's first mean pressure gradient, from the
    // loaded field (ntime 0: no controller term); the controller's integral
    // and the fcmfr.dat sums start from zero
    if (h->g.ich != 0) {
      HIPCHK(hipMemsetAsync(h->f.chan, 0, 32 * sizeof(double), h->s));
      launch_meanpg(h->s, h->g, h->f, h->ib[0], h->st, 0.0, 0);
      check_launches("meanpg");
      HIPCHK(hipStreamSynchronize(h->s));
    }
    return LW_OK;
  } catch (int code) {
    return code;
  }
}

int lw_get_fields(lw_handle *h, double *u, double *v, double *w, double *p, double *t) {
  if (!h) return fail(LW_ERR_ARG, "null handle");
  try {
    const size_t b = sizeof(double) * (size_t)h->g.NX * h->g.NY * h->g.NZ;
    if (u) HIPCHK(hipMemcpyAsync(u, h->f.u, b, hipMemcpyDeviceToHost, h->s));
    if (v) HIPCHK(hipMemcpyAsync(v, h->f.v, b, hipMemcpyDeviceToHost, h->s));
    if (w) HIPCHK(hipMemcpyAsync(w, h->f.w, b, hipMemcpyDeviceToHost, h->s));
    if (p) HIPCHK(hipMemcpyAsync(p, h->f.p, b, hipMemcpyDeviceToHost, h->s));
    if (t && h->heat) HIPCHK(hipMemcpyAsync(t, h->f.t, b, hipMemcpyDeviceToHost, h->s));
    HIPCHK(hipStreamSynchronize(h->s));
    return LW_OK;
  } catch (int code) {
    return code;
  }
}

int64_t lw_get_array(lw_handle *h, const char *name, double *out, int64_t n) {
  if (!h || !name) return fail(LW_ERR_ARG, "null argument");
  const int64_t np = (int64_t)h->g.NX * h->g.NY * h->g.NZ;
  const double *src = nullptr;
  std::string s(name);
  if (s == "pmi") {
    // the channel's mean pressure gradient pmi(0:4) (meanpg)
    if (h->g.ich == 0) return fail(LW_ERR_ARG, "pmi exists with ICH != 0 only");
    if (!out) return 5;
    if (n < 5) return fail(LW_ERR_ARG, "output buffer too small");
    try {
      HIPCHK(hipMemcpyAsync(out, h->f.chan, 5 * sizeof(double), hipMemcpyDeviceToHost, h->s));
      HIPCHK(hipStreamSynchronize(h->s));
    } catch (int code) {
      return code;
    }
    return 5;
  }
  if (s == "nusgs") src = h->f.nu;
  else if (s == "alsgs") src = h->f.al;
  else if (s == "divgsum") src = h->f.div;
  else if (s == "phi") src = h->f.phi;
  else if (s == "qmass") src = h->f.qmass;
  else if (s.size() == 8 && s.compare(0, 7, "nusgs1_") == 0 && s[7] >= '1' && s[7] <= '3') src = h->f.nu1[s[7] - '1'];
  else if (s.size() == 8 && s.compare(0, 7, "alsgs1_") == 0 && s[7] >= '1' && s[7] <= '3') src = h->f.al1[s[7] - '1'];
  else {
    // running sums of field_avg: interior arrays (n1m, n2m, n3m)
    static const char *names[kAvgArrays] = {"uavg", "vavg", "wavg", "uiujavg_1", "uiujavg_2", "uiujavg_3",
                                             "uiujavg_4", "uiujavg_5", "uiujavg_6", "pavg", "p2avg", "voravg_1",
                                             "voravg_2", "voravg_3", "vor2avg_1", "vor2avg_2", "vor2avg_3",
                                             "vor2avg_4", "vor2avg_5", "vor2avg_6", "ssavg", "tavg", "t2avg"};
    int l = 0;
    while (l < kAvgArrays && s != names[l]) ++l;
    if (l == kAvgArrays) return fail(LW_ERR_ARG, "unknown array name");
    if (l >= h->navg) return fail(LW_ERR_ARG, "statistics not kept (IAVG = 0, or T without IHTRANS)");
    const int64_t ni = (int64_t)h->g.n1m * h->g.n2m * h->g.n3m;
    if (!out) return ni;
    if (n < ni) return fail(LW_ERR_ARG, "output buffer too small");
    try {
      HIPCHK(hipMemcpyAsync(out, h->avg.a[l], sizeof(double) * ni, hipMemcpyDeviceToHost, h->s));
      HIPCHK(hipStreamSynchronize(h->s));
    } catch (int code) {
      return code;
    }
    return ni;
  }
  if (!out) return np;
  if (n < np) return fail(LW_ERR_ARG, "output buffer too small");
  try {
    HIPCHK(hipMemcpyAsync(out, src, sizeof(double) * np, hipMemcpyDeviceToHost, h->s));
    HIPCHK(hipStreamSynchronize(h->s));
  } catch (int code) {
    return code;
  }
  return np;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// the step
// ---------------------------------------------------------------------------
// the mixed-precision form: the multigrid level planes below the top level
// hold FP32 values (the coarse-grid corrections), while the top level's
// relaxation, residual, convergence test and solution stay FP64, so every
// V-cycle is a defect correction in FP64 with an FP32 coarse-grid
// correction. Switching clears the planes: the other form's bit patterns are
// not values of this one.
static void st_zero(lw_handle *h, void *p, size_t bytes) {
  if (g_rec) rec_zero(p, bytes);
  else HIPCHK(hipMemsetAsync(p, 0, bytes, h->s));
}
static void st_read(lw_handle *h, void *dst, const void *src, size_t bytes) {
  if (g_rec) rec_read(dst, src, bytes);
  else HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, h->s));
}
static void st_sync(lw_handle *h) {
  if (g_rec) batch_sync();
  else HIPCHK(hipStreamSynchronize(h->s));
}
static void st_event(lw_handle *h, hipEvent_t e) {
  if (!g_rec) HIPCHK(hipEventRecord(e, h->s));
}

static void set_mg_fp32(lw_handle *h, bool on) {
  MG &mg = h->mg;
  if (on && (h->st.poisson != 0 || !mg_mixed_supported(h->g, mg))) {
    g_err = "mg_fp32: needs the z-DFT multigrid path with its fused kernels and LDS-resident coarse levels";
    throw LW_ERR_ARG;
  }
  if ((int)on == mg.f32) return;
  const size_t b = sizeof(double2) * (size_t)mg.n3mh * mg.PLANE;
  HIPCHK(hipMemsetAsync(mg.gg, 0, b, h->s));
  HIPCHK(hipMemsetAsync(mg.resd, 0, b, h->s));
  if (mg.gd) HIPCHK(hipMemsetAsync(mg.gd, 0, b, h->s));
  if (mg.rsu) HIPCHK(hipMemsetAsync(mg.rsu, 0, b, h->s));
  HIPCHK(hipStreamSynchronize(h->s));
  mg.f32 = on ? 1 : 0;
}

static void stepinit(lw_handle *h) {
  const lw_settings &st = h->st;
  h->ntime += 1;
  h->m = h->ntime;
  for (int c = 0; c < h->ncomp; ++c)
    if (h->ib[c].nbody > 0)
      st_zero(h, h->ib[c].fcvavg, sizeof(double) * h->ib[c].nbody);
  if (st.idtopt == 0) h->dt = st.dt_size;
  st_zero(h, h->f.redmax, 8 * sizeof(unsigned long long));
  launch_cfl(h->s, h->g, h->f, h->dt);
  if (st.idtopt != 0) {
    // the new dt depends on the CFL number: one host round trip per step
    st_read(h, h->h_sumres, h->f.redmax, sizeof(unsigned long long));
    st_sync(h);
    double c;
    memcpy(&c, &h->h_sumres[0], 8);
    h->cflmax = c;
    if (h->cflmax != 0.) {
      const double dtcfl = fmin(h->dt * st.cflfac / h->cflmax, h->dt * (0.80 + 0.20 * st.cflfac / h->cflmax));
      if (st.idtopt == 1) h->dt = dtcfl;
    }
  }
  const double eps_t = 1.0e-12 * fmax(1.0, fabs(h->time));
  double dt_evt = h->dt;
  if (st.ptb_tst > h->time + eps_t) dt_evt = fmin(dt_evt, st.ptb_tst - h->time);
  if (st.avg_tst > h->time + eps_t) dt_evt = fmin(dt_evt, st.avg_tst - h->time);
  if (st.tend > h->time + eps_t) dt_evt = fmin(dt_evt, st.tend - h->time);
  if (dt_evt < h->dt) h->dt = dt_evt;
}

static void substepinit(lw_handle *h, int ms) {
  Sub &sb = h->sub;
  sb.msub = ms;
  sb.dt = h->dt;
  sb.gam = h->gamma[ms];
  sb.ro = h->ro[ms];
  sb.alpha = 0.5 * (h->gamma[ms] + h->ro[ms]);
  sb.dtconst = h->dt * (h->gamma[ms] + h->ro[ms]);
  sb.dtconsti = 1. / sb.dtconst;
  sb.acoef = sb.alpha * h->dt / h->st.re;
  sb.acoefi = 1. / sb.acoef;
  sb.re = h->st.re;
  sb.pr = h->st.pr;
  // pmiavg = pmiavg + pmi(0)*2*alpha (slv_auxi.f90:118), restarted by
  // stepinit (:13)
  if (h->g.ich != 0) launch_pmiavg(h->s, h->f, sb.alpha, ms == 1);
}

// addperturb with prdic_adj_uvw(0) (solver.f90:76-80, 153-160)
static void apply_perturb(lw_handle *h) {
  launch_addperturb(h->s, h->g, h->f, h->ptb);
  launch_prdic_uvw(h->s, h->g, h->f);
  check_launches("addperturb");
  h->ptb_applied = 1;
  h->ptb_pending = 0;
}

// after pcalc (solver.f90:136): the channel's mean pressure gradient for the
// next substep, with the controller term inside a step (ntime > 0, dt > 0)
static void after_projection(lw_handle *h) {
  if (h->g.ich != 0)
    launch_meanpg(h->s, h->g, h->f, h->ib[0], h->st, h->sub.dtconst, (int)(h->ntime > 0 && h->dt > 0.0));
}

static void phase_sgs(lw_handle *h) {
  if (!h->les || h->sub.msub != 1) return;
  const lw_settings &st = h->st;
  const Geo &g = h->g;
  Fld &f = h->f;
  const size_t np = (size_t)g.NX * g.NY * g.NZ;
  const bool zero_first = st.iread == 0 && h->ntime == 1;
  // sgscalc (sgs.f90:288-332)
  int slot = -1;
  if (zero_first) st_zero(h, f.nu, np * sizeof(double));
  else if (st.idvmon == 1) {
    if (h->sgs_fused) {
      launch_sgs_dvmg_fused(h->s, g, f, h->nwall, (int)st.filter, h->sgs_fp32);
    } else {
      launch_sgs_dvmg(h->s, g, f, h->sb, (int)st.filter);
      launch_sgs_dvmg_coef(h->s, g, f, h->sb, h->nwall);
    }
    slot = 4;
  } else {
    launch_sgs_cvm(h->s, g, f, h->csgsts);
  }
  launch_sgs_finish(h->s, g, f, f.nu, f.nu1, 0, slot, st);
  if (!h->heat) return;
  // sgscalc_t (sgs.f90:1203-1247)
  slot = -1;
  if (zero_first) st_zero(h, f.al, np * sizeof(double));
  else if (st.idvmon == 1) {
    if (h->sgs_fused) {
      launch_sgs_edm_fused(h->s, g, f, h->nwall, (int)st.filter, h->sgs_fp32);
    } else {
      launch_sgs_edm(h->s, g, f, h->sb, (int)st.filter);
      launch_sgs_edm_coef(h->s, g, f, h->sb, h->nwall);
    }
    slot = 5;
  } else {
    launch_sgs_cedm(h->s, g, f, h->csgshf);
  }
  launch_sgs_finish(h->s, g, f, f.al, f.al1, 1, slot, st);
}

static void phase_rhsnlhs(lw_handle *h) {
  const Geo &g = h->g;
  Body b{};
  launch_rhs_mom(h->s, g, h->sub, h->f, h->les);
  if (h->st.ibmon != 0) {
    // one launch per phase for the three components (LW_IBM_FUSED=0: one
    // per component)
    const char *e = getenv("LW_IBM_FUSED");
    if (e && atoi(e) == 0)
      for (int c = 0; c < 3; ++c) launch_ibm(h->s, g, h->sub, h->f, h->ib, c);
    else
      launch_ibm_mom(h->s, g, h->sub, h->f, h->ib);
  }
  launch_convbc_mom(h->s, g, h->sub, h->f, b, h->les);
  launch_adi_mom(h->s, g, h->sub, h->f, h->les);
  launch_retrv_mom(h->s, g, h->f);
}

// kv = 1 (the mean in z) needs several times the V-cycles of the other
// wavenumbers (C2: 55-110 against 12-22), and alone it is latency-bound. Its
// cycles therefore run on s2 beside the bandwidth-bound cycles of kv = 2..n3mh
// on s, each side with its own batches and flags; the two sides touch
// disjoint wavenumbers, and every wavenumber goes through the same kernels
// and cycles as in one stream. The host polls both streams and gives a side
// its next batch as soon as its flags say it needs one.
static void poisson_split(lw_handle *h, double test, bool fold) {
  const Geo &g = h->g;
  MG a = h->mg, b = h->mg;
  a.kv0 = 1;
  a.nkv = 1;
  for (int q = 0; q < 3; ++q) a.tile[q] = a.tile1[q];
  // kv = 1 as one launch of k_mg_kv1 when its shape allows (the check must
  // be folded: the launch starts with the test of the first toplevel)
  bool kv1 = false;
  if (h->mg_kv1 && fold) {
    MG p = a;
    for (int q = 0; q < 3; ++q) p.tile[q] = h->kv1_tiles[q];
    if (mg_kv1_supported(g, p)) {
      a = p;
      kv1 = true;
    }
  }
  b.kv0 = 2;
  b.nkv = h->mg.nkv - 1;
  const int itr = h->mg.mgitr;
  HIPCHK(hipEventRecord(h->ev_fork, h->s));
  HIPCHK(hipStreamWaitEvent(h->s2, h->ev_fork, 0));
  struct Side {
    const MG *m;
    hipStream_t s;
    int *flags;
    int launched, batch, pending;
    bool done;
  };
  Side sd[2] = {{&a, h->s2, h->h_flags2, 0, h->mg_hint1 > 1 ? h->mg_hint1 : 1, 0, false},
                {&b, h->s, h->h_flags, 0, h->mg_hint > 1 ? h->mg_hint : 1, 0, false}};
  auto prof_side = [&](const Side &x, int cycle) {
    if (!g_prof) return;
    g_prof->cycle = cycle;
    g_prof->kv0 = x.m->kv0;
    g_prof->nkv = x.m->nkv;
  };
  for (Side &x : sd) {
    prof_side(x, 0);
    launch_mg_toplevel(x.s, g, *x.m, 0, 0.0, test, !fold);
  }
  // enqueue the next batch of every side that is idle and not done, one
  // cycle of each side at a time, then its flags
  auto enqueue = [&]() {
    int n[2];
    for (int k = 0; k < 2; ++k) {
      Side &x = sd[k];
      const bool go = !x.pending && !x.done && x.launched < itr;
      n[k] = go ? (x.batch < itr - x.launched ? x.batch : itr - x.launched) : 0;
    }
    // k_mg_kv1 runs every cycle kv = 1 may need in its one launch, first
    if (kv1 && n[0] > 0) {
      n[0] = itr - sd[0].launched;
      prof_side(sd[0], 1);
      launch_mg_kv1(sd[0].s, g, a, test, n[0], h->kv1_tlim);
      HIPCHK(hipMemcpyAsync(sd[0].flags + 400, a.kerr, sizeof(int), hipMemcpyDeviceToHost, sd[0].s));
    }
    const int nmax = n[0] > n[1] ? n[0] : n[1];
    for (int c = 0; c < nmax; ++c)
      for (int k = 0; k < 2; ++k) {
        Side &x = sd[k];
        if (c >= n[k] || (k == 0 && kv1)) continue;
        const int m = x.launched + c + 1;
        prof_side(x, m);
        const int cyc = fold && !(c == 0 && x.launched > 0) ? m - 1 : -1;
        launch_mg_cycle(x.s, g, *x.m, test, cyc, !fold || c == n[k] - 1);
      }
    for (int k = 0; k < 2; ++k) {
      Side &x = sd[k];
      if (!n[k]) continue;
      x.launched += n[k];
      x.pending = 1;
      x.batch = 2;
      HIPCHK(hipMemcpyAsync(x.flags, h->mg.done, 256 * sizeof(int), hipMemcpyDeviceToHost, x.s));
    }
  };
  enqueue();
  while (sd[0].pending || sd[1].pending) {
    bool progressed = false;
    for (Side &x : sd) {
      if (!x.pending) continue;
      const hipError_t q = hipStreamQuery(x.s);
      if (q == hipErrorNotReady) continue;
      HIPCHK(q);
      x.pending = 0;
      if (kv1 && x.m == &a && x.flags[400] != 0) {
        g_err = "k_mg_kv1: a grid barrier wait ran out (the workgroups were not co-resident)";
        throw LW_ERR_DEVICE;
      }
      bool d = true;
      for (int i = x.m->kv0 - 1; i < x.m->kv0 - 1 + x.m->nkv; ++i) d = d && x.flags[i] != 0;
      x.done = d;
      progressed = true;
    }
    if (progressed) enqueue();
    else std::this_thread::yield();
  }
  HIPCHK(hipEventRecord(h->ev_join, h->s2));
  HIPCHK(hipStreamWaitEvent(h->s, h->ev_join, 0));
  HIPCHK(hipMemcpyAsync(h->h_flags, h->mg.done, 256 * sizeof(int), hipMemcpyDeviceToHost, h->s));
  HIPCHK(hipMemcpyAsync(h->h_rfin, h->mg.rfin, 3 * sizeof(double), hipMemcpyDeviceToHost, h->s));
  HIPCHK(hipStreamSynchronize(h->s));
}

// gsor2d (poiss_z_ft.f90:513-587) for kv = IMGSOR+1..n3mh, after the
// multigrid wavenumbers: the sweep before its loop, then batches of sweeps
// (the first as long as the previous solve needed), each with its test on
// the device, flags read between batches; then the sweeps' max norms for
// the RG rows of mgftresiduemax.dat, kv by kv as the reference's loop over
// kv writes them
static void poisson_sor(lw_handle *h, double test) {
  const Geo &g = h->g;
  MG m = h->mg;
  m.kv0 = h->mg.nkv + 1;
  m.nkv = h->mg.n3mh - h->mg.nkv;
  const double ww = h->st.wwsor;
  const int itr = m.mgitr, ld = m.n3mh;
  auto prof = [&](int sweep) {
    if (!g_prof) return;
    g_prof->cycle = sweep;
    g_prof->kv0 = m.kv0;
    g_prof->nkv = m.nkv;
  };
  prof(0);
  launch_sor_sweep(h->s, g, m, 0.0, ww, test, 0, h->sorlog, ld);
  int launched = 0, batch = h->sor_hint > 1 ? h->sor_hint : 1;
  for (;;) {
    const int n = std::min(batch, itr - launched);
    for (int c = 0; c < n; ++c) {
      prof(launched + c + 1);
      launch_sor_sweep(h->s, g, m, 1.0, ww, test, 1, h->sorlog, ld);
    }
    launched += n;
    st_read(h, h->h_flags, m.done, 256 * sizeof(int));
    st_read(h, h->h_rfin, m.rfin, 3 * sizeof(double));
    st_sync(h);
    bool all = true;
    for (int q = m.kv0 - 1; q < m.kv0 - 1 + m.nkv; ++q) all = all && h->h_flags[q] != 0;
    if (all || launched >= itr) break;
    batch = 4;
  }
  int kmax = 0;
  for (int q = m.kv0 - 1; q < m.kv0 - 1 + m.nkv; ++q) kmax = std::max(kmax, h->h_flags[128 + q]);
  h->sor_hint = kmax;
  if (kmax > 0) {
    h->h_sorlog.resize((size_t)kmax * ld);
    st_read(h, h->h_sorlog.data(), h->sorlog, h->h_sorlog.size() * sizeof(double));
    st_sync(h);
  }
  // RG = errmax * dtconst * float(n3mh) (poiss_z_ft.f90:584)
  for (int q = m.kv0 - 1; q < m.kv0 - 1 + m.nkv; ++q)
    for (int kk = 1; kk <= h->h_flags[128 + q]; ++kk) {
      h->rrow_kind.push_back(1);
      h->rrow_kv.push_back(q + 1);
      h->rrow_kk.push_back(kk);
      h->rrow_val.push_back(h->h_sorlog[(size_t)(kk - 1) * ld + q] * h->sub.dtconst * (double)m.n3mh);
    }
}

// poisson (poiss_z_ft.f90:114-299): z-DFT, then V-cycles of every wavenumber
// until its max-norm residual passes the test. Convergence is decided on the
// device; the host enqueues batches of cycles (the first as long as the
// previous solve needed) and reads the flags between batches. Cycles enqueued
// after a wavenumber converged leave it untouched.
// poisson of poiss_mgrd.f90 (25-224): BiCGSTAB iterations on the device,
// one host read of the iteration's max norm after each; the log rows of
// poiss_itr.dat and ftrpoittr.dat are kept for the host
static void phase_poisson_mgrd(lw_handle *h) {
  const Geo &g = h->g;
  const M3 &m = h->m3;
  const double test1 = h->st.resid1 * h->sub.dtconsti;
  if (h->sub.msub == 1) {
    h->bicg_sub.clear(); h->bicg_icg.clear(); h->bicg_rm.clear(); h->bicg_loc.clear();
    h->solve_icg.clear(); h->solve_rm.clear();
  }
  launch_mgrd_init(h->s, g, m, h->f, test1);
  const int64_t itr = h->st.mgitr;
  int64_t icg = 1;
  double resm = 0.0;
  for (; icg <= itr; ++icg) {
    launch_mgrd_iteration(h->s, g, m, h->f, icg == 1);
    HIPCHK(hipMemcpyAsync(h->h_rfin, m.sc + kM3Resm, 2 * sizeof(double), hipMemcpyDeviceToHost, h->s));
    HIPCHK(hipStreamSynchronize(h->s));
    resm = h->h_rfin[0];
    const int64_t at = (int64_t)h->h_rfin[1], ni = g.n1m, nij = ni * g.n2m;
    h->bicg_sub.push_back(h->sub.msub);
    h->bicg_icg.push_back(icg);
    h->bicg_rm.push_back(resm * h->sub.dtconst);
    h->bicg_loc.push_back(at % ni + 1);
    h->bicg_loc.push_back((at % nij) / ni + 1);
    h->bicg_loc.push_back(at / nij + 1);
    if (resm < test1) break;
  }
  if (icg > itr) fprintf(stderr, "leswht: === MUTLGRID : NOT CONVERGED ===\n");
  launch_mgrd_finish(h->s, g, m, h->f);
  h->solve_icg.push_back(icg);
  h->solve_rm.push_back(resm * h->sub.dtconst);
  if (icg > h->cyc_max) h->cyc_max = icg;
  h->cyc_sum += icg;
  h->mgres_n = 0;
}

// poisson of poiss_x_ft.f90 (115-232): the wavenumbers up to IMGSOR by
// mg2d, the rest by gsor2d, each running its own cycle count; the kernels
// skip wavenumbers whose flag is set, so the host checks the flags only
// every few cycles. Then the rows mg2d (wavenumbers 1 and 2) and gsor2d (at
// its limit) append to mgftresiduemax.dat, and the inverse DFT.
static void phase_poisson_xft(lw_handle *h) {
  const Geo &g = h->g;
  const Xft &x = h->xft;
  const lw_settings &st = h->st;
  const int n1mh = x.n1mh;
  launch_xft_forward(h->s, g, h->f, x);
  HIPCHK(hipMemsetAsync(x.pc, 0, sizeof(double2) * (size_t)n1mh * (g.n2m + 2) * g.n3m, h->s));
  HIPCHK(hipMemsetAsync(x.gg, 0, sizeof(double2) * (size_t)n1mh * (g.n2m + 2) * (x.n3md + 1), h->s));
  HIPCHK(hipMemsetAsync(x.resd, 0, sizeof(double2) * (size_t)n1mh * g.n2m * x.n3md, h->s));
  HIPCHK(hipMemsetAsync(x.done, 0, sizeof(int) * n1mh, h->s));
  HIPCHK(hipMemsetAsync(x.ii, 0, sizeof(int) * n1mh, h->s));
  HIPCHK(hipMemsetAsync(x.nact, 0, sizeof(int) * 2, h->s));
  HIPCHK(hipMemsetAsync(x.sres, 0, sizeof(unsigned long long) * n1mh, h->s));
  const double test1 = st.resid1 * h->sub.dtconsti;
  const double test = test1 / (double)n1mh * (double)g.n1m * 0.9;
  const int nmg = (int)std::max<int64_t>(0, std::min<int64_t>(st.imgsor, n1mh)), nsor = n1mh - nmg;
  const int mgitr = (int)st.mgitr;
  if (nmg > 0) launch_xft_toplevel(h->s, g, x, 1, nmg, 0, test, 0);
  if (nsor > 0) launch_xft_sor(h->s, g, x, nmg + 1, nsor, 0.0, st.wwsor, test, 0);
  int *nact = h->h_flags;
  auto poll = [&]() {
    HIPCHK(hipMemcpyAsync(nact, x.nact, 2 * sizeof(int), hipMemcpyDeviceToHost, h->s));
    HIPCHK(hipStreamSynchronize(h->s));
  };
  poll();
  int amg = nmg > 0 ? nact[0] : 0, asor = nsor > 0 ? nact[1] : 0;
  for (int cyc = 1; cyc <= mgitr && (amg > 0 || asor > 0); ++cyc) {
    if (amg > 0) {
      launch_xft_vcycle(h->s, g, x, 1, nmg);
      launch_xft_toplevel(h->s, g, x, 1, nmg, 1, test, cyc);
    }
    if (asor > 0) launch_xft_sor(h->s, g, x, nmg + 1, nsor, 1.0, st.wwsor, test, cyc);
    check_launches("poiss_x_ft");
    if (cyc % 4 == 0 || cyc == mgitr) {
      poll();
      amg = amg > 0 ? nact[0] : 0;
      asor = asor > 0 ? nact[1] : 0;
    }
  }
  std::vector<int> done(n1mh), ii(n1mh);
  std::vector<double> sfin(n1mh);
  HIPCHK(hipMemcpyAsync(done.data(), x.done, sizeof(int) * n1mh, hipMemcpyDeviceToHost, h->s));
  HIPCHK(hipMemcpyAsync(ii.data(), x.ii, sizeof(int) * n1mh, hipMemcpyDeviceToHost, h->s));
  HIPCHK(hipMemcpyAsync(sfin.data(), x.sfin, sizeof(double) * n1mh, hipMemcpyDeviceToHost, h->s));
  HIPCHK(hipStreamSynchronize(h->s));
  if (h->sub.msub == 1) {
    h->rrow_kind.clear(); h->rrow_kv.clear(); h->rrow_kk.clear(); h->rrow_val.clear();
  }
  bool limit = false;
  int cmax = 0;
  for (int q = 0; q < n1mh; ++q) {
    const int iv = q + 1;
    const int64_t kk = done[q] ? ii[q] : (int64_t)mgitr + 1;
    const double rg = sfin[q] * h->sub.dtconst * (double)n1mh / (double)g.n1m;
    if (!done[q]) limit = true;
    if (iv <= nmg && iv <= 2) {
      h->rrow_kind.push_back(2); h->rrow_kv.push_back(iv); h->rrow_kk.push_back(kk); h->rrow_val.push_back(rg);
    } else if (iv > nmg && !done[q]) {
      h->rrow_kind.push_back(3); h->rrow_kv.push_back(iv); h->rrow_kk.push_back(kk); h->rrow_val.push_back(rg);
    }
    const int c = done[q] ? ii[q] : mgitr;
    cmax = c > cmax ? c : cmax;
    h->cyc_sum += c;
  }
  if (limit) fprintf(stderr, "leswht: poiss_x_ft: ITERATION LIMIT EXCEEDED.\n");
  if (cmax > h->cyc_max) h->cyc_max = cmax;
  launch_xft_inverse(h->s, g, h->f, x, (int)g.ich);
}

static void phase_poisson(lw_handle *h) {
  if (h->st.poisson == 1) {
    phase_poisson_mgrd(h);
    return;
  }
  if (h->st.poisson == 3) {
    phase_poisson_xft(h);
    return;
  }
  if (h->st.poisson == 2) {
    // poiss_ftft is direct: no cycles, no residual rows
    launch_poisson_ftft(h->s, h->g, h->f, h->ftft);
    st_gate();  // where the z-DFT path's cases wait after their solve
    return;
  }
  const Geo &g = h->g;
  MG &mg = h->mg;
  launch_fft_forward(h->s, g, mg, h->f);
  launch_mg_zero(h->s, g, mg);
  const double test1 = h->st.resid1 * h->sub.dtconsti;
  const double test = test1 / (double)mg.n3mh * (double)g.n3m * 0.8;
  if (g_prof) {
    g_prof->mg_first = g_prof->recs.size();
    g_prof->cycle = 0;
  }
  // with the check folded, the test after a toplevel runs in the next
  // cycle's first launch, and a k_mg_check only ends each batch
  const bool fold = mg_folds_check(mg);
  const int nmg = mg.nkv;  // kv = 1..nmg by multigrid, the rest by line-SOR
  if (g_prof) g_prof->nkv = nmg;
  if (h->mg_split && !g_rec && mg.fused && nmg >= 2 && mg.mgitr > 0) {
    poisson_split(h, test, fold);
  } else {
    launch_mg_toplevel(h->s, g, mg, 0, 0.0, test, !fold);
    int launched = 0, batch = h->mg_hint > 1 ? h->mg_hint : 1;
    bool all_done = false;
    while (!all_done && launched < mg.mgitr) {
      if (batch > mg.mgitr - launched) batch = mg.mgitr - launched;
      for (int c = 0; c < batch; ++c) {
        const int m = launched + c + 1;  // this V-cycle's number in the solve
        if (g_prof) g_prof->cycle = m;
        const int cyc = fold && !(c == 0 && launched > 0) ? m - 1 : -1;
        launch_mg_cycle(h->s, g, mg, test, cyc, !fold || c == batch - 1);
      }
      launched += batch;
      st_read(h, h->h_flags, mg.done, 256 * sizeof(int));
      st_read(h, h->h_rfin, mg.rfin, 3 * sizeof(double));
      st_sync(h);
      all_done = true;
      for (int q = 0; q < nmg; ++q) all_done = all_done && h->h_flags[q] != 0;
      // further batches of 2 cycles; in a batch, where every read of the
      // flags is a rendezvous of all its cases, batches that grow with the
      // cycles done (half of them), so a long solve needs a few rendezvous
      // instead of one per 2 cycles (surplus cycles of a converged
      // wavenumber leave it untouched)
      batch = g_rec ? std::max(2, launched / 2) : 2;
    }
    if (launched == 0) {  // every wavenumber passed on the first toplevel
      st_read(h, h->h_flags, mg.done, 256 * sizeof(int));
      st_read(h, h->h_rfin, mg.rfin, 3 * sizeof(double));
      st_sync(h);
    }
  }
  if (h->sub.msub == 1) {
    h->rrow_kind.clear(); h->rrow_kv.clear(); h->rrow_kk.clear(); h->rrow_val.clear();
  }
  // mg2d's rows come first: its wavenumbers precede gsor2d's
  for (int q = 0; q < nmg && q < 3; ++q) {
    h->rrow_kind.push_back(0);
    h->rrow_kv.push_back(q + 1);
    h->rrow_kk.push_back(h->h_flags[q] == 2 ? (int64_t)mg.mgitr + 1 : h->h_flags[128 + q]);
    h->rrow_val.push_back(h->h_rfin[q] * h->sub.dtconst);
  }
  if (nmg < mg.n3mh) poisson_sor(h, test);
  int cmax = 0, crest = 0;
  bool limit = false;
  for (int q = 0; q < mg.n3mh; ++q) {
    const int c = h->h_flags[128 + q];
    cmax = c > cmax ? c : cmax;
    if (q > 0 && q < nmg) crest = c > crest ? c : crest;
    h->cyc_sum += c;
    limit = limit || h->h_flags[q] != 1;
  }
  if (limit) fprintf(stderr, "leswht: ITERATION LIMIT EXCEEDED.\n");
  // mgftresiduemax.dat: kv <= 3 with KK (the reference's loop index: the
  // cycles done, or MGITR + 1 when the limit ended the loop) and the final
  // max norm times dtconst
  if (h->sub.msub == 1) h->mgres_n = 0;
  for (int q = 0; q < nmg && q < 3; ++q) {
    if (h->mgres_n >= 9) break;
    const int64_t n = h->mgres_n++;
    h->mgres_kv[n] = q + 1;
    h->mgres_kk[n] = h->h_flags[q] == 2 ? (int64_t)mg.mgitr + 1 : h->h_flags[128 + q];
    h->mgres_rm[n] = h->h_rfin[q] * h->sub.dtconst;
  }
  if (g_prof) g_prof->resolve_mg(h->h_flags + 128, mg.n3mh);
  if (cmax > h->cyc_max) h->cyc_max = cmax;
  if (h->mg_split && !g_rec && mg.fused && nmg >= 2) {
    h->mg_hint1 = h->h_flags[128];
    h->mg_hint = crest;
  } else {
    h->mg_hint = cmax;
  }
  st_gate();
  launch_fft_inverse(h->s, g, mg, h->f);
}

static void phase_energy(lw_handle *h) {
  if (!h->heat) return;
  const Geo &g = h->g;
  Body b{};
  launch_rhs_t(h->s, g, h->sub, h->f, h->les);
  if (h->st.ibmon != 0 && h->st.iconjg == 0) launch_ibm(h->s, g, h->sub, h->f, h->ib, 3);
  launch_convbc_t(h->s, g, h->sub, h->f, b, h->les, h->st);
  launch_adi_t(h->s, g, h->sub, h->f, h->les);
  launch_energy_bcs(h->s, g, h->f, b, h->st);
}

// after the substeps (solver.f90:150-172): time update, convergence_check and
// masscheck, the history counter, draglift, heatflux and the bulk
// temperatures of writehistory
static void end_of_step(lw_handle *h) {
  const lw_settings &st = h->st;
  const double time_prev = h->time;
  h->time = h->time + h->dt;
  // addperturb once, in the step that reaches PTB_TST (solver.f90:150-160),
  // before the step's diagnostics
  if (h->ptb_armed && !h->ptb_applied && time_prev < st.ptb_tst && h->time >= st.ptb_tst) apply_perturb(h);
  launch_diag(h->s, h->g, h->f, (int)(st.masson == 1 && st.ibmon == 1));
  const int hstep = st.npin > 0 ? (h->ntime % st.npin) == 0 : 1;
  if (hstep) h->ihist += 1;
  h->hist_written = hstep;
  unsigned act = 0;
  if (st.ibmon == 1) {
    act |= (1u << D_IBU) | (1u << D_IBV) | (1u << D_IBW);
    if (st.bc_ybtm == 0) act |= (1u << D_YB_U) | (1u << D_YB_W);
    if (st.bc_ytop == 0) act |= (1u << D_YT_U) | (1u << D_YT_W);
    if (st.bc_zbtm == 0) act |= (1u << D_ZB_U) | (1u << D_ZB_V);
    if (st.bc_ztop == 0) act |= (1u << D_ZT_U) | (1u << D_ZT_V);
  }
  if (h->heat) {
    for (int q = D_HXM; q <= D_HZP; ++q) act |= 1u << q;
    if (hstep)
      for (int q = D_TFV; q <= D_TSI; ++q) act |= 1u << q;
  }
  launch_step_diag(h->s, h->g, h->f, h->ib, h->dplanes, act, st.re);
  launch_hist_update(h->s, h->f, h->hist, act, st.ibmon == 1, h->heat, hstep, h->time, h->dt, st.avg_tst,
                     h->g.xl * h->g.yl * h->g.zl);
  // writehistory's fcmfr.dat quantities (misc_hist.f90:112-136)
  if (h->g.ich != 0 && hstep) launch_chan_hist(h->s, h->g, h->f, h->ib[0], st, h->hist, h->time, h->dt);
}

// ---------------------------------------------------------------------------
// time-averaged statistics: field_avg and field_avg_finalize
// (misc_hist.f90:287-697); fav files in the reference's record layout
// ---------------------------------------------------------------------------
static void avg_zero(lw_handle *h) {
  const size_t ni = (size_t)h->g.n1m * h->g.n2m * h->g.n3m;
  for (int l = 0; l < h->navg; ++l) HIPCHK(hipMemsetAsync(h->avg.a[l], 0, ni * sizeof(double), h->s));
}

static std::string fav_name(int64_t i0, int64_t i1) {
  char b[32];
  snprintf(b, sizeof b, "fav%07lld-%07lld", (long long)i0, (long long)i1);
  return b;
}

// one fav file: n1m n2m n3m re / timeinit timeend ihistinit ihistend (reals) /
// u v w / uiuj(6) / p / p2 / vor(3) / vor2(6) / ss / [T / T2]
// (misc_hist.f90:484-499); arr holds navg interior arrays, i fastest
static void write_fav(const lw_handle *h, const std::string &path, const std::vector<double> &arr, double t0,
                      double t1, int64_t i0, int64_t i1) {
  FILE *fp = fopen(path.c_str(), "wb");
  if (!fp) {
    g_err = "cannot write " + path;
    throw LW_ERR_IO;
  }
  auto rec = [&](const void *d, size_t len) {
    const int32_t l = (int32_t)len;
    fwrite(&l, 4, 1, fp);
    fwrite(d, 1, len, fp);
    fwrite(&l, 4, 1, fp);
  };
  const size_t ni = (size_t)h->g.n1m * h->g.n2m * h->g.n3m;
  char hdr[32];
  const int64_t n[3] = {h->g.n1m, h->g.n2m, h->g.n3m};
  memcpy(hdr, n, 24);
  memcpy(hdr + 24, &h->st.re, 8);
  rec(hdr, 32);
  const double tt[4] = {t0, t1, (double)i0, (double)i1};
  rec(tt, 32);
  // records of 1, 1, 1, 6, 1, 1, 3, 6, 1 (, 1, 1) arrays
  static const int per[11] = {1, 1, 1, 6, 1, 1, 3, 6, 1, 1, 1};
  int l = 0;
  for (int r = 0; r < 11 && l < h->navg; ++r) {
    rec(arr.data() + (size_t)l * ni, (size_t)per[r] * ni * 8);
    l += per[r];
  }
  fclose(fp);
}

static void append_manifest(const lw_handle *h, const std::string &name) {
  const std::string path = h->avg_dir + "/fav_manifest.dat";
  FILE *fp = fopen(path.c_str(), "a");
  if (!fp) {
    g_err = "cannot append to " + path;
    throw LW_ERR_IO;
  }
  fprintf(fp, "%s\n", name.c_str());
  fclose(fp);
}

static std::vector<double> avg_download(lw_handle *h) {
  const size_t ni = (size_t)h->g.n1m * h->g.n2m * h->g.n3m;
  std::vector<double> a((size_t)h->navg * ni);
  for (int l = 0; l < h->navg; ++l)
    HIPCHK(hipMemcpyAsync(a.data() + (size_t)l * ni, h->avg.a[l], ni * 8, hipMemcpyDeviceToHost, h->s));
  HIPCHK(hipStreamSynchronize(h->s));
  return a;
}

// field_avg (misc_hist.f90:321-527), called after the step's time and ihist
// updates with the step's dt
static void field_avg(lw_handle *h) {
  const lw_settings &st = h->st;
  if (h->time <= st.avg_tst) return;
  if (!h->avg_started) {
    h->avg_started = true;
    h->avg_count = 0;
    h->avg_timeinit = st.avg_tst;
    h->avg_ihistinit = h->ihist;
    avg_zero(h);
  }
  launch_field_avg(h->s, h->g, h->f, h->avg, h->dt, h->heat);
  h->avg_count += 1;
  if (h->avg_count >= st.npriavg) {
    if (!h->avg_dir.empty()) {
      const std::string name = fav_name(h->avg_ihistinit, h->ihist);
      write_fav(h, h->avg_dir + "/" + name, avg_download(h), h->avg_timeinit, h->time, h->avg_ihistinit, h->ihist);
      append_manifest(h, name);
    }
    avg_zero(h);
    h->avg_count = 0;
    h->avg_timeinit = h->time;
    h->avg_ihistinit = h->ihist;
  }
}

static void read_stats(lw_handle *h, lw_step_stats *stats) {
  st_read(h, h->h_sumres + 8, h->f.redmax, 3 * sizeof(unsigned long long));
  st_read(h, h->h_red, h->f.red, 16 * sizeof(double));
  st_read(h, h->h_red + 16, h->hist, 24 * sizeof(double));
  st_sync(h);
  memcpy(h->h_hist, h->h_red + 16, 24 * sizeof(double));
  for (int q = 0; q < 4; ++q) {
    float a = 0.f, b = 0.f, c = 0.f;
    if (h->tev_valid) {
      HIPCHK(hipEventElapsedTime(&a, h->tev[0][2 * q], h->tev[0][2 * q + 1]));
      HIPCHK(hipEventElapsedTime(&b, h->tev[1][2 * q], h->tev[1][2 * q + 1]));
      HIPCHK(hipEventElapsedTime(&c, h->tev[2][2 * q], h->tev[2][2 * q + 1]));
    }
    h->s_phase[q] = ((double)a + (double)b + (double)c) * 1e-3;
  }
  double c[3];
  memcpy(c, h->h_sumres + 8, 3 * 8);
  h->cflmax = c[0];
  h->dvmax = c[1];
  h->qmmax = c[2];
  if (h->les && h->st.idvmon == 1) {
    h->csgsts = h->h_red[4];
    if (h->heat) h->csgshf = h->h_red[5];
  }
  if (!stats) return;
  stats->dt = h->dt;
  stats->time = h->time;
  stats->cflmax = h->cflmax;
  stats->dvmax = h->dvmax;
  stats->qmmax = h->qmmax;
  stats->csgsts = h->csgsts;
  stats->csgshf = h->csgshf;
  stats->ntime = h->ntime;
  stats->mg_cycles_max = h->cyc_max;
  stats->mg_cycles_sum = h->cyc_sum;
  stats->ms_sgs = h->s_phase[0] * 1e3;
  stats->ms_momentum = h->s_phase[1] * 1e3;
  stats->ms_poisson = h->s_phase[2] * 1e3;
  stats->ms_energy = h->s_phase[3] * 1e3;
}

// ---------------------------------------------------------------------------
// autotuner of the fused multigrid kernels' tile sizes (lines per workgroup
// of the down, up and top kernels). Tiles change only how lines are grouped
// into workgroups, never the arithmetic, so any choice gives the same bits.
// The tuner times whole Poisson solves of the divergence of the current
// field (the solve's own initial guess is zero, and divgs, the solve and
// phi are recomputed in every substep, so the state the steps see is
// unchanged), first for the main stream's tiles, then, with the kv = 1
// split, for the kv = 1 side's.
// ---------------------------------------------------------------------------
extern "C" int lw_autotune(lw_handle *h, int64_t reps, lw_tune_result *out) {
  if (!h || !out || reps < 1) return fail(LW_ERR_ARG, "lw_autotune: bad argument");
  memset(out, 0, sizeof *out);
  MG &mg = h->mg;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  try {
    for (int q = 0; q < 3; ++q) {
      out->tile[q] = mg.tile[q];
      out->tile1[q] = mg.tile1[q];
    }
    // separate launches, or the 3-D solver: no tiles to tune
    if (!mg.fused || h->st.poisson != 0) return LW_OK;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    const double dt0 = h->dt;
    if (!(h->dt > 0.0)) h->dt = h->st.dt_size;
    substepinit(h, 1);
    launch_divgs(h->s, h->g, h->sub, h->f, (int)(h->st.masson == 1 && h->st.imovingon == 0));
    phase_poisson(h);  // warm-up; sets the batch hints
    auto solve_ms = [&]() {
      float best = 1e30f;
      for (int64_t r = 0; r < reps; ++r) {
        HIPCHK(hipEventRecord(e0, h->s));
        phase_poisson(h);
        HIPCHK(hipEventRecord(e1, h->s));
        HIPCHK(hipEventSynchronize(e1));
        float t = 0.f;
        HIPCHK(hipEventElapsedTime(&t, e0, e1));
        best = t < best ? t : best;
      }
      return (double)best;
    };
    static const int cand[][3] = {{2, 2, 4}, {2, 2, 2}, {4, 4, 4}, {4, 4, 8}, {2, 2, 8},
                                  {8, 8, 8}, {2, 4, 4}, {4, 2, 4}};
    const int ncand = (int)(sizeof cand / sizeof cand[0]);
    out->ms_default = solve_ms();
    double best = out->ms_default;
    const int split = h->mg_split && mg.n3mh >= 2 && mg.mgitr > 0;
    for (int side = 0; side < 1 + split; ++side) {
      int *t = side == 0 ? mg.tile : mg.tile1;
      int keep[3] = {t[0], t[1], t[2]};
      for (int c = 0; c < ncand; ++c) {
        if (cand[c][0] == keep[0] && cand[c][1] == keep[1] && cand[c][2] == keep[2]) continue;
        for (int q = 0; q < 3; ++q) t[q] = cand[c][q];
        const double ms = solve_ms();
        out->candidates += 1;
        if (ms < best) {
          best = ms;
          for (int q = 0; q < 3; ++q) keep[q] = cand[c][q];
        }
      }
      for (int q = 0; q < 3; ++q) t[q] = keep[q];
    }
    out->ms_best = best;
    for (int q = 0; q < 3; ++q) {
      out->tile[q] = mg.tile[q];
      out->tile1[q] = mg.tile1[q];
    }
    h->dt = dt0;
    h->mg_hint = 4;
    h->mg_hint1 = 4;
    h->cyc_max = 0;
    h->cyc_sum = 0;
    HIPCHK(hipEventDestroy(e0));
    HIPCHK(hipEventDestroy(e1));
    return LW_OK;
  } catch (int code) {
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    return code;
  }
}

// records of a Fortran unformatted sequential file (4-byte markers)
static bool read_records(const std::string &path, std::vector<std::vector<char>> &recs) {
  FILE *fp = fopen(path.c_str(), "rb");
  if (!fp) return false;
  recs.clear();
  int32_t l = 0, l2 = 0;
  bool ok = true;
  while (fread(&l, 4, 1, fp) == 1) {
    std::vector<char> r((size_t)l);
    if (l < 0 || fread(r.data(), 1, (size_t)l, fp) != (size_t)l || fread(&l2, 4, 1, fp) != 1 || l2 != l) {
      ok = false;
      break;
    }
    recs.push_back(std::move(r));
  }
  fclose(fp);
  return ok;
}

extern "C" int lw_set_option(lw_handle *h, const char *name, int64_t value) {
  if (!h || !name) return fail(LW_ERR_ARG, "null argument");
  const std::string n(name);
  if (n == "mg_split") {
    h->mg_split = value != 0;
    return LW_OK;
  }
  if (n == "mg_kv1") {
    h->mg_kv1 = value != 0;
    return LW_OK;
  }
  if (n == "mg_fp32") {
    try {
      set_mg_fp32(h, value != 0);
    } catch (int code) {
      return code;
    }
    return LW_OK;
  }
  if (n == "sgs_fp32") {
    if (value != 0 && !h->sgs_fused)
      return fail(LW_ERR_ARG, "sgs_fp32: needs the one-pass dynamic SGS kernels (LW_SGS_FUSED=0 keeps the separate ones)");
    h->sgs_fp32 = value != 0;
    return LW_OK;
  }
  return fail(LW_ERR_ARG, "unknown option " + n);
}

extern "C" int lw_set_avg_dir(lw_handle *h, const char *dir) {
  if (!h) return fail(LW_ERR_ARG, "null handle");
  h->avg_dir = dir ? dir : "";
  return LW_OK;
}

// field_avg_finalize (misc_hist.f90:530-697): the open interval's file, then
// the sum of every file the manifest lists, written under combined/
extern "C" int lw_avg_finalize(lw_handle *h) {
  if (!h) return fail(LW_ERR_ARG, "null handle");
  if (h->st.iavg != 1 || !h->avg_started) return LW_OK;
  try {
    const bool files = !h->avg_dir.empty();
    if (h->avg_count > 0 && files) {
      const std::string name = fav_name(h->avg_ihistinit, h->ihist);
      write_fav(h, h->avg_dir + "/" + name, avg_download(h), h->avg_timeinit, h->time, h->avg_ihistinit, h->ihist);
      append_manifest(h, name);
    }
    avg_zero(h);
    HIPCHK(hipStreamSynchronize(h->s));
    if (!files) return LW_OK;
    const size_t ni = (size_t)h->g.n1m * h->g.n2m * h->g.n3m;
    std::vector<double> acc((size_t)h->navg * ni, 0.0);
    FILE *mf = fopen((h->avg_dir + "/fav_manifest.dat").c_str(), "r");
    if (!mf) return LW_OK;
    bool have = false;
    int64_t count = 0, ihistinit = 0, ihistend = 0;
    double timeinit = 0, timeend = h->st.avg_tst;
    char line[512];
    std::vector<std::vector<char>> recs;
    static const int per[11] = {1, 1, 1, 6, 1, 1, 3, 6, 1, 1, 1};
    while (fgets(line, sizeof line, mf)) {
      std::string nm(line);
      while (!nm.empty() && (nm.back() == '\n' || nm.back() == '\r' || nm.back() == ' ')) nm.pop_back();
      if (nm.empty()) continue;
      if (!read_records(h->avg_dir + "/" + nm, recs)) continue;
      if (recs.size() < 2 || recs[1].size() != 32) continue;
      const int nrec = h->heat ? 11 : 9;
      if ((int)recs.size() < 2 + nrec) {
        fclose(mf);
        g_err = "fav file " + nm + " has too few records";
        throw LW_ERR_IO;
      }
      int l = 0;
      for (int r = 0; r < nrec; ++r) {
        if (recs[2 + r].size() != (size_t)per[r] * ni * 8) {
          fclose(mf);
          g_err = "fav file " + nm + " does not match the grid";
          throw LW_ERR_IO;
        }
        const double *d = (const double *)recs[2 + r].data();
        double *a = acc.data() + (size_t)l * ni;
        for (size_t q = 0; q < (size_t)per[r] * ni; ++q) a[q] = a[q] + d[q];
        l += per[r];
      }
      double tt[4];
      memcpy(tt, recs[1].data(), 32);
      count += 1;
      if (!have) {
        timeinit = tt[0];
        ihistinit = (int64_t)llround(tt[2]);
        have = true;
      }
      timeend = tt[1];
      ihistend = (int64_t)llround(tt[3]);
    }
    fclose(mf);
    if (!have || count <= 0) return LW_OK;
    std::string dir = h->avg_dir + "/combined";
    std::error_code ec;
    std::filesystem::create_directories(dir, ec);
    if (ec) dir = h->avg_dir;
    write_fav(h, dir + "/" + fav_name(ihistinit, ihistend), acc, timeinit, timeend, ihistinit, ihistend);
    return LW_OK;
  } catch (int code) {
    return code;
  }
}

extern "C" int lw_step(lw_handle *h, int64_t nsteps, lw_step_stats *stats) {
  if (!h || nsteps < 0) return fail(LW_ERR_ARG, "bad argument");
  ProfScope ps(h->prof);
  try {
    (void)hipGetLastError();  // clear what earlier calls on this thread left
    for (int64_t n = 0; n < nsteps; ++n) {
      h->cyc_max = 0;
      h->cyc_sum = 0;
      if (h->ptb_pending) apply_perturb(h);
      stepinit(h);
      check_launches("stepinit");
      for (int ms = 1; ms <= 3; ++ms) {
        substepinit(h, ms);
        st_gate();
        hipEvent_t *ev = h->tev[ms - 1];
        st_event(h, ev[0]);
        phase_sgs(h);
        check_launches("sgscalc");
        st_event(h, ev[1]);
        st_event(h, ev[2]);
        phase_rhsnlhs(h);
        check_launches("rhsnlhs");
        st_event(h, ev[3]);
        st_event(h, ev[4]);
        launch_divgs(h->s, h->g, h->sub, h->f, (int)(h->st.masson == 1 && h->st.imovingon == 0));
        phase_poisson(h);
        check_launches("divgs/poisson");
        st_event(h, ev[5]);
        launch_projection(h->s, h->g, h->sub, h->f);
        after_projection(h);
        if (h->st.ibmon == 1) launch_lagforce(h->s, h->g, h->sub, h->f, h->ib);
        check_launches("ucalc/pcalc");
        st_event(h, ev[6]);
        phase_energy(h);
        check_launches("rhsnlhs_t");
        st_event(h, ev[7]);
        if (ms == 3) h->tev_valid = !g_rec;
      }
      end_of_step(h);
      if (h->st.iavg == 1) field_avg(h);
      check_launches("end of step");
      if (g_prof) {
        HIPCHK(hipStreamSynchronize(h->s));
        g_prof->flush();
      }
    }
    read_stats(h, stats);
    if (g_prof) g_prof->flush();
    if (!(h->dvmax == h->dvmax)) return fail(LW_ERR_NUMERIC, "non-finite divergence after the step");
    return LW_OK;
  } catch (int code) {
    return code;
  }
}

extern "C" int lw_phase(lw_handle *h, int phase) {
  if (!h) return fail(LW_ERR_ARG, "null handle");
  ProfScope ps(h->prof);
  try {
    (void)hipGetLastError();
    if (phase == 0) {
      h->cyc_max = 0;
      h->cyc_sum = 0;
      if (h->ptb_pending) apply_perturb(h);
      stepinit(h);
    } else if (phase >= 11 && phase <= 13) substepinit(h, phase - 10);
    else if (phase == 1) phase_sgs(h);
    else if (phase == 2) phase_rhsnlhs(h);
    else if (phase == 3) launch_divgs(h->s, h->g, h->sub, h->f, (int)(h->st.masson == 1 && h->st.imovingon == 0));
    else if (phase == 4) phase_poisson(h);
    else if (phase == 5) {
      launch_projection(h->s, h->g, h->sub, h->f);
      after_projection(h);
      if (h->st.ibmon == 1) launch_lagforce(h->s, h->g, h->sub, h->f, h->ib);
    } else if (phase == 6) phase_energy(h);
    else if (phase == 7) {
      end_of_step(h);
      read_stats(h, nullptr);
    } else return fail(LW_ERR_ARG, "unknown phase");
    check_launches("lw_phase");
    HIPCHK(hipStreamSynchronize(h->s));
    if (g_prof) g_prof->flush();
    return LW_OK;
  } catch (int code) {
    return code;
  }
}

extern "C" int lw_get_history(const lw_handle *h, lw_history *out) {
  if (!h || !out) return fail(LW_ERR_ARG, "null argument");
  memset(out, 0, sizeof(*out));
  out->ntime = h->ntime;
  out->ihist = h->ihist;
  out->written = h->hist_written;
  out->time = h->time;
  out->dt = h->dt;
  out->cflmax = h->cflmax;
  out->dvmax = h->dvmax;
  out->qmmax = h->qmmax;
  for (int l = 0; l < 3; ++l) {
    out->cd[l] = h->h_hist[l];
    out->cd_avg[l] = h->h_hist[3 + l];
    out->hflux[l] = h->h_hist[6 + l];
    out->hflux_avg[l] = h->h_hist[9 + l];
  }
  out->tbulk[0] = h->h_hist[12];
  out->tbulk[1] = h->h_hist[13];
  out->tbulk_avg[0] = h->h_hist[14];
  out->tbulk_avg[1] = h->h_hist[15];
  out->s_sgs = h->s_phase[0];
  out->s_rhsnlhs = h->s_phase[1];
  out->s_poisson = h->s_phase[2];
  out->s_energy = h->s_phase[3];
  out->qflux = h->h_hist[16];
  out->ubulk = h->h_hist[17];
  out->pmiavg = h->h_hist[18];
  for (int l = 0; l < 3; ++l) out->cmfravg[l] = h->h_hist[19 + l];
  return LW_OK;
}

// tracer cells from the TRACE block of settings.input (misc_init.f90:47-52):
// the reference keeps at most 64 (trpts(64, 3), mod_common.f90:9) and reads
// u(i+1), v(j+1), w(k+1), so each cell must have 0 <= i < n1, 0 <= j < n2,
// 0 <= k < n3
extern "C" int lw_set_trace(lw_handle *h, int64_t ntrace, int64_t ntr, const int64_t *pts) {
  if (!h) return fail(LW_ERR_ARG, "null handle");
  if (ntrace < 0 || ntrace > 64) return fail(LW_ERR_ARG, "lw_set_trace: NTRACE must be in 0..64");
  if (ntrace > 0 && (!pts || ntr < 1)) return fail(LW_ERR_ARG, "lw_set_trace: points and NTR >= 1 required");
  int ip[64 * 3];
  for (int64_t n = 0; n < ntrace; ++n) {
    const int64_t i = pts[3 * n], j = pts[3 * n + 1], k = pts[3 * n + 2];
    if (i < 0 || i >= h->g.n1 || j < 0 || j >= h->g.n2 || k < 0 || k >= h->g.n3)
      return fail(LW_ERR_ARG, "lw_set_trace: trace cell " + std::to_string(n + 1) + " outside the grid");
    ip[3 * n] = (int)i; ip[3 * n + 1] = (int)j; ip[3 * n + 2] = (int)k;
  }
  try {
    if (ntrace > 0) {
      HIPCHK(hipMemcpyAsync(h->d_trpts, ip, sizeof(int) * 3 * ntrace, hipMemcpyHostToDevice, h->s));
      HIPCHK(hipStreamSynchronize(h->s));
    }
  } catch (int code) {
    return code;
  }
  h->ntrace = ntrace;
  h->ntr = ntrace > 0 ? ntr : 1;
  for (int64_t q = 0; q < 3 * ntrace; ++q) h->trpts[q] = pts[q];
  return LW_OK;
}

extern "C" int lw_trace_info(const lw_handle *h, int64_t *ntrace, int64_t *ntr, int64_t *pts) {
  if (!h || !ntrace || !ntr) return fail(LW_ERR_ARG, "null argument");
  *ntrace = h->ntrace;
  *ntr = h->ntr;
  if (pts)
    for (int64_t q = 0; q < 3 * h->ntrace; ++q) pts[q] = h->trpts[q];
  return LW_OK;
}

// the values tracer writes (misc_hist.f90:198-205) for the current fields:
// out[c ntrace + n], c = u, v, w at the cell centre, then p
extern "C" int lw_get_trace(lw_handle *h, double *out, int64_t cap) {
  if (!h || (!out && h->ntrace > 0)) return fail(LW_ERR_ARG, "null argument");
  if (cap < 4 * h->ntrace) return fail(LW_ERR_ARG, "lw_get_trace: output needs 4 * NTRACE values");
  if (h->ntrace == 0) return LW_OK;
  try {
    (void)hipGetLastError();
    launch_trace(h->s, h->g, h->f, h->d_trpts, (int)h->ntrace, h->d_trace);
    check_launches("tracer");
    HIPCHK(hipMemcpyAsync(out, h->d_trace, sizeof(double) * 4 * h->ntrace, hipMemcpyDeviceToHost, h->s));
    HIPCHK(hipStreamSynchronize(h->s));
    return LW_OK;
  } catch (int code) {
    return code;
  }
}

// the mgftresiduemax.dat rows of the last step (poiss_z_ft.f90:293-297):
// returns their number (3 substeps x min(3, n3m/2+1)); cap limits the copy
extern "C" int64_t lw_get_mg_residuals(const lw_handle *h, int64_t *kv, int64_t *kk, double *rm, int64_t cap) {
  if (!h) return fail(LW_ERR_ARG, "null handle");
  const int64_t n = h->mgres_n < cap ? h->mgres_n : (cap > 0 ? cap : 0);
  for (int64_t q = 0; q < n; ++q) {
    if (kv) kv[q] = h->mgres_kv[q];
    if (kk) kk[q] = h->mgres_kk[q];
    if (rm) rm[q] = h->mgres_rm[q];
  }
  return h->mgres_n;
}

extern "C" int64_t lw_get_residual_rows(const lw_handle *h, int64_t *kind, int64_t *kv, int64_t *kk, double *val,
                                        int64_t cap) {
  if (!h) return fail(LW_ERR_ARG, "null handle");
  const int64_t tot = (int64_t)h->rrow_kv.size();
  const int64_t n = tot < cap ? tot : (cap > 0 ? cap : 0);
  for (int64_t q = 0; q < n; ++q) {
    if (kind) kind[q] = h->rrow_kind[q];
    if (kv) kv[q] = h->rrow_kv[q];
    if (kk) kk[q] = h->rrow_kk[q];
    if (val) val[q] = h->rrow_val[q];
  }
  return tot;
}

extern "C" int64_t lw_get_bicg_log(const lw_handle *h, int64_t *sub, int64_t *icg, double *rm, int64_t *loc,
                                   int64_t cap) {
  if (!h) return fail(LW_ERR_ARG, "null handle");
  const int64_t n = (int64_t)h->bicg_icg.size();
  for (int64_t q = 0; q < n && q < cap; ++q) {
    if (sub) sub[q] = h->bicg_sub[q];
    if (icg) icg[q] = h->bicg_icg[q];
    if (rm) rm[q] = h->bicg_rm[q];
    if (loc)
      for (int c = 0; c < 3; ++c) loc[3 * q + c] = h->bicg_loc[3 * q + c];
  }
  return n;
}

extern "C" int64_t lw_get_bicg_solves(const lw_handle *h, int64_t *icg, double *rm, int64_t cap) {
  if (!h) return fail(LW_ERR_ARG, "null handle");
  const int64_t n = (int64_t)h->solve_icg.size();
  for (int64_t q = 0; q < n && q < cap; ++q) {
    if (icg) icg[q] = h->solve_icg[q];
    if (rm) rm[q] = h->solve_rm[q];
  }
  return n;
}

extern "C" int lw_get_settings(const lw_handle *h, lw_settings *out) {
  if (!h || !out) return fail(LW_ERR_ARG, "null argument");
  *out = h->st;
  return LW_OK;
}

extern "C" int lw_profile_begin(lw_handle *h, const char *kernels) {
  if (!h) return fail(LW_ERR_ARG, "null handle");
  try {
    HIPCHK(hipStreamSynchronize(h->s));
    delete h->prof;
    h->prof = new Prof();
    h->prof->on = true;
    h->prof->all = kernels == nullptr || kernels[0] == 0;
    if (!h->prof->all) {
      std::string k(kernels), cur;
      for (char c : k + ",") {
        if (c == ',') {
          if (!cur.empty()) h->prof->sel.push_back(cur);
          cur.clear();
        } else if (c != ' ') cur += c;
      }
    }
    return LW_OK;
  } catch (int code) {
    return code;
  }
}

extern "C" int lw_profile_end(lw_handle *h) {
  if (!h) return fail(LW_ERR_ARG, "null handle");
  if (h->prof) h->prof->on = false;
  return LW_OK;
}

extern "C" int64_t lw_profile_count(const lw_handle *h) {
  if (!h) return fail(LW_ERR_ARG, "null handle");
  return h->prof ? (int64_t)h->prof->names.size() : 0;
}

extern "C" int lw_profile_get(const lw_handle *h, int64_t idx, char *name, int64_t cap, int64_t *launches,
                              double *ms_total, double *bytes_total) {
  if (!h || !h->prof || idx < 0 || idx >= (int64_t)h->prof->names.size())
    return fail(LW_ERR_ARG, "no such profile entry");
  const Prof &p = *h->prof;
  if (name && cap > 0) {
    strncpy(name, p.names[idx].c_str(), (size_t)cap - 1);
    name[cap - 1] = 0;
  }
  if (launches) *launches = p.launches[idx];
  if (ms_total) *ms_total = p.ms[idx];
  if (bytes_total) *bytes_total = p.bytes[idx];
  return LW_OK;
}

extern "C" int lw_write_field(lw_handle *h, const char *path) {
  if (!h || !path) return fail(LW_ERR_ARG, "null argument");
  const size_t np = (size_t)h->g.NX * h->g.NY * h->g.NZ;
  std::vector<double> u(np), v(np), w(np), p(np), t(h->heat ? np : 0);
  int rc = lw_get_fields(h, u.data(), v.data(), w.data(), p.data(), h->heat ? t.data() : nullptr);
  if (rc) return rc;
  const lw_settings &s = h->st;
  const int64_t n[3] = {h->g.n1, h->g.n2, h->g.n3};
  const int64_t prdic[3] = {s.xprdic, s.yprdic, s.zprdic};
  const int64_t bcs[6] = {s.xprdic ? 4 : 2, s.xprdic ? 4 : 3, s.bc_ybtm, s.bc_ytop, s.zprdic ? 4 : s.bc_zbtm,
                          s.zprdic ? 4 : s.bc_ztop};
  const int64_t ich[2] = {s.ich, s.iconjg};
  return lw_write_fld(path, n, s.re, s.pr, s.gr, h->ihist, h->m, h->time, h->dt, prdic, bcs, ich, u.data(), v.data(),
                      w.data(), p.data(), h->heat ? t.data() : nullptr);
}

// ---------------------------------------------------------------------------
// batched stepping (lw_batch.hip): each case's lw_step runs on a thread of the
// batch that records its launches; the batch launches every kernel once for
// all the cases that issue it
// ---------------------------------------------------------------------------
struct lw_batch {
  std::vector<lw_handle *> hs;
  std::vector<lw_step_stats> stats;
  std::vector<std::string> errs;
  Batch *core = nullptr;
};

static int batch_member(void *ctx, int i, int64_t nsteps) {
  lw_batch *b = (lw_batch *)ctx;
  const int rc = lw_step(b->hs[i], nsteps, &b->stats[i]);
  if (rc != LW_OK) b->errs[i] = g_err;
  return rc;
}

extern "C" int lw_batch_create(lw_handle *const *hs, int64_t n, lw_batch **out) {
  if (!out) return fail(LW_ERR_ARG, "lw_batch_create: null output");
  *out = nullptr;
  if (!hs || n < 1 || n > 4096) return fail(LW_ERR_ARG, "lw_batch_create: 1..4096 handles required");
  for (int64_t i = 0; i < n; ++i) {
    const lw_handle *h = hs[i];
    const std::string c = "lw_batch_create: case " + std::to_string(i) + ": ";
    if (!h) return fail(LW_ERR_ARG, c + "null handle");
    for (int64_t k = 0; k < i; ++k)
      if (hs[k] == h) return fail(LW_ERR_ARG, c + "the same handle twice");
    if (h->batch) return fail(LW_ERR_ARG, c + "the handle already belongs to a batch");
    if (h->st.poisson == 1 || h->st.poisson == 3)
      return fail(LW_ERR_ARG, c + "batched stepping takes the z-DFT or poiss_ftft Poisson solvers");
    if (h->st.iavg == 1) return fail(LW_ERR_ARG, c + "batched stepping takes IAVG = 0");
    if (h->prof) return fail(LW_ERR_ARG, c + "a profile is running on the handle");
  }
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return fail(LW_ERR_DEVICE, "lw_batch_create: no device");
  lw_batch *b = new lw_batch;
  b->hs.assign(hs, hs + n);
  b->stats.assign((size_t)n, lw_step_stats{});
  b->errs.assign((size_t)n, std::string());
  b->core = batch_new(dev, (int)n, batch_member, b);
  if (!b->core) {
    delete b;
    return fail(LW_ERR_DEVICE, "lw_batch_create: the batch's stream could not be created");
  }
  for (lw_handle *h : b->hs) h->batch = b;
  *out = b;
  return LW_OK;
}

extern "C" int lw_batch_step(lw_batch *b, int64_t nsteps, lw_step_stats *stats) {
  if (!b || nsteps < 0) return fail(LW_ERR_ARG, "lw_batch_step: bad argument");
  // work a case enqueued on its own streams comes first
  for (lw_handle *h : b->hs) {
    if (h->s && hipStreamSynchronize(h->s) != hipSuccess) return fail(LW_ERR_DEVICE, "lw_batch_step: stream error");
    if (h->s2 && hipStreamSynchronize(h->s2) != hipSuccess) return fail(LW_ERR_DEVICE, "lw_batch_step: stream error");
  }
  std::vector<int> rc(b->hs.size(), 0);
  const int r = batch_run(b->core, nsteps, rc.data());
  if (stats)
    for (size_t i = 0; i < b->hs.size(); ++i) stats[i] = b->stats[i];
  if (r != LW_OK) {
    for (size_t i = 0; i < rc.size(); ++i)
      if (rc[i] != LW_OK) return fail(rc[i], "lw_batch_step: case " + std::to_string(i) + ": " + b->errs[i]);
    return fail(r, "lw_batch_step: the batch failed earlier: " + batch_error(b->core));
  }
  return LW_OK;
}

extern "C" int lw_batch_info(const lw_batch *b, int64_t info[6]) {
  if (!b || !info) return fail(LW_ERR_ARG, "lw_batch_info: null argument");
  batch_counts(b->core, info);
  return LW_OK;
}

extern "C" int lw_batch_destroy(lw_batch *b) {
  if (!b) return LW_OK;
  batch_free(b->core);
  for (lw_handle *h : b->hs) h->batch = nullptr;
  delete b;
  return LW_OK;
}
