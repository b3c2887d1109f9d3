/*
 * leswht.h - C ABI of the MI355X (gfx950) LESwHT time step.
 *
 * The reference solver has no library boundary: solver.f90 calls external
 * Fortran subroutines with implicit interfaces on module-global state
 * (solver.f90:26-196; state in mod_common.f90, mod_flowarray.f90,
 * mod_poiss.f90). This header replaces that in-process operator API with a
 * handle that owns the state in HBM:
 *
 *   reference (solver.f90)                         here
 *   readsettings/readbcs/readgeom/alloinit/         lw_create (arrays) or
 *     allo_array/ibmpreread/nutzeroread             lw_create_from_dir (files)
 *     (misc_init.f90:19-440)
 *   prefld (misc_init.f90:610-786)                  lw_set_fields / lw_create_from_dir
 *   dttimeinit/rk3coefinit/lhsinit/poisinit/        inside lw_create
 *     sgsfilterinit (solver.f90:44-60)
 *   one pass of the time loop solver.f90:85-181     lw_step
 *     (stepinit, sgscalc(_t), rhsnlhs, divgs,
 *      poisson, ucalc, pcalc, lagforce, rhsnlhs_t,
 *      convergence_check, masscheck)
 *   writefield (misc_hist.f90:212-252)              lw_get_fields / lw_write_field
 *
 * Conventions: every function returns 0 on success and a negative code on
 * failure (the message is available from lw_last_error); nothing throws across
 * the ABI. Arrays are passed as plain host pointers in the reference's
 * Fortran layouts: padded fields (0:n1, 0:n2, 0:n3) with i fastest, 1-D
 * metrics at their Fortran index (x[1..n1], c2cx[0..n1], ipv[1..n1m]),
 * interior masks (n1m, n2m, n3m) with i fastest. Integers are int64 like the
 * reference's -fdefault-integer-8 build. A handle is used from one host
 * thread at a time; separate handles may run concurrently on their own
 * streams.
 */
#ifndef LESWHT_H
#define LESWHT_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LW_OK 0
#define LW_ERR_ARG -1      /* invalid argument or unsupported configuration */
#define LW_ERR_DEVICE -2   /* no usable gfx950 device or HIP runtime error */
#define LW_ERR_IO -3       /* file contract violated */
#define LW_ERR_NUMERIC -4  /* non-finite values detected */

/* settings.input + boundary.input + ibmpre_prdic.bin
 * (misc_init.f90:19-158; mod_common.f90:2-22) */
typedef struct lw_settings {
  double re, pr, gr, tflu_i, tsol_i, eps_ptr, udrv_i, dt_size, cflfac, resid1,
      wwsor, csgsts, csgshf, tend, ptb_tst, avg_tst;
  int64_t grdir, ireset, iread, iavg, ipzero, nend, nprint, npriavg, npin,
      idtopt, nlev, nbli, ioldv, mgitr, imgsor, iles, insmdl, itemdl, idvmon,
      filter, ibmon, masson, imovingon, ihtrans, idecomp;
  int64_t bc_ybtm, bc_ytop, bc_zbtm, bc_ztop, bc_t_ybtm, bc_t_ytop, bc_t_zbtm,
      bc_t_ztop, ich, iconjg, xprdic, yprdic, zprdic, iintp;
  double val_t_ybtm, val_t_ytop, val_t_zbtm, val_t_ztop;
  /* the Poisson solver; the reference links one (02_solver/Makefile:23-34)
   * and settings.input does not name it: 0 = poiss_z_ft (z-DFT, then a 2-D
   * multigrid per wavenumber), 1 = poiss_mgrd (BiCGSTAB preconditioned by
   * one 3-D multigrid V-cycle; lw_create_from_dir takes it from the
   * environment variable LW_POISSON = "mgrd"), 2 = poiss_ftft (x and z
   * DFTs and a y tridiagonal solve; the one for x-periodic grids, which
   * lw_create_from_dir picks for them), 3 = poiss_x_ft (an x DFT and a
   * periodic-z multigrid per x wavenumber; LW_POISSON = "x_ft" for an
   * x-periodic case directory) */
  int64_t poisson;
} lw_settings;

/* grid.bin as readgeom holds it (misc_init.f90:204-251) */
typedef struct lw_grid {
  int64_t n1, n2, n3, n1m, n2m, n3m;
  double xl, yl, zl;
  const double *x, *y, *z;
  const int64_t *ipv, *jpv, *kpv, *imv, *jmv, *kmv;
  const int64_t *fixil, *fixjl, *fixkl, *fixiu, *fixju, *fixku;
  const double *c2cx, *c2cy, *c2cz, *c2cxi, *c2cyi, *c2czi;
  const double *f2fx, *f2fy, *f2fz, *f2fxi, *f2fyi, *f2fzi;
  const double *xmp, *ymp, *zmp;
} lw_grid;

/* ibmpre tables as ibmpreread/nutzeroread hold them (misc_init.f90:312-409);
 * component c = 0..3 is u, v, w, T. intpindx is 3 per point, geomfac 27 per
 * point ordered [x-offset][y-offset][z-offset]. inz/jnz/knz are the values
 * on disk (0-based, used by the solver as 1-based indices). */
typedef struct lw_tables {
  int64_t nintp[4], nbody[4];
  const int64_t *ifc[4], *jfc[4], *kfc[4], *intpindx[4];
  const double *geomfac[4];
  int64_t nzero;
  const int64_t *inz, *jnz, *knz;
  const int64_t *nwall_dvm;
} lw_tables;

/* per-step results (stdout / fhist.dat quantities of the reference) */
typedef struct lw_step_stats {
  double dt, time, cflmax, dvmax, qmmax, csgsts, csgshf;
  int64_t ntime;
  int64_t mg_cycles_max;   /* largest V-cycle count of any wavenumber, step */
  int64_t mg_cycles_sum;   /* sum over wavenumbers and substeps */
  double ms_sgs, ms_momentum, ms_poisson, ms_energy; /* device time, ms */
} lw_step_stats;

/* the per-step history quantities of the driver loop (solver.f90:150-172),
 * for the step lw_step completed last:
 *   fhist.dat  time dt cflmax dvmax qmmax      writehistory, misc_hist.f90:113
 *   ftemp.dat  time tbulk[0..1] tbulk_avg[0..1] (IHTRANS; on history steps)
 *                                              misc_hist.f90:145-178
 *   fcdcl.dat  time cd[0..2] cd_avg[0..2]       draglift, slv_auxi.f90:669-840
 *   fheatflux.dat time hflux[0..2] hflux_avg[0..2] heatflux, slv_auxi.f90:573-666
 *   ftime.dat  the device seconds of the step's phases (the reference writes
 *              its own wall clock, misc_hist.f90:255-284)
 * written = 1 when mod(ntime, NPIN) = 0: the reference then advances ihist
 * and writes the fhist/ftemp/fcdcl/fheatflux rows. Averages run from
 * AVG_TST on, as in the reference, and are 0 before it. */
typedef struct lw_history {
  int64_t ntime, ihist, written;
  double time, dt, cflmax, dvmax, qmmax;
  double tbulk[2], tbulk_avg[2];
  double cd[3], cd_avg[3];
  double hflux[3], hflux_avg[3];
  double s_sgs, s_rhsnlhs, s_poisson, s_energy;
  /* the channel's fcmfr.dat row (misc_hist.f90:112-136), ICH != 0 */
  double qflux, ubulk, pmiavg, cmfravg[3];
} lw_history;

typedef struct lw_handle lw_handle;

/* body level set used inside the step (convbc, nutzero, wallbc_t), > 0 in
 * fluid; the reference compiles one funcbody into the solver (Makefile:104-105):
 * kind 0 = circular cylinder sqrt(x^2+y^2)-p0 (funcbody_cylinder.f90:13);
 * kind 1 = square max(|x'|,|y'|)-p0 with (x',y') rotated by angle p1;
 * kind 2 = ellipse (sqrt((x'/p0)^2+(y'/p1)^2)-1)*min(p0,p1), angle p2 */
typedef struct lw_body {
  int64_t kind;
  double p[4];
} lw_body;

/* the level set of a body at one point (NULL: the cylinder of radius 0.5) */
double lw_body_eval(const lw_body *b, double x, double y, double z);

int lw_create(const lw_settings *s, const lw_grid *g, const lw_tables *t,
              const lw_body *body, lw_handle **out);
/* lw_create without device allocation. No counterpart in the reference: it
 * is for callers that cannot call hipMalloc (harnesses that forbid it, tight
 * serving loops that must not allocate). Every device array of the case is
 * carved from a static area of lw_static_area_bytes() (256 MiB) that each
 * GPU holds from its first such call until the process ends; host copies
 * live in the handle, not in pinned memory, and the Poisson solve runs on one
 * stream. Cases that need more return LW_ERR_ARG. One static handle per GPU
 * at a time: a second returns LW_ERR_ARG until the first is destroyed. The
 * area keeps the last case's bytes; the handle clears each array it takes.
 * Same steps, same results as lw_create's handle. */
int lw_create_static(const lw_settings *s, const lw_grid *g, const lw_tables *t,
                     const lw_body *body, lw_handle **out);
int64_t lw_static_area_bytes(void);
/* readsettings .. prefld from a reference case directory; solver_dir is the
 * directory holding settings.input (the reference runs with it as cwd) */
int lw_create_from_dir(const char *solver_dir, const lw_body *body, lw_handle **out);
int lw_destroy(lw_handle *h);

int lw_set_fields(lw_handle *h, const double *u, const double *v, const double *w,
                  const double *p, const double *t, int64_t ihist, int64_t m,
                  double time, double dt);
int lw_get_fields(lw_handle *h, double *u, double *v, double *w, double *p, double *t);
/* nsteps passes of solver.f90:85-181; stats of the last step */
int lw_step(lw_handle *h, int64_t nsteps, lw_step_stats *stats);
/* writefield: fld file with the reference's record layout */
int lw_write_field(lw_handle *h, const char *path);
/* the fld writer without a handle: n = (n1, n2, n3), padded host arrays */
int lw_write_fld(const char *path, const int64_t n[3], double re, double pr, double gr,
                 int64_t ihist, int64_t m, double time, double dt, const int64_t prdic[3],
                 const int64_t bcs[6], const int64_t ich[2], const double *u, const double *v,
                 const double *w, const double *p, const double *t);

/* Time-averaged statistics, field_avg (misc_hist.f90:287-527): with
 * IAVG = 1, every step after AVG_TST adds dt-weighted cell-centre sums of
 * u, v, w, u_i u_j, p, p^2, vorticity, its products, S:S (and T, T^2) on the
 * device; every NPRIAVG steps they are written as fav%07d-%07d with the
 * reference's records, listed in fav_manifest.dat, and restarted.
 * lw_set_avg_dir names the directory (lw_create_from_dir sets
 * <solver_dir>/../output/field_avg; NULL or "" writes no files, the sums still
 * restart). lw_get_array reads the current sums ("uavg", "vavg", "wavg",
 * "uiujavg_1".."uiujavg_6", "pavg", "p2avg", "voravg_1".."voravg_3",
 * "vor2avg_1".."vor2avg_6", "ssavg", "tavg", "t2avg"; n1m*n2m*n3m values).
 * lw_avg_finalize is field_avg_finalize (misc_hist.f90:530-697): the open
 * interval's file, then the sum of every file in the manifest under
 * combined/. */
int lw_set_avg_dir(lw_handle *h, const char *dir);
int lw_avg_finalize(lw_handle *h);

/* sizes: n[0..2] = n1, n2, n3 */
int lw_get_dims(const lw_handle *h, int64_t n[3]);
/* history quantities of the last completed step (see lw_history) */
int lw_get_history(const lw_handle *h, lw_history *out);
/* tracer cells, the TRACE block of settings.input (misc_init.f90:47-52;
 * lw_create_from_dir sets it from the file): ntrace <= 64 cells (i, j, k)
 * each, traced every ntr steps by the host (solver.f90:171) */
int lw_set_trace(lw_handle *h, int64_t ntrace, int64_t ntr, const int64_t *pts);
/* ntrace, ntr and (pts not NULL) the 3*ntrace cell indices */
int lw_trace_info(const lw_handle *h, int64_t *ntrace, int64_t *ntr, int64_t *pts);
/* what tracer (misc_hist.f90:184-209) writes for the current fields:
 * out[c*ntrace + n] for c = 0..3: u, v, w at the cell centre, then p;
 * cap >= 4*ntrace */
int lw_get_trace(lw_handle *h, double *out, int64_t cap);
/* rows mg2d appends to mgftresiduemax.dat during the last step
 * (poiss_z_ft.f90:293-297): for each substep, kv = 1..min(3, n3m/2+1) with
 * KK (V-cycles done; MGITR+1 when the limit ended the solve) and RM (the
 * final max-norm residual times dtconst). Returns the row count; copies at
 * most cap rows into the arrays that are not NULL. */
int64_t lw_get_mg_residuals(const lw_handle *h, int64_t *kv, int64_t *kk, double *rm, int64_t cap);
/* every row the z-DFT solver appends to mgftresiduemax.dat during the last
 * step, in the file's order: per substep, mg2d's rows (kind 0: kv <= 3, KK,
 * RM as lw_get_mg_residuals gives them; 'KV=',i4,3x,'KK=',i4,3x,'RM=',es24.16)
 * and then, for every kv above IMGSOR in turn, gsor2d's row of each sweep of
 * its loop (kind 1: KK the sweep, RG its max-norm residual times dtconst
 * times n3m/2+1; 'KV=',i3,3x,'KK=',i3,3x,'RG=',es24.16;
 * poiss_z_ft.f90:293-297, 584). With poisson = 3 (poiss_x_ft) the rows are
 * mg2d's for x wavenumbers 1 and 2 (kind 2: KV the wavenumber, KK its
 * V-cycles, MGITR+1 at the limit, and the max norm times dtconst times
 * (n1m/2+1)/n1m; ' mg, time = ',f13.5,' IV=',i5,' II=',i5,' RG=',es16.8 with
 * the time the step began) and gsor2d's of each wavenumber above IMGSOR
 * that reached its limit (kind 3; 'SOR','  IV=',i5,'  II=',i5,'  RG=',
 * es23.15; poiss_x_ft.f90:293, 606). Returns the row count; copies at most
 * cap rows into the arrays that are not NULL. */
int64_t lw_get_residual_rows(const lw_handle *h, int64_t *kind, int64_t *kv, int64_t *kk, double *val, int64_t cap);
/* with poisson = 1: the rows poiss_mgrd.f90 writes during the last step,
 * per BiCGSTAB iteration to poiss_itr.dat (:185-186: the substep, the
 * iteration, max |residual| * dtconst and the last cell i, j, k holding it,
 * loc[3 q .. 3 q + 2]) and per solve to ftrpoittr.dat (:221-222: the
 * iteration count, MGITR+1 when the limit ended it, and the residual).
 * Return the row counts; copy at most cap rows. */
int64_t lw_get_bicg_log(const lw_handle *h, int64_t *sub, int64_t *icg, double *rm, int64_t *loc, int64_t cap);
int64_t lw_get_bicg_solves(const lw_handle *h, int64_t *icg, double *rm, int64_t cap);
/* the settings the handle runs with (settings.input + boundary.input +
 * ibmpre_prdic.bin as readsettings/readbcs hold them, misc_init.f90:19-158) */
int lw_get_settings(const lw_handle *h, lw_settings *out);
/* named internal array for phase checks ("nusgs", "nusgs1_1".."nusgs1_3",
 * "alsgs", "alsgs1_1".."alsgs1_3", "divgsum", "phi", "qmass");
 * n = capacity in doubles. Returns the element count or a negative code. */
int64_t lw_get_array(lw_handle *h, const char *name, double *out, int64_t n);
/* phase stepping for kernel-level parity checks, in solver.f90 order:
 * 0 stepinit, 10+m substepinit(m), 1 sgs, 2 rhsnlhs, 3 divgs, 4 poisson,
 * 5 ucalc+pcalc+lagforce, 6 energy, 7 end-of-step diagnostics */
int lw_phase(lw_handle *h, int phase);
/* per-handle switches, no counterpart in the reference: "mg_split" 1 solves
 * kv = 1 of the Poisson equation on a second stream beside the other
 * wavenumbers, 0 on the handle's one stream (default: 1 when n1m*n2m >= 65536);
 * results are the same bits either way. "mg_fp32" 1 selects the
 * mixed-precision form of the multigrid: the level planes below the top
 * level hold FP32 values while the top level's relaxation, residual,
 * convergence test and solution stay FP64 (also set at creation by the
 * environment variable LW_MG_FP32=1); results then differ from the FP64
 * form by about the solver tolerance. */
int lw_set_option(lw_handle *h, const char *name, int64_t value);
/* Autotuner of the fused multigrid kernels' tile sizes (lines per workgroup
 * of the down, up and top kernels; tile1 for kv = 1 on its own stream). Times
 * reps Poisson solves of the current field's divergence per candidate and
 * keeps the fastest set in the handle. Tiles never change the arithmetic, and
 * the solve's inputs and outputs are recomputed in every substep, so steps
 * after tuning give the same bits. No counterpart in the reference. */
typedef struct lw_tune_result {
  int64_t tile[3], tile1[3];  /* chosen (down, up, top) */
  int64_t candidates;         /* sets timed besides the starting one */
  double ms_default, ms_best; /* best solve time of the starting and chosen sets */
} lw_tune_result;
int lw_autotune(lw_handle *h, int64_t reps, lw_tune_result *out);
/* device scratch footprint in bytes */
int64_t lw_device_bytes(const lw_handle *h);

/* Per-kernel device time. No counterpart in the reference, whose only timer is
 * the per-step wall clock written to ftime.dat (misc_hist.f90:255-284).
 * lw_profile_begin starts a fresh profile of the following lw_step / lw_phase
 * calls: kernels = NULL or "" times every kernel, otherwise a comma-separated
 * list of the names lw_profile_get reports. Each timed launch is bracketed by
 * two events on the handle's stream. lw_profile_get returns, per kernel, the
 * launch count, the summed event time and the summed algorithmic HBM bytes
 * (each kernel's minimal reads and writes of its arrays). */
int lw_profile_begin(lw_handle *h, const char *kernels);
int lw_profile_end(lw_handle *h);
int64_t lw_profile_count(const lw_handle *h);
int lw_profile_get(const lw_handle *h, int64_t idx, char *name, int64_t cap,
                   int64_t *launches, double *ms_total, double *bytes_total);

const char *lw_last_error(void);

/* Upper bound of the device bytes lw_create allocates for a case of
 * n1m x n2m x n3m cells with these settings (lw_device_bytes gives the
 * actual figure once created). Negative on bad sizes. No counterpart in the
 * reference, which allocates host arrays in alloinit (misc_init.f90:285-309). */
int64_t lw_estimate_device_bytes(const lw_settings *s, int64_t n1m, int64_t n2m, int64_t n3m);
/* settings.input, boundary.input and ibmpre_prdic.bin (readsettings/readbcs,
 * misc_init.f90:19-158) and the cell counts of the grid file of a case
 * directory, without reading tables or fields */
int lw_probe_dir(const char *solver_dir, lw_settings *s, int64_t nm[3]);

/* ---- immersed-boundary tables built in the process ------------------------
 * The pre-processor's table routines (ibm_body.f90:14-399 as called from
 * preprocessing.py:124-276, IBM-INTERP ON) for a body chosen at run time:
 * find_inout, findbdy_intp, geomfac_preset, geomfac_intp and, with les,
 * find_zero_nu_sgs. Inputs are the grid lines of each direction (nx, ny, nz
 * lines, the values grid.dat holds) and the periodic flags; the axes are
 * formed as lib_setgrid.py:1-79 forms them. heat adds the T component.
 * lw_ib_tables fills an lw_tables for lw_create (pointers valid while the
 * lw_ib lives; geomfac as the solver reads it back from the %15.8f text);
 * lw_ib_geomfac_raw gives the unrounded weights of component 0..3;
 * lw_ib_write writes ibmpre_fcpts.bin, ibmpre_fcpts_t.bin, ibmpre_nutzero.bin
 * and ibmpre_wallfdvm.bin with the pre-processor's formats
 * (lib_preprocessing.py:223-338). */
typedef struct lw_ib lw_ib;
int lw_ib_build(const double *xlines, int64_t nx, const double *ylines, int64_t ny,
                const double *zlines, int64_t nz, const int64_t periodic[3], const lw_body *body,
                int64_t heat, int64_t les, lw_ib **out);
int lw_ib_tables(const lw_ib *ib, lw_tables *t);
int64_t lw_ib_geomfac_raw(const lw_ib *ib, int64_t comp, double *out, int64_t n);
int lw_ib_write(const lw_ib *ib, const char *ibmpre_dir);
int lw_ib_free(lw_ib *ib);

/* ---- request loop over independent cases on one GPU ----------------------
 * The reference runs one case per solver_exec process (run_solver.sh:41); it
 * has no counterpart to this. A server owns `slots` worker threads. Each
 * request is a reference case directory (as for lw_create_from_dir) and a
 * step count; requests are admitted first-come first-served when a slot is
 * free and lw_estimate_device_bytes of the case fits what is left of the
 * budget. An admitted case is created on the device, stepped on its own
 * streams beside the other admitted cases, its final field is written with
 * lw_write_field to fld_out (skipped when NULL or ""), and its memory freed.
 * Times are seconds since lw_server_create. */
#define LW_CASE_DONE 0
#define LW_CASE_QUEUED 1
#define LW_CASE_RUNNING 2
typedef struct lw_case_result {
  int64_t ticket;
  int64_t status;          /* LW_CASE_*, or the negative code the case failed with */
  int64_t cells, steps_done, mg_cycles_sum;
  int64_t reserved_bytes;  /* the admission estimate */
  int64_t device_bytes;    /* what lw_create allocated */
  double t_submit, t_admit, t_done;
  double time, dt;         /* simulated time and dt after the last step */
} lw_case_result;
typedef struct lw_server_info {
  int64_t submitted, queued, running, slots;
  int64_t reserved_bytes, peak_reserved_bytes, peak_running, budget_bytes;
  double elapsed;
} lw_server_info;
typedef struct lw_server lw_server;
int lw_server_create(int64_t mem_budget_bytes, int64_t slots, lw_server **out);
/* The same server with continuous batching: an admitted case of at most
 * batch_max_cells cells is created by a worker and then joins the server's
 * batch (lw_batch_*), which steps all its cases together, one step at a
 * time; cases join and leave it between those steps, as they arrive and
 * finish, so a slot is held only while its case is created. Larger cases
 * run on their own streams as in lw_server_create. batch_max_cells = 0 is
 * lw_server_create. Every case gives the same results as alone. */
int lw_server_create_batched(int64_t mem_budget_bytes, int64_t slots, int64_t batch_max_cells,
                             lw_server **out);
int lw_server_submit(lw_server *s, const char *solver_dir, const lw_body *body, int64_t nsteps,
                     const char *fld_out, int64_t *ticket);
/* current state of a request (non-blocking) */
int lw_server_poll(lw_server *s, int64_t ticket, lw_case_result *res);
/* blocks until the request has finished or failed */
int lw_server_wait(lw_server *s, int64_t ticket, lw_case_result *res);
int lw_server_stats(lw_server *s, lw_server_info *info);
/* runs every queued request to the end, then stops the workers */
int lw_server_destroy(lw_server *s);

/* Batched stepping (no counterpart in the reference, which runs one case
 * per solver_exec process; it serves the independent cases of a request
 * stream). The handles of a batch step together: each kernel of the step is
 * launched once for every case that issues it, its workgroups divided among
 * the cases, so many small cases fill the GPU that one of them leaves
 * mostly idle. Each case runs exactly the arithmetic of lw_step and gives
 * the same bits as stepped alone. The handles must live on the current
 * device, use the z-DFT (default) or poiss_ftft Poisson solver, have
 * IAVG = 0 and no profile running; a handle belongs to one batch at a time
 * and cannot be destroyed while it does. A batch is driven from one thread
 * at a time. After a device error every later lw_batch_step of the batch
 * returns LW_ERR_DEVICE with the same message. lw_batch_step runs nsteps steps of
 * every case; stats (NULL or n entries) get each case's last step (phase
 * times are not measured: 0). lw_batch_info: info[0] rendezvous (host
 * synchronisation points), [1] batched launches, [2] per-case launches they
 * carried, [3] per-case read-backs, [4] host microseconds spent merging the
 * recordings and filling the staging buffers, [5] host microseconds waiting
 * for the device. */
typedef struct lw_batch lw_batch;
int lw_batch_create(lw_handle *const *handles, int64_t n, lw_batch **out);
int lw_batch_step(lw_batch *b, int64_t nsteps, lw_step_stats *stats);
int lw_batch_info(const lw_batch *b, int64_t info[6]);
int lw_batch_destroy(lw_batch *b);

#ifdef __cplusplus
}
#endif
#endif
