/*
 * CPU restatement of the LESwHT time step (02_solver/f90/solver.f90:85-183).
 *
 * TEST INFRASTRUCTURE ONLY. This library is the oracle the HIP product is
 * checked against; nothing in the product links, loads or calls it. Only the
 * tests, the smoke check and the CPU-baseline leg of the benchmark use it.
 *
 * Arrays use the reference's Fortran layouts: padded fields (0:n1,0:n2,0:n3)
 * with i fastest; 1-D metrics stored at their Fortran index (x[1..n1],
 * f2fx[0..n1], ipv[1..n1m]); interior arrays (n1m,n2m,n3m) 1-based.
 */
#ifndef LW_ORACLE_H
#define LW_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ora_settings {
  double re, pr, gr, tflu_i, tsol_i, eps_ptr, udrv_i, dt_size, cflfac, resid1,
      wwsor, csgsts, csgshf, tend, ptb_tst, avg_tst;
  int64_t grdir, ireset, iread, iavg, ipzero, nend, nprint, npriavg, npin,
      idtopt, nlev, nbli, ioldv, mgitr, imgsor, iles, insmdl, itemdl, idvmon,
      filter, ibmon, masson, imovingon, ihtrans, idecomp;
  /* boundary.input + ibmpre_prdic.bin (misc_init.f90:95-158) */
  int64_t bc_ybtm, bc_ytop, bc_zbtm, bc_ztop, bc_t_ybtm, bc_t_ytop, bc_t_zbtm,
      bc_t_ztop, ich, iconjg, xprdic, yprdic, zprdic, iintp;
  double val_t_ybtm, val_t_ytop, val_t_zbtm, val_t_ztop;
} ora_settings;

/* grid.bin contents (misc_init.f90:204-251); arrays at Fortran indices. */
typedef struct ora_grid {
  int64_t n1, n2, n3, n1m, n2m, n3m;
  double xl, yl, zl;
  const double *x, *y, *z;                 /* [1..n] */
  const int64_t *ipv, *jpv, *kpv, *imv, *jmv, *kmv;        /* [1..nm] */
  const int64_t *fixil, *fixjl, *fixkl, *fixiu, *fixju, *fixku;
  const double *c2cx, *c2cy, *c2cz, *c2cxi, *c2cyi, *c2czi; /* [0..n] */
  const double *f2fx, *f2fy, *f2fz, *f2fxi, *f2fyi, *f2fzi;
  const double *xmp, *ymp, *zmp;
} ora_grid;

/* ibmpre tables (misc_init.f90:254-409): per component c = 0..3 (u,v,w,t),
 * ifc/jfc/kfc have nbody[c] entries, intpindx 3*nintp[c] (point-major),
 * geomfac 27*nintp[c] ordered [x-off][y-off][z-off]. */
typedef struct ora_tables {
  int64_t nintp[4], nbody[4];
  const int64_t *ifc[4], *jfc[4], *kfc[4], *intpindx[4];
  const double *geomfac[4];
  int64_t nzero;
  const int64_t *inz, *jnz, *knz;  /* as read: 0-based values (D7 quirk) */
  const int64_t *nwall_dvm;        /* (n1m,n2m,n3m) i fastest */
} ora_tables;

typedef struct ora_diag {
  double dt, time, cflmax, dvmax, qmmax, csgsts, csgshf, cd[3], heatflux[3];
  int64_t ntime, mg_cycles_max, mg_cycles_sum;
} ora_diag;

typedef struct ora ora;

ora *ora_create(const ora_settings *s, const ora_grid *g, const ora_tables *t);
void ora_destroy(ora *o);
/* body level set used inside the step: kind 0 cylinder (radius p[0], the
 * default 0.5), 1 square (half side p[0], angle p[1]), 2 ellipse (semi-axes
 * p[0], p[1], angle p[2]) */
void ora_set_body(ora *o, int64_t kind, const double p[4]);
double ora_funcbody(const ora *o, double x, double y, double z);
/* fields: padded (0:n1,0:n2,0:n3) arrays after prefld; t may be NULL. Also
 * sets ihist, m, time, dt as read from the restart file. */
void ora_set_fields(ora *o, const double *u, const double *v, const double *w,
                    const double *p, const double *t, int64_t ihist, int64_t m,
                    double time, double dt);
void ora_get_fields(const ora *o, double *u, double *v, double *w, double *p,
                    double *t);
/* dttimeinit/rk3coefinit/lhsinit/poisinit (solver.f90:57-61) */
void ora_init(ora *o);
/* one full time step of solver.f90:85-181 */
void ora_step(ora *o, ora_diag *d);
/* named internal arrays for phase checks: "nusgs", "nusgs1", "alsgs",
 * "alsgs1", "qmass", "divgsum", "phi", "rhs1". Returns element count. */
int64_t ora_get_array(const ora *o, const char *name, double *out);

/* phase entry points for kernel-level parity tests */
void ora_substep_begin(ora *o, int msub);
void ora_phase_sgs(ora *o);
void ora_phase_rhsnlhs(ora *o);
void ora_phase_divgs(ora *o);
void ora_phase_poisson(ora *o);
void ora_phase_ucalc_pcalc(ora *o);
void ora_phase_energy(ora *o);
void ora_stepinit(ora *o);

/* building blocks exposed for unit tests */
/* Poisson path: 0 = poiss_z_ft (default), 1 = poiss_mgrd (BiCGSTAB with a
 * 3-D multigrid preconditioner; IOLDV = 0, z periodic or not, x not
 * periodic), 2 = poiss_ftft and 3 = poiss_x_ft (x periodic; x_ft also z
 * periodic with n3m divisible by 2**NLEV). Returns 0, or -1 when the case
 * is outside that scope. */
int ora_set_poisson(ora *o, int64_t kind);
/* poiss_x_ft's mgftresiduemax.dat rows of the last step: kind 0 for mg2d
 * (wavenumbers 1 and 2, poiss_x_ft.f90:293), 1 for gsor2d at its iteration
 * limit (:606); the wavenumber, the cycle counter and the residual times
 * dtconst * n1mh / n1m. Returns the number of rows. */
int64_t ora_xft_rows(const ora *o, int64_t *kind, int64_t *iv, int64_t *ii, double *rg, int64_t cap);
/* the BiCGSTAB log of the last step: per iteration (substep, iteration,
 * max |residual| * dtconst, its cell i, j, k as poiss_itr.dat holds them);
 * per solve (the final iteration number, MGITR+1 if not converged, and
 * residual, as ftrpoittr.dat holds them) */
int64_t ora_bicg_log(const ora *o, int64_t *sub, int64_t *icg, double *rm, int64_t *loc, int64_t cap);
int64_t ora_bicg_solves(const ora *o, int64_t *icg, double *rm);

void ora_fft_forward(int n, double *re, double *im);   /* e^{-2 pi i jk/n} */
void ora_fft_inverse(int n, double *re, double *im);   /* conj trick, 1/n */
void ora_thomas(int64_t n, const double *a, const double *b, const double *c,
                double *r);  /* trdiag1/2/3 on one line, in place */
void ora_thomas_periodic(int64_t n, const double *a, const double *b,
                         const double *c, double *f); /* trdiag1p/3p */

#ifdef __cplusplus
}
#endif
#endif
