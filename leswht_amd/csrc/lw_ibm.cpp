// Immersed-boundary tables built in the process, for bodies chosen at run
// time: the pre-processor's table routines (ibm_body.f90:14-399, called from
// preprocessing.py:124-276) and its text writers (lib_preprocessing.py:223-338),
// so a served case needs no Python and no recompiled funcbody.
//
// Inputs are the grid lines of the three directions (the values grid.dat
// holds) and the periodic flags; the axes (coordinates, cell centres, one-
// sided flags) are formed as the pre-processor's lib_setgrid.py:1-79 forms
// them. Every value is computed with the same operations in the same order
// as the pre-processor, so the tables, and the text written from them, are
// the same bits.
#include <math.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

#include "../../include/leswht.h"

void lw_set_last_error(const std::string &msg);

namespace {

// one direction (lib_setgrid.py: readcoord, setindex, setmesh, physcenter):
// coord and center are indexed 0..N (N grid lines), fixes per cell 1..ncell
// stored at 0..ncell-1
struct Axis {
  int64_t ncell = 0;
  std::vector<double> coord, center;
  std::vector<int64_t> lowerfix, upperfix;
};

Axis make_axis(const double *lines, int64_t n, bool periodic) {
  Axis a;
  const int64_t nc = n - 1;
  a.ncell = nc;
  a.coord.assign((size_t)n + 1, 0.0);
  for (int64_t q = 0; q < n; ++q) a.coord[q + 1] = lines[q];
  a.coord[0] = lines[0];
  a.lowerfix.assign((size_t)nc, 0);
  a.upperfix.assign((size_t)nc, 0);
  a.lowerfix[0] = 1;
  a.upperfix[nc - 1] = 1;
  std::vector<double> f2fd((size_t)n + 1, 0.0);
  if (nc != 1) {
    for (int64_t i = 1; i <= nc; ++i) f2fd[i] = a.coord[i + 1] - a.coord[i];
  } else {
    a.coord[2] = a.coord[1] + 1.0;
    f2fd[1] = 1.0;
    f2fd[2] = 1.0;
  }
  f2fd[0] = 0.0;
  f2fd[nc + 1] = 0.0;
  if (periodic) {
    a.lowerfix[0] = 0;
    a.upperfix[nc - 1] = 0;
    a.coord[0] = a.coord[1] - f2fd[nc];
  }
  a.center.assign((size_t)n + 1, 0.0);
  for (int64_t i = 1; i <= nc; ++i) a.center[i] = a.coord[i] + 0.5 * f2fd[i];
  a.center[nc + 1] = a.coord[nc + 1];
  a.center[0] = a.coord[0];
  return a;
}

struct Comp {
  int64_t nintp = 0, ninner = 0;
  std::vector<int64_t> ifc, jfc, kfc, indx;  // indx: 3 per interpolation point
  std::vector<double> raw, gf;                // 27 per point, [x][y][z]; gf as read back from %15.8f
};

}  // namespace

struct lw_ib {
  Axis ax, ay, az;
  lw_body body{};
  int ncomp = 3;
  int les = 0;
  Comp c[4];
  int64_t nzero = 0;
  std::vector<int64_t> iszero;      // (ncell_x, ncell_y, ncell_z), i fastest; 0 = zero nu_t
  std::vector<int64_t> inz, jnz, knz;  // 0-based, i outer .. k inner

  double f(double x, double y, double z) const { return lw_body_eval(&body, x, y, z); }
};

namespace {

void add(std::string &s, const char *fmt, ...) __attribute__((format(printf, 2, 3)));

// find_inout (ibm_body.f90:14-46) and findbdy_intp (49-123) for one
// component on its staggered grid (xc, yc, zc indexed 0..N): body points in
// k -> j -> i order, interpolation points first
void find_points(const lw_ib &b, int dir, const std::vector<double> &xc, const std::vector<double> &yc,
                 const std::vector<double> &zc, Comp &cp) {
  const int64_t nx = (int64_t)xc.size() - 1, ny = (int64_t)yc.size() - 1, nz = (int64_t)zc.size() - 1;
  const int64_t sx = nx + 1, sy = ny + 1;
  std::vector<uint8_t> inout((size_t)(sx * sy * (nz + 1)));
  auto IO = [&](int64_t i, int64_t j, int64_t k) -> int64_t { return inout[(size_t)(i + sx * (j + sy * k))]; };
  for (int64_t k = 0; k <= nz; ++k)
    for (int64_t j = 0; j <= ny; ++j)
      for (int64_t i = 0; i <= nx; ++i)
        inout[(size_t)(i + sx * (j + sy * k))] = b.f(xc[i], yc[j], zc[k]) <= 1.0e-10 ? 0 : 1;
  const int64_t i0 = dir == 1 ? 2 : 1, j0 = dir == 2 ? 2 : 1, k0 = dir == 3 ? 2 : 1;
  std::vector<int64_t> inner;
  for (int64_t k = k0; k < nz; ++k)
    for (int64_t j = j0; j < ny; ++j)
      for (int64_t i = i0; i < nx; ++i) {
        if (IO(i, j, k) != 0) continue;
        const int64_t dx = IO(i + 1, j, k) - IO(i - 1, j, k);
        const int64_t dy = IO(i, j + 1, k) - IO(i, j - 1, k);
        const int64_t dz = IO(i, j, k + 1) - IO(i, j, k - 1);
        const int64_t fx = (1 - b.ax.upperfix[i - 1]) * (1 - b.ax.lowerfix[i - 1]);
        const int64_t fy = (1 - b.ay.upperfix[j - 1]) * (1 - b.ay.lowerfix[j - 1]);
        const int64_t fz = (1 - b.az.upperfix[k - 1]) * (1 - b.az.lowerfix[k - 1]);
        const int64_t ssum = llabs(dx) * fx + llabs(dy) * fy + llabs(dz) * fz;
        if (ssum > 0) {
          cp.ifc.push_back(i); cp.jfc.push_back(j); cp.kfc.push_back(k);
          cp.indx.push_back(dx); cp.indx.push_back(dy); cp.indx.push_back(dz);
        } else {
          inner.push_back(i); inner.push_back(j); inner.push_back(k);
        }
      }
  cp.nintp = (int64_t)cp.ifc.size();
  cp.ninner = (int64_t)inner.size() / 3;
  for (size_t q = 0; q < inner.size(); q += 3) {
    cp.ifc.push_back(inner[q]); cp.jfc.push_back(inner[q + 1]); cp.kfc.push_back(inner[q + 2]);
  }
}

// geomfac_preset (ibm_body.f90:126-157): i_adj(-1:ni+1), stored from index
// -1 at 0; the periodic ghost rows are never set (the option string is
// compared with 'on' while the pre-processor passes 'ON')
std::vector<double> preset(const std::vector<double> &v) {
  std::vector<double> out(v.size() + 2, 0.0);
  for (size_t q = 0; q < v.size(); ++q) out[q + 1] = v[q];
  return out;
}

// left-to-right product -(inv * f1 * f2 ...) (ibm_body.f90:289-317)
double negprod(double inv, std::initializer_list<double> fs) {
  double acc = inv;
  for (double f : fs) acc = acc * f;
  return -acc;
}

// geomfac_intp (ibm_body.f90:160-322): the surface point between the body
// point and its first fluid neighbour by three levels of 20-step scans, then
// 1-D linear weights per axis and their products
void geomfac(const lw_ib &b, const std::vector<double> &xp, const std::vector<double> &yp,
             const std::vector<double> &zp, Comp &cp) {
  cp.raw.assign((size_t)cp.nintp * 27, 0.0);
  for (int64_t n = 0; n < cp.nintp; ++n) {
    const int64_t fi = cp.ifc[n], fj = cp.jfc[n], fk = cp.kfc[n];
    const int64_t di = cp.indx[3 * n], dj = cp.indx[3 * n + 1], dk = cp.indx[3 * n + 2];
    const double x1 = xp[fi + 1], y1 = yp[fj + 1], z1 = zp[fk + 1];
    const double x2 = xp[fi + di + 1], y2 = yp[fj + dj + 1], z2 = zp[fk + dk + 1];
    double x0 = x1, y0 = y1, z0 = z1;
    double *g = cp.raw.data() + 27 * n;
    const double ffs = b.f(x1, y1, z1), ffe = b.f(x2, y2, z2);
    if (ffs * ffe > 0.0) {  // no sign change: identity weights
      g[0] = 1.0;
      continue;
    }
    double xt1 = x1, yt1 = y1, zt1 = z1, xt2 = x2, yt2 = y2, zt2 = z2;
    double xx1 = 0, yy1 = 0, zz1 = 0, xx2 = 0, yy2 = 0, zz2 = 0;
    bool done = false;
    for (int m = 1; m <= 3 && !done; ++m) {
      if (m > 1) {
        xt1 = xx1; yt1 = yy1; zt1 = zz1;
        xt2 = xx2; yt2 = yy2; zt2 = zz2;
      }
      for (int s = 0; s < 20; ++s) {
        const double ddx = xt2 - xt1, ddy = yt2 - yt1, ddz = zt2 - zt1;
        const double a1 = xt1 + ddx * (double)s / 20.0, a2 = xt1 + ddx * (double)(s + 1) / 20.0;
        const double b1 = yt1 + ddy * (double)s / 20.0, b2 = yt1 + ddy * (double)(s + 1) / 20.0;
        const double c1 = zt1 + ddz * (double)s / 20.0, c2 = zt1 + ddz * (double)(s + 1) / 20.0;
        xx1 = a1; xx2 = a2; yy1 = b1; yy2 = b2; zz1 = c1; zz2 = c2;
        const double ff1 = b.f(a1, b1, c1), ff2 = b.f(a2, b2, c2);
        if (ff1 == 0.0) {
          x0 = a1; y0 = b1; z0 = c1;
          done = true;
          break;
        }
        if (ff2 == 0.0) {
          x0 = a2; y0 = b2; z0 = c2;
          done = true;
          break;
        }
        if (ff1 * ff2 < 0.0) {
          x0 = 0.5 * (a1 + a2); y0 = 0.5 * (b1 + b2); z0 = 0.5 * (c1 + c2);
          if (m == 3) done = true;
          break;
        }
      }
    }
    const double x3 = xp[fi + 2 * di + 1], y3 = yp[fj + 2 * dj + 1], z3 = zp[fk + 2 * dk + 1];
    auto weights = [](int64_t ind, double d1, double d2, double d3, double &w0, double &w1) {
      if (ind == 0) {
        w0 = 1.0; w1 = 1.0;
      } else if (d2 >= d1) {
        w0 = d2 / (d1 + d2); w1 = 1.0;
      } else {
        w0 = 0.5; w1 = (d3 - d1) / (d3 - d2);
      }
    };
    double a0, a1, b0, b1, c0, c1;
    weights(di, fabs(x1 - x0), fabs(x2 - x0), fabs(x3 - x0), a0, a1);
    weights(dj, fabs(y1 - y0), fabs(y2 - y0), fabs(y3 - y0), b0, b1);
    weights(dk, fabs(z1 - z0), fabs(z2 - z0), fabs(z3 - z0), c0, c1);
    const double inv = 1.0 / (a0 * b0 * c0);
    const double nA0 = 1.0 - a0, nB0 = 1.0 - b0, nC0 = 1.0 - c0;
    const double nA1 = 1.0 - a1, nB1 = 1.0 - b1, nC1 = 1.0 - c1;
#define GF(p, q, r) g[9 * (p) + 3 * (q) + (r)]
    GF(0, 0, 0) = inv;
    GF(0, 0, 1) = negprod(inv, {a0, b0, nC0, c1});
    GF(0, 0, 2) = negprod(inv, {a0, b0, nC0, nC1});
    GF(0, 1, 0) = negprod(inv, {a0, nB0, c0, b1});
    GF(0, 1, 1) = negprod(inv, {a0, nB0, nC0, b1, c1});
    GF(0, 1, 2) = negprod(inv, {a0, nB0, nC0, b1, nC1});
    GF(0, 2, 0) = negprod(inv, {a0, nB0, c0, nB1});
    GF(0, 2, 1) = negprod(inv, {a0, nB0, nC0, nB1, c1});
    GF(0, 2, 2) = negprod(inv, {a0, nB0, nC0, nB1, nC1});
    GF(1, 0, 0) = negprod(inv, {nA0, b0, c0, a1});
    GF(1, 0, 1) = negprod(inv, {nA0, b0, nC0, a1, c1});
    GF(1, 0, 2) = negprod(inv, {nA0, b0, nC0, a1, nC1});
    GF(1, 1, 0) = negprod(inv, {nA0, nB0, c0, a1, b1});
    GF(1, 1, 1) = negprod(inv, {nA0, nB0, nC0, a1, b1, c1});
    GF(1, 1, 2) = negprod(inv, {nA0, nB0, nC0, a1, b1, nC1});
    GF(1, 2, 0) = negprod(inv, {nA0, nB0, c0, a1, nB1});
    GF(1, 2, 1) = negprod(inv, {nA0, nB0, nC0, a1, nB1, c1});
    GF(1, 2, 2) = negprod(inv, {nA0, nB0, nC0, a1, nB1, nC1});
    GF(2, 0, 0) = negprod(inv, {nA0, b0, c0, nA1});
    GF(2, 0, 1) = negprod(inv, {nA0, b0, nC0, nA1, c1});
    GF(2, 0, 2) = negprod(inv, {nA0, b0, nC0, nA1, nC1});
    GF(2, 1, 0) = negprod(inv, {nA0, nB0, c0, nA1, b1});
    GF(2, 1, 1) = negprod(inv, {nA0, nB0, nC0, nA1, b1, c1});
    GF(2, 1, 2) = negprod(inv, {nA0, nB0, nC0, nA1, b1, nC1});
    GF(2, 2, 0) = negprod(inv, {nA0, nB0, c0, nA1, nB1});
    GF(2, 2, 1) = negprod(inv, {nA0, nB0, nC0, nA1, nB1, c1});
    GF(2, 2, 2) = negprod(inv, {nA0, nB0, nC0, nA1, nB1, nC1});
#undef GF
  }
  // the solver reads the weights back from the %15.8f text
  // (lib_preprocessing.py:261; misc_init.f90:341-349)
  cp.gf.resize(cp.raw.size());
  std::string txt;
  for (size_t q = 0; q < cp.raw.size(); ++q) {
    txt.clear();
    add(txt, " %15.8f ", cp.raw[q]);
    cp.gf[q] = strtod(txt.c_str(), nullptr);
  }
}

// find_zero_nu_sgs (ibm_body.f90:357-399): cells whose centre or one of its
// six neighbours' centres is solid
void zero_nu(lw_ib &b) {
  const std::vector<double> &xm = b.ax.center, &ym = b.ay.center, &zm = b.az.center;
  const int64_t nx = (int64_t)xm.size() - 1, ny = (int64_t)ym.size() - 1, nz = (int64_t)zm.size() - 1;
  const int64_t sx = nx + 1, sy = ny + 1;
  std::vector<uint8_t> solid((size_t)(sx * sy * (nz + 1)));
  for (int64_t k = 0; k <= nz; ++k)
    for (int64_t j = 0; j <= ny; ++j)
      for (int64_t i = 0; i <= nx; ++i) solid[(size_t)(i + sx * (j + sy * k))] = b.f(xm[i], ym[j], zm[k]) <= 1.0e-10;
  auto S = [&](int64_t i, int64_t j, int64_t k) { return solid[(size_t)(i + sx * (j + sy * k))] != 0; };
  const int64_t cx = nx - 1, cy = ny - 1, cz = nz - 1;
  b.iszero.assign((size_t)(cx * cy * cz), 1);
  b.nzero = 0;
  for (int64_t k = 1; k < nz; ++k)
    for (int64_t j = 1; j < ny; ++j)
      for (int64_t i = 1; i < nx; ++i) {
        const bool s = S(i, j, k) || S(i - 1, j, k) || S(i + 1, j, k) || S(i, j - 1, k) || S(i, j + 1, k) ||
                       S(i, j, k - 1) || S(i, j, k + 1);
        if (s) {
          b.iszero[(size_t)((i - 1) + cx * ((j - 1) + cy * (k - 1)))] = 0;
          b.nzero += 1;
        }
      }
  // the nutzero file lists them with i outermost and k innermost, 0-based
  // (lib_preprocessing.py:306-330)
  for (int64_t i = 0; i < cx; ++i)
    for (int64_t j = 0; j < cy; ++j)
      for (int64_t k = 0; k < cz; ++k)
        if (b.iszero[(size_t)(i + cx * (j + cy * k))] == 0) {
          b.inz.push_back(i); b.jnz.push_back(j); b.knz.push_back(k);
        }
}

bool write_text(const std::string &path, const std::string &text) {
  FILE *fp = fopen(path.c_str(), "w");
  if (!fp) return false;
  const bool ok = fwrite(text.data(), 1, text.size(), fp) == text.size();
  return fclose(fp) == 0 && ok;
}

void add(std::string &s, const char *fmt, ...) {
  char buf[256];
  va_list ap, ap2;
  va_start(ap, fmt);
  va_copy(ap2, ap);
  const int n = vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (n >= 0 && (size_t)n < sizeof buf) {
    s.append(buf, (size_t)n);
  } else if (n > 0) {  // a value wider than the buffer (e.g. a huge weight in %15.8f)
    std::vector<char> big((size_t)n + 1);
    vsnprintf(big.data(), big.size(), fmt, ap2);
    s.append(big.data(), (size_t)n);
  }
  va_end(ap2);
}

// ibmpre_fcpts.bin (lib_preprocessing.py:223-277) for comps u, v, w, or
// ibmpre_fcpts_t.bin (279-303) for T alone
std::string fcpts_text(const lw_ib &b, int c0, int nc) {
  std::string s;
  if (nc == 1) {
    add(s, "%lld\n", (long long)b.c[c0].nintp);
    add(s, "%lld\n", (long long)b.c[c0].ninner);
  } else {
    add(s, "%lld %lld %lld\n", (long long)b.c[0].nintp, (long long)b.c[1].nintp, (long long)b.c[2].nintp);
    add(s, "%lld %lld %lld\n", (long long)b.c[0].ninner, (long long)b.c[1].ninner, (long long)b.c[2].ninner);
  }
  for (int c = c0; c < c0 + nc; ++c)
    for (size_t q = 0; q < b.c[c].ifc.size(); ++q)
      add(s, "%lld %lld %lld\n", (long long)b.c[c].ifc[q], (long long)b.c[c].jfc[q], (long long)b.c[c].kfc[q]);
  for (int c = c0; c < c0 + nc; ++c)
    for (int64_t q = 0; q < b.c[c].nintp; ++q)
      add(s, "%lld %lld %lld\n", (long long)b.c[c].indx[3 * q], (long long)b.c[c].indx[3 * q + 1],
          (long long)b.c[c].indx[3 * q + 2]);
  for (int c = c0; c < c0 + nc; ++c)
    for (int64_t q = 0; q < b.c[c].nintp; ++q) {
      for (int r = 0; r < 27; ++r) add(s, " %15.8f ", b.c[c].raw[27 * q + r]);
      s += "\n";
    }
  return s;
}

}  // namespace

extern "C" {

int lw_ib_build(const double *xlines, int64_t nx, const double *ylines, int64_t ny, const double *zlines,
                int64_t nz, const int64_t periodic[3], const lw_body *body, int64_t heat, int64_t les,
                lw_ib **out) {
  if (!xlines || !ylines || !zlines || !periodic || !out || nx < 3 || ny < 3 || nz < 3) {
    lw_set_last_error("lw_ib_build: null argument or fewer than 2 cells in a direction");
    return LW_ERR_ARG;
  }
  auto *b = new lw_ib();
  if (body) b->body = *body;
  else b->body.p[0] = 0.5;
  b->ax = make_axis(xlines, nx, periodic[0] != 0);
  b->ay = make_axis(ylines, ny, periodic[1] != 0);
  b->az = make_axis(zlines, nz, periodic[2] != 0);
  b->ncomp = heat ? 4 : 3;
  b->les = les ? 1 : 0;
  const std::vector<double> &X = b->ax.coord, &XM = b->ax.center, &Y = b->ay.coord, &YM = b->ay.center,
                            &Z = b->az.coord, &ZM = b->az.center;
  // component grids (preprocessing.py:124-200) and the i_adj columns
  const std::vector<double> *grids[4][3] = {{&X, &YM, &ZM}, {&XM, &Y, &ZM}, {&XM, &YM, &Z}, {&XM, &YM, &ZM}};
  const std::vector<double> px = preset(X), pxm = preset(XM), py = preset(Y), pym = preset(YM), pz = preset(Z),
                            pzm = preset(ZM);
  const std::vector<double> *pre[4][3] = {{&px, &pym, &pzm}, {&pxm, &py, &pzm}, {&pxm, &pym, &pz}, {&pxm, &pym, &pzm}};
  for (int c = 0; c < b->ncomp; ++c) {
    find_points(*b, c + 1, *grids[c][0], *grids[c][1], *grids[c][2], b->c[c]);
    geomfac(*b, *pre[c][0], *pre[c][1], *pre[c][2], b->c[c]);
  }
  if (b->les) zero_nu(*b);
  *out = b;
  return LW_OK;
}

int lw_ib_tables(const lw_ib *b, lw_tables *t) {
  if (!b || !t) return LW_ERR_ARG;
  memset(t, 0, sizeof *t);
  for (int c = 0; c < b->ncomp; ++c) {
    t->nintp[c] = b->c[c].nintp;
    t->nbody[c] = b->c[c].nintp + b->c[c].ninner;
    t->ifc[c] = b->c[c].ifc.data();
    t->jfc[c] = b->c[c].jfc.data();
    t->kfc[c] = b->c[c].kfc.data();
    t->intpindx[c] = b->c[c].indx.data();
    t->geomfac[c] = b->c[c].gf.data();
  }
  t->nzero = b->nzero;
  t->inz = b->inz.data();
  t->jnz = b->jnz.data();
  t->knz = b->knz.data();
  t->nwall_dvm = b->les ? b->iszero.data() : nullptr;
  return LW_OK;
}

int64_t lw_ib_geomfac_raw(const lw_ib *b, int64_t comp, double *out, int64_t n) {
  if (!b || comp < 0 || comp >= b->ncomp) return LW_ERR_ARG;
  const int64_t need = (int64_t)b->c[comp].raw.size();
  if (!out) return need;
  if (n < need) return LW_ERR_ARG;
  memcpy(out, b->c[comp].raw.data(), (size_t)need * 8);
  return need;
}

int lw_ib_write(const lw_ib *b, const char *ibmpre_dir) {
  if (!b || !ibmpre_dir) return LW_ERR_ARG;
  const std::string d(ibmpre_dir);
  bool ok = write_text(d + "/ibmpre_fcpts.bin", fcpts_text(*b, 0, 3));
  if (b->ncomp == 4) ok = ok && write_text(d + "/ibmpre_fcpts_t.bin", fcpts_text(*b, 3, 1));
  std::string nz;
  if (b->les) {
    add(nz, "%lld \n", (long long)b->nzero);
    for (const std::vector<int64_t> *v : {&b->inz, &b->jnz, &b->knz}) {
      for (int64_t q : *v) add(nz, " %lld ", (long long)q);
      nz += "\n";
    }
    std::string wf;
    for (int64_t q : b->iszero) add(wf, " %lld ", (long long)q);
    ok = ok && write_text(d + "/ibmpre_wallfdvm.bin", wf);
  } else {
    nz = "0\n\n\n\n";
  }
  ok = ok && write_text(d + "/ibmpre_nutzero.bin", nz);
  if (!ok) {
    lw_set_last_error("lw_ib_write: cannot write the tables under " + d);
    return LW_ERR_IO;
  }
  return LW_OK;
}

int lw_ib_free(lw_ib *b) {
  delete b;
  return LW_OK;
}

}  // extern "C"
