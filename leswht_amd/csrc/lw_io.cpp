// Disk contract of the reference solver, so a case directory prepared for
// solver_exec runs unchanged: list-directed text inputs (misc_init.f90:19-440),
// the fld restart file and its ghost-cell rules (prefld, misc_init.f90:610-786)
// and the fld writer (writefield, misc_hist.f90:212-252).
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/leswht.h"

// the message lw_last_error returns (lw_solver.hip)
void lw_set_last_error(const std::string &msg);

namespace {

thread_local std::string io_err;

struct IoError {
  std::string msg;
};

// Fortran list-directed READ: each READ starts at a new record (line), takes
// blank- or comma-separated values across lines until satisfied, and drops
// the rest of its last line; a quoted string is one value.
class ListReader {
 public:
  explicit ListReader(const std::string &path) : path_(path) {
    std::ifstream in(path);
    if (!in) throw IoError{"cannot open " + path};
    std::string ln;
    while (std::getline(in, ln)) lines_.push_back(ln);
  }
  std::vector<std::string> read(size_t n) {
    std::vector<std::string> v;
    while (v.size() < n) {
      if (pos_ >= lines_.size()) throw IoError{"unexpected end of " + path_};
      tokens(lines_[pos_++], v);
    }
    v.resize(n);
    return v;
  }
  std::vector<double> reals(size_t n) {
    auto t = read(n);
    std::vector<double> r(n);
    for (size_t q = 0; q < n; ++q) r[q] = num(t[q]);
    return r;
  }
  std::vector<int64_t> ints(size_t n) {
    auto t = read(n);
    std::vector<int64_t> r(n);
    for (size_t q = 0; q < n; ++q) r[q] = (int64_t)llround(num(t[q]));
    return r;
  }
  void skip() { read(1); }

 private:
  double num(const std::string &s) {
    std::string t = s;
    for (auto &c : t)
      if (c == 'd' || c == 'D') c = 'e';
    char *end = nullptr;
    const double v = strtod(t.c_str(), &end);
    if (end == t.c_str()) throw IoError{"bad number '" + s + "' in " + path_};
    return v;
  }
  static void tokens(const std::string &ln, std::vector<std::string> &out) {
    std::string cur;
    char q = 0;
    bool have = false;
    for (char ch : ln) {
      if (q) {
        if (ch == q) {
          out.push_back(cur);
          cur.clear();
          q = 0;
          have = false;
        } else cur += ch;
      } else if (ch == '"' || ch == '\'') {
        q = ch;
      } else if (ch == ' ' || ch == '\t' || ch == ',' || ch == '\r') {
        if (have) {
          out.push_back(cur);
          cur.clear();
          have = false;
        }
      } else {
        cur += ch;
        have = true;
      }
    }
    if (have) out.push_back(cur);
  }
  std::string path_;
  std::vector<std::string> lines_;
  size_t pos_ = 0;
};

std::string join(const std::string &a, const std::string &b) {
  if (!b.empty() && b[0] == '/') return b;
  return a + "/" + b;
}

std::string trim25(std::string s) {
  // gridfile and prev_fld are character*25 (mod_common.f90:10-11)
  if (s.size() > 25) s.resize(25);
  while (!s.empty() && s.back() == ' ') s.pop_back();
  return s;
}

struct Case {
  lw_settings s{};
  std::string gridfile, prev_fld;
  // TRACE block (misc_init.f90:47-52)
  int64_t ntrace = 0, ntr = 1;
  std::vector<int64_t> trpts;
  // grid
  int64_t n1 = 0, n2 = 0, n3 = 0, n1m = 0, n2m = 0, n3m = 0;
  double xl = 0, yl = 0, zl = 0;
  std::vector<double> x, y, z, c2cx, c2cy, c2cz, c2cxi, c2cyi, c2czi, f2fx, f2fy, f2fz, f2fxi, f2fyi, f2fzi, xmp,
      ymp, zmp;
  std::vector<int64_t> ipv, jpv, kpv, imv, jmv, kmv, fixil, fixjl, fixkl, fixiu, fixju, fixku;
  // tables
  int64_t nintp[4] = {0, 0, 0, 0}, nbody[4] = {0, 0, 0, 0};
  std::vector<int64_t> ifc[4], jfc[4], kfc[4], indx[4];
  std::vector<double> gf[4];
  int64_t nzero = 0;
  std::vector<int64_t> inz, jnz, knz, nwall;
  // fields
  std::vector<double> u, v, w, p, t;
  int64_t ihist = 0, m = 0;
  double time = 0, dt = 0;
};

void read_settings(Case &c, const std::string &dir) {
  ListReader r(join(dir, "settings.input"));
  lw_settings &s = c.s;
  r.skip();
  auto a = r.reals(6);
  s.re = a[0]; s.pr = a[1]; s.gr = a[2]; s.grdir = (int64_t)a[3]; s.tflu_i = a[4]; s.tsol_i = a[5];
  r.skip();
  a = r.reals(6);
  s.ireset = (int64_t)a[0]; s.iread = (int64_t)a[1]; s.iavg = (int64_t)a[2]; s.ipzero = (int64_t)a[3];
  s.eps_ptr = a[4]; s.udrv_i = a[5];
  r.skip();
  auto b = r.ints(4);
  s.nend = b[0]; s.nprint = b[1]; s.npriavg = b[2]; s.npin = b[3];
  r.skip();
  a = r.reals(3);
  s.tend = a[0]; s.ptb_tst = a[1]; s.avg_tst = a[2];
  r.skip();
  a = r.reals(3);
  s.idtopt = (int64_t)a[0]; s.dt_size = a[1]; s.cflfac = a[2];
  r.skip();
  a = r.reals(7);
  s.resid1 = a[0]; s.nlev = (int64_t)a[1]; s.nbli = (int64_t)a[2]; s.ioldv = (int64_t)a[3];
  s.mgitr = (int64_t)a[4]; s.imgsor = (int64_t)a[5]; s.wwsor = a[6];
  r.skip();
  a = r.reals(7);
  s.iles = (int64_t)a[0]; s.insmdl = (int64_t)a[1]; s.itemdl = (int64_t)a[2]; s.idvmon = (int64_t)a[3];
  s.csgsts = a[4]; s.csgshf = a[5]; s.filter = (int64_t)a[6];
  r.skip();
  b = r.ints(5);
  s.ibmon = b[0]; s.masson = b[1]; s.imovingon = b[2]; s.ihtrans = b[3]; s.idecomp = b[4];
  r.skip();
  c.gridfile = trim25(r.read(1)[0]);
  c.prev_fld = trim25(r.read(1)[0]);
  r.skip();
  b = r.ints(2);
  c.ntrace = b[0];
  c.ntr = b[1];
  if (c.ntrace < 0 || c.ntrace > 64) throw IoError{"settings.input: NTRACE must be in 0..64 (trpts(64, 3))"};
  for (int64_t n = 0; n < c.ntrace; ++n) {
    b = r.ints(3);
    c.trpts.insert(c.trpts.end(), b.begin(), b.end());
  }
  ListReader bc(join(dir, "boundary.input"));
  bc.skip(); b = bc.ints(2); s.bc_ybtm = b[0]; s.bc_ytop = b[1];
  bc.skip(); b = bc.ints(2); s.bc_zbtm = b[0]; s.bc_ztop = b[1];
  bc.skip(); b = bc.ints(2); s.bc_t_ybtm = b[0]; s.bc_t_ytop = b[1];
  bc.skip(); a = bc.reals(2); s.val_t_ybtm = a[0]; s.val_t_ytop = a[1];
  bc.skip(); b = bc.ints(2); s.bc_t_zbtm = b[0]; s.bc_t_ztop = b[1];
  bc.skip(); a = bc.reals(2); s.val_t_zbtm = a[0]; s.val_t_ztop = a[1];
  bc.skip(); b = bc.ints(2); s.ich = b[0]; s.iconjg = b[1];
  ListReader pr(join(dir, "../output/ibmpre/ibmpre_prdic.bin"));
  b = pr.ints(4);
  s.xprdic = b[0]; s.yprdic = b[1]; s.zprdic = b[2]; s.iintp = b[3];
  // the Poisson solver is the reference's link-time choice (Makefile:23-34),
  // named by no input file: LW_POISSON = mgrd selects poiss_mgrd
  // and an x-periodic grid takes poiss_ftft, the Makefile's default, or
  // poiss_x_ft with LW_POISSON = x_ft
  const char *ps = getenv("LW_POISSON");
  if (s.xprdic == 1) s.poisson = (ps && strcmp(ps, "x_ft") == 0) ? 3 : 2;
  else s.poisson = (ps && strcmp(ps, "mgrd") == 0) ? 1 : 0;
}

template <class T>
std::vector<T> at_index(const std::vector<T> &vals, int64_t lo) {
  std::vector<T> out(vals.size() + lo + 1, T(0));
  for (size_t q = 0; q < vals.size(); ++q) out[q + lo] = vals[q];
  return out;
}

void read_grid(Case &c, const std::string &dir) {
  ListReader r(join(dir, c.gridfile));
  auto n = r.ints(3);
  c.n1 = n[0]; c.n2 = n[1]; c.n3 = n[2];
  n = r.ints(3);
  c.n1m = n[0]; c.n2m = n[1]; c.n3m = n[2];
  auto l = r.reals(3);
  c.xl = l[0]; c.yl = l[1]; c.zl = l[2];
  c.x = at_index(r.reals(c.n1), 1);
  c.y = at_index(r.reals(c.n2), 1);
  c.z = at_index(r.reals(c.n3), 1);
  c.ipv = at_index(r.ints(c.n1m), 1); c.jpv = at_index(r.ints(c.n2m), 1); c.kpv = at_index(r.ints(c.n3m), 1);
  c.imv = at_index(r.ints(c.n1m), 1); c.jmv = at_index(r.ints(c.n2m), 1); c.kmv = at_index(r.ints(c.n3m), 1);
  c.fixil = at_index(r.ints(c.n1m), 1); c.fixjl = at_index(r.ints(c.n2m), 1); c.fixkl = at_index(r.ints(c.n3m), 1);
  c.fixiu = at_index(r.ints(c.n1m), 1); c.fixju = at_index(r.ints(c.n2m), 1); c.fixku = at_index(r.ints(c.n3m), 1);
  c.c2cx = r.reals(c.n1 + 1); c.c2cy = r.reals(c.n2 + 1); c.c2cz = r.reals(c.n3 + 1);
  c.c2cxi = r.reals(c.n1 + 1); c.c2cyi = r.reals(c.n2 + 1); c.c2czi = r.reals(c.n3 + 1);
  c.f2fx = r.reals(c.n1 + 1); c.f2fy = r.reals(c.n2 + 1); c.f2fz = r.reals(c.n3 + 1);
  c.f2fxi = r.reals(c.n1 + 1); c.f2fyi = r.reals(c.n2 + 1); c.f2fzi = r.reals(c.n3 + 1);
  c.xmp = r.reals(c.n1 + 1); c.ymp = r.reals(c.n2 + 1); c.zmp = r.reals(c.n3 + 1);
}

void read_fcpts(Case &c, const std::string &path, int c0, int nc) {
  ListReader r(path);
  auto ni = r.ints(nc);
  auto nn = r.ints(nc);
  for (int q = 0; q < nc; ++q) {
    c.nintp[c0 + q] = ni[q];
    c.nbody[c0 + q] = ni[q] + nn[q];
  }
  for (int q = 0; q < nc; ++q)
    for (int64_t n = 0; n < c.nbody[c0 + q]; ++n) {
      auto v = r.ints(3);
      c.ifc[c0 + q].push_back(v[0]);
      c.jfc[c0 + q].push_back(v[1]);
      c.kfc[c0 + q].push_back(v[2]);
    }
  for (int q = 0; q < nc; ++q)
    for (int64_t n = 0; n < c.nintp[c0 + q]; ++n) {
      auto v = r.ints(3);
      c.indx[c0 + q].insert(c.indx[c0 + q].end(), v.begin(), v.end());
    }
  for (int q = 0; q < nc; ++q)
    for (int64_t n = 0; n < c.nintp[c0 + q]; ++n) {
      auto v = r.reals(27);
      c.gf[c0 + q].insert(c.gf[c0 + q].end(), v.begin(), v.end());
    }
}

void read_tables(Case &c, const std::string &dir) {
  read_fcpts(c, join(dir, "../output/ibmpre/ibmpre_fcpts.bin"), 0, 3);
  if (c.s.ihtrans == 1) read_fcpts(c, join(dir, "../output/ibmpre/ibmpre_fcpts_t.bin"), 3, 1);
  const size_t ni = (size_t)(c.n1m * c.n2m * c.n3m);
  c.nwall.assign(ni, 1);
  if (c.s.iles == 1) {
    ListReader r(join(dir, "../output/ibmpre/ibmpre_nutzero.bin"));
    c.nzero = r.ints(1)[0];
    if (c.nzero > 0) {
      c.inz = r.ints(c.nzero);
      c.jnz = r.ints(c.nzero);
      c.knz = r.ints(c.nzero);
    }
    ListReader wr(join(dir, "../output/ibmpre/ibmpre_wallfdvm.bin"));
    c.nwall = wr.ints(ni);
  }
}

// unformatted sequential records: 4-byte length, payload, 4-byte length
std::vector<std::vector<char>> read_records(const std::string &path) {
  std::ifstream in(path, std::ios::binary);
  if (!in) throw IoError{"cannot open " + path};
  std::vector<std::vector<char>> recs;
  while (true) {
    int32_t len = 0;
    if (!in.read((char *)&len, 4)) break;
    if (len < 0) throw IoError{"corrupt record marker in " + path};
    std::vector<char> buf((size_t)len);
    if (!in.read(buf.data(), len)) throw IoError{"truncated record in " + path};
    int32_t len2 = 0;
    if (!in.read((char *)&len2, 4) || len2 != len) throw IoError{"corrupt record marker in " + path};
    recs.push_back(std::move(buf));
  }
  return recs;
}

void read_field(Case &c, const std::string &dir) {
  auto recs = read_records(join(dir, c.prev_fld));
  if (recs.size() < 9) throw IoError{"fld file has too few records"};
  int64_t nn[3];
  memcpy(nn, recs[0].data(), 24);
  if (nn[0] != c.n1 || nn[1] != c.n2 || nn[2] != c.n3)
    throw IoError{"PREVIOUS FIELD DOES NOT MATCH WITH THE GRID SYSTEM."};
  memcpy(&c.ihist, recs[1].data(), 8);
  memcpy(&c.m, recs[1].data() + 8, 8);
  memcpy(&c.time, recs[1].data() + 16, 8);
  memcpy(&c.dt, recs[1].data() + 24, 8);
  int64_t pr[3], bcs[6], ich[2];
  memcpy(pr, recs[2].data(), 24);
  memcpy(bcs, recs[3].data(), 48);
  memcpy(ich, recs[4].data(), 16);
  lw_settings &s = c.s;
  s.xprdic = pr[0]; s.yprdic = pr[1]; s.zprdic = pr[2];
  s.bc_ybtm = bcs[2]; s.bc_ytop = bcs[3]; s.bc_zbtm = bcs[4]; s.bc_ztop = bcs[5];
  s.ich = ich[0]; s.iconjg = ich[1];
  const int64_t NX = c.n1 + 1, NY = c.n2 + 1, NZ = c.n3 + 1;
  const size_t np = (size_t)(NX * NY * NZ);
  c.u.assign(np, 0.0); c.v.assign(np, 0.0); c.w.assign(np, 0.0); c.p.assign(np, 0.0);
  auto P = [&](int64_t i, int64_t j, int64_t k) { return (size_t)(i + NX * (j + NY * k)); };
  auto unpack = [&](const std::vector<char> &rec, std::vector<double> &a, int64_t i0, int64_t i1, int64_t j0, int64_t j1,
                    int64_t k0, int64_t k1) {
    const size_t need = (size_t)((i1 - i0 + 1) * (j1 - j0 + 1) * (k1 - k0 + 1)) * 8;
    if (rec.size() != need) throw IoError{"fld record size mismatch"};
    const double *d = (const double *)rec.data();
    size_t q = 0;
    for (int64_t k = k0; k <= k1; ++k)
      for (int64_t j = j0; j <= j1; ++j)
        for (int64_t i = i0; i <= i1; ++i) a[P(i, j, k)] = d[q++];
  };
  unpack(recs[5], c.u, 1, c.n1, 0, c.n2, 0, c.n3);
  unpack(recs[6], c.v, 0, c.n1, 1, c.n2, 0, c.n3);
  unpack(recs[7], c.w, 0, c.n1, 0, c.n2, 1, c.n3);
  unpack(recs[8], c.p, 1, c.n1m, 1, c.n2m, 1, c.n3m);
  if (s.ihtrans == 1) {
    if (recs.size() < 10) throw IoError{"fld file lacks the temperature record"};
    c.t.assign(np, s.tsol_i);  // prefld: t = tsol_i, then the interior is read
    unpack(recs[9], c.t, 1, c.n1m, 1, c.n2m, 1, c.n3m);
  }
  if (s.ipzero == 1) std::fill(c.p.begin(), c.p.end(), 0.0);
  // periodic ghosts (misc_init.f90:653-783), x before z; y-periodic grids
  // are out of scope
  if (s.xprdic == 1) {
    for (int64_t k = 0; k <= c.n3; ++k)
      for (int64_t j = 0; j <= c.n2; ++j) {
        c.u[P(0, j, k)] = c.u[P(c.n1m, j, k)];
        c.u[P(c.n1, j, k)] = c.u[P(1, j, k)];
        if (j >= 1) {
          c.v[P(0, j, k)] = c.v[P(c.n1m, j, k)];
          c.v[P(c.n1, j, k)] = c.v[P(1, j, k)];
        }
        if (k >= 1) {
          c.w[P(0, j, k)] = c.w[P(c.n1m, j, k)];
          c.w[P(c.n1, j, k)] = c.w[P(1, j, k)];
        }
        if (s.ihtrans == 1) {
          c.t[P(0, j, k)] = c.t[P(c.n1m, j, k)];
          c.t[P(c.n1, j, k)] = c.t[P(1, j, k)];
        }
      }
  }
  if (s.zprdic == 1) {
    for (int64_t j = 0; j <= c.n2; ++j)
      for (int64_t i = 0; i <= c.n1; ++i) {
        if (i >= 1) {
          c.u[P(i, j, 0)] = c.u[P(i, j, c.n3m)];
          c.u[P(i, j, c.n3)] = c.u[P(i, j, 1)];
        }
        if (j >= 1) {
          c.v[P(i, j, 0)] = c.v[P(i, j, c.n3m)];
          c.v[P(i, j, c.n3)] = c.v[P(i, j, 1)];
        }
        c.w[P(i, j, 0)] = c.w[P(i, j, c.n3m)];
        c.w[P(i, j, c.n3)] = c.w[P(i, j, 1)];
        if (s.ihtrans == 1) {
          c.t[P(i, j, 0)] = c.t[P(i, j, c.n3m)];
          c.t[P(i, j, c.n3)] = c.t[P(i, j, 1)];
        }
      }
  }
}

// makefld (misc_init.f90:443-607), the initial field when IREAD = 0: u =
// the ICH's profile at fluid u points, v = w = p = 0, no-slip
// rows at wall boundaries; T = TSOL_I, then TFLU_I at fluid cell centres of
// rows 1..n2m, the inlet ghost by the level set and the y (and z) ghosts by
// the thermal BCs. The u loop starts at i = 0 and reads x(0), which the
// reference never allocated (x is x(1:n1), mod_common.f90:124); the u ghost
// column i = 0 is read by no routine of the step, and here x(0) is taken as 0.
void make_field(Case &c, const lw_body *body) {
  const lw_settings &s = c.s;
  const int64_t NX = c.n1 + 1, NY = c.n2 + 1, NZ = c.n3 + 1;
  const size_t np = (size_t)(NX * NY * NZ);
  c.u.assign(np, 0.0); c.v.assign(np, 0.0); c.w.assign(np, 0.0); c.p.assign(np, 0.0);
  auto P = [&](int64_t i, int64_t j, int64_t k) { return (size_t)(i + NX * (j + NY * k)); };
  auto fb = [&](double x, double y, double z) { return lw_body_eval(body, x, y, z); };
  c.ihist = 0;
  c.m = 0;
  c.time = 0.0;
  c.dt = s.dt_size;
  // the profile by ICH (misc_init.f90:471-481): uniform UDRV_I, the
  // laminar channel 1.5 UDRV_I (1 - y^2), or that of bulk velocity 12
  for (int64_t k = 0; k <= c.n3; ++k)
    for (int64_t j = 0; j <= c.n2; ++j) {
      const double y = c.ymp[j];
      const double prof = s.ich == 1 ? 1.5 * s.udrv_i * (1.0 - y * y)
                          : s.ich == 2 ? (1.5 * (1.0 - y * y)) * 12.0
                                       : s.udrv_i;
      for (int64_t i = 0; i <= c.n1; ++i)
        c.u[P(i, j, k)] = fb(c.x[i], c.ymp[j], c.zmp[k]) >= 1.e-10 ? prof : 0.0;
    }
  for (int64_t k = 0; k <= c.n3; ++k)
    for (int64_t i = 1; i <= c.n1; ++i) {
      if (s.bc_ybtm == 0) c.u[P(i, 0, k)] = c.u[P(i, 1, k)] = 0.0;
      if (s.bc_ytop == 0) c.u[P(i, c.n2m, k)] = c.u[P(i, c.n2, k)] = 0.0;
    }
  for (int64_t j = 0; j <= c.n2; ++j)
    for (int64_t i = 1; i <= c.n1; ++i) {
      if (s.bc_zbtm == 0) c.u[P(i, j, 0)] = c.u[P(i, j, 1)] = 0.0;
      if (s.bc_ztop == 0) c.u[P(i, j, c.n3m)] = c.u[P(i, j, c.n3)] = 0.0;
    }
  if (s.ihtrans != 1) return;
  c.t.assign(np, s.tsol_i);
  for (int64_t k = 0; k <= c.n3; ++k)
    for (int64_t j = 1; j <= c.n2m; ++j)
      for (int64_t i = 0; i <= c.n1; ++i)
        if (fb(c.xmp[i], c.ymp[j], c.zmp[k]) >= 1.e-10) c.t[P(i, j, k)] = s.tflu_i;
  if (s.xprdic == 0)
    for (int64_t k = 0; k <= c.n3; ++k)
      for (int64_t j = 0; j <= c.n2; ++j)
        c.t[P(0, j, k)] = fb(c.xmp[0], c.ymp[j], c.zmp[k]) >= 1.e-10 ? s.tflu_i : s.tsol_i;
  if (s.yprdic == 0)
    for (int64_t k = 0; k <= c.n3; ++k)
      for (int64_t i = 0; i <= c.n1; ++i) {
        if (s.bc_t_ybtm == 0) c.t[P(i, 0, k)] = s.val_t_ybtm;
        else if (s.bc_t_ybtm == 1) c.t[P(i, 0, k)] = c.t[P(i, 1, k)] - s.val_t_ybtm / c.c2cyi[1];
        if (s.bc_t_ytop == 0) c.t[P(i, c.n2, k)] = s.val_t_ytop;
        else if (s.bc_t_ytop == 1) c.t[P(i, c.n2, k)] = c.t[P(i, c.n2m, k)] + s.val_t_ytop / c.c2cyi[c.n2];
      }
  if (s.zprdic == 0)
    for (int64_t j = 0; j <= c.n2; ++j)
      for (int64_t i = 0; i <= c.n1; ++i) {
        if (s.bc_t_zbtm == 0) c.t[P(i, j, 0)] = s.val_t_zbtm;
        else if (s.bc_t_zbtm == 1) c.t[P(i, j, 0)] = c.t[P(i, j, 1)] - s.val_t_zbtm / c.c2czi[1];
        if (s.bc_t_ztop == 0) c.t[P(i, j, c.n3)] = s.val_t_ztop;
        else if (s.bc_t_ztop == 1) c.t[P(i, j, c.n3)] = c.t[P(i, j, c.n3m)] + s.val_t_ztop / c.c2czi[c.n3];
      }
}

}  // namespace

extern "C" int lw_create_from_dir(const char *solver_dir, const lw_body *body, lw_handle **out) {
  if (!solver_dir || !out) return -1;
  try {
    Case c;
    const std::string dir(solver_dir);
    read_settings(c, dir);
    read_grid(c, dir);
    read_tables(c, dir);
    lw_body cyl{};
    cyl.p[0] = 0.5;
    if (c.s.iread == 1) read_field(c, dir);
    else make_field(c, body ? body : &cyl);
    lw_grid g{};
    g.n1 = c.n1; g.n2 = c.n2; g.n3 = c.n3; g.n1m = c.n1m; g.n2m = c.n2m; g.n3m = c.n3m;
    g.xl = c.xl; g.yl = c.yl; g.zl = c.zl;
    g.x = c.x.data(); g.y = c.y.data(); g.z = c.z.data();
    g.ipv = c.ipv.data(); g.jpv = c.jpv.data(); g.kpv = c.kpv.data();
    g.imv = c.imv.data(); g.jmv = c.jmv.data(); g.kmv = c.kmv.data();
    g.fixil = c.fixil.data(); g.fixjl = c.fixjl.data(); g.fixkl = c.fixkl.data();
    g.fixiu = c.fixiu.data(); g.fixju = c.fixju.data(); g.fixku = c.fixku.data();
    g.c2cx = c.c2cx.data(); g.c2cy = c.c2cy.data(); g.c2cz = c.c2cz.data();
    g.c2cxi = c.c2cxi.data(); g.c2cyi = c.c2cyi.data(); g.c2czi = c.c2czi.data();
    g.f2fx = c.f2fx.data(); g.f2fy = c.f2fy.data(); g.f2fz = c.f2fz.data();
    g.f2fxi = c.f2fxi.data(); g.f2fyi = c.f2fyi.data(); g.f2fzi = c.f2fzi.data();
    g.xmp = c.xmp.data(); g.ymp = c.ymp.data(); g.zmp = c.zmp.data();
    lw_tables t{};
    for (int q = 0; q < 4; ++q) {
      t.nintp[q] = c.nintp[q];
      t.nbody[q] = c.nbody[q];
      t.ifc[q] = c.ifc[q].data(); t.jfc[q] = c.jfc[q].data(); t.kfc[q] = c.kfc[q].data();
      t.intpindx[q] = c.indx[q].data();
      t.geomfac[q] = c.gf[q].data();
    }
    t.nzero = c.nzero;
    t.inz = c.inz.data(); t.jnz = c.jnz.data(); t.knz = c.knz.data();
    t.nwall_dvm = c.nwall.data();
    int rc = lw_create(&c.s, &g, &t, body, out);
    if (rc != 0) return rc;
    // field_avg writes under ../output/field_avg relative to the solver's
    // working directory (misc_hist.f90:484)
    if (c.s.iavg == 1) lw_set_avg_dir(*out, join(dir, "../output/field_avg").c_str());
    rc = lw_set_fields(*out, c.u.data(), c.v.data(), c.w.data(), c.p.data(), c.s.ihtrans == 1 ? c.t.data() : nullptr,
                       c.ihist, c.m, c.time, c.dt);
    if (rc == 0) rc = lw_set_trace(*out, c.ntrace, c.ntr, c.trpts.data());
    if (rc != 0) {
      lw_destroy(*out);
      *out = nullptr;
    }
    return rc;
  } catch (const IoError &e) {
    io_err = e.msg;
    lw_set_last_error(e.msg);
    fprintf(stderr, "leswht: %s\n", e.msg.c_str());
    return LW_ERR_IO;
  }
}

// settings.input, boundary.input and the grid header of a case directory,
// without the tables and the field: what an admission decision needs
extern "C" int lw_probe_dir(const char *solver_dir, lw_settings *s, int64_t nm[3]) {
  if (!solver_dir || !s || !nm) return LW_ERR_ARG;
  try {
    Case c;
    const std::string dir(solver_dir);
    read_settings(c, dir);
    ListReader r(join(dir, c.gridfile));
    r.ints(3);
    auto n = r.ints(3);
    *s = c.s;
    nm[0] = n[0]; nm[1] = n[1]; nm[2] = n[2];
    return LW_OK;
  } catch (const IoError &e) {
    io_err = e.msg;
    lw_set_last_error(e.msg);
    return LW_ERR_IO;
  }
}

// writefield (misc_hist.f90:212-252); settings needed for the header come
// from the arguments the caller keeps
extern "C" int lw_write_fld(const char *path, const int64_t n[3], double re, double pr, double gr, int64_t ihist,
                            int64_t m, double time, double dt, const int64_t prdic[3], const int64_t bcs[6],
                            const int64_t ich[2], const double *u, const double *v, const double *w,
                            const double *p, const double *t) {
  FILE *fp = fopen(path, "wb");
  if (!fp) return LW_ERR_IO;
  auto rec = [&](const void *d, size_t len) {
    const int32_t l = (int32_t)len;
    fwrite(&l, 4, 1, fp);
    fwrite(d, 1, len, fp);
    fwrite(&l, 4, 1, fp);
  };
  const int64_t n1 = n[0], n2 = n[1], n3 = n[2], NX = n1 + 1, NY = n2 + 1;
  char h0[48];
  memcpy(h0, n, 24);
  memcpy(h0 + 24, &re, 8); memcpy(h0 + 32, &pr, 8); memcpy(h0 + 40, &gr, 8);
  rec(h0, 48);
  char h1[32];
  memcpy(h1, &ihist, 8); memcpy(h1 + 8, &m, 8); memcpy(h1 + 16, &time, 8); memcpy(h1 + 24, &dt, 8);
  rec(h1, 32);
  rec(prdic, 24);
  rec(bcs, 48);
  rec(ich, 16);
  auto pack = [&](const double *a, int64_t i0, int64_t i1, int64_t j0, int64_t j1, int64_t k0, int64_t k1) {
    std::vector<double> b;
    b.reserve((size_t)((i1 - i0 + 1) * (j1 - j0 + 1) * (k1 - k0 + 1)));
    for (int64_t k = k0; k <= k1; ++k)
      for (int64_t j = j0; j <= j1; ++j)
        for (int64_t i = i0; i <= i1; ++i) b.push_back(a[i + NX * (j + NY * k)]);
    rec(b.data(), b.size() * 8);
  };
  pack(u, 1, n1, 0, n2, 0, n3);
  pack(v, 0, n1, 1, n2, 0, n3);
  pack(w, 0, n1, 0, n2, 1, n3);
  pack(p, 1, n1 - 1, 1, n2 - 1, 1, n3 - 1);
  if (t) pack(t, 1, n1 - 1, 1, n2 - 1, 1, n3 - 1);
  const int bad = ferror(fp);
  fclose(fp);
  return bad ? LW_ERR_IO : LW_OK;
}
