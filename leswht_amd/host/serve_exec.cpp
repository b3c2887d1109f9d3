// serve_exec: runs many reference case directories through the request loop
// (lw_server_*) on one GPU. The reference runs one case per solver_exec
// process (run_solver.sh:41); this driver takes a manifest instead:
//
//   serve_exec MANIFEST [--budget-gb G] [--slots S]
//
// Each non-empty manifest line is "SOLVER_DIR NSTEPS [KIND P1 P2 P3]": the
// case's solver directory (what solver_exec runs in), its step count and
// optionally its body (lw_body kind 0 cylinder, 1 square, 2 ellipse, with
// its parameters; default the cylinder of radius 0.5). Lines starting with
// '#' are comments. Cases are submitted in manifest order; each writes its
// final field to SOLVER_DIR/../output/field/fld%07d (the step count).
// One line per case is printed when all are done: status, cells, steps,
// queue wait, run time and time-to-solution in seconds. The exit status is
// 0 when every case succeeded, 1 otherwise, 2 on a bad manifest.
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/leswht.h"

struct Entry {
  std::string dir;
  int64_t nsteps = 0;
  lw_body body{};
  int64_t ticket = -1;
};

int main(int argc, char **argv) {
  if (argc < 2) {
    fprintf(stderr, "usage: serve_exec MANIFEST [--budget-gb G] [--slots S]\n");
    return 2;
  }
  double budget_gb = 200.0;
  int64_t slots = 8;
  for (int a = 2; a + 1 < argc; a += 2) {
    if (!strcmp(argv[a], "--budget-gb")) budget_gb = atof(argv[a + 1]);
    else if (!strcmp(argv[a], "--slots")) slots = atoll(argv[a + 1]);
    else {
      fprintf(stderr, "serve_exec: unknown option %s\n", argv[a]);
      return 2;
    }
  }
  std::ifstream in(argv[1]);
  if (!in) {
    fprintf(stderr, "serve_exec: cannot open %s\n", argv[1]);
    return 2;
  }
  std::vector<Entry> cases;
  std::string ln;
  int lineno = 0;
  while (std::getline(in, ln)) {
    ++lineno;
    const size_t p = ln.find_first_not_of(" \t");
    if (p == std::string::npos || ln[p] == '#') continue;
    std::istringstream ss(ln);
    Entry e;
    e.body.p[0] = 0.5;
    if (!(ss >> e.dir >> e.nsteps) || e.nsteps < 0) {
      fprintf(stderr, "serve_exec: %s:%d: expected SOLVER_DIR NSTEPS\n", argv[1], lineno);
      return 2;
    }
    int64_t kind;
    if (ss >> kind) {
      e.body.kind = kind;
      e.body.p[0] = 0.0;
      for (int q = 0; q < 4 && (ss >> e.body.p[q]); ++q) {
      }
    }
    cases.push_back(e);
  }
  lw_server *srv = nullptr;
  if (lw_server_create((int64_t)(budget_gb * (double)(1LL << 30)), slots, &srv) != LW_OK) {
    fprintf(stderr, "serve_exec: %s\n", lw_last_error());
    return 2;
  }
  int failed = 0;
  std::vector<std::string> outs(cases.size());
  for (size_t q = 0; q < cases.size(); ++q) {
    Entry &e = cases[q];
    char name[32];
    snprintf(name, sizeof name, "fld%07lld", (long long)e.nsteps);
    outs[q] = e.dir + "/../output/field/" + name;
    if (lw_server_submit(srv, e.dir.c_str(), &e.body, e.nsteps, outs[q].c_str(), &e.ticket) != LW_OK) {
      fprintf(stderr, "serve_exec: %s: %s\n", e.dir.c_str(), lw_last_error());
      ++failed;
    }
  }
  printf("# dir status cells steps wait_s run_s time_to_solution_s\n");
  for (Entry &e : cases) {
    if (e.ticket < 0) continue;
    lw_case_result r{};
    lw_server_wait(srv, e.ticket, &r);
    if (r.status != LW_CASE_DONE) {
      fprintf(stderr, "serve_exec: %s failed (%lld): %s\n", e.dir.c_str(), (long long)r.status, lw_last_error());
      ++failed;
    }
    printf("%s %lld %lld %lld %.3f %.3f %.3f\n", e.dir.c_str(), (long long)r.status, (long long)r.cells,
           (long long)r.steps_done, r.t_admit - r.t_submit, r.t_done - r.t_admit, r.t_done - r.t_submit);
  }
  lw_server_destroy(srv);
  return failed ? 1 : 0;
}
