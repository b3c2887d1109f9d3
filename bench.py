"""Benchmark of the LESwHT time step on one MI355X.

Workload: the 384x256x64 heated-cylinder LES (Re=3900, Pr=0.71, dynamic
Vreman SGS with a global Germano coefficient, dynamic eddy diffusivity, box
test filter in x and z, CFL-adaptive dt at CFL 0.5, seeded noise 1e-2 on the
initial field). A step is one full RK3 time step (three substeps) of
solver.f90:85-181 on data resident in HBM; the metric is cell-updates per
second = n1m*n2m*n3m*steps / wall time.

Prints one JSON line. ``roofline`` reports the dominant kernel's achieved HBM
bandwidth from its algorithmic bytes and its event-timed average duration;
``cpu_baseline`` times the reference solver (built from its own sources with
amdflang, oracle/_ref/solver_exec) on the host cores of the same machine for
two steps of the same case, read from its own per-step timer file.
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import os
import shutil
import statistics
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

from leswht_amd import case as C  # noqa: E402
from leswht_amd.solver import Solver  # noqa: E402

HBM_PEAK_GBS = 8000.0

PROFILES = os.path.join(HERE, "profiles")


def traffic_from_profile(kernel: str, config: str):
    """HBM bytes per launch of ``kernel`` measured by the FETCH_SIZE and
    WRITE_SIZE counters (rocprofv3 --pmc, scaled by the same counters on a
    device copy of known size), from the committed profile summary of this
    workload (a file ``*_<workload>.md``), whose line ``traffic <kernel>:
    <bytes> bytes per launch`` the summary's own commands produce. None when
    no summary holds it."""
    tag = "_%s.md" % WORKLOAD_TAGS[config]
    names = sorted(n for n in os.listdir(PROFILES) if n.endswith(tag)) if os.path.isdir(PROFILES) else []
    for name in names:
        path = os.path.join(PROFILES, name)
        for ln in open(path):
            t = ln.strip().split()
            if len(t) >= 3 and t[0] == "traffic" and t[1] == kernel + ":":
                return int(t[2]), os.path.relpath(path, HERE)
    return None, None


def c2_spec(nend: int = 2, layout: str = "tanh") -> C.CaseSpec:
    s = C.Settings(re=3900.0, pr=0.71, iles=1, insmdl=1, itemdl=0, idvmon=1, filter=3, idtopt=1,
                   cflfac=0.5, dt_size=0.005, resid1=1.0e-6, nlev=5, nbli=3, mgitr=500, imgsor=170,
                   nend=nend)
    return C.CaseSpec(384, 256, 64, dmin=0.01, noise=1.0e-2, settings=s, layout=layout)


def c1_spec(nend: int = 2, layout: str = "tanh") -> C.CaseSpec:
    """The plain FP64 case: Re 100, 192x128x16, fixed dt 0.005, no SGS model."""
    s = C.Settings(re=100.0, pr=0.71, iles=0, idtopt=0, dt_size=0.005, resid1=1.0e-6, nlev=5, nbli=3,
                   mgitr=500, nend=nend)
    return C.CaseSpec(192, 128, 16, dmin=0.01, noise=1.0e-3, settings=s, layout=layout)


WORKLOAD_TAGS = {"c2": "c2_384x256x64", "c1": "c1_192x128x16"}

WORKLOADS = {
    "c2": "heated circular cylinder LES Re=3900 Pr=0.71, 384x256x64 on [-15,30]x[-15,15]x[0,pi], dynamic "
          "Vreman + dynamic eddy diffusivity, box test filter x/z, CFL 0.5, NLEV 5; x and y lines stretched by "
          "hyperbolic-tangent segments, spacing 0.01D at the cylinder surface",
    "c1": "heated circular cylinder Re=100 Pr=0.71, 192x128x16 on [-15,30]x[-15,15]x[0,pi], no SGS model, "
          "dt 0.005, NLEV 5; x and y lines stretched by hyperbolic-tangent segments, spacing 0.01D at the "
          "cylinder surface",
}
GRID_SINH = ("; the earlier layout instead: x and y lines clustered by a two-sided sinh at the body centre "
             "(spacing 0.01 there, 0.0155-0.046 at the surface)")


SERIAL_REF_FILES = "slv_mmtm.f90, slv_engy.f90 and slv_auxi.f90"


def _mg_file_rows(path):
    rows = []
    for ln in open(path):
        t = ln.replace("=", " ").split()
        if len(t) == 6:
            rows.append((int(t[1]), int(t[3])))
    return rows


def cpu_baseline(spec: C.CaseSpec, case, fld_path: str, nsteps: int, threads: int, gpu: dict) -> dict:
    """The reference solver on the host over the same steps the GPU timed
    first: it restarts (IREAD = 1) from the fld file the GPU wrote at the
    start of its timed window and runs ``nsteps`` steps; per-step wall time
    and the SGS / RHSNLHS / POISSON split from its own timer file ftime.dat
    (misc_hist.f90:255-284), its multigrid cycle counts of kv = 1..3 from
    mgftresiduemax.dat (poiss_z_ft.f90:293-297) against the GPU's."""
    import oracle  # checker-side: the reference binary lives under oracle/_ref

    root = tempfile.mkdtemp(prefix="lw_cpu_")
    try:
        st = dataclasses.replace(spec.settings, iread=1, ireset=0, nend=nsteps, nprint=100000)
        if st.iles == 0:
            # the reference's ILES = 0 path reads LES arrays it never
            # allocated; "no SGS" runs there as ILES = 1 with the model off
            # and zero coefficients, which gives the same fields
            st = dataclasses.replace(st, iles=1, idvmon=0, csgsts=0.0, csgshf=0.0)
            case = C.build_case(dataclasses.replace(spec, settings=st))
        else:
            case = dataclasses.replace(case, spec=dataclasses.replace(spec, settings=st))
        C.write_case(case, root)
        shutil.copyfile(fld_path, os.path.join(root, "output/field/fld000000"))
        t0 = time.time()
        oracle.run_reference(root, threads=threads, timeout=1500)
        wall = time.time() - t0
        rows = [[float(x) for x in ln.split()] for ln in open(os.path.join(root, "output/ftr/ftime.dat"))
                if ln.strip()]
        kk = _mg_file_rows(os.path.join(root, "output/ftr/mgftresiduemax.dat"))
    finally:
        shutil.rmtree(root, ignore_errors=True)
    step_s = [r[4] for r in rows]
    med = statistics.median(step_s)
    n = spec.n1m * spec.n2m * spec.n3m
    gpu_kk = [(kv, k) for kv, k, _ in gpu["mg_rows"][:len(kk)]]
    return {
        "value": n / med, "unit": "cell-updates/s", "cores": threads, "kind": "reference",
        "sample": "reference solver built from its own sources (amdflang -O3; OpenMP with %d threads, except "
                  "%s, compiled serial: flang rejects or mis-handles their OpenMP clauses), restarted from "
                  "the fld the GPU wrote at the start of its timed window, over the same %d steps; median of "
                  "its own per-step timer (ftime.dat); process wall %.1f s incl. reading the text inputs"
                  % (threads, SERIAL_REF_FILES, len(step_s), wall),
        "s_per_step": med,
        "phase_s_per_step": {"sgs": statistics.median(r[1] for r in rows),
                             "rhsnlhs": statistics.median(r[2] for r in rows),
                             "poisson": statistics.median(r[3] for r in rows)},
        "same_steps": {"steps": len(step_s), "cpu_s": sum(step_s), "gpu_s": gpu["s"],
                       "ratio": sum(step_s) / gpu["s"] if gpu["s"] else None,
                       "mg_cycles_kv123_rows": len(kk),
                       "mg_cycles_kv123_rows_differing": sum(a != b for a, b in zip(kk, gpu_kk)) + abs(len(kk) - len(gpu_kk)),
                       "mg_cycles_kv1_per_step_cpu": sum(k for kv, k in kk if kv == 1) / max(len(step_s), 1),
                       "mg_cycles_kv1_per_step_gpu": sum(k for kv, k in gpu_kk if kv == 1) / max(len(step_s), 1)},
        "host": {"nproc": os.cpu_count(), "affinity_cpus": len(os.sched_getaffinity(0))},
    }


def serve_leg(args) -> dict:
    """A short served stream beside the time step: seeded Poisson arrivals of
    small cases (config 4's sizes, bodies, Re and Pr) and a share of
    384x256x64 LES cases through the native request loop; p50/p99
    time-to-solution and aggregate cell-updates/s of the cases that finished."""
    from leswht_amd import serve
    cases = serve.draw_cases(args.serve_leg, seed=args.seed, rate=args.rate,
                             steps=tuple(int(round(v * args.serve_leg_scale)) for v in (200, 2000)),
                             large_share=args.large_share,
                             large_steps=tuple(int(round(v * args.serve_leg_scale)) for v in (200, 2000)))
    root = tempfile.mkdtemp(prefix="lw_serve_")
    try:
        t0 = time.time()
        dirs = serve.prepare(cases, root)
        t_prep = time.time() - t0
        budget = int(args.budget_gb * (1 << 30))
        out = serve.run_stream(cases, dirs, budget, args.slots, honour_arrivals=True, write_fields=False)
        bat = serve.run_stream(cases, dirs, budget, args.slots, honour_arrivals=True, write_fields=False,
                               batch_max_cells=args.batch_cells)
    finally:
        shutil.rmtree(root, ignore_errors=True)
    return {"cases": len(cases), "large_cases": sum(c.large for c in cases), "rate_per_s": args.rate,
            "step_budgets": "U{200..2000} x %g" % args.serve_leg_scale, "slots": args.slots,
            "p50_s": out["p50"], "p99_s": out["p99"], "cell_updates_per_s": out["throughput"],
            "failed": len(out["failed"]) + len(bat["failed"]), "prep_s": round(t_prep, 1),
            "continuous_batching": {"batch_max_cells": args.batch_cells, "p50_s": bat["p50"], "p99_s": bat["p99"],
                                    "cell_updates_per_s": bat["throughput"]}}


def serve_bench(args) -> dict:
    """The served form: a seeded Poisson-arrival stream of independent cases
    (sizes 64..256 x 64..192 x 8..32, Re 40..400, Pr 0.71 or 7, circle,
    square or ellipse, plus a share of 384x256x64 LES cases) through the
    native request loop with a memory budget. Then the same cases submitted
    together, once through one slot (one case at a time, the plain form) and
    once through all slots, for the aggregate-throughput ratio."""
    from leswht_amd import serve
    cases = serve.draw_cases(args.serve, seed=args.seed, rate=args.rate,
                             steps=tuple(int(round(v * args.step_scale)) for v in (200, 2000)),
                             large_share=args.large_share,
                             large_steps=tuple(int(round(v * args.step_scale)) for v in (200, 2000)))
    root = tempfile.mkdtemp(prefix="lw_serve_")
    try:
        t0 = time.time()
        dirs = serve.prepare(cases, root)
        t_prep = time.time() - t0
        print("prepared %d case directories in %.1f s" % (len(cases), t_prep), file=sys.stderr, flush=True)
        budget = int(args.budget_gb * (1 << 30))
        stream = serve.run_stream(cases, dirs, budget, args.slots, honour_arrivals=True)
        print("stream: p50 %.3f s p99 %.3f s, %.3e cell-updates/s" % (
            stream["p50"], stream["p99"], stream["throughput"]), file=sys.stderr, flush=True)
        closed = serve.run_stream(cases, dirs, budget, args.slots, honour_arrivals=False, write_fields=False)
        print("all at once, %d slots: %.3e cell-updates/s" % (args.slots, closed["throughput"]),
              file=sys.stderr, flush=True)
        single = serve.run_stream(cases, dirs, budget, 1, honour_arrivals=False, write_fields=False)
        print("all at once, 1 slot: %.3e cell-updates/s" % single["throughput"], file=sys.stderr, flush=True)
        # continuous batching of the cases of at most batch_cells cells
        bstream = serve.run_stream(cases, dirs, budget, args.slots, honour_arrivals=True, write_fields=False,
                                   batch_max_cells=args.batch_cells)
        print("stream, batched: p50 %.3f s p99 %.3f s, %.3e cell-updates/s" % (
            bstream["p50"], bstream["p99"], bstream["throughput"]), file=sys.stderr, flush=True)
        bclosed = serve.run_stream(cases, dirs, budget, args.slots, honour_arrivals=False, write_fields=False,
                                   batch_max_cells=args.batch_cells)
        print("all at once, batched: %.3e cell-updates/s" % bclosed["throughput"], file=sys.stderr, flush=True)
        mixed = None
        if args.mixed:
            # the served mixed-precision form: the same batch through all
            # slots with the FP32 multigrid levels wherever a case allows them
            os.environ["LW_MG_FP32"] = "1"
            try:
                mixed = serve.run_stream(cases, dirs, budget, args.slots, honour_arrivals=False, write_fields=False)
            finally:
                del os.environ["LW_MG_FP32"]
            print("all at once, %d slots, mixed precision: %.3e cell-updates/s" % (args.slots, mixed["throughput"]),
                  file=sys.stderr, flush=True)
    finally:
        shutil.rmtree(root, ignore_errors=True)
    n_large = sum(c.large for c in cases)
    return {
        "metric": "p50/p99 time-to-solution of served cases; aggregate cell-updates/s (1 GPU)",
        "value": stream["throughput"],
        "unit": "cell-updates/s",
        "p50_s": stream["p50"],
        "p99_s": stream["p99"],
        "n_gpus": 1,
        "higher_is_better": True,
        "dtype": "f64",
        "data": "synthetic (seeded case stream: grids, bodies, IB tables and initial fields generated)",
        "config": {"workload": "Poisson-arrival stream of %d cases at %.1f/s, %d of them 384x256x64 LES; "
                               "step budgets U{200..2000} x %.3g" % (len(cases), args.rate, n_large,
                                                                    args.step_scale),
                   "slots": args.slots, "budget_gb": args.budget_gb, "seed": args.seed,
                   "hw_queues": int(os.environ.get("GPU_MAX_HW_QUEUES", "4")),
                   "cell_updates": stream["cell_updates"], "prep_s": round(t_prep, 1),
                   "peak_running": stream["info"]["peak_running"],
                   "peak_reserved_gb": round(stream["info"]["peak_reserved_bytes"] / (1 << 30), 2)},
        "closed_batch": {"slots": args.slots, "value": closed["throughput"], "one_slot": single["throughput"],
                         "ratio": closed["throughput"] / single["throughput"]},
        "continuous_batching": {"batch_max_cells": args.batch_cells, "stream_p50_s": bstream["p50"],
                                "stream_p99_s": bstream["p99"], "stream_value": bstream["throughput"],
                                "closed_value": bclosed["throughput"],
                                "closed_ratio_vs_one_slot": bclosed["throughput"] / single["throughput"]},
        "closed_batch_mixed": None if mixed is None else {
            "slots": args.slots, "value": mixed["throughput"],
            "ratio_vs_one_slot_fp64": mixed["throughput"] / single["throughput"]},
        "failed": len(stream["failed"]) + len(closed["failed"]) + len(single["failed"]) + len(bstream["failed"]) +
                  len(bclosed["failed"]) +
                  (len(mixed["failed"]) if mixed else 0),
    }


def batch_bench(args) -> dict:
    """The batched form (config 4): independent cases drawn as the served
    stream's small cases (cells 64..256 x 64..192 x 8..32, Re log-uniform in
    40..400, Pr 0.71 or 7, circle, square or ellipse at a random angle, CFL
    0.5), resident in HBM together. Timed three ways over the same steps of
    the same cases: one case at a time (the plain form), every case on its
    own streams from its own host thread (the concurrent form of the request
    loop), and one batch (lw_batch: each kernel launched once for all the
    cases that issue it)."""
    import threading

    from leswht_amd import serve
    from leswht_amd.solver import Batch
    cases = serve.draw_cases(args.batch, seed=args.seed, large_share=0.0)
    t0 = time.time()
    views = [C.case_view(C.build_case(c.spec)) for c in cases]
    t_prep = time.time() - t0
    print("built %d cases in %.1f s" % (len(cases), t_prep), file=sys.stderr, flush=True)
    sols = [Solver.from_case(v) for v in views]
    n_mixed = 0
    if args.mixed:
        # the FP32 multigrid levels wherever a case's levels allow them
        for sv in sols:
            try:
                sv.set_option("mg_fp32", 1)
                n_mixed += 1
            except Exception:
                pass
    cells = [c.cells for c in cases]
    k, w = args.steps, args.warmup
    t_plain = t_conc = None
    if not args.batch_only:
        # plain form: one case at a time
        t_plain = 0.0
        for sv in sols:
            sv.step(w)
            t0 = time.perf_counter()
            sv.step(k)
            t_plain += time.perf_counter() - t0
        print("one at a time: %.3f s" % t_plain, file=sys.stderr, flush=True)

    # concurrent form: every case on its own streams from its own thread
    def run(sv, n):
        sv.step(n)

    def concurrent(n):
        th = [threading.Thread(target=run, args=(sv, n)) for sv in sols]
        t0 = time.perf_counter()
        for t in th:
            t.start()
        for t in th:
            t.join()
        return time.perf_counter() - t0

    if not args.batch_only:
        concurrent(w)
        t_conc = concurrent(k)
        print("concurrent streams: %.3f s" % t_conc, file=sys.stderr, flush=True)
    # batched form
    b = Batch(sols)
    b.step(w)
    i0 = b.info()
    t0 = time.perf_counter()
    st = b.step(k)
    t_bat = time.perf_counter() - t0
    i1 = b.info()
    b.close()
    print("batched: %.3f s" % t_bat, file=sys.stderr, flush=True)
    upd = float(sum(cells)) * k
    info = {q: i1[q] - i0[q] for q in i1}
    for sv in sols:
        sv.close()
    return {
        "metric": "aggregate cell-updates/s of a batch of independent cases (1 GPU)",
        "value": upd / t_bat,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": k,
        "warmup": w,
        "higher_is_better": True,
        "dtype": "f64+f32 multigrid levels" if args.mixed else "f64",
        "data": "synthetic (seeded cases: grids, bodies, IB tables and initial fields generated)",
        "config": {"workload": "%d independent heated-body cases drawn as config 4 (cells 64..256 x 64..192 x "
                               "8..32, Re 40..400, Pr 0.71/7, circle/square/ellipse, CFL 0.5)" % len(cases),
                   "cases": len(cases), "cells_total": int(sum(cells)), "seed": args.seed, "prep_s": round(t_prep, 1)},
        "one_at_a_time": upd / t_plain if t_plain else None,
        "concurrent_streams": upd / t_conc if t_conc else None,
        "ratio_vs_one_at_a_time": t_plain / t_bat if t_plain else None,
        "ratio_vs_concurrent_streams": t_conc / t_bat if t_conc else None,
        "launch_sharing": {"batched_launches": info["launches"], "case_launches": info["case_launches"],
                           "rendezvous": info["rendezvous"],
                           "cases_per_launch": info["case_launches"] / max(info["launches"], 1)},
        "mg_cycles_sum_last_step": int(sum(x["mg_cycles_sum"] for x in st)),
        "cases_with_fp32_levels": n_mixed,
        "host_ms_per_step": {"flush_prepare": info["prepare_us"] / 1e3 / k, "device_wait": info["wait_us"] / 1e3 / k},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int,
                    default=int(os.environ.get("LESWHT_CPU_THREADS", os.environ.get("OMP_NUM_THREADS", "16"))),
                    help="OpenMP threads of the reference solver (default: OMP_NUM_THREADS, else 16)")
    ap.add_argument("--cpu-steps", type=int, default=5,
                    help="steps of the reference run: the first steps of the GPU's timed window")
    ap.add_argument("--later-at", type=int, default=200,
                    help="step after which a later window is timed (0: none)")
    ap.add_argument("--later-steps", type=int, default=10, help="steps of the later window")
    ap.add_argument("--serve-leg", type=int, default=16,
                    help="cases of the short served stream reported beside the step (0: none)")
    ap.add_argument("--serve-leg-scale", type=float, default=0.05,
                    help="short served stream: factor on the step budgets U{200..2000}")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run a device-to-device copy of this many MiB five times (the known-size "
                         "calibration of the FETCH_SIZE / WRITE_SIZE counters), then exit")
    ap.add_argument("--kernel-steps", type=int, default=3,
                    help="steps after the timed region with every launch timed, for the roofline kernel's "
                         "average duration (0: use the timed region's own timing of that kernel)")
    ap.add_argument("--config", choices=("c2", "c1"), default="c2",
                    help="c2: the 384x256x64 LES the metric is quoted on; c1: the plain 192x128x16 case")
    ap.add_argument("--nlev", type=int, default=0, help="multigrid levels (0: the workload's own)")
    ap.add_argument("--grid", choices=("tanh", "sinh"), default="tanh",
                    help="tanh: the configurations' lines (0.01D at the body surface); sinh: the earlier "
                         "layout, clustered at the body centre")
    ap.add_argument("--mixed", action="store_true",
                    help="the mixed-precision form: the multigrid level planes below the top level stored in FP32 "
                         "(arithmetic, the top level and the convergence test stay FP64) and the dynamic SGS "
                         "models' test filter in FP32; not the default, and not the metric's form")
    ap.add_argument("--autotune", action="store_true",
                    help="pick the multigrid tile sizes with the library's autotuner before the warmup")
    ap.add_argument("--batch", type=int, default=0,
                    help="run the batched form on this many independent small cases instead of the time step")
    ap.add_argument("--batch-only", action="store_true",
                    help="batched form: time the batch alone, without the one-at-a-time and concurrent forms "
                         "(for profiling it)")
    ap.add_argument("--batch-leg", type=int, default=32,
                    help="cases of the short batched-form comparison reported beside the step (0: none)")
    ap.add_argument("--batch-cells", type=int, default=2000000,
                    help="served forms: cases of at most this many cells join the continuous batch")
    ap.add_argument("--serve", type=int, default=0,
                    help="run the served form on a stream of this many cases instead of the time step")
    ap.add_argument("--rate", type=float, default=10.0, help="served form: arrivals per second")
    ap.add_argument("--step-scale", type=float, default=0.1,
                    help="served form: factor on the step budgets U{200..2000}")
    ap.add_argument("--slots", type=int, default=8, help="served form: cases on the device at once")
    ap.add_argument("--large-share", type=float, default=0.1,
                    help="served form: share of 384x256x64 LES cases in the stream (0: the batch of small cases)")
    ap.add_argument("--budget-gb", type=float, default=200.0, help="served form: admission budget")
    ap.add_argument("--seed", type=int, default=20261203)
    ap.add_argument("--hw-queues", type=int, default=0,
                    help="served form: hardware queues per process (GPU_MAX_HW_QUEUES, at most 32); "
                         "every case runs on two streams, which share the queues")
    args = ap.parse_args()
    if args.hw_queues:
        # read by the HIP runtime when it initialises, before any device call
        os.environ["GPU_MAX_HW_QUEUES"] = str(min(max(args.hw_queues, 1), 32))
    if args.serve:
        print(json.dumps(serve_bench(args)))
        return
    if args.batch:
        print(json.dumps(batch_bench(args)))
        return
    if args.copy_probe_mib:
        import torch
        n = args.copy_probe_mib * (1 << 20) // 8
        a = torch.rand(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_each_way": n * 8, "launches": 5}))
        return
    if args.gpus != 1:
        print("bench.py runs on one GPU only: this build measures the single-GPU time step "
              "(multi-GPU sharding is a design in DESIGN.md, not implemented)", file=sys.stderr)
        sys.exit(2)

    spec = c1_spec(layout=args.grid) if args.config == "c1" else c2_spec(layout=args.grid)
    if args.nlev:
        spec = dataclasses.replace(spec, settings=dataclasses.replace(spec.settings, nlev=args.nlev))
    t0 = time.time()
    case = C.build_case(spec)
    view = C.case_view(case)
    t_build = time.time() - t0
    sol = Solver.from_case(view)
    if args.mixed:
        sol.set_option("mg_fp32", 1)
        sol.set_option("sgs_fp32", 1)
    tune = sol.autotune(reps=2) if args.autotune else None
    # untimed warmup; its last step times every kernel once to find the one
    # with the most device time
    sol.step(max(args.warmup - 1, 0))
    sol.profile_begin()
    sol.step(1)
    table = sorted(sol.profile(), key=lambda r: -r["ms_total"])
    # the roofline kernel: the one with the most device time among those with
    # a byte model (k_mg_kv1, when switched on, runs a data-dependent number
    # of cycles and has none)
    dom_name = next((r["name"] for r in table if r["bytes"] > 0), table[0]["name"])
    # the state at the start of the timed window, for the reference run over
    # the same steps (written before the clock starts)
    work = tempfile.mkdtemp(prefix="lw_bench_")
    fld_start = os.path.join(work, "fld_start")
    sol.write_field(fld_start)
    # timed region: K steps, with the dominant kernel's launches timed by the
    # start and end timestamps of their own dispatches on the streams they run on
    sol.profile_begin([dom_name])
    step_s, mg_rows = [], []
    pois_ms, cycles = 0.0, 0
    t0 = time.perf_counter()
    for q in range(args.steps):
        ts = time.perf_counter()
        st = sol.step(1)
        step_s.append(time.perf_counter() - ts)
        pois_ms += st["ms_poisson"]
        cycles += st["mg_cycles_sum"]
        if q < args.cpu_steps:
            mg_rows += sol.mg_residuals()
    elapsed = time.perf_counter() - t0
    prof_region = sol.profile()
    sol.profile_end()
    # then a few steps with every launch timed: a timed dispatch that follows
    # untimed ones on a stream also counts the time the stream spends
    # reaching it, so the timed region's figure for a kernel of the two-stream
    # Poisson solve runs high; timing every launch gives the kernel's own
    # durations (rocprofv3's kernel trace agrees) at the cost of a slower step
    prof = prof_region
    if args.kernel_steps > 0:
        sol.profile_begin()
        sol.step(args.kernel_steps)
        prof = sol.profile()
        sol.profile_end()
    n = spec.n1m * spec.n2m * spec.n3m
    value = n * args.steps / elapsed
    done_steps = args.warmup + args.steps + max(args.kernel_steps, 0)
    later = None
    if args.later_at > 0 and args.later_steps > 0:
        # out of the impulsive start: the same case further on
        if args.later_at > done_steps:
            sol.step(args.later_at - done_steps)
        cyc, t1 = 0, time.perf_counter()
        for _ in range(args.later_steps):
            cyc += sol.step(1)["mg_cycles_sum"]
        lt = time.perf_counter() - t1
        later = {"after_step": max(args.later_at, done_steps), "steps": args.later_steps,
                 "ms_per_step": lt / args.later_steps * 1e3, "value": n * args.later_steps / lt,
                 "mg_cycles_per_step": cyc / args.later_steps}
    roof = None
    dom = next((r for r in prof if r["name"] == dom_name), None)
    if dom and dom["count"] and dom["bytes"]:
        avg_s = dom["ms_total"] / dom["count"] / 1e3
        per_launch = dom["bytes"] / dom["count"]
        ach = per_launch / avg_s / 1e9
        traffic, traffic_src = traffic_from_profile(dom_name, args.config)
        roof = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": dom_name,
                "launches": dom["count"], "avg_us": round(avg_s * 1e6, 2),
                "bytes_per_launch": round(per_launch),
                "timing": "every launch timed, %d steps after the timed region" % args.kernel_steps
                if args.kernel_steps > 0 else "timed region",
                "traffic_source": traffic_src}
        reg = next((r for r in prof_region if r["name"] == dom_name), None)
        if reg and reg["count"]:
            roof["avg_us_timed_region"] = round(reg["ms_total"] / reg["count"] * 1e3, 2)
    # Poisson HBM GB/s: the survey's byte model of divgs, the z-DFT pair and
    # the multigrid (16 n1m n2m (11 n3mh + 22 cycles) per substep: the initial
    # toplevel and 22 complex-plane passes per V-cycle), over the device time
    # of divgs + poisson (phase events of every substep)
    ncell = spec.n1m * spec.n2m * spec.n3m
    n3mh = spec.n3m // 2 + 1
    cpass = (spec.n3m + 2) / spec.n3m
    pois_bytes = args.steps * 3 * 8.0 * ncell * (5 + 2 + 2 * cpass) + \
        16.0 * spec.n1m * spec.n2m * (11 * n3mh * 3 * args.steps + 22 * cycles)
    poisson = {"gbs": round(pois_bytes / (pois_ms / 1e3) / 1e9, 1), "ms_per_step": round(pois_ms / args.steps, 3),
               "bytes_per_step": round(pois_bytes / args.steps), "mg_cycles_per_step": cycles / args.steps,
               "frac": round(pois_bytes / (pois_ms / 1e3) / 1e9 / HBM_PEAK_GBS, 4)}
    step_ms = sum(r["ms_total"] for r in table)
    print("kernel table of one step (event time, algorithmic bytes):", file=sys.stderr)
    for r in table[:25]:
        gbs = r["bytes"] / (r["ms_total"] / 1e3) / 1e9 if r["ms_total"] else 0.0
        print("  %-24s %6d launches %9.3f ms %5.1f%% %8.1f GB/s" % (
            r["name"], r["count"], r["ms_total"], 100 * r["ms_total"] / step_ms, gbs), file=sys.stderr)
    sol.close()
    out = {
        "metric": "cell-updates/s per RK3 step (1 GPU)",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+f32 multigrid levels and test filter" if args.mixed else "f64",
        "data": "synthetic (seeded initial field, generated grid and IB tables)",
        "config": {"workload": WORKLOADS[args.config] + (GRID_SINH if args.grid == "sinh" else ""),
                   "cells": n, "nlev": spec.settings.nlev, "resid1": spec.settings.resid1,
                   "mg_cycles_max_last_step": st["mg_cycles_max"], "dt_last": st["dt"],
                   "case_build_s": round(t_build, 2), "mg_tiles": tune,
                   "mg_level_storage": "f32" if args.mixed else "f64"},
        "roofline": roof,
        "poisson": poisson,
        "later_window": later,
        "serve": None,
        "batch": None,
        "cpu_baseline": None,
    }
    if args.serve_leg > 0:
        try:
            out["serve"] = serve_leg(args)
        except Exception as e:  # reported beside the step, never required
            out["serve"] = {"error": str(e)[:300]}
    if args.batch_leg > 0:
        try:
            leg = argparse.Namespace(**vars(args))
            leg.batch, leg.steps, leg.warmup, leg.mixed, leg.batch_only = args.batch_leg, 5, 2, False, False
            bb = batch_bench(leg)
            out["batch"] = {q: bb[q] for q in ("value", "unit", "steps", "one_at_a_time", "concurrent_streams",
                                               "ratio_vs_one_at_a_time", "ratio_vs_concurrent_streams",
                                               "launch_sharing")}
            out["batch"]["workload"] = bb["config"]["workload"]
        except Exception as e:  # reported beside the step, never required
            out["batch"] = {"error": str(e)[:300]}
    if not args.no_cpu_baseline:
        k = min(args.cpu_steps, args.steps)
        gpu_same = {"s": sum(step_s[:k]), "mg_rows": mg_rows}
        try:
            cb = cpu_baseline(spec, case, fld_start, k, args.cpu_threads, gpu_same)
            out["cpu_baseline"] = {q: cb[q] for q in ("value", "unit", "cores", "kind", "sample", "phase_s_per_step",
                                                      "same_steps", "host")}
            out["speedup_vs_cpu_baseline"] = cb["same_steps"]["ratio"]
        except Exception as e:  # baseline is reported, never required
            out["cpu_baseline"] = {"error": str(e)[:300]}
    shutil.rmtree(work, ignore_errors=True)
    print(json.dumps(out))

if __name__ == "__main__":
    main()
