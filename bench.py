#!/usr/bin/env python3
"""Benchmark: Wang-Teter energy + potential evaluations per second at 128^3 (FP64).

One step = one complete CalculatePotentialPlus evaluation (TF + vW + WT, LDA,
Hartree, ion-electron, all nine eTable entries) of fcc Al 4x4x4 (256 atoms,
Gaussian displacements sigma = 0.1 bohr, synthetic V_loc Z = 3, rc = 1.2 bohr)
on a 128^3 grid, density rho0 (1 + 0.3 f) already resident in HBM; evaluations
are enqueued back to back on the library's stream and timed with HIP events on
that stream. Prints one JSON line (see the keys below).

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--grid 128] [--batch B]
                  [--no-cpu] [--no-serving]
  python bench.py --probe [--steps K]     # K evaluations + 3 device copies of
                                          # 256 MiB, for rocprofv3 --pmc runs
  python bench.py --pmc-summary CSV...    # per-kernel FETCH_SIZE/WRITE_SIZE of
                                          # such runs, scaled by the copy kernel
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def b_wt(n1, n2, n3):
    """Algorithmic HBM bytes of one WT evaluation (alpha == beta): rho read
    twice, V_ion read once, the potential written once (4 real grids) and
    three half-complex spectra each written once and read once (6 complex)."""
    nr = n1 * n2 * n3
    nc = (n1 // 2 + 1) * n2 * n3
    return 4 * 8 * nr + 6 * 16 * nc


def b_wgc(n1, n2, n3):
    """Algorithmic HBM bytes of one second-order WGC evaluation: the four
    real grids as for WT and eight half-complex spectra (rho, sqrt rho, and
    theta^k rho^alpha, theta^k rho^beta for k = 0, 1, 2) each written once and
    read once (16 complex)."""
    nr = n1 * n2 * n3
    nc = (n1 // 2 + 1) * n2 * n3
    return 4 * 8 * nr + 16 * 16 * nc


class Hip:
    def __init__(self):
        self.L = C.CDLL("/opt/rocm/lib/libamdhip64.so")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.L.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        self.L.hipEventSynchronize.argtypes = [C.c_void_p]
        self.L.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        self.L.hipEventDestroy.argtypes = [C.c_void_p]
        self.L.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]

    def check(self, rc, what):
        if rc != 0:
            raise RuntimeError("%s failed (%d)" % (what, rc))

    def malloc(self, nbytes):
        p = C.c_void_p()
        self.check(self.L.hipMalloc(C.byref(p), nbytes), "hipMalloc")
        return p.value

    def h2d(self, dst, arr):
        self.check(self.L.hipMemcpy(C.c_void_p(dst), arr.ctypes.data_as(C.c_void_p), arr.nbytes, 1), "hipMemcpy")

    def d2h(self, arr, src):
        self.check(self.L.hipMemcpy(arr.ctypes.data_as(C.c_void_p), C.c_void_p(src), arr.nbytes, 2), "hipMemcpy")

    def event(self):
        e = C.c_void_p()
        self.check(self.L.hipEventCreate(C.byref(e)), "hipEventCreate")
        return e.value

    def record(self, e, stream):
        self.check(self.L.hipEventRecord(C.c_void_p(e), C.c_void_p(stream)), "hipEventRecord")

    def elapsed_ms(self, e0, e1):
        self.check(self.L.hipEventSynchronize(C.c_void_p(e1)), "hipEventSynchronize")
        ms = C.c_float()
        self.check(self.L.hipEventElapsedTime(C.byref(ms), C.c_void_p(e0), C.c_void_p(e1)), "hipEventElapsedTime")
        return float(ms.value)


def make_case(n):
    from profess_amd import inputs
    reps = max(1, n // 32)
    return inputs.al_fcc_case((reps, reps, reps), (n, n, n), sigma=0.1, seed_ion=7, seed_rho=11)


def time_single(pa, hip, case, steps, warmup, precision=0):
    L = pa.lib()
    s = pa.OFDSystem(case["lattice"], case["grid"], case["num_electrons"], kinetic=4, precision=precision)
    mg, vals = case["recpot"]
    s.build_ion_potential([(mg, vals, case["frac"])])
    rho = np.ascontiguousarray(case["rho"])
    d_rho = hip.malloc(rho.nbytes)
    d_pot = hip.malloc(rho.nbytes)
    d_sum = hip.malloc(9 * 8)
    hip.h2d(d_rho, rho)
    stream = L.ofd_get_stream(s._h)

    def enq():
        s._check(L.ofd_calpot_enqueue(s._h, C.c_void_p(d_rho), 1, C.c_void_p(d_pot), C.c_void_p(d_sum)))

    for _ in range(warmup):
        enq()
    s._check(L.ofd_sync(s._h))
    e0, e1 = hip.event(), hip.event()
    hip.record(e0, stream)
    for _ in range(steps):
        enq()
    hip.record(e1, stream)
    ms = hip.elapsed_ms(e0, e1)
    sums = np.zeros(9)
    hip.d2h(sums, d_sum)
    et = np.zeros(9)
    L.ofd_sums_to_etable(s._h, sums.ctypes.data_as(C.POINTER(C.c_double)), et.ctypes.data_as(C.POINTER(C.c_double)))
    if not np.all(np.isfinite(et)):
        raise RuntimeError("non-finite energies")
    vion = s.ion_potential()
    return s, ms / steps, et, vion, (d_rho, d_pot, d_sum)


def time_wgc(pa, hip, n=256, steps=10, warmup=2):
    """One second-order WGC evaluation (gamma = 2.7) of hcp Mg with 2000
    atoms on n^3, device-resident rho: microseconds per evaluation by HIP
    events on the context's stream, and the eTable."""
    from profess_amd import inputs
    L = pa.lib()
    reps = 10 if n == 256 else max(1, round(n / 25.6))
    c = inputs.mg_hcp_case((reps, reps, reps), (n, n, n), sigma=0.1, seed_ion=53)
    s = pa.OFDSystem(c["lattice"], c["grid"], c["num_electrons"], kinetic=5, gamma=2.7)
    s.build_ion_potential([(c["recpot"][0], c["recpot"][1], c["frac"])])
    p = s.params
    try:
        from tests import wgc_tables
        s.set_wgc_table(*wgc_tables.for_params(p["alpha"], p["beta"], p["gamma"]))
        table = "the reference's makeKernel2 table (tests/golden fixture)"
    except (ImportError, KeyError, OSError):
        s.build_wgc_table()
        table = "ofd_build_wgc_table"
    rho = np.ascontiguousarray(c["rho"])
    d_rho, d_pot, d_sum = hip.malloc(rho.nbytes), hip.malloc(rho.nbytes), hip.malloc(9 * 8)
    hip.h2d(d_rho, rho)
    stream = L.ofd_get_stream(s._h)

    def enq():
        s._check(L.ofd_calpot_enqueue(s._h, C.c_void_p(d_rho), 1, C.c_void_p(d_pot), C.c_void_p(d_sum)))

    for _ in range(warmup):
        enq()
    s._check(L.ofd_sync(s._h))
    e0, e1 = hip.event(), hip.event()
    hip.record(e0, stream)
    for _ in range(steps):
        enq()
    hip.record(e1, stream)
    ms = hip.elapsed_ms(e0, e1) / steps
    sums = np.zeros(9)
    hip.d2h(sums, d_sum)
    et = np.zeros(9)
    L.ofd_sums_to_etable(s._h, sums.ctypes.data_as(C.POINTER(C.c_double)), et.ctypes.data_as(C.POINTER(C.c_double)))
    for b in (d_rho, d_pot, d_sum):
        hip.L.hipFree(C.c_void_p(b))
    s.close()
    if not np.all(np.isfinite(et)):
        raise RuntimeError("non-finite WGC energies")
    b = b_wgc(n, n, n)
    return {"workload": "WGC second order (gamma 2.7) + LDA + Hartree + ion-electron energy and potential, hcp Mg "
                        "%d atoms (sigma 0.1 bohr), %d^3 grid, FP64, device-resident rho; kernel table: %s"
                        % (len(c["frac"]), n, table),
            "grid": [n, n, n], "us_per_eval": ms * 1e3, "evals_per_s": 1000.0 / ms, "steps": steps,
            "algorithmic_bytes": b, "achieved_GBs": b / (ms * 1e-3) / 1e9,
            "roofline_frac": b / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, "etable": list(et)}


def time_batched(pa, hip, s0, case, vion, nb, steps, warmup):
    L = pa.lib()
    systems = [s0]
    for _ in range(nb - 1):
        t = pa.OFDSystem(case["lattice"], case["grid"], case["num_electrons"], kinetic=4)
        t.set_ion_terms(vion, 0.0)
        systems.append(t)
    rho = np.ascontiguousarray(case["rho"])
    rp, pp = [], []
    for _ in range(nb):
        r = hip.malloc(rho.nbytes)
        hip.h2d(r, rho)
        rp.append(r)
        pp.append(hip.malloc(rho.nbytes))
    hs = (C.c_void_p * nb)(*[x._h.value for x in systems])
    ra = (C.c_void_p * nb)(*rp)
    pa_ = (C.c_void_p * nb)(*pp)
    stream = L.ofd_get_stream(s0._h)

    def run():
        rc = L.ofd_calpot_many(hs, nb, ra, pa_, 1, None)
        if rc != 0:
            raise RuntimeError(L.ofd_last_error(s0._h).decode())

    for _ in range(warmup):
        run()
    s0._check(L.ofd_sync(s0._h))
    e0, e1 = hip.event(), hip.event()
    hip.record(e0, stream)
    for _ in range(steps):
        run()
    hip.record(e1, stream)
    ms = hip.elapsed_ms(e0, e1)
    return ms / steps, systems


COPY_BYTES = 256 << 20


def probe(pa, hip, n, steps):
    """Evaluations followed by plain device-to-device copies of known size:
    the copy kernel calibrates the FETCH_SIZE / WRITE_SIZE counters."""
    case = make_case(n)
    s, ms, et, vion, bufs = time_single(pa, hip, case, steps, 3)
    a, b = hip.malloc(COPY_BYTES), hip.malloc(COPY_BYTES)
    hip.check(hip.L.hipMemset(C.c_void_p(a), 1, COPY_BYTES), "hipMemset")
    for _ in range(3):
        hip.check(hip.L.hipMemcpy(C.c_void_p(b), C.c_void_p(a), COPY_BYTES, 3), "hipMemcpy")
    hip.check(hip.L.hipDeviceSynchronize(), "hipDeviceSynchronize")
    print("probe: %d evaluations at %d^3 (%.1f us each), 3 copies of %d bytes" % (steps, n, ms * 1e3, COPY_BYTES),
          flush=True)


def pmc_summary(paths, n=128):
    """Per-kernel average of each counter (KiB per dispatch) in a rocprofv3
    counter_collection.csv, and the evaluation's HBM bytes scaled by the copy
    kernel's reading: the probe's device copies of COPY_BYTES read and write
    COPY_BYTES each, so COPY_BYTES / reading is the counter's scale factor."""
    import csv
    from collections import defaultdict
    rows = [r for path in paths for r in csv.DictReader(open(path))]
    copies = [r for r in rows if "copyBuffer" in r["Kernel_Name"]]
    gmax = max(int(r["Grid_Size"]) for r in copies)
    copy_val = defaultdict(list)
    for r in copies:
        if int(r["Grid_Size"]) == gmax:
            copy_val[r["Counter_Name"]].append(float(r["Counter_Value"]))
    vals = defaultdict(list)
    for r in rows:
        if "rocclr" in r["Kernel_Name"]:
            continue
        vals[(r["Kernel_Name"], r["Counter_Name"])].append(float(r["Counter_Value"]))
    ndisp = defaultdict(int)
    for (kn, c), v in vals.items():
        ndisp[kn] = max(ndisp[kn], len(v))
    nmax = max(ndisp.values())
    ev_kernels = sorted(kn for kn, v in ndisp.items() if v == nmax)
    lines = ["evaluation kernels: the %d kernels dispatched %d times (one per evaluation)" % (len(ev_kernels), nmax)]
    total = {}
    for c in sorted(copy_val):
        cv = sum(copy_val[c]) / len(copy_val[c])
        factor = COPY_BYTES / 1024.0 / cv
        lines.append("%s: copy of %d B reads as %.1f KiB -> scale factor %.4f" % (c, COPY_BYTES, cv, factor))
        evsum = 0.0
        for kn in ev_kernels:
            v = vals.get((kn, c))
            if not v:
                continue
            avg = sum(v) / len(v)
            evsum += avg
            lines.append("  %-72s %11.1f KiB/dispatch  %12.0f B scaled" % (kn[:72], avg, avg * 1024 * factor))
        total[c] = evsum * 1024 * factor
        lines.append("  %s per evaluation, scaled: %.0f B" % (c, total[c]))
    if "FETCH_SIZE" in total and "WRITE_SIZE" in total:
        lines.append("traffic_bytes_per_eval = %.0f" % (total["FETCH_SIZE"] + total["WRITE_SIZE"]))
    print("\n".join(lines))


def cpu_model():
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_cores():
    """Cores this process may use: the affinity mask, capped by OMP_NUM_THREADS
    when that is set (the CPU share given to the job)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return n


def cpu_baseline(case, vion, budget_s=10.0, threads=1):
    """The CPU restatement of the WT evaluation (oracle/cpu/ofd_cpu.cpp: the
    reference's dataflow of four forward and four inverse 3D transforms, FP64,
    its own mixed-radix FFT, OpenMP) on the same inputs and `threads` host
    threads, for about budget_s seconds. Falls back to the numpy oracle (one
    thread) when the library is not built."""
    from oracle import profess_oracle as O
    n1, n2, n3 = case["grid"]
    cell = O.Cell(case["lattice"], n1, n2, n3)
    f = O.Functional(kinetic=4).resolve(case["num_electrons"], cell.vol)
    rho = case["rho"]
    try:
        from oracle import cpubuild
        ce = cpubuild.CpuEval(case["lattice"], case["grid"], f.rho0)

        def run():
            return ce.calpot(rho, vion, 0.0, True, threads)[1]
        kind = "C++ restatement (oracle/cpu/ofd_cpu.cpp), %d thread%s" % (threads, "" if threads == 1 else "s")
    except OSError:
        sysm = O.System(cell, f, vion)

        def run():
            return O.calculate_potential_plus(rho, sysm, True, True)[1]
        kind = "numpy oracle, 1 thread (oracle/cpu not built)"
        threads = 1
    run()   # warm-up (pages, plans)
    t0 = time.perf_counter()
    n = 0
    et = None
    while True:
        et = run()
        n += 1
        if time.perf_counter() - t0 > budget_s or n >= 200:
            break
    dt = time.perf_counter() - t0
    return n / dt, n, dt, et, kind, threads


def read_traffic():
    """HBM bytes per evaluation measured by rocprofv3 --pmc (FETCH_SIZE and
    WRITE_SIZE, scaled by the copy-kernel factor), from profiles/."""
    p = os.path.join(ROOT, "profiles", "wt128_traffic.txt")
    if not os.path.exists(p):
        return None
    for ln in open(p):
        if ln.startswith("traffic_bytes_per_eval"):
            try:
                return float(ln.split("=")[1].split()[0])
            except ValueError:
                return None
    return None


SERVE_TOL_EV_PER_ATOM = 1e-5


def serving_sample(pa, nreq=512, seed=4):
    """Batched serving on the mixed-grid workload (fcc Al / hcp Mg cells of
    4-500 atoms on 32^3 ... 128^3 grids, each a full NTN solve to a potential
    norm of 1e-5):
    - a burst (Poisson arrivals at 1000/s, above the pool's capacity): the
      FP32-then-FP64 pool and, on the same requests, the plain FP64 drop-in
      path (ProjNormRhoMinimization one request at a time); solves/s, p50/p99
      latency (arrival to result, so mostly queueing here), their ratio, and
      the largest |E_pool - E_fp64| per atom over all requests, which must
      stay below 1e-5 eV/atom (the run fails otherwise);
    - the same requests arriving at 60/s, below capacity, through the pool:
      p50/p99 are then the service latency of a loaded server."""
    from profess_amd import inputs, serve
    grids = (32, 48, 64, 80, 96, 120, 128)
    mk = lambda rate: serve.make_requests(nreq, seed=seed, grids=grids, max_atoms=500, rate=rate)   # noqa: E731
    pool_reqs = mk(1000.0)
    pool = serve.run_workload(pool_reqs, mixed=True)
    plain_reqs = mk(1000.0)
    plain = serve.run_sequential(plain_reqs)
    de = []
    for a, b in zip(sorted(pool_reqs, key=lambda r: r.rid), sorted(plain_reqs, key=lambda r: r.rid)):
        de.append(abs(a.energy[0] - b.energy[0]) * inputs.HARTREE_TO_EV / serve.atom_count(a))
    out = {k: pool[k] for k in ("requests", "converged", "solves_per_s", "p50_ms", "p99_ms", "grids")}
    out["arrival_rate_per_s"] = 1000.0
    out["mixed_precision"] = True
    out["max_dE_eV_per_atom_vs_fp64"] = max(de)
    out["plain_fp64_sequential"] = {k: plain[k] for k in ("converged", "solves_per_s", "p50_ms", "p99_ms")}
    out["speedup_vs_plain_fp64"] = pool["solves_per_s"] / plain["solves_per_s"]
    steady = serve.run_workload(mk(60.0), mixed=True)
    out["below_capacity"] = {"arrival_rate_per_s": 60.0, "converged": steady["converged"],
                             "solves_per_s": steady["solves_per_s"], "p50_ms": steady["p50_ms"],
                             "p99_ms": steady["p99_ms"]}
    if max(de) >= SERVE_TOL_EV_PER_ATOM:
        raise RuntimeError("serving: pool energy differs from the FP64 solve by %.3e eV/atom" % max(de))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--grid", type=int, default=128)
    ap.add_argument("--batch", type=int, default=4)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-wgc", action="store_true", help="skip the 256^3 WGC evaluation leg")
    ap.add_argument("--wgc-only", action="store_true", help="only the 256^3 WGC evaluation leg")
    ap.add_argument("--probe", action="store_true")
    ap.add_argument("--serve", type=int, default=0, metavar="NREQ",
                    help="only the mixed-grid serving workload with NREQ requests (pool alone, no plain run)")
    ap.add_argument("--fp32", action="store_true", help="FP32 FFT passes (mixed-precision form), reported apart")
    ap.add_argument("--pmc-summary", nargs="+", default=None)
    args = ap.parse_args()
    if args.pmc_summary:
        pmc_summary(args.pmc_summary, args.grid)
        return
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (the evaluation is a single-GPU path; "
              "multi-GPU sharding is a design note, not a measured configuration)", file=sys.stderr)
        sys.exit(2)

    import profess_amd as pa
    pa.lib()
    hip = Hip()
    n = args.grid
    if args.probe:
        probe(pa, hip, n, args.steps)
        return
    if args.wgc_only:
        print(json.dumps({"wgc256": time_wgc(pa, hip, steps=args.steps, warmup=args.warmup)}), flush=True)
        return
    case = make_case(n)
    if args.serve:
        from profess_amd import serve
        reqs = serve.make_requests(args.serve, seed=4, grids=(32, 48, 64, 80, 96, 120, 128), max_atoms=500,
                                   rate=1000.0)
        out = serve.run_workload(reqs, mixed=True)
        out["evals_mean"] = sum(q.stats["num_evals"] for q in reqs) / len(reqs)
        out["fp32_evals_mean"] = sum(q.stats["fp32_evals"] for q in reqs) / len(reqs)
        print(json.dumps(out), flush=True)
        return
    if args.fp32:
        _, ms32, et32, _, _ = time_single(pa, hip, case, args.steps, args.warmup, precision=1)
        print(json.dumps({"fp32_passes": True, "grid": n, "us_per_eval": ms32 * 1e3, "evals_per_s": 1000.0 / ms32,
                          "etable": list(et32)}), flush=True)
        return
    s, ms1, et, vion, bufs = time_single(pa, hip, case, args.steps, args.warmup)
    evals_s = 1000.0 / ms1
    bw = b_wt(n, n, n)
    achieved = bw * evals_s / 1e9
    out = {
        "metric": "WT energy+potential evals/s at 128\u00b3; batched converged solves/s + p99 latency",
        "value": evals_s,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms1,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "WT (TF+vW+WT a=b=5/6) + LDA + Hartree + ion-electron energy and potential, "
                               "fcc Al %dx%dx%d (%d atoms, sigma 0.1 bohr), %d^3 grid, FP64, device-resident rho"
                               % (max(1, n // 32), max(1, n // 32), max(1, n // 32), len(case["frac"]), n),
                   "grid": [n, n, n], "atoms": len(case["frac"]), "batch": 1},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": read_traffic(),
                     "kernel": "one evaluation (5 fused passes + partial-sum finalize)",
                     "algorithmic_bytes": bw},
    }
    if args.batch > 1:
        # the batch sizes 2, 4, 8 (and --batch), each in one launch set; the line's
        # entry is --batch, the others are listed under "by_batch"
        sizes = sorted({2, 4, 8, args.batch})
        rates = {}
        for b in sizes:
            msb, _ = time_batched(pa, hip, s, case, vion, b, max(10, args.steps // b), args.warmup)
            rates[b] = (msb, b * 1000.0 / msb)
        msb, rate = rates[args.batch]
        out["batched"] = {"batch": args.batch, "evals_per_s": rate, "ms_per_batch": msb,
                          "roofline_frac": bw * rate / 1e9 / HBM_PEAK_GBS,
                          "by_batch": {str(b): {"evals_per_s": r, "roofline_frac": bw * r / 1e9 / HBM_PEAK_GBS}
                                       for b, (_, r) in rates.items()}}
    if not args.no_wgc:
        try:
            out["wgc256"] = time_wgc(pa, hip)
        except Exception as e:   # reported, never hidden
            out["wgc256"] = {"error": str(e)}
    if not args.no_serving:
        try:
            out["serving"] = serving_sample(pa)
        except Exception as e:   # reported, never hidden
            out["serving"] = {"error": str(e)}
    if not args.no_cpu:
        rate, nev, dt, et_cpu, kind, _ = cpu_baseline(case, vion)
        ncores = host_cores()
        out["cpu_baseline"] = {"value": rate, "unit": "evals/s", "cores": 1, "kind": "port",
                               "sample": "%d full WT evaluations of the same %d^3 workload in %.1f s by the %s"
                                         % (nev, n, dt, kind),
                               "cpu_model": cpu_model(), "host_cores_available": ncores}
        rel = max(abs(et[k] - et_cpu[k]) / max(abs(et_cpu[k]), 1e-300) for k in range(9) if abs(et_cpu[k]) > 1e-12)
        out["cpu_baseline"]["max_term_rel_diff_vs_gpu"] = rel
        if ncores > 1:
            rate_n, nev_n, dt_n, _, kind_n, th = cpu_baseline(case, vion, budget_s=5.0, threads=ncores)
            out["cpu_baseline"]["all_cores"] = {
                "value": rate_n, "cores": th,
                "sample": "%d evaluations in %.1f s by the %s" % (nev_n, dt_n, kind_n)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
