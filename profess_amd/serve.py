"""Continuous-batching server for density optimizations of uneven cells/grids.

Requests (fcc Al / hcp Mg supercells, synthetic V_loc, uniform start density)
arrive over time. A request is admitted as soon as (a) it has arrived, (b) the
pool for its grid shape has a free slot and (c) its device footprint
(``ofd_request_bytes``) fits in the HBM budget; its rho, ion potential, WGC
kernels and optimizer state then live on the device until it converges. A
native progress thread (``Server``) drives every pool on its own stream: one
phase at a time per pool, each phase one batched launch set per (optimizer
phase, precision) over all of the pool's active requests, so requests that
arrive while others run share those launches, and pools of different grids
never wait for each other. Latency is arrival -> converged result.
"""
from __future__ import annotations

import ctypes as C
import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import inputs
from ._lib import OFDError, dptr, lib, ofd_opt_stats
from .calpot import OFDSystem, opt_params


@dataclass
class Request:
    rid: int
    arrival: float
    species: str          # "Al" (fcc) or "Mg" (hcp)
    reps: tuple
    grid: tuple
    sigma: float
    seed: int
    # filled by the server
    admitted: float = -1.0
    finished: float = -1.0
    energy: np.ndarray = field(default=None, repr=False)
    outcome: int = -1
    stats: dict = field(default=None)
    rho: np.ndarray = field(default=None, repr=False)


def _reps_for(species, n, max_atoms):
    if species == "Al":
        r = max(1, int(round(n * 0.24 / inputs.AL_A)))
        while 4 * r ** 3 > max_atoms and r > 1:
            r -= 1
        return (r, r, r)
    r = max(1, int(round(n * 0.24 / inputs.MG_A)))
    rc = max(1, int(round(n * 0.24 / inputs.MG_C)))
    while 2 * r * r * rc > max_atoms and r > 1:
        r -= 1
    return (r, r, rc)


def atom_count(req: Request):
    """Atoms in the request's supercell (4 per fcc cell, 2 per hcp cell)."""
    r = req.reps
    return (4 if req.species == "Al" else 2) * r[0] * r[1] * r[2]


def make_requests(n, seed=0, grids=(32, 48, 64, 80, 96, 120, 128), max_atoms=500, rate=50.0):
    """Poisson arrivals at `rate` per second; species, grid and displacement
    amplitude (0.05-0.15 bohr) drawn from a seeded generator."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(n):
        t += rng.exponential(1.0 / rate)
        sp = "Al" if rng.random() < 0.5 else "Mg"
        g = int(rng.choice(grids))
        out.append(Request(i, t, sp, _reps_for(sp, g, max_atoms), (g, g, g), float(rng.uniform(0.05, 0.15)),
                           int(rng.integers(1 << 30))))
    return out


def build_system(req: Request, precision=1, with_rho=True):
    """The request's system (cell, ions, ionic potential on the device) and,
    with_rho, its uniform starting density on the host."""
    if req.species == "Al":
        lat, frac = inputs.fcc_supercell(inputs.AL_A, req.reps)
        Z, rc = inputs.AL_Z, inputs.AL_RC
    else:
        lat, frac = inputs.hcp_supercell(inputs.MG_A, inputs.MG_C, req.reps)
        Z, rc = inputs.MG_Z, inputs.MG_RC
    frac = inputs.displace(lat, frac, req.sigma, req.seed)
    ne = Z * frac.shape[0]
    s = OFDSystem(lat, req.grid, ne, kinetic=4, precision=precision)
    mg, vals = inputs.synthetic_recpot(Z, rc)
    s.build_ion_potential([(mg, vals, frac)])
    rho = np.full(s.shape, ne / s.vol) if with_rho else None
    return s, rho


class Pool:
    def __init__(self, grid, capacity, precision):
        L = lib()
        h = C.c_void_p()
        if L.ofd_pool_create(grid[0], grid[1], grid[2], capacity, 4, precision, C.byref(h)) != 0:
            raise OFDError("pool create failed")
        self.h = h
        self.grid = grid
        self.capacity = capacity
        self.slots = {}   # slot -> Request

    def close(self):
        if self.h is not None:
            lib().ofd_pool_destroy(self.h)
            self.h = None

    def admit(self, req, system, rho, params):
        """Admit `req` with its system and starting density (None: uniform).
        Returns the slot, or -1 when the pool did not take it."""
        slot = self.admit_native(system, rho, params)
        if slot >= 0:
            self.slots[slot] = req
        return slot

    def admit_native(self, system, rho, params):
        """The native admission alone (safe from a worker thread); the caller
        records the slot in `slots`. Returns -1 when the pool is full."""
        r = None if rho is None else np.ascontiguousarray(rho)
        slot = lib().ofd_pool_admit(self.h, system._h, dptr(r), C.byref(params))
        if slot == -2:
            return -1
        if slot < 0:
            raise OFDError("admission failed: " + lib().ofd_pool_last_error(self.h).decode())
        system.detach()
        return slot

    def step(self):
        n = lib().ofd_pool_step(self.h)
        if n < 0:
            raise OFDError(lib().ofd_pool_last_error(self.h).decode())
        return n

    def collect(self, now, want_rho=True):
        L = lib()
        done = []
        fin = (C.c_int * self.capacity)()
        nfin = L.ofd_pool_finished(self.h, fin, self.capacity)
        if nfin <= 0:
            return done
        for slot in list(fin[:nfin]):
            req = self.slots.get(slot)
            if req is None:   # admitted by a worker, not registered yet
                continue
            n = self.grid[0] * self.grid[1] * self.grid[2]
            rho = np.empty(n) if want_rho else None
            et = np.zeros(9)
            oc = C.c_int(-1)
            st = ofd_opt_stats()
            got = L.ofd_pool_collect(self.h, slot, dptr(rho), dptr(et), C.byref(oc), C.byref(st), None)
            if got == 1:
                req.finished = now()
                req.energy = et
                req.outcome = oc.value
                req.stats = {k: getattr(st, k) for k, _ in st._fields_}
                req.rho = rho.reshape(self.grid[2], self.grid[1], self.grid[0]) if want_rho else None
                del self.slots[slot]
                done.append(req)
        return done


def step_pools(pools):
    """One step of every pool at once (ofd_pools_step: concurrent streams,
    one host wait per round)."""
    arr = (C.c_void_p * len(pools))(*[p.h.value for p in pools])
    n = lib().ofd_pools_step(arr, len(pools))
    if n < 0:
        raise OFDError(lib().ofd_pool_last_error(pools[0].h).decode())
    return n


class Server:
    """A native progress thread (ofd_server) that drives the pools added to it
    asynchronously: every pool keeps one phase in flight on its own stream and
    launches the next as soon as it completes, while this thread admits and
    collects requests."""

    def __init__(self):
        h = C.c_void_p()
        if lib().ofd_server_create(C.byref(h)) != 0:
            raise OFDError("server create failed")
        self.h = h

    def add(self, pool):
        if lib().ofd_server_add_pool(self.h, pool.h) != 0:
            raise OFDError("pool cannot join the server")

    def wait(self, timeout):
        """Requests finished since the previous call (0 after `timeout` s)."""
        k = lib().ofd_server_wait(self.h, float(timeout))
        if k < 0:
            raise OFDError(lib().ofd_server_last_error(self.h).decode())
        return k

    def close(self):
        if self.h is not None:
            lib().ofd_server_destroy(self.h)
            self.h = None


def capacity_for(grid):
    """Slots of a pool: 64 concurrent requests, so one batched launch set
    fills the device even at 32^3 and a burst of arrivals on one grid runs in
    few generations; memory is bounded separately by the HBM budget at
    admission."""
    return 64


def run_workload(requests, mixed=True, hbm_budget=160e9, pot_tol=1e-5, mixed_switch=1e-3, verbose=False,
                 keep_rho=False, fused_dots=True, tune=True, threaded=True, builders=8):
    """Serve `requests` (arrival times in seconds from the start) to completion.
    Returns throughput and latency statistics. fused_dots: the optimizer's
    inner products from the evaluation's sums and the CG recurrences (one
    host round trip per inner iteration). threaded: the pools are driven by
    a native progress thread (Server) and `builders` worker threads build and
    admit arriving requests (system, ionic potential, uniform start density
    on the device) while the device works; otherwise this loop builds, admits
    and steps the pools (ofd_pools_step) in lock step."""
    precision = 1 if mixed else 0
    params = opt_params(pot_tol=pot_tol, conv_check="POT", mixed_switch=mixed_switch if mixed else 0.0,
                        fused_dots=fused_dots)
    run = _serve_threaded if threaded else _serve_lockstep
    t0 = time.perf_counter()
    finished, host = run(requests, precision, params, hbm_budget, keep_rho, tune, verbose, t0, builders)
    makespan = max((r.finished for r in finished), default=0.0)
    lat = np.array([r.finished - r.arrival for r in finished])
    ok = sum(1 for r in finished if r.outcome == 0)
    return {"requests": len(requests), "converged": ok, "makespan_s": makespan,
            "solves_per_s": len(requests) / makespan,
            "p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "mixed_precision": bool(mixed),
            "grids": sorted({r.grid[0] for r in requests}),
            "host_s": {k: round(v, 4) for k, v in host.items()}}


def _new_pool(req, precision, tune):
    """A pool for the grid of `req`: its pass tiles are tuned first on the
    request's own system (the pool copies the tuned tiles)."""
    s, _ = build_system(req, precision, with_rho=False)
    if tune:
        s.autotune(reps=3)
    s.close()
    return Pool(req.grid, capacity_for(req.grid), precision)


def _finish(req, keep_rho, verbose, finished):
    if not keep_rho:
        req.rho = None
    finished.append(req)
    if verbose:
        print("req %d grid %s done in %.3f s (E = %.8f Ha, outcome %d)"
              % (req.rid, req.grid, req.finished - req.arrival, req.energy[0], req.outcome), flush=True)


def _serve_threaded(requests, precision, params, hbm_budget, keep_rho, tune, verbose, t0, builders):
    from concurrent.futures import ThreadPoolExecutor
    now = lambda: time.perf_counter() - t0   # noqa: E731
    tic = time.perf_counter
    host = {"pool_setup_s": 0.0, "submit_s": 0.0, "wait_s": 0.0, "collect_s": 0.0}
    pending = sorted(requests, key=lambda r: r.arrival)
    pools, reserved = {}, {}
    in_use = 0.0
    finished, waiting = [], []
    building = {}   # future -> (req, pool, need)
    qi = 0

    def build_admit(req, pool):
        s, _ = build_system(req, precision, with_rho=False)
        try:
            slot = pool.admit_native(s, None, params)
        except Exception:
            s.close()
            raise
        if slot < 0:
            s.close()
        else:
            req.admitted = now()
        return slot

    server = Server()
    ex = ThreadPoolExecutor(max_workers=builders)
    try:
        while len(finished) < len(requests):
            t = now()
            while qi < len(pending) and pending[qi].arrival <= t:
                waiting.append(pending[qi])
                qi += 1
            t1 = tic()
            still = []
            for req in waiting:   # FIFO among arrived requests that fit
                g = req.grid
                if g not in pools:
                    t2 = tic()
                    pools[g] = _new_pool(req, precision, tune)
                    reserved[g] = 0
                    server.add(pools[g])
                    host["pool_setup_s"] += tic() - t2
                pool = pools[g]
                need = float(lib().ofd_request_bytes(g[0], g[1], g[2], 4, precision))
                if len(pool.slots) + reserved[g] < pool.capacity and in_use + need <= hbm_budget:
                    reserved[g] += 1
                    in_use += need
                    building[ex.submit(build_admit, req, pool)] = (req, pool, need)
                else:
                    still.append(req)
            waiting = still
            for fut in [f for f in building if f.done()]:
                req, pool, need = building.pop(fut)
                reserved[req.grid] -= 1
                slot = fut.result()
                if slot >= 0:
                    pool.slots[slot] = req
                else:
                    in_use -= need
                    waiting.insert(0, req)
            host["submit_s"] += tic() - t1
            # wake on a finished request, in time for the next arrival, or soon
            # when admissions are under way
            nxt = pending[qi].arrival - now() if qi < len(pending) else 0.005
            tmo = 0.0005 if building else max(0.0, min(0.005, nxt))
            t1 = tic()
            server.wait(tmo)
            host["wait_s"] += tic() - t1
            t1 = tic()
            for g, pool in pools.items():
                for req in pool.collect(now, want_rho=keep_rho):
                    in_use -= float(lib().ofd_request_bytes(g[0], g[1], g[2], 4, precision))
                    _finish(req, keep_rho, verbose, finished)
            host["collect_s"] += tic() - t1
    finally:
        ex.shutdown(wait=True)
        server.close()
        for p in pools.values():
            p.close()
    return finished, host


def _serve_lockstep(requests, precision, params, hbm_budget, keep_rho, tune, verbose, t0, builders):
    now = lambda: time.perf_counter() - t0   # noqa: E731
    tic = time.perf_counter
    host = {"pool_setup_s": 0.0, "build_s": 0.0, "admit_s": 0.0, "step_s": 0.0, "collect_s": 0.0}
    pending = sorted(requests, key=lambda r: r.arrival)
    pools = {}
    in_use = 0.0
    finished, waiting = [], []
    qi = 0
    try:
        while len(finished) < len(requests):
            t = now()
            while qi < len(pending) and pending[qi].arrival <= t:
                waiting.append(pending[qi])
                qi += 1
            still = []
            for req in waiting:   # FIFO among arrived requests that fit
                g = req.grid
                if g not in pools:
                    t1 = tic()
                    pools[g] = _new_pool(req, precision, tune)
                    host["pool_setup_s"] += tic() - t1
                pool = pools[g]
                need = float(lib().ofd_request_bytes(g[0], g[1], g[2], 4, precision))
                if len(pool.slots) < pool.capacity and in_use + need <= hbm_budget:
                    t1 = tic()
                    s, rho = build_system(req, precision)
                    t2 = tic()
                    ok = pool.admit(req, s, rho, params) >= 0
                    host["build_s"] += t2 - t1
                    host["admit_s"] += tic() - t2
                    if ok:
                        req.admitted = now()
                        in_use += need
                        continue
                    s.close()
                still.append(req)
            waiting = still
            busy = [pool for pool in pools.values() if pool.slots]
            t1 = tic()
            active = step_pools(busy) if busy else 0
            host["step_s"] += tic() - t1
            t1 = tic()
            for g, pool in pools.items():
                for req in pool.collect(now, want_rho=keep_rho):
                    in_use -= float(lib().ofd_request_bytes(g[0], g[1], g[2], 4, precision))
                    _finish(req, keep_rho, verbose, finished)
            host["collect_s"] += tic() - t1
            if active == 0 and not waiting and qi < len(pending):
                time.sleep(max(0.0, min(0.01, pending[qi].arrival - now())))
    finally:
        for p in pools.values():
            p.close()
    return finished, host


def run_sequential(requests, pot_tol=1e-5, precision=0, mixed_switch=0.0, keep_rho=False):
    """The plain drop-in path as a server: each request is solved alone with
    ProjNormRhoMinimization (ofd_rho_opt), first come first served, one at a
    time. Returns the same statistics as run_workload."""
    from .calpot import ProjNormRhoMinimization
    pending = sorted(requests, key=lambda r: r.arrival)
    t0 = time.perf_counter()
    now = lambda: time.perf_counter() - t0   # noqa: E731
    for req in pending:
        wait = req.arrival - now()
        if wait > 0:
            time.sleep(wait)
        s, rho = build_system(req, precision)
        req.admitted = now()
        r, et, oc, st = ProjNormRhoMinimization(s, rho, conv_check="POT", pot_tol=pot_tol, mixed_switch=mixed_switch)
        req.finished = now()
        req.energy = et
        req.outcome = oc
        req.stats = st
        req.rho = r if keep_rho else None
        s.close()
    makespan = now()
    lat = np.array([r.finished - r.arrival for r in pending])
    ok = sum(1 for r in pending if r.outcome == 0)
    return {"requests": len(requests), "converged": ok, "makespan_s": makespan,
            "solves_per_s": len(requests) / makespan,
            "p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "mixed_precision": precision == 1, "grids": sorted({r.grid[0] for r in requests})}
