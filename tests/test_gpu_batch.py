"""GPU: batched evaluation and serving. A result must be the same alone and in
a batch (bitwise for single evaluations; converged energies identical for the
pool), and FP32-then-FP64 solves must land within 1e-5 eV/atom of FP64."""
import numpy as np
import pytest

from oracle import profess_oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pa():
    import profess_amd
    profess_amd.lib()
    return profess_amd


def _system(pa, grid, seed, precision=0, kinetic=4):
    from profess_amd import inputs
    rng = np.random.default_rng(seed)
    lat = np.array([[7.3, 0.2, 0.1], [0.3, 7.5, 0.0], [-0.2, 0.4, 7.6]]).T * (1.0 + 0.05 * rng.random())
    frac = rng.random((4, 3))
    s = pa.OFDSystem(lat, grid, 12.0, kinetic=kinetic, precision=precision)
    mg, v = inputs.synthetic_recpot(3, 1.2)
    s.build_ion_potential([(mg, v, frac)])
    rho = 12.0 / s.vol * (1.0 + 0.3 * (rng.random(s.shape) - 0.5))
    return s, rho


@pytest.mark.parametrize("grid", [(64, 32, 64), (24, 20, 18)])
def test_batched_evaluation_equals_alone(pa, grid):
    from profess_amd.devmem import DeviceArray
    systems, rhos = zip(*[_system(pa, grid, seed) for seed in range(3)])
    alone = [pa.CalculatePotentialPlus(s, r) for s, r in zip(systems, rhos)]
    dr = [DeviceArray.from_host(r) for r in rhos]
    dp = [DeviceArray(r.shape) for r in rhos]
    for _ in range(2):   # twice: the reduction counters must reset themselves
        et = pa.calpot.CalculatePotentialPlusMany(list(systems), [d.ptr for d in dr], [d.ptr for d in dp])
        for i in range(3):
            assert np.array_equal(et[i], alone[i][1])
            assert np.array_equal(dp[i].download(), alone[i][0])


def test_pool_results_same_alone_and_mixed(pa):
    from profess_amd import serve
    reqs = serve.make_requests(6, seed=3, grids=(32, 48), max_atoms=8, rate=1e6)
    mixed = serve.run_workload(reqs, mixed=False, keep_rho=True)
    assert mixed["converged"] == len(reqs)
    e_mixed = {r.rid: r.energy.copy() for r in reqs}
    for r in reqs[:3]:
        one = serve.make_requests(6, seed=3, grids=(32, 48), max_atoms=8, rate=1e6)[r.rid]
        serve.run_workload([one], mixed=False)
        assert np.array_equal(one.energy, e_mixed[r.rid]), (one.energy, e_mixed[r.rid])


def test_mixed_precision_solve_within_1e5_ev_per_atom(pa):
    from profess_amd import inputs
    c = inputs.al_fcc_case((2, 2, 2), (64, 64, 64), sigma=0.1, seed_ion=5)
    out = {}
    for prec in (0, 1):
        s = pa.OFDSystem(c["lattice"], c["grid"], c["num_electrons"], kinetic=4, precision=prec)
        mg, v = c["recpot"]
        s.build_ion_potential([(mg, v, c["frac"])])
        rho0 = np.full(s.shape, c["num_electrons"] / s.vol)
        rho, et, oc, st = pa.ProjNormRhoMinimization(s, rho0, conv_check="POT", pot_tol=1e-5,
                                                       mixed_switch=1e-3 if prec else 0.0)
        assert oc == 0
        out[prec] = (et, st)
    natom = len(c["frac"])
    de = abs(out[1][0][0] - out[0][0][0]) * O.HARTREE_TO_EV / natom
    assert de < 1e-5, de
    assert out[1][1]["fp32_evals"] > 0


def test_serving_mix_within_1e5_ev_per_atom_of_fp64(pa):
    """Sixteen requests over all seven serving grids (32 ... 128^3, fcc Al and
    hcp Mg): each converged energy of the FP32-then-FP64 pool lies within
    1e-5 eV/atom of the same request's plain FP64 solve (ofd_rho_opt alone)."""
    from profess_amd import serve
    grids = (32, 48, 64, 80, 96, 120, 128)
    def mk():
        reqs = serve.make_requests(16, seed=12, grids=grids, max_atoms=64, rate=1e6)
        for i, r in enumerate(reqs):   # every grid at least twice, both species
            g = grids[i % len(grids)]
            r.grid = (g, g, g)
            r.species = "Al" if i % 2 == 0 else "Mg"
            r.reps = serve._reps_for(r.species, g, 64)
        return reqs
    pool_reqs = mk()
    out = serve.run_workload(pool_reqs, mixed=True)
    assert out["converged"] == len(pool_reqs)
    plain = mk()
    ref = serve.run_sequential(plain)
    assert ref["converged"] == len(plain)
    worst = 0.0
    for a, b in zip(sorted(pool_reqs, key=lambda r: r.rid), sorted(plain, key=lambda r: r.rid)):
        de = abs(a.energy[0] - b.energy[0]) * O.HARTREE_TO_EV / serve.atom_count(a)
        worst = max(worst, de)
        assert de < 1e-5, (a.rid, a.grid, de)
    print("serving mix: max |E_pool - E_fp64| = %.2e eV/atom over %d requests" % (worst, len(plain)), flush=True)


def test_graph_replays_match_eager(pa, monkeypatch):
    """A density optimization replays the captured evaluation graphs many
    times with eager vector work and synchronizations between the replays;
    every result must equal the eager (uncaptured) run bit for bit."""
    res = {}
    monkeypatch.delenv("OFD_NO_GRAPH", raising=False)
    for mode in ("1", "0"):   # OFD_GRAPH=1: captured and replayed; 0: eager
        monkeypatch.setenv("OFD_GRAPH", mode)
        s, rho = _system(pa, (32, 32, 32), 11)
        r, et, oc, st = pa.ProjNormRhoMinimization(s, rho, conv_check="POT", pot_tol=1e-6)
        e2 = [pa.CalculatePotentialPlus(s, r)[1] for _ in range(3)]
        res[mode] = (et, r, st["num_evals"], e2)
        s.close()
    assert res["0"][2] == res["1"][2] and res["0"][2] > 10
    assert np.array_equal(res["0"][0], res["1"][0])
    assert np.array_equal(res["0"][1], res["1"][1])
    for a, b in zip(res["0"][3], res["1"][3]):
        assert np.array_equal(a, b)


def test_autotune_keeps_results(pa):
    """Tuned tiles change launch shapes only: results stay bitwise the same."""
    s, rho = _system(pa, (64, 64, 64), 4)
    pot0, et0 = pa.CalculatePotentialPlus(s, rho)
    r = s.autotune(reps=3)
    assert r["configs_tried"] >= 3 and r["us_best"] <= r["us_default"] * 1.0001
    pot1, et1 = pa.CalculatePotentialPlus(s, rho)
    assert np.array_equal(et1, et0)     # tuning is numerically neutral
    assert np.array_equal(pot1, pot0)
    s2, _ = _system(pa, (64, 64, 64), 5)   # a new system of the same shape starts tuned
    pot2, et2 = pa.CalculatePotentialPlus(s2, rho)
    assert np.all(np.isfinite(et2))


def test_pool_replays_after_requests_leave(pa, monkeypatch):
    """Requests admitted while others finish (the first admitted context is
    destroyed at collection while launch graphs of the same batch width are
    replayed later): every result equals the request's solve run alone."""
    from profess_amd import serve
    monkeypatch.delenv("OFD_NO_GRAPH", raising=False)
    monkeypatch.setenv("OFD_GRAPH", "1")
    reqs = serve.make_requests(6, seed=9, grids=(32,), max_atoms=8, rate=1e6)
    params = pa.calpot.opt_params(pot_tol=1e-5, conv_check="POT", fused_dots=True)   # run_workload's settings
    pool = serve.Pool((32, 32, 32), 4, 0)
    clock = iter(range(10 ** 9))
    now = lambda: float(next(clock))   # noqa: E731
    done = []
    queue = list(reqs)
    try:
        for _ in range(2):
            r = queue.pop(0)
            s, rho = serve.build_system(r, 0)
            assert pool.admit(r, s, rho, params) >= 0
        while len(done) < len(reqs):
            pool.step()
            got = pool.collect(now)
            done += got
            for _ in got:   # refill: the batch width goes back to an earlier value
                if queue:
                    r = queue.pop(0)
                    s, rho = serve.build_system(r, 0)
                    assert pool.admit(r, s, rho, params) >= 0
    finally:
        pool.close()
    assert all(r.outcome == 0 for r in done)
    for r in reqs:
        one = serve.make_requests(6, seed=9, grids=(32,), max_atoms=8, rate=1e6)[r.rid]
        serve.run_workload([one], mixed=False)
        assert np.array_equal(one.energy, r.energy), (one.energy, r.energy)


def test_server_thread_matches_lock_step(pa):
    """Pools driven by the native progress thread (one phase in flight per
    pool, no global step) give every request the same operations in the same
    order as lock-step ofd_pools_step: energies, densities and evaluation
    counts are bitwise equal, in FP64 and in the FP32-then-FP64 mode."""
    from profess_amd import serve
    for mixed in (False, True):
        res = {}
        for threaded in (False, True):
            reqs = serve.make_requests(8, seed=21, grids=(32, 48, 64), max_atoms=16, rate=2000.0)
            out = serve.run_workload(reqs, mixed=mixed, keep_rho=True, threaded=threaded, tune=False)
            assert out["converged"] == len(reqs)
            res[threaded] = {r.rid: (r.energy.copy(), r.rho.copy(), r.stats["num_evals"],
                                     r.stats["fp32_evals"]) for r in reqs}
        for rid, (e0, rho0, n0, f0) in res[False].items():
            e1, rho1, n1, f1 = res[True][rid]
            assert (n0, f0) == (n1, f1), (rid, n0, n1, f0, f1)
            assert np.array_equal(e0, e1), (rid, e0, e1)
            assert np.array_equal(rho0, rho1), rid
