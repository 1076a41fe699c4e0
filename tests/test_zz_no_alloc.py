"""GPU: ofd_calpot_noalloc, the evaluation of ofd_calpot_enqueue with its
device scratch in a static device array (no allocation per call), against
ofd_calpot_device on the same inputs: two calls in a row, the second on a
smaller grid, bitwise the same potential and the same eTable; a grid whose
scratch exceeds the array's limit is refused with an error."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pa():
    import profess_amd
    profess_amd.lib()
    return profess_amd


def _system(pa, grid, seed):
    from profess_amd import inputs
    rng = np.random.default_rng(seed)
    lat = np.array([[7.3, 0.2, 0.1], [0.3, 7.5, 0.0], [-0.2, 0.4, 7.6]]).T * (grid[0] / 32.0)
    frac = rng.random((6, 3))
    s = pa.OFDSystem(lat, grid, 18.0, kinetic=4)
    mg, v = inputs.synthetic_recpot(3, 1.2)
    s.build_ion_potential([(mg, v, frac)], e_ionion=-1.5)
    rho = 18.0 / s.vol * (1.0 + 0.3 * (rng.random(s.shape) - 0.5))
    return s, rho


def test_noalloc_matches_calpot_device(pa):
    from profess_amd.devmem import DeviceArray
    L = pa.lib()
    for grid, seed in (((128, 128, 128), 1), ((64, 48, 40), 2)):
        print("noalloc grid %s" % (grid,), flush=True)
        s, rho = _system(pa, grid, seed)
        d_rho = DeviceArray.from_host(rho)
        d_ref = DeviceArray(rho.shape)
        d_pot = DeviceArray(rho.shape)
        d_et = DeviceArray((9,))
        for opt in (1, 0):
            et_ref = np.zeros(9)
            s._check(L.ofd_calpot_device(s._h, C.c_void_p(d_rho.ptr), opt, C.c_void_p(d_ref.ptr),
                                         et_ref.ctypes.data_as(C.POINTER(C.c_double))))
            s._check(L.ofd_calpot_noalloc(s._h, C.c_void_p(d_rho.ptr), opt, C.c_void_p(d_pot.ptr),
                                          C.c_void_p(d_et.ptr)))
            s._check(L.ofd_sync(s._h))
            assert np.array_equal(d_pot.download(), d_ref.download())
            et = d_et.download()
            np.testing.assert_allclose(et, et_ref, rtol=1e-13, atol=1e-14)
        for d in (d_rho, d_ref, d_pot, d_et):
            d.free()
        s.close()


def test_noalloc_refuses_grids_above_its_limit(pa):
    L = pa.lib()
    s, rho = _system(pa, (192, 192, 192), 3)
    assert L.ofd_noalloc_limit_bytes() == 64 << 20
    from profess_amd.devmem import DeviceArray
    d_rho = DeviceArray.from_host(rho)
    d_pot = DeviceArray(rho.shape)
    rc = L.ofd_calpot_noalloc(s._h, C.c_void_p(d_rho.ptr), 1, C.c_void_p(d_pot.ptr), None)
    assert rc != 0
    assert b"limit" in L.ofd_last_error(s._h)
    d_rho.free()
    d_pot.free()
    s.close()
