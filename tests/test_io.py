"""Input and density files through the library's host readers and writers
(ofd_read_geometry, ofd_read_recpot, ofd_write_density, ofd_read_density;
no device work, so these run without a GPU)."""
import os

import numpy as np
import pytest

from oracle import profess_oracle as O
from tests import sample_cases as S

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def io():
    from profess_amd import inputs
    return inputs


def test_reference_ion_files(io):
    """The reference's test/optIon/displaced.ion and test/optDen/Al4.ion."""
    for name, lat_a, frac in (("displaced.ion", S.OPTION_LATTICE_A, S.OPTION_FRAC),
                              ("Al4.ion", S.OPTDEN_LATTICE_A, S.OPTDEN_FRAC)):
        g = io.read_geometry(os.path.join(GOLDEN, name))
        assert np.array_equal(g["lattice"], np.asarray(lat_a).T / O.BOHR)
        assert np.array_equal(g["frac"], frac)
        assert g["type_start"] == [0, 4] and g["species"] == ["Al"]
        assert g["pseudo_files"] == ["al_HC.lda.recpot"]


def test_geometry_forms(io, tmp_path):
    """LATTICE_ABC (hexagonal), POSITIONS_CART / POSITIONS_BOHR, wrapping into
    [0,1), species grouped in order of first appearance (stable), comments."""
    a, c = 3.2, 5.2
    p = tmp_path / "hcp.ion"
    p.write_text("# hcp cell\n%%BLOCK LATTICE_ABC\n  %.1f %.1f %.1f 90 90 120\n%%END BLOCK LATTICE_ABC\n"
                 "%%BLOCK POSITIONS_CART\n  Mg 0.0 0.0 0.0\n  Al 1.0 0.5 2.0\n  # comment\n  Mg 0.5 1.0 -1.0\n"
                 "%%END BLOCK POSITIONS_CART\n%%BLOCK SPECIES_POT\n  Mg mg.recpot\n  Al al.recpot\n"
                 "%%END BLOCK SPECIES_POT\n" % (a, a, c))
    g = io.read_geometry(p)
    lat_a = np.array([[a, 0, 0], [a * np.cos(2 * np.pi / 3), a * np.sin(2 * np.pi / 3), 0], [0, 0, c]]).T
    assert np.allclose(g["lattice"], lat_a / O.BOHR, rtol=0, atol=1e-13)
    cart = np.array([[0.0, 0.0, 0.0], [0.5, 1.0, -1.0], [1.0, 0.5, 2.0]])   # Mg, Mg, Al
    want = np.linalg.solve(lat_a, cart.T).T
    want -= np.floor(want)
    assert np.allclose(g["frac"], want, atol=1e-14)
    assert g["species"] == ["Mg", "Al"] and g["type_start"] == [0, 2, 3] and g["file_index"] == [0, 2, 1]
    assert g["pseudo_files"] == ["mg.recpot", "al.recpot"]
    q = tmp_path / "bohr.ion"
    q.write_text("%%BLOCK LATTICE_CART\n %.1f 0 0\n 0 %.1f 0\n 0 0 %.1f\n%%END BLOCK LATTICE_CART\n"
                 "%%BLOCK POSITIONS_BOHR\n Al 1.0 2.0 3.0\n%%END BLOCK POSITIONS_BOHR\n" % (4.0, 5.0, 6.0))
    g = io.read_geometry(q)
    assert np.allclose(g["frac"][0], np.array([1.0, 2.0, 3.0]) * O.BOHR / np.array([4.0, 5.0, 6.0]), atol=1e-15)
    bad = tmp_path / "left.ion"
    bad.write_text("%BLOCK LATTICE_CART\n 1 0 0\n 0 0 1\n 0 1 0\n%END BLOCK LATTICE_CART\n"
                   "%BLOCK POSITIONS_FRAC\n Al 0 0 0\n%END BLOCK POSITIONS_FRAC\n")
    from profess_amd import OFDError
    with pytest.raises(OFDError):
        io.read_geometry(bad)


def test_recpot_reader_matches_reading_rules(io, tmp_path):
    """al_HC.lda.recpot as ReadPseudo reads it; a table whose last data line is
    short reads the terminator as 0 (ReadIonFile.f90:885-888)."""
    mg, v = io.read_recpot_file(os.path.join(GOLDEN, "al_HC.lda.recpot"))
    with open(os.path.join(GOLDEN, "al_HC.lda.recpot")) as fh:
        mg2, v2 = O.read_recpot_text(fh.read())
    assert mg == mg2 and np.array_equal(v, v2)
    psp = O.make_psp(mg, v)
    assert psp.charge == 3
    p = tmp_path / "short.recpot"
    p.write_text("COMMENT\nEND COMMENT\n3 5\n10.0\n1.0 2.0 3.0\n4.0 5.0\n1000\n")
    mg, v = io.read_recpot_file(p)
    assert mg == 10.0 and list(v) == [1.0, 2.0, 3.0, 4.0, 5.0, 0.0]


def test_density_round_trip_and_format(io, tmp_path):
    rng = np.random.default_rng(5)
    rho = rng.random((6, 5, 4)) * 0.03 + 1e-3
    p = tmp_path / "a.den"
    io.write_density(p, rho)
    raw = p.read_bytes()
    assert len(raw) == 27 * (4 + rho.size)
    lines = raw.decode().split("\n")
    # FORMAT(A14 I12.1) headers
    assert lines[0] == "x-dimension:  " + "%12d" % 4 and lines[3] == "# of spins:   " + "%12d" % 1
    assert all(len(ln) == 26 for ln in lines[:-1]) and lines[4].startswith("0.") and "E-0" in lines[4]
    assert np.array_equal(io.read_density(p, (4, 5, 6)), rho)   # E26.20 keeps every bit


def test_density_repeat_regrid_and_scaling(io, tmp_path):
    rng = np.random.default_rng(6)
    rho = rng.random((3, 4, 5)) + 0.5          # file grid (5, 4, 3)
    p = tmp_path / "b.den"
    io.write_density(p, rho)
    # repeated cell: periodic copies
    r2 = io.read_density(p, (10, 4, 6), repeat=(2, 1, 2))
    assert np.array_equal(r2, np.tile(rho, (2, 1, 2)))
    # a constant stays constant on any grid
    c = tmp_path / "c.den"
    io.write_density(c, np.full((3, 4, 5), 0.0268))
    assert np.allclose(io.read_density(c, (7, 9, 11)), 0.0268, rtol=1e-14, atol=0)
    # trilinear interpolation at the cell centres of the new grid: a field
    # linear in the file's x index is reproduced wherever no wrap is involved
    lin = np.broadcast_to(0.1 + 0.02 * np.arange(8)[None, None, :], (4, 4, 8)).copy()
    d = tmp_path / "d.den"
    io.write_density(d, lin)
    out = io.read_density(d, (16, 4, 4))
    fx = (8.0 / 16.0 * (1.0 + 2.0 * np.arange(16)) - 1.0) * 0.5
    inside = (fx >= 0.0) & (fx <= 7.0 - 1e-12)
    assert np.allclose(out[0, 0, inside], 0.1 + 0.02 * fx[inside], rtol=0, atol=1e-15)
    # periodic wrap between the last and the first file point
    f0 = fx[0] % 8.0
    w = f0 - np.floor(f0)
    assert abs(out[0, 0, 0] - (lin[0, 0, 7] * (1 - w) + lin[0, 0, 0] * w)) < 1e-15
    # scaled to the electron count
    lat = np.diag([7.0, 8.0, 9.0])
    s = io.read_density(p, (5, 4, 3), lattice=lat, num_electrons=12.0)
    assert abs(s.sum() * np.prod(np.diag(lat)) / s.size - 12.0) < 1e-12


def test_regrid_wrap_rounding_and_bad_inputs(io, tmp_path):
    """A cell centre whose position rounds to a tiny negative number wraps to
    the first file point (file x-length 2 onto 98 points: i = 24 sits at
    exactly 0). Malformed inputs are errors, not crashes: degenerate or
    inconsistent lattices, headers too large to hold, values the E26.20e2
    records cannot carry (magnitudes below 1e-99 are written as zero)."""
    from profess_amd._lib import OFDError
    rng = np.random.default_rng(9)
    rho = rng.random((2, 2, 2)) + 0.5
    p = tmp_path / "w.den"
    io.write_density(p, rho)
    out = io.read_density(p, (98, 2, 2))
    assert np.array_equal(out[:, :, 24], rho[:, :, 0])
    assert np.all(out >= rho.min()) and np.all(out <= rho.max())
    # lattices
    bad = tmp_path / "bad.ion"
    bad.write_text("%BLOCK LATTICE_CART\n1 0 0\n2 0 0\n0 0 1\n%END LATTICE_CART\n"
                   "%BLOCK POSITIONS_FRAC\nAl 0 0 0\n%END POSITIONS_FRAC\n")
    with pytest.raises(OFDError, match="degenerate"):
        io.read_geometry(bad)
    bad.write_text("%BLOCK LATTICE_ABC\n3 3 3\n10 10 120\n%END LATTICE_ABC\n"
                   "%BLOCK POSITIONS_FRAC\nAl 0 0 0\n%END POSITIONS_FRAC\n")
    with pytest.raises(OFDError, match="do not form a cell"):
        io.read_geometry(bad)
    # a header promising more than can be held
    big = tmp_path / "big.den"
    big.write_text("x-dimension:       100000\ny-dimension:       100000\n"
                   "z-dimension:       100000\n# of spins:             1\n")
    with pytest.raises(OFDError, match="too large"):
        io.read_density(big, (4, 4, 4))
    # values outside the record format
    tiny = np.full((1, 1, 2), 0.5)
    tiny[0, 0, 1] = 1e-120
    io.write_density(p, tiny)
    assert np.array_equal(io.read_density(p, (2, 1, 1))[0, 0], [0.5, 0.0])
    for v in (1e120, np.inf, np.nan):
        tiny[0, 0, 1] = v
        with pytest.raises(OFDError, match="not representable"):
            io.write_density(p, tiny)
