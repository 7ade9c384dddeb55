"""File formats around the analysis, through the library's host-side C-ABI
(no GPU): GridObservation's text format (GridObservation.hpp:401-476),
SimpleState's text format (SimpleState.hpp:162-188, 248-263) and the FP64
state sidecar -- on the reference's tutorial files and on hand-made files
whose expected contents are written out in the test."""
import ctypes as C
import os
import struct

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TUT = os.path.join(ROOT, "tests", "golden", "tutorial")

HEADER = ("Time    Z    Y    X        ZNU     XLONG_U     XLAT_U      Value\n"
          "-----------------------------------------------------------------\n")


@pytest.fixture(scope="module")
def mda():
    import metada_amd
    return metada_amd


def test_tutorial_observation_file(mda):
    obs = mda.Observation.from_grid_file(os.path.join(TUT, "obs_format.txt"), error=0.1)
    assert obs.size() == 28
    # first data line of the file: "0  1  3  4  ...  0.2052"
    assert (obs.loc_a[0], obs.loc_b[0], obs.loc_c[0]) == (4.0, 3.0, 1.0)
    assert obs.values[0] == 0.2052
    assert np.all(obs.variances == 0.1 * 0.1)
    assert obs.coord == mda._lib.COORD_GRID


def test_observation_file_rules(mda, tmp_path):
    f = tmp_path / "obs.txt"
    f.write_text("a preamble line that is skipped\n"
                 "0 9 9 9 0.5 1.0 2.0 7.0\n"            # before the header: skipped
                 + HEADER +
                 "0 1 2 3 0.25 100.5 -30.25 1.5\n"
                 "\n"
                 "0 2 3 4 0.75 101.5 -31.25 -999\n"     # missing value: dropped
                 "garbage line\n"                       # not 8 fields: skipped
                 "0 0 5 6 0.125 102.5 -32.25 -2.5\n")
    obs = mda.Observation.from_grid_file(str(f), error=0.5)
    assert obs.size() == 2
    np.testing.assert_array_equal(obs.values, [1.5, -2.5])
    np.testing.assert_array_equal(obs.variances, [0.25, 0.25])
    np.testing.assert_array_equal(obs.loc_a, [3, 6])   # X
    np.testing.assert_array_equal(obs.loc_b, [2, 5])   # Y
    np.testing.assert_array_equal(obs.loc_c, [1, 0])   # Z
    geo = mda.Observation.from_grid_file(str(f), error=0.5, coordinate="geographic")
    np.testing.assert_array_equal(geo.loc_a, [-30.25, -32.25])  # XLAT_U
    np.testing.assert_array_equal(geo.loc_b, [100.5, 102.5])    # XLONG_U
    np.testing.assert_array_equal(geo.loc_c, [0.25, 0.125])     # ZNU
    assert geo.coord == mda._lib.COORD_GEOGRAPHIC
    # another missing-value marker keeps the -999 line
    assert mda.Observation.from_grid_file(str(f), error=0.5, missing_value=-1.0).size() == 3


def test_observation_file_errors(mda, tmp_path):
    with pytest.raises(RuntimeError, match="Could not open observation file"):
        mda.Observation.from_grid_file(str(tmp_path / "absent.txt"), error=0.1)
    f = tmp_path / "empty.txt"
    f.write_text(HEADER + "0 1 1 1 0.5 1.0 2.0 -999\n")
    with pytest.raises(RuntimeError, match="No valid observations"):
        mda.Observation.from_grid_file(str(f), error=0.1)
    with pytest.raises(RuntimeError, match="Unsupported coordinate system"):
        mda.Observation.from_grid_file(str(f), error=0.1, coordinate="cartesian")


def test_state_text_reader_on_tutorial_state(mda):
    x = mda.read_simple_state(os.path.join(TUT, "ens_1.txt"))
    assert x.shape == (648,)
    np.testing.assert_array_equal(x, np.loadtxt(os.path.join(TUT, "ens_1.txt")).ravel())


def test_state_text_format(mda, tmp_path):
    """saveToFile: rows of nx values, '%12.6f' joined by one space."""
    from metada_amd.analysis import _write_state
    v = np.array([1.0, -2.5, 1234.5678905, 0.0, -0.0000004, 3e5])
    p = tmp_path / "s.txt"
    _write_state(str(p), v, 3, 2)
    assert p.read_text() == ("    1.000000    -2.500000  1234.567890\n"
                             "    0.000000    -0.000000 300000.000000\n")
    np.testing.assert_allclose(mda.read_simple_state(str(p)), v, atol=5e-7)


def test_state_sidecar_exact_round_trip(mda, tmp_path):
    rng = np.random.default_rng(5)
    v = np.concatenate([rng.standard_normal(60) * 10.0 ** rng.integers(-300, 300, 60),
                        [0.0, -0.0, 5e-324, np.pi, -1e308, 1.0 / 3.0]])
    p = tmp_path / "s.f64"
    mda.write_state_sidecar(str(p), v, 11, 3, 2)
    raw = p.read_bytes()
    assert len(raw) == 64 + 8 * v.size
    assert raw[:8] == b"MDASTATE"
    assert struct.unpack("<4q", raw[8:40]) == (1, 11, 3, 2)
    assert raw[40:64] == bytes(24)
    back, nx, ny, nlev = mda.read_state_sidecar(str(p))
    assert (nx, ny, nlev) == (11, 3, 2)
    assert back.tobytes() == v.tobytes()  # bit-exact, signed zeros and subnormals included


def test_state_sidecar_errors(mda, tmp_path):
    v = np.arange(6.0)
    with pytest.raises(ValueError):
        mda.write_state_sidecar(str(tmp_path / "x.f64"), v, 4, 2)
    p = tmp_path / "s.f64"
    mda.write_state_sidecar(str(p), v, 3, 2)
    bad = tmp_path / "bad.f64"
    bad.write_bytes(b"NOTSTATE" + p.read_bytes()[8:])
    with pytest.raises(RuntimeError, match="not a state sidecar"):
        mda.read_state_sidecar(str(bad))
    short = tmp_path / "short.f64"
    short.write_bytes(p.read_bytes()[:-8])
    with pytest.raises(RuntimeError, match="short"):
        mda.read_state_sidecar(str(short))
    with pytest.raises(RuntimeError, match="Cannot open"):
        mda.read_state_sidecar(str(tmp_path / "absent.f64"))
    # a buffer smaller than the file's state is refused
    lib = mda._lib.load()
    out = np.empty(5)
    nx, ny, nl = C.c_int64(), C.c_int64(), C.c_int64()
    rc = lib.mda_state_read_binary(None, os.fsencode(str(p)), out.ctypes.data, 5, C.byref(nx),
                                   C.byref(ny), C.byref(nl))
    assert rc == mda._lib.MDA_E_ARG
    assert b"more values than the buffer" in lib.mda_last_error(None)
