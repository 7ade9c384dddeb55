"""CPU-side checks of the drop-in boundary: the library loads without a GPU
and exports every entry point include/metada_amd.h declares; host-side
helpers (config narrowing, readers, stencils) behave as the reference."""
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    hdr = open(os.path.join(ROOT, "include", "metada_amd.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    return set(re.findall(r"\b(mda_[a-z0-9_]+)\s*\(", hdr))


def test_library_exports_every_declared_symbol():
    from metada_amd import _lib
    lib = _lib.load()
    decl = declared_symbols()
    assert decl, "no declarations parsed"
    assert decl == set(_lib.SIGNATURES), decl ^ set(_lib.SIGNATURES)
    for name in decl:
        assert hasattr(lib, name), name
    # the shared object's dynamic symbol table names them too
    import subprocess
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True).stdout
    exported = set(re.findall(r"\b(mda_[a-z0-9_]+)\b", out))
    assert decl <= exported, decl - exported


def test_context_without_gpu_reports_device_error():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from metada_amd import _lib
    import ctypes as C
    h = C.c_void_p()
    assert _lib.load().mda_ctx_create(0, C.byref(h)) == -2


def test_param_presets():
    import ctypes as C
    from metada_amd import _lib as L
    p = L.LetkfParams()
    L.load().mda_letkf_params_standard(C.byref(p), 1.05, 500.0)
    assert (p.localization, p.rweight, p.inflation_form, p.sqrt_form, p.combine) == (
        L.LOC_GASPARI_COHN, L.RW_REAL, L.INFLFORM_PRIOR, L.SQRT_SYMMETRIC, L.COMBINE_MATRIX)
    L.load().mda_letkf_params_refcompat(C.byref(p), 1.05, 10.0)
    assert (p.localization, p.rweight, p.inflation_form, p.sqrt_form, p.combine) == (
        L.LOC_CUTOFF, L.RW_IDENTITY, L.INFLFORM_POSTERIOR, L.SQRT_CHOLESKY,
        L.COMBINE_ELEMENTWISE)
    assert p.inflation == 1.05 and p.radius == 10.0


def test_config_float_narrowing():
    from metada_amd import Config
    c = Config({"analysis": {"inflation": 1.05, "n": 1, "error": 0.1}})
    a = c.GetSubsection("analysis")
    assert a.Get("inflation").asFloat() == 1.0499999523162842
    assert a.Get("error").asFloat() == 0.10000000149011612
    with pytest.raises(TypeError):
        a.Get("n").asFloat()
    with pytest.raises(RuntimeError):
        c.GetSubsection("missing")


def test_tutorial_readers():
    from metada_amd import Observation, read_simple_state
    tut = os.path.join(ROOT, "tests", "golden", "tutorial")
    x = read_simple_state(os.path.join(tut, "ens_1.txt"))
    assert x.shape == (648,)
    obs = Observation.from_grid_file(os.path.join(tut, "obs_format.txt"), error=0.1)
    assert obs.size() == 28
    assert obs.loc_a[0] == 4 and obs.loc_b[0] == 3 and obs.loc_c[0] == 1
    assert obs.values[0] == 0.2052
    assert np.all(obs.variances == 0.1 * 0.1)


def test_trilinear_reproduces_affine_fields():
    from metada_amd import ObsOperator
    nlon, nlat, nlev = 16, 8, 4
    rng = np.random.default_rng(1)
    lat = rng.uniform(-60, 60, 50)
    lon = rng.uniform(30, 300, 50)
    lev = rng.uniform(0, 3, 50)
    H = ObsOperator.trilinear_latlon(lat, lon, lev, nlon, nlat, nlev)
    LON = (np.arange(nlon) + 0.5) * 360 / nlon
    LAT = -90 + (np.arange(nlat) + 0.5) * 180 / nlat
    f = (2.0 * LAT[None, :, None] + 0.1 * LON[None, None, :] + 3.0 * np.arange(nlev)[:, None, None])
    hx = H.apply_host(f.reshape(1, -1))[0]
    np.testing.assert_allclose(hx, 2 * lat + 0.1 * lon + 3 * lev, rtol=1e-12, atol=1e-10)
    assert np.allclose(np.diff(H.row_ptr), 8)


def build_adapter_mock(out_dir):
    """Compiles and links tests/cpp/adapter_mock.cpp (the C++ adapter over
    stand-ins with the reference API's signatures) against the library and
    the oracle (its IDW4 observation operator).  Returns the executable."""
    import shutil
    import subprocess
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no g++")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    from oracle import oracle as orc
    orc.build()
    libdir, orcdir = os.path.join(root, "metada_amd"), os.path.join(root, "oracle")
    exe = os.path.join(str(out_dir), "adapter_mock")
    r = subprocess.run([gxx, "-std=c++20", "-O1", "-Wall", "-Wextra", "-Werror",
                        "-I", os.path.join(root, "include"), "-I", orcdir,
                        os.path.join(root, "tests", "cpp", "adapter_mock.cpp"), "-o", exe,
                        "-L", libdir, "-lmetada_amd", "-L", orcdir, "-loracle",
                        f"-Wl,-rpath,{libdir}", f"-Wl,-rpath,{orcdir}"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def test_cpp_adapter_builds_against_reference_shaped_types(tmp_path):
    """include/metada_amd.hpp instantiated and linked with stand-ins that
    reproduce the reference adapter API's signatures (getDataPtr, the
    void*-backed getData, ConfigValue-returning Get, GetSubsection, the
    private Location coordinate system)."""
    if not os.path.exists(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(
            __file__))), "metada_amd", "libmetada_amd.so")):
        pytest.skip("library not built")
    assert os.access(build_adapter_mock(tmp_path), os.X_OK)
