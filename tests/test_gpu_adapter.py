"""The C++ drop-in adapter (include/metada_amd.hpp) run end to end through the
C-ABI on the GPU, over stand-ins with the reference API's signatures
(tests/cpp/adapter_mock.cpp): LETKF in the reference's semantics on the
reference's tutorial data, ETKF and EnKF on a Lorenz-96 case -- analyses,
refreshed ensemble mean, EnKF AnalysisResults and saveEnsemble() files --
against the oracle."""
import os
import subprocess

import numpy as np
import pytest

from cases import l96_case, rel_err
from test_abi_cpu import build_adapter_mock

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    return build_adapter_mock(tmp_path_factory.mktemp("adapter"))


def _run(exe, tmp_path, algo, form, nx, ny, nlev, X, ox, oy, oz, vals, errs, infl, radius,
         Z=None):
    k = X.shape[0]
    inp = tmp_path / f"{algo}.in"
    with open(inp, "w") as f:
        f.write(f"{algo} {form} {nx} {ny} {nlev} {k} {len(vals)} {infl!r} {radius!r}\n")
        f.write("\n".join(repr(float(v)) for v in X.ravel()) + "\n")
        for x, y, z, v, e in zip(ox, oy, oz, vals, errs):
            f.write(f"{int(x)} {int(y)} {int(z)} {float(v)!r} {float(e)!r}\n")
        if Z is not None:
            f.write("\n".join(repr(float(v)) for v in Z.ravel()) + "\n")
    out = str(tmp_path / algo)
    r = subprocess.run([exe, str(inp), out], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    diag = dict(line.split(" ", 1) for line in r.stdout.splitlines())
    Xa = np.loadtxt(out + ".analysis").reshape(X.shape)
    mean = np.loadtxt(out + ".mean")
    return Xa, mean, diag, out


def _idw4_hx(orc, X, nx, ny, nz, ox, oy, oz):
    return np.array([[orc.idw4(X[j], nx, ny, nz, float(x), float(y), float(z))
                      for x, y, z in zip(ox, oy, oz)] for j in range(X.shape[0])])


def test_adapter_letkf_tutorial_refcompat(mda_lib_loaded, exe, orc, tmp_path):
    """letkf.cpp:62-65 with the drop-in class: the tutorial's 36 x 18 grid, 9
    members and 28 grid observations (error 0.1, radius 10, inflation 1.0)."""
    import metada_amd as mda
    tut = os.path.join(ROOT, "tests", "golden", "tutorial")
    X = np.stack([mda.read_simple_state(os.path.join(tut, f"ens_{i}.txt")) for i in range(1, 10)])
    obs = mda.Observation.from_grid_file(os.path.join(tut, "obs_format.txt"), error=0.1)
    err = np.full(obs.size(), 0.1)
    Xa, mean, diag, out = _run(exe, tmp_path, "letkf", "refcompat", 36, 18, 1, X, obs.loc_a,
                               obs.loc_b, obs.loc_c, obs.values, err, 1.0, 10.0)
    var = np.full(obs.size(), float(np.float32(0.1)) ** 2)
    prob = orc.Problem(9, 36, 18, 1, orc.COORD_GRID, np.tile(np.arange(36.0), 18),
                       np.repeat(np.arange(18.0), 36), obs.loc_a, obs.loc_b, obs.values, var,
                       1.0, 10.0)
    Xo = orc.letkf_refcompat_snapshot(prob, X, _idw4_hx(orc, X, 36, 18, 1, obs.loc_a,
                                                         obs.loc_b, obs.loc_c))
    assert rel_err(Xa, Xo) < 1e-10
    # Ensemble::RecomputeMean after the analysis (LETKF.hpp:116)
    assert rel_err(mean, orc.ensemble_mean(Xo)) < 1e-10
    # saveEnsemble: <base>_mean.txt and <base>_member_<i>.txt, 6-decimal rows of nx
    saved = np.loadtxt(out + "_saved_mean.txt")
    assert saved.shape == (18, 36)
    np.testing.assert_allclose(saved.ravel(), orc.ensemble_mean(Xo), atol=6e-7)
    for i in (0, 8):
        np.testing.assert_allclose(np.loadtxt(f"{out}_saved_member_{i}.txt").ravel(), Xo[i],
                                   atol=6e-7)
    assert int(diag["local_obs"]) > 0


def test_adapter_etkf_and_enkf_l96(mda_lib_loaded, exe, orc, tmp_path):
    case = l96_case(orc, N=40, k=20, seed=13)
    N, k, p = case["N"], case["k"], case["p"]
    X = case["X"]
    ox = case["obs_x"].astype(int)
    errs = np.ones(p)
    infl = 1.1
    rho = float(np.float32(infl))
    Yb = _idw4_hx(orc, X, N, 1, 1, ox, np.zeros(p), np.zeros(p))
    # ETKF (etkf.cpp:62-65)
    Xa, mean, _, _ = _run(exe, tmp_path, "etkf", "-", N, 1, 1, X, ox, np.zeros(p), np.zeros(p),
                          case["yo"], errs, infl, 0.0)
    Xo = orc.etkf_analyse(X, Yb, case["yo"], np.ones(p), rho)
    assert rel_err(Xa, Xo) < 1e-10
    assert rel_err(mean, orc.ensemble_mean(Xo)) < 1e-10
    # EnKF (enkf.cpp:80-136): root config, injected draws, AnalysisResults
    Xa, mean, diag, _ = _run(exe, tmp_path, "enkf", "-", N, 1, 1, X, ox, np.zeros(p),
                             np.zeros(p), case["yo"], errs, infl, 0.0, Z=case["Z"])
    Xo, d = orc.enkf_analyse(X, Yb, case["yo"], np.ones(p), case["Z"], rho)
    assert rel_err(Xa, Xo) < 1e-10
    assert rel_err(mean, orc.ensemble_mean(Xo)) < 1e-10
    for key in ("innovation_norm", "background_spread", "analysis_spread", "max_kalman_gain",
                "min_kalman_gain", "condition_number"):
        assert abs(float(diag[key]) - d[key]) <= 1e-10 * abs(d[key]), key
    assert diag["inflation_method"] == "relaxation"  # a known name, kept (EnKF.hpp:116-128)
    assert float(diag["inflation_factor"]) == rho
    assert int(diag["ensemble_size"]) == k and int(diag["observation_count"]) == p
    assert int(diag["perturbations"]) == k  # RecomputePerturbations (EnKF.hpp:238)


@pytest.fixture(scope="module")
def mda_lib_loaded():
    import metada_amd  # noqa: F401  (the library loads; the adapter binds to the same .so)
    return True
