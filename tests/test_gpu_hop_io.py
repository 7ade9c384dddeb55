"""IdentityObsOperator built on the device (mda_hop_idw4_build) and
saveEnsemble with FP64 sidecars (mda_ensemble_save), against the oracle's
IDW4 stencil (IdentityObsOperator.hpp:594-676), a brute-force nearest-column
search for GEOGRAPHIC observations (IdentityObsOperator.hpp:483-530) and the
closed form of the (0, 0, 0) fallback (IdentityObsOperator.hpp:532-587)."""
import ctypes as C
import math
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mda():
    import metada_amd
    return metada_amd


@pytest.fixture(scope="module")
def ctx(mda):
    c = mda.Context(0)
    yield c
    c.close()


def _oracle_rows(orc, nx, ny, nz, pts):
    """The oracle's four taps per point, weights normalised by their sum taken
    in tap order."""
    cols, ws = [], []
    for x, y, z in pts:
        idx, w = orc.idw4_stencil(nx, ny, nz, float(x), float(y), float(z))
        s = 0.0
        for t in range(4):
            s += w[t]
        cols.append(np.asarray(idx, np.int64))
        ws.append(np.asarray(w) / s)
    return cols, ws


@pytest.mark.parametrize("nx,ny,nz", [(9, 7, 1), (9, 7, 3), (1, 5, 1), (6, 1, 4)])
def test_idw4_grid_stencil_matches_oracle(mda, ctx, orc, nx, ny, nz):
    rng = np.random.default_rng(nx * 100 + nz)
    p = 60
    pts = rng.uniform(-0.5, 1.2, (p, 3)) * [max(nx - 1, 1), max(ny - 1, 1), max(nz - 1, 1)]
    pts[:10] = np.round(pts[:10])          # exact hits: weight 1e12
    pts[10:14] = [[nx - 1, ny - 1, nz - 1], [0, 0, 0], [nx + 3, -2, nz], [0.5, 0.5, 0.5]]
    ob = mda.Observation(np.zeros(p), np.ones(p), pts[:, 0], pts[:, 1], pts[:, 2])
    H = mda.ObsOperator.identity_idw4(ob, nx, ny, nz, ctx=ctx)
    np.testing.assert_array_equal(H.row_ptr, np.arange(p + 1) * 4)
    cols, ws = _oracle_rows(orc, nx, ny, nz, pts)
    np.testing.assert_array_equal(H.col, np.concatenate(cols))
    np.testing.assert_array_equal(H.weight, np.concatenate(ws))  # same operation order
    st = rng.standard_normal(nx * ny * nz)
    hx = H.apply_host(st[None, :])[0]
    want = [orc.idw4(st, nx, ny, nz, *q) for q in pts]  # the reference's sum(w x) / sum(w)
    np.testing.assert_allclose(hx, want, rtol=1e-14, atol=1e-14)


def test_idw4_invalid_observations_get_empty_rows(mda, ctx, orc):
    nx, ny = 8, 6
    pts = np.array([[1.5, 2.0], [3.0, 4.0], [6.2, 0.7], [0.0, 5.0]])
    var = np.array([1.0, np.inf, 0.5, np.inf])
    ob = mda.Observation(np.zeros(4), var, pts[:, 0], pts[:, 1], np.zeros(4))
    H = mda.ObsOperator.identity_idw4(ob, nx, ny, 1, ctx=ctx)
    np.testing.assert_array_equal(H.row_ptr, [0, 4, 4, 8, 8])
    cols, ws = _oracle_rows(orc, nx, ny, 1, [(1.5, 2.0, 0), (6.2, 0.7, 0)])
    np.testing.assert_array_equal(H.col, np.concatenate(cols))
    np.testing.assert_array_equal(H.weight, np.concatenate(ws))


def _nearest_column(lat, lon, glat, glon):
    # the reference's loop: sqrt((lon-glon)^2 + (lat-glat)^2), first strict minimum
    best, bd = 0, np.finfo(float).max
    d = np.sqrt((lon - glon) * (lon - glon) + (lat - glat) * (lat - glat))
    for i, v in enumerate(d):
        if v < bd:
            best, bd = i, v
    return best


def _nearest_level(lev, vcoord):
    best, bd = 0, np.finfo(float).max
    for z, v in enumerate(vcoord):
        if abs(lev - v) < bd:
            best, bd = z, abs(lev - v)
    return best


@pytest.mark.parametrize("nz", [1, 4])
def test_idw4_geographic_nearest_column(mda, ctx, orc, nz):
    rng = np.random.default_rng(40 + nz)
    nx, ny = 37, 23                      # 851 columns: > 64 per lane visit
    glat = (np.linspace(20, 50, ny)[:, None] + rng.uniform(-0.3, 0.3, (ny, nx))).ravel()
    glon = (np.linspace(-100, -60, nx)[None, :] + rng.uniform(-0.3, 0.3, (ny, nx))).ravel()
    vcoord = np.sort(rng.uniform(0.0, 1.0, nz))[::-1].copy()
    p = 90
    lat, lon = rng.uniform(18, 52, p), rng.uniform(-102, -58, p)
    lev = rng.uniform(-0.1, 1.1, p)
    # a tie: halfway between two columns of a regular patch gives the lower index
    glat[:2], glon[:2] = 10.0, [0.0, 1.0]
    lat[0], lon[0] = 10.0, 0.5
    # exactly on a column
    lat[1], lon[1] = glat[500], glon[500]
    geom = mda.ColumnGeometry(mda._lib.COORD_GEOGRAPHIC, glat, glon, None)
    ob = mda.Observation(np.zeros(p), np.ones(p), lat, lon, lev, mda._lib.COORD_GEOGRAPHIC)
    H = mda.ObsOperator.identity_idw4(ob, nx, ny, nz, geometry=geom,
                                      vcoord=vcoord if nz > 1 else None, ctx=ctx)
    pts = []
    for o in range(p):
        g = _nearest_column(lat[o], lon[o], glat, glon)
        kz = _nearest_level(lev[o], vcoord) if nz > 1 else 0
        pts.append((g % nx, g // nx, kz))
    assert pts[0][:2] == (0, 0) and pts[1][:2] == (500 % nx, 500 // nx)
    cols, ws = _oracle_rows(orc, nx, ny, nz, pts)
    np.testing.assert_array_equal(H.col, np.concatenate(cols))
    np.testing.assert_array_equal(H.weight, np.concatenate(ws))


def test_idw4_geographic_without_coordinates_falls_back_to_origin(mda, ctx):
    """SimpleGeometry holds no geographic locations: the fallback search of
    convertGeographicToGrid keeps (0, 0, 0), so every row is the stencil at
    the origin: distances 0, 1, 1, sqrt 2 -> weights (1e12, 1, 1, 1/sqrt 2)
    normalised."""
    p = 5
    ob = mda.Observation(np.zeros(p), np.ones(p), np.linspace(-40, 40, p),
                         np.linspace(0, 300, p), np.zeros(p), mda._lib.COORD_GEOGRAPHIC)
    H = mda.ObsOperator.identity_idw4(ob, 10, 8, 1, ctx=ctx)
    w = np.array([1e12, 1.0, 1.0, 1.0 / math.sqrt(2.0)])
    np.testing.assert_array_equal(H.col, np.tile([0, 1, 10, 11], p))
    np.testing.assert_allclose(H.weight, np.tile(w / w.sum(), p), rtol=1e-15)


def test_idw4_device_outputs_match_host_outputs(mda, ctx):
    import torch
    rng = np.random.default_rng(3)
    nx, ny, nz, p = 12, 9, 2, 40
    pts = rng.uniform(0, 1, (p, 3)) * [nx - 1, ny - 1, nz - 1]
    var = np.where(rng.uniform(size=p) < 0.2, np.inf, 1.0)
    ob = mda.Observation(np.zeros(p), var, pts[:, 0], pts[:, 1], pts[:, 2])
    H = mda.ObsOperator.identity_idw4(ob, nx, ny, nz, ctx=ctx)
    dev = torch.device("cuda:0")
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    dob = mda.Observation(t(np.zeros(p)), t(var), t(pts[:, 0]), t(pts[:, 1]), t(pts[:, 2]))
    os_, keep = mda.analysis._obs_struct(dob)
    gs, gk = mda.ColumnGeometry(mda._lib.COORD_GRID).struct()
    rp = torch.zeros(p + 1, dtype=torch.int64, device=dev)
    col = torch.full((4 * p,), -1, dtype=torch.int64, device=dev)
    w = torch.full((4 * p,), np.nan, dtype=torch.float64, device=dev)
    ctx.check(ctx.lib.mda_hop_idw4_build(ctx.h, C.byref(os_), mda._lib.COORD_GRID, C.byref(gs),
                                         nx, ny, nz, None, rp.data_ptr(), col.data_ptr(),
                                         w.data_ptr(), 1))
    torch.cuda.synchronize()
    nnz = int(H.row_ptr[-1])
    np.testing.assert_array_equal(rp.cpu().numpy(), H.row_ptr)
    np.testing.assert_array_equal(col.cpu().numpy()[:nnz], H.col)
    np.testing.assert_array_equal(w.cpu().numpy()[:nnz], H.weight)


def test_device_hop_drives_the_analysis(mda, ctx, orc):
    """A LETKF with the device-built H equals the oracle fed the oracle's own
    IDW4 equivalents."""
    rng = np.random.default_rng(8)
    nx, ny, k, p = 16, 10, 8, 50
    X = rng.standard_normal((k, nx * ny))
    ox, oy = rng.uniform(0, nx - 1, p), rng.uniform(0, ny - 1, p)
    yo, var = rng.standard_normal(p), rng.uniform(0.5, 1.5, p)
    ob = mda.Observation(yo, var, ox, oy, np.zeros(p))
    H = mda.ObsOperator.identity_idw4(ob, nx, ny, 1, ctx=ctx)
    ens = mda.Ensemble(X.copy(), nx, ny, 1)
    cfg = mda.Config({"inflation": 1.05, "localization_radius": 4.0})
    f = mda.LETKF(ens, ob, H, cfg, ctx=ctx)
    f.Analyse()
    Xa = ens.to_host()
    Hx = np.array([[orc.idw4(X[j], nx, ny, 1, ox[o], oy[o], 0.0) for o in range(p)]
                   for j in range(k)])
    prob = orc.Problem(k, nx, ny, 1, orc.COORD_GRID, np.tile(np.arange(nx, dtype=float), ny),
                       np.repeat(np.arange(ny, dtype=float), nx), ox, oy, yo, var,
                       float(np.float32(1.05)), float(np.float32(4.0)))
    Xo, _, _ = orc.letkf_standard(prob, X, Hx)
    assert np.sqrt(np.mean((Xa - Xo) ** 2)) / np.sqrt(np.mean(Xo ** 2)) < 1e-10


def test_ensemble_save_with_sidecars(mda, ctx, orc, tmp_path):
    import torch
    rng = np.random.default_rng(12)
    nx, ny, nlev, k = 7, 5, 3, 4
    X = rng.standard_normal((k, nx * ny * nlev)) * 1e3
    dX = torch.from_numpy(X.copy()).to("cuda:0")
    ens = mda.Ensemble(dX, nx, ny, nlev)
    es, keep = ens.struct()
    base = str(tmp_path / "out" / "ens")
    os.makedirs(os.path.dirname(base))
    ctx.check(ctx.lib.mda_ensemble_save(ctx.h, C.byref(es), base.encode(), b"txt", 1))
    for j in range(k):
        v, a, b, c = mda.read_state_sidecar(f"{base}_member_{j}.f64")
        assert (a, b, c) == (nx, ny, nlev)
        assert v.tobytes() == X[j].tobytes()
        txt = np.loadtxt(f"{base}_member_{j}.txt")
        assert txt.shape == (ny * nlev, nx)
        np.testing.assert_allclose(txt.ravel(), X[j], atol=5e-7)
    m, *_ = mda.read_state_sidecar(f"{base}_mean.f64")
    # Ensemble::RecomputeMean: members summed in order, times 1/k
    want = np.zeros(X.shape[1])
    for j in range(k):
        want = want + X[j]
    want = want * (1.0 / k)
    assert m.tobytes() == want.tobytes()
    np.testing.assert_allclose(np.loadtxt(f"{base}_mean.txt").ravel(), want, atol=5e-7)
    assert not os.path.exists(f"{base}_member_{k}.txt")
