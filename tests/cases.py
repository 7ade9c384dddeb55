"""Seeded synthetic inputs shared by the parity tests (numpy, host side)."""
import numpy as np


def rel_err(a, b):
    a = np.asarray(a); b = np.asarray(b)
    scale = np.max(np.abs(b))
    return float(np.max(np.abs(a - b)) / (scale if scale > 0 else 1.0))


def l96_case(orc, N=40, k=20, seed=7, spin=1000, dt=0.05):
    """Lorenz-96 twin case: truth from x=F with x_19 += 0.001 spun up with RK4,
    members truth + N(0, 0.5^2), every other variable observed through the
    reference IdentityObsOperator on an N x 1 grid, y = x_t + N(0, 1)."""
    rng = np.random.default_rng(seed)
    x = np.full(N, 8.0)
    x[19 % N] += 0.001
    for _ in range(spin):
        x = orc.l96_step(x, 8.0, dt, "rk4")
    truth = x
    X = truth + 0.5 * rng.standard_normal((k, N))
    obs_idx = np.arange(0, N, 2)
    p = len(obs_idx)
    rp, cols, ws = [0], [], []
    for i in obs_idx:
        idx, w = orc.idw4_stencil(N, 1, 1, float(i), 0.0, 0.0)
        cols += list(idx); ws += list(w / w.sum()); rp.append(len(cols))
    rp, cols, ws = np.array(rp), np.array(cols), np.array(ws)
    yo = orc.csr_apply(rp, cols, ws, truth) + rng.standard_normal(p)
    var = np.ones(p)
    Z = rng.standard_normal((p, k))
    return dict(truth=truth, X=X, rp=rp, cols=cols, ws=ws, yo=yo, var=var, Z=Z,
                obs_x=obs_idx.astype(float), obs_y=np.zeros(p), N=N, k=k, p=p)


def smooth_fields(rng, m, nlev, nlat, nlon, nmodes=12):
    """m smooth random fields on a lat-lon-level grid from a few random
    Fourier modes (unit-ish variance)."""
    lon = (np.arange(nlon) + 0.5) * 2 * np.pi / nlon
    lat = -np.pi / 2 + (np.arange(nlat) + 0.5) * np.pi / nlat
    lev = np.arange(nlev)
    out = np.zeros((m, nlev, nlat, nlon))
    for j in range(m):
        for _ in range(nmodes):
            a = rng.standard_normal() / np.sqrt(nmodes / 2)
            kx, ky, kz = rng.integers(1, 5), rng.uniform(0.5, 3), rng.uniform(0, 0.6)
            ph = rng.uniform(0, 2 * np.pi, 3)
            out[j] += a * (np.cos(kx * lon + ph[0])[None, None, :]
                           * np.cos(ky * lat + ph[1])[None, :, None]
                           * np.cos(kz * lev + ph[2])[:, None, None])
    return out.reshape(m, -1)


def latlon_case(nlon=64, nlat=32, nlev=4, k=20, p=3000, sigma=0.5, seed=11, cluster=False,
                nmodes=12):
    """Scaled-down global LETKF case: background = smooth random fields,
    truth = one more draw, obs area-uniform with trilinear H."""
    from metada_amd import ObsOperator
    rng = np.random.default_rng(seed)
    F = smooth_fields(rng, k + 1, nlev, nlat, nlon, nmodes)
    truth, X = F[0], F[1:]
    if cluster:
        c_lat = np.degrees(np.arcsin(rng.uniform(-1, 1, 5)))
        c_lon = rng.uniform(0, 360, 5)
        which = rng.integers(0, 5, p)
        lat = np.clip(c_lat[which] + rng.normal(0, 6, p), -89.9, 89.9)
        lon = np.mod(c_lon[which] + rng.normal(0, 6, p), 360)
    else:
        lat = np.degrees(np.arcsin(rng.uniform(-1, 1, p)))
        lon = rng.uniform(0, 360, p)
    lev = rng.uniform(0, nlev - 1, p)
    H = ObsOperator.trilinear_latlon(lat, lon, lev, nlon, nlat, nlev)
    yo = H.apply_host(truth[None])[0] + sigma * rng.standard_normal(p)
    var = np.full(p, sigma * sigma)
    lonc = (np.arange(nlon) + 0.5) * 360 / nlon
    latc = -90 + (np.arange(nlat) + 0.5) * 180 / nlat
    LON, LAT = np.meshgrid(lonc, latc)
    return dict(truth=truth, X=X, H=H, yo=yo, var=var, lat=lat, lon=lon, lev=lev,
                col_lat=LAT.ravel().copy(), col_lon=LON.ravel().copy(),
                nlon=nlon, nlat=nlat, nlev=nlev, k=k, p=p)
