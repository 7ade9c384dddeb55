"""Seeded synthetic LETKF workloads on a global regular latitude-longitude grid.

Background perturbations are Gaussian random fields: white noise smoothed by
separable Gaussian filters (vertical, latitude, and longitude with a
latitude-dependent width so that the filter is isotropic in km), normalised to
unit variance.  A correlation length L means corr(d) = exp(-d^2 / L^2), i.e.
a filter standard deviation of L / 2.  Observations are area-uniform on the
sphere (or clustered in Gaussian blobs), observed through trilinear H.

Generation runs on the GPU through torch (device memory plumbing only); it is
never part of a timed analysis.
"""
from __future__ import annotations

import math

import numpy as np

from .analysis import ColumnGeometry, ObsOperator, Observation
from . import _lib as L

EARTH_KM = 6371.0


def _smoothing_matrix(n: int, sigma_cells: float, device, dtype):
    """(n x n) Gaussian smoothing with reflecting boundaries, applied as a
    GEMM (no convolution library involved)."""
    import torch
    half = max(1, int(math.ceil(4.0 * sigma_cells)))
    offs = np.arange(-half, half + 1)
    g = np.exp(-0.5 * (offs / sigma_cells) ** 2)
    g /= g.sum()
    S = np.zeros((n, n))
    for i in range(n):
        for o, w in zip(offs, g):
            j = i + o
            while j < 0 or j >= n:  # reflect (edge not repeated)
                j = -j if j < 0 else 2 * (n - 1) - j
                if n == 1:
                    j = 0
            S[i, j] += w
    return torch.from_numpy(S).to(device=device, dtype=dtype)


class GRF:
    """Unit-variance Gaussian random fields on an nlon x nlat x nlev grid."""

    def __init__(self, nlon, nlat, nlev, L_km=200.0, L_lev=3.0, device="cuda"):
        import torch
        self.nlon, self.nlat, self.nlev, self.device = nlon, nlat, nlev, device
        dtype = torch.float64
        dy_km = math.pi * EARTH_KM / nlat  # meridional spacing
        self.Sz = _smoothing_matrix(nlev, L_lev / 2.0, device, dtype)
        self.Sy = _smoothing_matrix(nlat, (L_km / 2.0) / dy_km, device, dtype)
        lat = -90.0 + (torch.arange(nlat, device=device, dtype=dtype) + 0.5) * 180.0 / nlat
        dx_km = (2 * math.pi * EARTH_KM / nlon) * torch.cos(torch.deg2rad(lat)).clamp_min(1e-3)
        sig_x = (L_km / 2.0) / dx_km                                  # cells, per latitude row
        f = torch.fft.rfftfreq(nlon, d=1.0, device=device).to(dtype)  # cycles per cell
        self.hx = torch.exp(-2.0 * (math.pi ** 2) * (sig_x[:, None] ** 2) * (f[None, :] ** 2))
        # variance of filtered unit white noise: prod_d sum g_d^2 (Parseval along lon)
        w = torch.full_like(f, 2.0)
        w[0] = 1.0
        if nlon % 2 == 0:
            w[-1] = 1.0
        vx = (w[None, :] * self.hx ** 2).sum(-1) / nlon              # (nlat,)
        vz = (self.Sz ** 2).sum(1)                                    # (nlev,)
        vy = (self.Sy ** 2).sum(1)                                    # (nlat,)
        self.scale = 1.0 / torch.sqrt(vz[:, None] * vy[None, :] * vx[None, :])  # (nlev, nlat)

    def sample(self, gen, out=None):
        import torch
        z = torch.randn((self.nlev, self.nlat, self.nlon), generator=gen, device=self.device,
                        dtype=torch.float64)
        z = torch.matmul(self.Sz, z.reshape(self.nlev, -1)).reshape(z.shape)
        z = torch.matmul(self.Sy, z)                                   # batched over levels
        Z = torch.fft.rfft(z, dim=2)
        Z *= self.hx[None, :, :]
        z = torch.fft.irfft(Z, n=self.nlon, dim=2)
        z *= self.scale[:, :, None]
        if out is not None:
            out.copy_(z.reshape(-1))
            return out
        return z.reshape(-1)


def latlon_workload(nlon=1024, nlat=512, nlev=40, k=64, p=1_000_000, sigma=0.5,
                    L_km=200.0, L_lev=3.0, seed=2026, clustered=False, nblobs=50,
                    blob_km=300.0, background_frac=0.10, device="cuda", progress=None):
    """Config-2-shaped workload on the device: returns a dict with the
    ensemble (k, n) and truth (n) tensors, the Observation, the trilinear
    ObsOperator (device CSR) and the column geometry (device)."""
    import torch
    gen = torch.Generator(device=device)
    gen.manual_seed(seed)
    grf = GRF(nlon, nlat, nlev, L_km, L_lev, device)
    n = nlon * nlat * nlev
    X = torch.empty((k, n), dtype=torch.float64, device=device)
    truth = torch.empty(n, dtype=torch.float64, device=device)
    grf.sample(gen, truth)
    for j in range(k):
        grf.sample(gen, X[j])
        if progress and (j + 1) % 16 == 0:
            torch.cuda.synchronize()
            progress(f"generated {j + 1}/{k} members")
    rng = np.random.default_rng(seed)
    if clustered:
        nb = int(round(p * background_frac))
        lat_b = np.degrees(np.arcsin(rng.uniform(-1, 1, nb)))
        lon_b = rng.uniform(0, 360, nb)
        c_lat = np.degrees(np.arcsin(rng.uniform(-1, 1, nblobs)))
        c_lon = rng.uniform(0, 360, nblobs)
        m = p - nb
        which = rng.integers(0, nblobs, m)
        # Gaussian blob in local tangent-plane km, mapped back to lat/lon
        dn = rng.normal(0, blob_km, m)
        de = rng.normal(0, blob_km, m)
        lat_c = np.clip(c_lat[which] + np.degrees(dn / EARTH_KM), -89.999, 89.999)
        lon_c = c_lon[which] + np.degrees(de / (EARTH_KM * np.cos(np.radians(lat_c))))
        lat = np.concatenate([lat_b, lat_c])
        lon = np.mod(np.concatenate([lon_b, lon_c]), 360.0)
    else:
        lat = np.degrees(np.arcsin(rng.uniform(-1, 1, p)))
        lon = rng.uniform(0, 360, p)
    lev = rng.uniform(0, nlev - 1, p)
    H = ObsOperator.trilinear_latlon(lat, lon, lev, nlon, nlat, nlev)
    Hd = ObsOperator(torch.from_numpy(H.row_ptr).to(device), torch.from_numpy(H.col).to(device),
                     torch.from_numpy(H.weight).to(device))
    cols = Hd.col.view(p, 8)
    hxt = (truth[cols] * Hd.weight.view(p, 8)).sum(1)
    noise = torch.randn(p, generator=gen, device=device, dtype=torch.float64) * sigma
    yo = hxt + noise
    obs = Observation(yo, torch.full((p,), sigma * sigma, dtype=torch.float64, device=device),
                      torch.from_numpy(lat).to(device), torch.from_numpy(lon).to(device), None,
                      L.COORD_GEOGRAPHIC)
    geom = ColumnGeometry.latlon(nlon, nlat, device=device)
    return dict(X=X, truth=truth, obs=obs, H=Hd, H_host=H, geom=geom, nlon=nlon, nlat=nlat,
                nlev=nlev, k=k, p=p, lat=lat, lon=lon, lev=lev)
