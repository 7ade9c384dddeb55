"""Host-side mirror of metada's analysis interface over the C-ABI.

The reference exposes, per algorithm (src/framework/algorithms):
    EnKF<Tag>(Ensemble&, Observation&, const ObsOperator&, const Config&)
        Analyse(), saveEnsemble(), getAnalysisResults()       EnKF.hpp:105-301
    ETKF<Tag>(...)  Analyse(), saveEnsemble()                 ETKF.hpp:86-200
    LETKF<Tag>(...) Analyse(), saveEnsemble()                 LETKF.hpp:47-142
The classes below keep those names, argument meanings and error behaviour
(errors surface as RuntimeError, as the reference throws std::runtime_error),
and route the arithmetic through libmetada_amd.so.  Ensembles may be host
numpy arrays (copied per call, the reference's ownership model) or device
torch tensors (updated in place without host round trips).
"""
from __future__ import annotations

import ctypes as C
import math
import os
from dataclasses import dataclass, field
from typing import Callable, Optional, Sequence

import numpy as np

from . import _lib as L


# ---------------------------------------------------------------------------
# configuration values (ConfigValue.hpp:35-108, YamlConfig.cpp:232-268)
# ---------------------------------------------------------------------------
class ConfigValue:
    """A config scalar with the reference's typed accessors.  Reals are stored
    as float32 (ConfigValue.hpp:108), so asFloat() returns the float-narrowed
    value, and an integer literal for a real key raises, as the reference's
    std::get<float> does (bad_variant_access)."""

    def __init__(self, v):
        self.v = v

    def asFloat(self) -> float:
        if isinstance(self.v, bool) or isinstance(self.v, int):
            raise TypeError("bad_variant_access: integer literal where a float is expected")
        return float(np.float32(self.v))

    def asInt(self) -> int:
        if not isinstance(self.v, int) or isinstance(self.v, bool):
            raise TypeError("bad_variant_access: not an integer")
        return int(self.v)

    def asBool(self) -> bool:
        if not isinstance(self.v, bool):
            raise TypeError("bad_variant_access: not a bool")
        return self.v

    def asString(self) -> str:
        if not isinstance(self.v, str):
            raise TypeError("bad_variant_access: not a string")
        return self.v


class Config:
    """Dot-path config over a mapping (Config.hpp:203-313): Get() returns a
    ConfigValue (a missing key yields an empty value, as Config::Get swallows
    errors into a default), GetSubsection() raises on a missing key."""

    def __init__(self, data: dict):
        self.data = data or {}

    @classmethod
    def from_yaml(cls, path: str) -> "Config":
        import yaml
        with open(path) as f:
            return cls(yaml.safe_load(f))

    def _lookup(self, key):
        cur = self.data
        for part in key.split("."):
            if not isinstance(cur, dict) or part not in cur:
                return None
            cur = cur[part]
        return cur

    def Get(self, key: str) -> ConfigValue:
        return ConfigValue(self._lookup(key))

    def GetSubsection(self, key: str) -> "Config":
        v = self._lookup(key)
        if not isinstance(v, dict):
            raise RuntimeError(f"Config subsection not found: {key}")
        return Config(v)

    def Has(self, key: str) -> bool:
        return self._lookup(key) is not None


# ---------------------------------------------------------------------------
# device context and pointer plumbing
# ---------------------------------------------------------------------------
class Context:
    """One mda_ctx (one HIP stream, one reusable device workspace)."""

    def __init__(self, device: int = 0):
        self.lib = L.load()
        h = C.c_void_p()
        rc = self.lib.mda_ctx_create(device, C.byref(h))
        if rc != L.MDA_OK:
            raise L.MdaError(rc, f"mda_ctx_create(device={device}) failed")
        self.h = h
        self.device = device

    def close(self):
        if getattr(self, "h", None):
            self.lib.mda_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, rc):
        return L.check(self.h, rc)

    @property
    def stream(self) -> int:
        return self.lib.mda_ctx_stream(self.h)

    def synchronize(self):
        self.check(self.lib.mda_ctx_synchronize(self.h))

    def workspace_bytes(self) -> int:
        return int(self.lib.mda_ctx_workspace_bytes(self.h))


_default_ctx: Optional[Context] = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


def _is_torch(x):
    return type(x).__module__.startswith("torch")


def _buf(x, dtype=np.float64):
    """(address, on_device, keepalive) for a numpy array or a torch tensor."""
    if x is None:
        return None, 0, None
    if _is_torch(x):
        if not x.is_contiguous():
            raise ValueError("device tensors must be contiguous")
        return x.data_ptr(), int(x.is_cuda), x
    a = np.ascontiguousarray(x, dtype=dtype)
    return a.ctypes.data, 0, a


# ---------------------------------------------------------------------------
# ensemble, observations, observation operator
# ---------------------------------------------------------------------------
class Ensemble:
    """k members of an nx x ny x nlev state, member-major (Ensemble.hpp).
    `data` is a (k, n) float64 numpy array or torch tensor (device: updated in
    place by the analysis)."""

    def __init__(self, data, nx: int, ny: int, nlev: int = 1):
        if _is_torch(data):
            if data.dim() != 2:
                raise ValueError("ensemble tensor must be (k, n)")
            self.data = data
        else:
            self.data = np.ascontiguousarray(data, dtype=np.float64)
            if self.data.ndim != 2:
                raise ValueError("ensemble array must be (k, n)")
        k, n = self.data.shape
        if n != nx * ny * nlev:
            raise ValueError(f"ensemble has n={n}, geometry says {nx}*{ny}*{nlev}")
        self.nx, self.ny, self.nlev = nx, ny, nlev
        self._mean = None

    def Size(self) -> int:
        return int(self.data.shape[0])

    def GetMember(self, i: int):
        if i < 0 or i >= self.Size():
            raise IndexError("Member index out of range")  # Ensemble.hpp:77-79
        return self.data[i]

    def struct(self):
        ptr, dev, keep = _buf(self.data)
        s = L.Ensemble(self.Size(), self.nx, self.ny, self.nlev, ptr, None, dev)
        return s, keep

    def RecomputeMean(self, ctx: Optional[Context] = None):
        ctx = ctx or default_context()
        s, keep = self.struct()
        if s.on_device:
            import torch
            out = torch.empty(self.data.shape[1], dtype=torch.float64, device=self.data.device)
            ctx.check(ctx.lib.mda_ensemble_mean_perturbations(ctx.h, C.byref(s), 1.0,
                                                              out.data_ptr(), None, 1))
        else:
            out = np.empty(self.data.shape[1])
            ctx.check(ctx.lib.mda_ensemble_mean_perturbations(ctx.h, C.byref(s), 1.0,
                                                              out.ctypes.data, None, 0))
        self._mean = out
        return out

    def Mean(self, ctx: Optional[Context] = None):
        if self._mean is None:
            return self.RecomputeMean(ctx)
        return self._mean

    def Perturbations(self, scale: float = 1.0, ctx: Optional[Context] = None):
        ctx = ctx or default_context()
        s, keep = self.struct()
        if s.on_device:
            import torch
            out = torch.empty_like(self.data)
            ctx.check(ctx.lib.mda_ensemble_mean_perturbations(ctx.h, C.byref(s), scale, None,
                                                              out.data_ptr(), 1))
        else:
            out = np.empty_like(self.data)
            ctx.check(ctx.lib.mda_ensemble_mean_perturbations(ctx.h, C.byref(s), scale, None,
                                                              out.ctypes.data, 0))
        return out

    def to_host(self) -> np.ndarray:
        return self.data.detach().cpu().numpy() if _is_torch(self.data) else self.data


@dataclass
class Observation:
    """Point observations (GridObservation / PointObservation): values, R
    diagonal (variances, +inf marks invalid), locations in the coordinate
    system `coord` (grid x,y | latitude,longitude | x,y,z)."""
    values: object
    variances: object
    loc_a: object
    loc_b: object
    loc_c: object = None
    coord: int = L.COORD_GRID

    def size(self) -> int:
        return int(len(self.values))

    def getObservationValues(self):
        return self.values

    def getCovariance(self):
        return self.variances

    @classmethod
    def from_grid_file(cls, path: str, error: float, missing_value: float = -999.0,
                       coordinate: str = "grid") -> "Observation":
        """GridObservation::loadFromFile (GridObservation.hpp:401-476) through
        the library (mda_obs_read_grid_file): the 'Time Z Y X ZNU XLONG_U
        XLAT_U Value' text format; values equal to missing_value are dropped;
        R diagonal = error^2, with error as given (the reference's config layer
        hands it over narrowed through float, GridObservation.hpp:85-86; pass
        Config(...).Get("error").asFloat() for that)."""
        coords = {"grid": L.COORD_GRID, "geographic": L.COORD_GEOGRAPHIC}
        if coordinate not in coords:
            raise RuntimeError(f"Unsupported coordinate system: {coordinate}")
        lib = L.load()
        h = C.c_void_p()
        rc = lib.mda_obs_read_grid_file(None, os.fsencode(path), float(error),
                                        float(missing_value), coords[coordinate], C.byref(h))
        if rc != L.MDA_OK:
            raise RuntimeError(lib.mda_last_io_error().decode())
        try:
            v = L.Obs()
            cs = C.c_int32()
            lib.mda_obs_set_view(h, C.byref(v), C.byref(cs))
            p = v.p
            arr = lambda ptr: np.ctypeslib.as_array(
                C.cast(ptr, C.POINTER(C.c_double)), (p,)).copy()
            return cls(arr(v.value), arr(v.variance), arr(v.loc_a), arr(v.loc_b), arr(v.loc_c),
                       cs.value)
        finally:
            lib.mda_obs_set_free(h)


class ObsOperator:
    """A linear H in CSR form over state linear indices, or a precomputed
    H(x_j) (member-major [k][p]) for operators that exist only on the host."""

    def __init__(self, row_ptr=None, col=None, weight=None, hx=None):
        if hx is None and (row_ptr is None or col is None or weight is None):
            raise ValueError("ObsOperator needs a CSR stencil or hx")
        self.row_ptr, self.col, self.weight, self.hx = row_ptr, col, weight, hx

    def struct(self):
        if self.hx is not None:
            p, dev, keep = _buf(self.hx)
            return L.Hop(L.HOP_HX, None, None, None, 0, p, dev), [keep]
        rp, dev, k1 = _buf(self.row_ptr, np.int64)
        cp, _, k2 = _buf(self.col, np.int64)
        wp, _, k3 = _buf(self.weight)
        return L.Hop(L.HOP_CSR, rp, cp, wp, int(len(self.col)), None, dev), [k1, k2, k3]

    def apply_host(self, X: np.ndarray) -> np.ndarray:
        """H x_j for every member on the host (member-major [k][p])."""
        if self.hx is not None:
            return np.asarray(self.hx)
        rp, col, w = (np.asarray(v) for v in (self.row_ptr, self.col, self.weight))
        rows = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
        out = np.zeros((X.shape[0], len(rp) - 1))
        for j in range(X.shape[0]):
            np.add.at(out[j], rows, w * X[j, col])
        return out

    @classmethod
    def identity_idw4(cls, obs: Observation, nx: int, ny: int, nz: int = 1,
                      geometry: "ColumnGeometry" = None, vcoord=None,
                      ctx: Optional[Context] = None) -> "ObsOperator":
        """IdentityObsOperator (IdentityObsOperator.hpp:154-281,483-676) as a
        CSR stencil built on the device (mda_hop_idw4_build): the four
        neighbours find4NearestGridPoints selects, weights 1/dist (1e12 on an
        exact hit) normalised; GEOGRAPHIC observations mapped to the nearest
        grid point of `geometry` (nearest level of vcoord), or to (0, 0, 0)
        without geographic column coordinates; invalid observations (+inf
        variance) get empty rows.  The stencil comes back to the host."""
        ctx = ctx or default_context()
        os_, ok = _obs_struct(obs)
        geom = geometry or ColumnGeometry(L.COORD_GRID)
        gs, gk = geom.struct()
        p = obs.size()
        rp = np.zeros(p + 1, np.int64)
        col = np.zeros(4 * p, np.int64)
        w = np.zeros(4 * p)
        vp, _, vk = _buf(vcoord)
        ctx.check(ctx.lib.mda_hop_idw4_build(ctx.h, C.byref(os_), obs.coord, C.byref(gs), nx, ny,
                                             nz, vp, rp.ctypes.data, col.ctypes.data,
                                             w.ctypes.data, 0))
        nnz = int(rp[-1])
        return cls(rp, col[:nnz].copy(), w[:nnz].copy())

    @classmethod
    def trilinear_latlon(cls, lat, lon, lev, nlon: int, nlat: int, nlev: int) -> "ObsOperator":
        """Trilinear interpolation on the global regular grid with centres at
        lon=(i+1/2)*360/nlon, lat=-90+(j+1/2)*180/nlat, periodic in longitude,
        clamped at the outermost latitude rows and levels (8 taps)."""
        lat, lon, lev = (np.asarray(v, dtype=np.float64) for v in (lat, lon, lev))
        fi = np.mod(lon, 360.0) / (360.0 / nlon) - 0.5
        fj = (lat + 90.0) / (180.0 / nlat) - 0.5
        fj = np.clip(fj, 0.0, nlat - 1.0)
        fk = np.clip(lev, 0.0, nlev - 1.0)
        i0 = np.floor(fi).astype(np.int64); ti = fi - i0
        j0 = np.minimum(np.floor(fj).astype(np.int64), nlat - 2 if nlat > 1 else 0); tj = fj - j0
        k0 = np.minimum(np.floor(fk).astype(np.int64), nlev - 2 if nlev > 1 else 0); tk = fk - k0
        i0m, i1m = np.mod(i0, nlon), np.mod(i0 + 1, nlon)
        j1 = np.minimum(j0 + 1, nlat - 1); k1 = np.minimum(k0 + 1, nlev - 1)
        p = len(lat)
        cols = np.empty((p, 8), np.int64); ws = np.empty((p, 8))
        t = 0
        for kk, wk in ((k0, 1 - tk), (k1, tk)):
            for jj, wj in ((j0, 1 - tj), (j1, tj)):
                for ii, wi in ((i0m, 1 - ti), (i1m, ti)):
                    cols[:, t] = kk * nlat * nlon + jj * nlon + ii
                    ws[:, t] = wk * wj * wi
                    t += 1
        rp = np.arange(0, 8 * p + 1, 8, dtype=np.int64)
        return cls(rp, cols.ravel(), ws.ravel())


@dataclass
class ColumnGeometry:
    """Horizontal column locations used for localisation (Location.hpp)."""
    coord: int = L.COORD_GRID
    col_a: object = None
    col_b: object = None
    col_c: object = None

    def struct(self):
        a, dev, k1 = _buf(self.col_a)
        b, _, k2 = _buf(self.col_b)
        c, _, k3 = _buf(self.col_c)
        return L.Geometry(self.coord, a, b, c, dev), [k1, k2, k3]

    @classmethod
    def latlon(cls, nlon: int, nlat: int, device=None):
        lon = (np.arange(nlon) + 0.5) * 360.0 / nlon
        lat = -90.0 + (np.arange(nlat) + 0.5) * 180.0 / nlat
        LON, LAT = np.meshgrid(lon, lat)
        a, b = LAT.ravel().copy(), LON.ravel().copy()
        if device is not None:
            import torch
            a = torch.from_numpy(a).to(device)
            b = torch.from_numpy(b).to(device)
        return cls(L.COORD_GEOGRAPHIC, a, b, None)


def _obs_struct(obs: Observation):
    v, dev, k1 = _buf(obs.values)
    r, _, k2 = _buf(obs.variances)
    a, _, k3 = _buf(obs.loc_a)
    b, _, k4 = _buf(obs.loc_b)
    c, _, k5 = _buf(obs.loc_c)
    return L.Obs(obs.size(), v, r, a, b, c, dev), [k1, k2, k3, k4, k5]


# ---------------------------------------------------------------------------
# algorithms
# ---------------------------------------------------------------------------
def _write_state(path: str, values: np.ndarray, nx: int, nrows: int):
    """SimpleState::saveToFile (SimpleState.hpp:162-188) through the library:
    nrows rows of nx values, '%12.6f' separated by one space."""
    d = os.path.dirname(path)
    if d:
        os.makedirs(d, exist_ok=True)
    v = np.ascontiguousarray(values, np.float64)
    lib = L.load()
    if lib.mda_state_write_text(None, os.fsencode(path), v.ctypes.data, nx, nrows) != L.MDA_OK:
        raise RuntimeError(lib.mda_last_io_error().decode())


class _AlgorithmBase:
    def __init__(self, ensemble: Ensemble, obs: Observation, obs_op: ObsOperator,
                 ctx: Optional[Context]):
        self.ensemble, self.obs, self.obs_op = ensemble, obs, obs_op
        self.ctx = ctx or default_context()
        self.output_base_file = "analysis"
        self.format = "txt"

    def saveEnsemble(self, sidecar: bool = False):
        """saveEnsemble (EnKF.hpp:261-282): <base>_mean.<format> and
        <base>_member_<i>.<format> in the 6-decimal text format, plus FP64-exact
        .f64 sidecars with sidecar=True (mda_ensemble_save)."""
        d = os.path.dirname(self.output_base_file)
        if d:
            os.makedirs(d, exist_ok=True)
        es, ek = self.ensemble.struct()
        self.ctx.check(self.ctx.lib.mda_ensemble_save(self.ctx.h, C.byref(es),
                                                      os.fsencode(self.output_base_file),
                                                      self.format.encode(), int(sidecar)))


class EnKF(_AlgorithmBase):
    """EnKF<Tag> (EnKF.hpp:105-301).  Reads the `analysis` subsection of the
    root config itself (EnKF.hpp:109).  The N(0,1) draws the reference takes
    from std::random_device are passed in (`Z`, shape (p, k)) or drawn from a
    seeded numpy Generator."""

    METHODS = {"multiplicative": L.INFL_MULTIPLICATIVE, "additive": L.INFL_ADDITIVE,
               "relaxation": L.INFL_RELAXATION}

    def __init__(self, ensemble, obs, obs_op, config: Config, ctx=None, seed: int = 0):
        super().__init__(ensemble, obs, obs_op, ctx)
        ac = config.GetSubsection("analysis")
        self.inflation_factor = ac.Get("inflation").asFloat()
        self.output_base_file = ac.Get("output_base_file").v or "analysis"
        self.format = ac.Get("format").v or "txt"
        m = ac.Get("inflation_method").v
        self.inflation_method = self.METHODS.get(m, L.INFL_MULTIPLICATIVE)
        self.inflation_method_name = m if m in self.METHODS else "multiplicative"
        self.rng = np.random.default_rng(seed)
        self._diag = None

    def Analyse(self, Z=None):
        k, p = self.ensemble.Size(), self.obs.size()
        if Z is None:
            Z = self.rng.standard_normal((p, k))
        es, ek = self.ensemble.struct()
        os_, ok = _obs_struct(self.obs)
        hs, hk = self.obs_op.struct()
        zp, zdev, zk = _buf(Z)
        dg = L.EnkfDiag()
        self.ctx.check(self.ctx.lib.mda_enkf_analyse(
            self.ctx.h, self.inflation_factor, self.inflation_method, C.byref(es),
            C.byref(os_), C.byref(hs), zp, zdev, C.byref(dg)))
        self._diag = dg

    def getAnalysisResults(self) -> dict:
        if self._diag is None:
            raise RuntimeError("Analyse() has not been called")
        d = {f: getattr(self._diag, f) for f, _ in L.EnkfDiag._fields_}
        d["inflation_method"] = self.inflation_method_name
        return d


class LWEnKF(_AlgorithmBase):
    """LWEnKF<Tag> (LWEnKF.hpp:207-334).  Reads the `analysis` subsection of
    the root config (LWEnKF.hpp:103-145): inflation, localization_radius,
    inflation_method, localization_function and weighting_scheme, unknown
    names falling back to multiplicative / gaussian / uniform as the
    reference's do.  Every inflation method scales X' by sqrt(inflation)
    (LWEnKF.hpp:648-668).  gc_form "reference" keeps the reference's
    "gaspari_cohn" polynomial (4 at distance 0); "standard" evaluates Gaspari
    & Cohn (1999) eq. (4.10) in its place.  The N(0,1) draws the reference
    takes from std::random_device are passed in (`Z`, shape (p, k)) or drawn
    from a seeded numpy Generator."""

    METHODS = EnKF.METHODS
    LOCS = {"gaussian": L.LW_LOC_GAUSSIAN, "exponential": L.LW_LOC_EXPONENTIAL,
            "cutoff": L.LW_LOC_CUTOFF, "gaspari_cohn": L.LW_LOC_GASPARI_COHN}
    WEIGHTS = {"uniform": L.LW_W_UNIFORM, "adaptive": L.LW_W_ADAPTIVE,
               "inverse_var": L.LW_W_INVERSE_VAR, "likelihood": L.LW_W_LIKELIHOOD}

    def __init__(self, ensemble, obs, obs_op, config: Config, ctx=None, seed: int = 0,
                 gc_form: str = "reference"):
        super().__init__(ensemble, obs, obs_op, ctx)
        ac = config.GetSubsection("analysis")
        self.inflation_factor = ac.Get("inflation").asFloat()
        self.localization_radius = ac.Get("localization_radius").asFloat()
        self.output_base_file = ac.Get("output_base_file").v or "analysis"
        self.format = ac.Get("format").v or "txt"
        m, f, w = (ac.Get(key).v for key in
                   ("inflation_method", "localization_function", "weighting_scheme"))
        self.inflation_method_name = m if m in self.METHODS else "multiplicative"
        self.localization_function = f if f in self.LOCS else "gaussian"
        self.weighting_scheme = w if w in self.WEIGHTS else "uniform"
        if gc_form not in ("reference", "standard"):
            raise ValueError("gc_form is 'reference' or 'standard'")
        self.gc_form = gc_form
        self.rng = np.random.default_rng(seed)
        self.weights = None
        self._diag = None

    def params(self) -> "L.LwenkfParams":
        return L.LwenkfParams(self.inflation_factor, self.localization_radius,
                              self.METHODS[self.inflation_method_name],
                              self.LOCS[self.localization_function],
                              L.LW_GC_STANDARD if self.gc_form == "standard"
                              else L.LW_GC_REFERENCE,
                              self.WEIGHTS[self.weighting_scheme])

    def Analyse(self, Z=None):
        k, p = self.ensemble.Size(), self.obs.size()
        if Z is None:
            Z = self.rng.standard_normal((p, k))
        es, ek = self.ensemble.struct()
        os_, ok = _obs_struct(self.obs)
        hs, hk = self.obs_op.struct()
        zp, zdev, zk = _buf(Z)
        prm = self.params()
        w = np.zeros(k)
        dg = L.LwenkfDiag()
        self.ctx.check(self.ctx.lib.mda_lwenkf_analyse(
            self.ctx.h, C.byref(prm), C.byref(es), C.byref(os_), C.byref(hs), zp, zdev,
            w.ctypes.data, C.byref(dg)))
        self.weights = w
        self._diag = dg
        self.ensemble._mean = None

    def getAnalysisResults(self) -> dict:
        """LWEnKF::AnalysisResults (LWEnKF.hpp:66-84, 340-362)."""
        if self._diag is None:
            raise RuntimeError("Analyse() has not been called")
        d = {f: getattr(self._diag, f) for f, _ in L.LwenkfDiag._fields_}
        d["inflation_method"] = self.inflation_method_name
        d["localization_function"] = self.localization_function
        d["weighting_scheme"] = self.weighting_scheme
        del d["localization"], d["weighting"], d["ms_gain_solve"], d["ms_update"]
        return d

    def stage_ms(self) -> dict:
        """Device time of the last Analyse()'s S^-T Y' solve and fused update."""
        if self._diag is None:
            raise RuntimeError("Analyse() has not been called")
        return {"ms_gain_solve": self._diag.ms_gain_solve, "ms_update": self._diag.ms_update}


class ETKF(_AlgorithmBase):
    """ETKF<Tag> (ETKF.hpp:86-178); config is the analysis subsection."""

    def __init__(self, ensemble, obs, obs_op, config: Config, ctx=None):
        super().__init__(ensemble, obs, obs_op, ctx)
        self.inflation = config.Get("inflation").asFloat()
        self.output_base_file = config.Get("output_base_file").v or "analysis"
        self.format = config.Get("format").v or "txt"

    def Analyse(self):
        es, ek = self.ensemble.struct()
        os_, ok = _obs_struct(self.obs)
        hs, hk = self.obs_op.struct()
        self.ctx.check(self.ctx.lib.mda_etkf_analyse(self.ctx.h, self.inflation, C.byref(es),
                                                     C.byref(os_), C.byref(hs)))


class LETKF(_AlgorithmBase):
    """LETKF<Tag> (LETKF.hpp:47-243); config is the analysis subsection
    (letkf.cpp:62-63).  `form` selects the standard Hunt et al. (2007) LETKF
    with Gaspari-Cohn R-localisation ("standard", the default here) or the
    reference's own semantics with snapshot H ("refcompat": R = I, hard cutoff,
    Pa *= rho, Cholesky square root, element-wise combination)."""

    def __init__(self, ensemble, obs, obs_op, config: Config, geometry: ColumnGeometry = None,
                 ctx=None, form: str = "standard", precision: int = L.PREC_FP64,
                 rho_column=None, eig_method: int = L.EIG_AUTO,
                 transform_precision: int = L.PREC_FP64, keep_transforms: bool = False):
        super().__init__(ensemble, obs, obs_op, ctx)
        self.inflation = config.Get("inflation").asFloat()
        self.localization_radius = config.Get("localization_radius").asFloat()
        self.output_base_file = config.Get("output_base_file").v or "analysis"
        self.format = config.Get("format").v or "nc"
        self.geometry = geometry or ColumnGeometry(L.COORD_GRID)
        self.params = L.LetkfParams()
        if form == "standard":
            self.ctx.lib.mda_letkf_params_standard(C.byref(self.params), self.inflation,
                                                   self.localization_radius)
        elif form == "refcompat":
            self.ctx.lib.mda_letkf_params_refcompat(C.byref(self.params), self.inflation,
                                                    self.localization_radius)
        else:
            raise ValueError(f"unknown LETKF form {form!r}")
        self.params.precision = precision
        self.params.eig_method = eig_method
        self.params.transform_precision = transform_precision
        # transforms() needs T kept; otherwise the eigenvector kernel may apply
        # X_a = X T' itself without storing T
        self.params.keep_transforms = 1 if keep_transforms else 0
        self._rho_keep = None
        if rho_column is not None:
            rp, rdev, self._rho_keep = _buf(rho_column)
            self.params.rho_column = rp
            self.params.rho_on_device = rdev
        self.diag = None

    def Analyse(self):
        es, ek = self.ensemble.struct()
        gs, gk = self.geometry.struct()
        os_, ok = _obs_struct(self.obs)
        hs, hk = self.obs_op.struct()
        dg = L.LetkfDiag()
        self.ctx.check(self.ctx.lib.mda_letkf_analyse(self.ctx.h, C.byref(self.params),
                                                      C.byref(es), C.byref(gs), C.byref(os_),
                                                      C.byref(hs), C.byref(dg)))
        self.diag = {f: getattr(dg, f) for f, _ in L.LetkfDiag._fields_}
        return self.diag

    def transforms(self, cols: Sequence[int]) -> np.ndarray:
        cols = np.ascontiguousarray(cols, np.int64)
        k = self.ensemble.Size()
        out = np.empty((len(cols), k, k))
        self.ctx.check(self.ctx.lib.mda_letkf_last_transforms(self.ctx.h, cols.ctypes.data,
                                                              len(cols), out.ctypes.data))
        return out


def _local_counts(self) -> np.ndarray:
    out = np.empty(self.ensemble.nx * self.ensemble.ny, np.int32)
    self.ctx.check(self.ctx.lib.mda_letkf_last_local_counts(self.ctx.h, out.ctypes.data))
    return out


def _column_flags(self) -> np.ndarray:
    out = np.empty(self.ensemble.nx * self.ensemble.ny, np.int32)
    self.ctx.check(self.ctx.lib.mda_letkf_last_column_flags(self.ctx.h, out.ctypes.data))
    return out


LETKF.local_counts = _local_counts
LETKF.column_flags = _column_flags


def metrics(ensemble: Ensemble, truth, ctx: Optional[Context] = None) -> dict:
    """Metrics::CalculateAll scalars (Metrics.hpp:74-94) on the device."""
    ctx = ctx or default_context()
    es, ek = ensemble.struct()
    tp, tdev, tk = _buf(truth)
    out = L.Metrics()
    ctx.check(ctx.lib.mda_ensemble_metrics(ctx.h, C.byref(es), tp, tdev, C.byref(out)))
    return {f: getattr(out, f) for f, _ in L.Metrics._fields_}


def read_simple_state(path: str) -> np.ndarray:
    """SimpleState::readFromFile (SimpleState.hpp:248-263) through the
    library: all whitespace-separated numbers in file order."""
    lib = L.load()
    n = C.c_int64(0)
    if lib.mda_state_read_text(None, os.fsencode(path), None, 0, C.byref(n)) != L.MDA_OK:
        raise RuntimeError(lib.mda_last_io_error().decode())
    out = np.empty(n.value)
    if lib.mda_state_read_text(None, os.fsencode(path), out.ctypes.data, n.value,
                               C.byref(n)) != L.MDA_OK:
        raise RuntimeError(lib.mda_last_io_error().decode())
    return out


def read_state_sidecar(path: str):
    """An FP64-exact state sidecar (mda_state_read_binary): (values, nx, ny, nlev)."""
    lib = L.load()
    nx, ny, nl = C.c_int64(), C.c_int64(), C.c_int64()
    if lib.mda_state_read_binary(None, os.fsencode(path), None, 0, C.byref(nx), C.byref(ny),
                                 C.byref(nl)) != L.MDA_OK:
        raise RuntimeError(lib.mda_last_io_error().decode())
    out = np.empty(nx.value * ny.value * nl.value)
    if lib.mda_state_read_binary(None, os.fsencode(path), out.ctypes.data, out.size,
                                 C.byref(nx), C.byref(ny), C.byref(nl)) != L.MDA_OK:
        raise RuntimeError(lib.mda_last_io_error().decode())
    return out, nx.value, ny.value, nl.value


def write_state_sidecar(path: str, values, nx: int, ny: int, nlev: int = 1):
    """Writes an FP64-exact state sidecar (mda_state_write_binary)."""
    lib = L.load()
    v = np.ascontiguousarray(values, np.float64)
    if v.size != nx * ny * nlev:
        raise ValueError("sidecar values do not match nx * ny * nlev")
    if lib.mda_state_write_binary(None, os.fsencode(path), v.ctypes.data, nx, ny,
                                  nlev) != L.MDA_OK:
        raise RuntimeError(lib.mda_last_io_error().decode())


def letkf_batch(filters, flags=L.BATCH_GRAPH | L.BATCH_QUEUE | L.BATCH_AUTOTUNE):
    """Runs independent LETKF analyses (e.g. different ensemble sizes) as one
    batch (mda_letkf_analyse_batch): concurrent branches of one hipGraph,
    work-queue Gram, autotuned launch.  Inputs must be device-resident.
    Returns the per-analysis diagnostics and the batch info."""
    ctx = filters[0].ctx
    n = len(filters)
    P = (L.LetkfParams * n)(*[f.params for f in filters])
    E, G, O, H = (L.Ensemble * n)(), (L.Geometry * n)(), (L.Obs * n)(), (L.Hop * n)()
    keep = []
    for i, f in enumerate(filters):
        assert f.ctx is ctx, "all analyses of a batch share one context"
        es, ek = f.ensemble.struct()
        gs, gk = f.geometry.struct()
        os_, ok = _obs_struct(f.obs)
        hs, hk = f.obs_op.struct()
        E[i], G[i], O[i], H[i] = es, gs, os_, hs
        keep += [ek, gk, ok, hk]
    D = (L.LetkfDiag * n)()
    info = L.BatchInfo()
    ctx.check(ctx.lib.mda_letkf_analyse_batch(ctx.h, n, P, E, G, O, H, D, flags, C.byref(info)))
    del keep
    diags = [{fld: getattr(D[i], fld) for fld, _ in L.LetkfDiag._fields_} for i in range(n)]
    for f, d in zip(filters, diags):
        f.diag = d
    return diags, {"ms_total": info.ms_total, "graph_replayed": info.graph_replayed,
                   "n_graphs": info.n_graphs, "gram_blocks": list(info.gram_blocks[:n]),
                   "gram_queue": list(info.gram_queue[:n])}
