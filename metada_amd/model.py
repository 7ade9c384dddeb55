"""The forecast step between analyses: Model<BackendTag> (Model.hpp:57-177)
over a Lorenz-96 backend that runs in the library (mda_l96_forecast).

Lorenz96Model is a model backend with the members ModelBackendImpl asks for
(ModelConcepts.hpp:51-71): construction from a config, initialize, reset,
finalize, isInitialized and run(initial, final).  Model forwards to a backend
and wraps its errors as Model.hpp does ("Model run failed: ...").

Config keys, each optional: forcing (8.0, lorenz95.py:11), dt (0.05),
num_steps (1), scheme ("euler": the forward step of lorenz95.py:21-31 with
its clipping to +-1e10; "rk4": classical Runge-Kutta).  Reals come
float-narrowed, as every config real does (ConfigValue.hpp:108).
"""
from __future__ import annotations

from typing import Optional

import numpy as np

from .analysis import Config, Context, Ensemble, _is_torch, default_context

SCHEMES = {"rk4": 0, "euler": 1}


class Lorenz96Model:
    """dx_i/dt = (x_{i+1} - x_{i-2}) x_{i-1} - x_i + F on a ring of N points
    (lorenz95.py:3-37), 4 <= N <= 6144, one state or a whole ensemble per
    launch."""

    def __init__(self, config: Config, ctx: Optional[Context] = None):
        self._ctx = ctx
        self._initialized = False
        self.initialize(config)

    @property
    def ctx(self) -> Context:
        if self._ctx is None:
            self._ctx = default_context()
        return self._ctx

    def initialize(self, config: Config):
        if self._initialized:
            return

        def real(key, default):
            if not config.Has(key):
                return default
            v = config.Get(key)
            return float(v.v) if isinstance(v.v, int) and not isinstance(v.v, bool) \
                else v.asFloat()
        self.forcing = real("forcing", 8.0)
        self.dt = real("dt", 0.05)
        self.num_steps = config.Get("num_steps").asInt() if config.Has("num_steps") else 1
        scheme = config.Get("scheme").asString() if config.Has("scheme") else "euler"
        if scheme not in SCHEMES:
            raise RuntimeError(f"Lorenz96Model: unknown scheme '{scheme}' (euler or rk4)")
        if self.num_steps < 0:
            raise RuntimeError("Lorenz96Model: num_steps must be >= 0")
        if not self.dt > 0.0:
            raise RuntimeError("Lorenz96Model: dt must be positive")
        self.scheme = scheme
        self._initialized = True

    def isInitialized(self) -> bool:
        return self._initialized

    def reset(self):
        if not self._initialized:
            raise RuntimeError("Lorenz96Model not initialized")

    def finalize(self):
        self._initialized = False

    def _advance(self, X, nsteps=None):
        """X: (nstates, N) float64, numpy (host) or torch (device), in place."""
        nsteps = self.num_steps if nsteps is None else nsteps
        if nsteps == 0:
            return
        if _is_torch(X):
            if not X.is_contiguous():
                raise ValueError("device states must be contiguous")
            ptr, dev = X.data_ptr(), int(X.is_cuda)
        else:
            if not (X.flags.c_contiguous and X.dtype == np.float64):
                raise ValueError("host states must be contiguous float64")
            ptr, dev = X.ctypes.data, 0
        self.ctx.check(self.ctx.lib.mda_l96_forecast(self.ctx.h, ptr, X.shape[0], X.shape[1],
                                                     self.forcing, self.dt, nsteps,
                                                     SCHEMES[self.scheme], dev))

    def run(self, initial, final):
        """final = M^num_steps(initial); 1-D states of equal size (host or
        device); final may be initial."""
        if not self._initialized:
            raise RuntimeError("Lorenz96Model not initialized")
        if initial.shape != final.shape or initial.ndim != 1:
            raise ValueError("states must be 1-D and of equal size")
        contiguous = final.is_contiguous() if _is_torch(final) else final.flags.c_contiguous
        if not contiguous:
            raise ValueError("the final state must be contiguous (it is advanced in place)")
        if final is not initial:
            final[...] = initial
        self._advance(final.reshape(1, -1))

    def runEnsemble(self, ensemble: Ensemble):
        """Every member of the ensemble, in place, in one launch."""
        if not self._initialized:
            raise RuntimeError("Lorenz96Model not initialized")
        self._advance(ensemble.data)
        ensemble._mean = None


class Model:
    """Model<BackendTag> (Model.hpp:57-177): forwards to its backend."""

    def __init__(self, config: Config, backend=Lorenz96Model, ctx: Optional[Context] = None):
        self._backend = backend(config, ctx=ctx)

    def initialize(self, config: Config):
        if self._backend.isInitialized():
            return
        try:
            self._backend.initialize(config)
        except Exception as e:  # Model.hpp:116-122
            raise RuntimeError(f"Model initialization failed: {e}") from e

    def isInitialized(self) -> bool:
        return self._backend.isInitialized()

    def reset(self):
        if not self._backend.isInitialized():
            raise RuntimeError("Model not initialized")
        try:
            self._backend.reset()
        except Exception as e:
            raise RuntimeError(f"Model reset failed: {e}") from e

    def finalize(self):
        if not self._backend.isInitialized():
            return
        try:
            self._backend.finalize()
        except Exception as e:
            raise RuntimeError(f"Model finalization failed: {e}") from e

    def run(self, initial, final):
        if not self._backend.isInitialized():
            raise RuntimeError("Model not initialized")
        try:
            self._backend.run(initial, final)
        except Exception as e:
            raise RuntimeError(f"Model run failed: {e}") from e

    def backend(self):
        return self._backend
