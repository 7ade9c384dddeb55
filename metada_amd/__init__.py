"""metada_amd: MI355X (gfx950) ensemble Kalman analysis (EnKF / ETKF / LETKF).

The arithmetic lives in libmetada_amd.so (HIP kernels behind the C-ABI in
include/metada_amd.h); this package is the host-side mirror of metada's
analysis interface on top of it.  Importing it does not touch the GPU.
"""
from . import _lib
from .model import Lorenz96Model, Model
from .analysis import (ColumnGeometry, Config, ConfigValue, Context, EnKF, Ensemble, ETKF,
                       LETKF, LWEnKF, ObsOperator, Observation, default_context, letkf_batch,
                       metrics, read_simple_state, read_state_sidecar, write_state_sidecar)

__all__ = ["ColumnGeometry", "Config", "ConfigValue", "Context", "EnKF", "Ensemble", "ETKF",
           "LETKF", "LWEnKF", "Lorenz96Model", "Model", "ObsOperator", "Observation", "default_context",
           "letkf_batch", "metrics", "read_simple_state", "read_state_sidecar",
           "write_state_sidecar", "_lib"]
