"""Config handling and lifecycle of the Lorenz-96 model backend and the Model
wrapper (ModelConcepts.hpp:51-71, Model.hpp:57-177) -- no GPU call."""
import numpy as np
import pytest


def test_lorenz96_config_and_lifecycle():
    from metada_amd import Config, Lorenz96Model, Model
    m = Lorenz96Model(Config({"forcing": 10, "dt": 0.01, "num_steps": 7, "scheme": "rk4"}))
    assert m.forcing == 10.0 and m.num_steps == 7 and m.scheme == "rk4"
    assert m.dt == float(np.float32(0.01))  # config reals are float-narrowed
    d = Lorenz96Model(Config({}))
    assert (d.forcing, d.dt, d.num_steps, d.scheme) == (8.0, 0.05, 1, "euler")
    with pytest.raises(RuntimeError, match="unknown scheme"):
        Lorenz96Model(Config({"scheme": "leapfrog"}))
    with pytest.raises(RuntimeError, match="num_steps"):
        Lorenz96Model(Config({"num_steps": -1}))
    with pytest.raises(RuntimeError, match="dt must be positive"):
        Lorenz96Model(Config({"dt": 0.0}))
    with pytest.raises(TypeError):
        Lorenz96Model(Config({"num_steps": 2.5}))  # asInt on a real
    model = Model(Config({"scheme": "euler"}))
    assert model.isInitialized()
    model.reset()
    model.finalize()
    assert not model.isInitialized()
    with pytest.raises(RuntimeError, match="Model not initialized"):
        model.run(np.zeros(8), np.zeros(8))
    with pytest.raises(RuntimeError, match="Model not initialized"):
        model.reset()
    model.initialize(Config({"scheme": "rk4"}))
    assert model.isInitialized() and model.backend().scheme == "rk4"


def test_model_run_rejects_bad_states():
    from metada_amd import Config, Model
    model = Model(Config({}))
    with pytest.raises(RuntimeError, match="Model run failed: states must be 1-D"):
        model.run(np.zeros(8), np.zeros(9))
    X = np.zeros((8, 2))
    with pytest.raises(RuntimeError, match="Model run failed: the final state must be contiguous"):
        model.run(np.zeros(8), X[:, 0])
