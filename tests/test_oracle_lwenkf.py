"""Pins of the oracle's LWEnKF restatement (LWEnKF.hpp:207-334, 413-668)
by known answers and properties that do not restate its formulas: the
special values of the reference's taper polynomial (4 at 0, -0.125 below 1,
0.125 at 1, 1.0 below 2), a k = 2, p = 1 case worked by hand, weight ratios
of the adaptive, inverse-variance and likelihood schemes, the support of the
gain taper, cond(S) of a diagonal S and the large-R limit."""
import math

import numpy as np
import pytest


def test_taper_special_values(orc):
    t = orc.lw_taper
    below1, below2 = np.nextafter(1.0, 0.0), np.nextafter(2.0, 0.0)
    # the reference's "gaspari_cohn" polynomial
    assert t(0.0, 1.0, "gaspari_cohn") == 4.0
    assert abs(t(below1, 1.0, "gaspari_cohn") - (-0.125)) < 1e-12
    assert t(1.0, 1.0, "gaspari_cohn") == 0.125
    assert abs(t(below2, 1.0, "gaspari_cohn") - 1.0) < 1e-12
    assert t(2.0, 1.0, "gaspari_cohn") == 0.0
    # Gaspari & Cohn (1999) eq. (4.10): 1, 5/24, 0
    assert t(0.0, 1.0, "gaspari_cohn", gc_standard=True) == 1.0
    assert abs(t(1.0, 1.0, "gaspari_cohn", gc_standard=True) - 5.0 / 24.0) < 1e-15
    assert t(2.0, 1.0, "gaspari_cohn", gc_standard=True) == 0.0
    # the distance is normalised by the radius
    assert t(0.5, 0.5, "gaussian") == math.exp(-0.5)
    assert t(0.25, 0.5, "exponential") == math.exp(-0.5)
    assert t(0.5, 0.5, "cutoff") == 1.0 and t(np.nextafter(0.5, 1.0), 0.5, "cutoff") == 0.0


def test_k2_p1_worked_by_hand(orc):
    """x = (1, 3), H = I, yo = 2.5, r = 1, z = (0.5, -1), uniform weights:
    xbar = 2, X' = (-1, 1), Pf = (1 + 1)/2 = 1, S = 1 + 1 = 2,
    K = ((-1)(-1) + (1)(1)) / 2 / (k-1) = 1, D = (2.5 + 0.5 - 1, 2.5 - 1 - 3)
    = (2, -1.5), x_a = (2 - 1 + 2, 2 + 1 - 1.5) = (3, 1.5).  With rho = 4
    only X' is inflated: K = ((-2)(-1) + (2)(1)) / 2 = 2 and
    x_a = (2 - 2 + 4, 2 + 2 - 3) = (4, 1)."""
    X = np.array([[1.0], [3.0]])
    Yb = X.copy()
    Z = np.array([[0.5, -1.0]])
    Xa, w, d, K = orc.lwenkf_analyse(X, Yb, [2.5], [1.0], Z, 1.0, 0.5, want_K=True)
    np.testing.assert_array_equal(Xa[:, 0], [3.0, 1.5])
    np.testing.assert_array_equal(w, [0.5, 0.5])
    assert K[0, 0] == 1.0 and d["condition_number"] == 1.0
    assert d["innovation_norm"] == 0.5 and d["background_spread"] == 1.0
    assert d["analysis_spread"] == 0.75
    assert d["max_weight"] == d["min_weight"] == 0.5 and d["weight_variance"] == 0.0
    Xa, _, d, K = orc.lwenkf_analyse(X, Yb, [2.5], [1.0], Z, 4.0, 0.5, want_K=True)
    np.testing.assert_array_equal(Xa[:, 0], [4.0, 1.0])
    assert K[0, 0] == 2.0 and d["background_spread"] == 2.0


def test_weight_schemes(orc):
    # perturbations (-2, -1, 3): norms 2, 1, 3
    X = np.array([[1.0], [2.0], [6.0]])
    Yb = X.copy()
    Z = np.zeros((1, 3))
    args = (X, Yb, [2.0], [1.0], Z, 1.0, 1.0)
    _, w, _, _ = orc.lwenkf_analyse(*args, weighting="adaptive")
    np.testing.assert_allclose(w / w[1], [0.5, 1.0, 1.0 / 3.0], rtol=1e-7)
    assert abs(w.sum() - 1.0) < 1e-15
    _, w, _, _ = orc.lwenkf_analyse(*args, weighting="inverse_var")
    np.testing.assert_allclose(w / w[1], [0.25, 1.0, 1.0 / 9.0], rtol=1e-7)
    # likelihood: chi2 = (2-1)^2, (2-2)^2, (2-6)^2 = 1, 0, 16
    _, w, _, _ = orc.lwenkf_analyse(*args, weighting="likelihood")
    np.testing.assert_allclose(w / w[1], [math.exp(-0.5), 1.0, math.exp(-8.0)], rtol=1e-14)
    _, w, _, _ = orc.lwenkf_analyse(*args, weighting="uniform")
    np.testing.assert_allclose(w, [1 / 3] * 3, rtol=1e-15)
    # members equally far from the observations share the weight
    X2 = np.array([[0.0], [2.0]])
    _, w, _, _ = orc.lwenkf_analyse(X2, X2.copy(), [1.0], [1.0], np.zeros((1, 2)), 1.0, 1.0,
                                    weighting="likelihood")
    np.testing.assert_array_equal(w, [0.5, 0.5])


def test_weights_change_the_covariance(orc):
    """Non-uniform weights change S and hence the analysis (unless every
    perturbation has the same norm)."""
    rng = np.random.default_rng(3)
    X = rng.standard_normal((6, 8)) * np.arange(1, 7)[:, None]
    Yb = X[:, ::2].copy()
    yo, var, Z = rng.standard_normal(4), np.ones(4), rng.standard_normal((4, 6))
    a, _, _, _ = orc.lwenkf_analyse(X, Yb, yo, var, Z, 1.0, 0.5, weighting="uniform")
    b, _, _, _ = orc.lwenkf_analyse(X, Yb, yo, var, Z, 1.0, 0.5, weighting="inverse_var")
    assert np.max(np.abs(a - b)) > 1e-3


def test_gain_taper_support(orc):
    """cutoff with radius 0.1 on n = p = 3: |i - j| / 3 / 0.1 > 1 off the
    diagonal, so K is diagonal, and observation b moves state b only; S is
    diagonal too, with cond(S) = max / min of Pf_aa + r_a."""
    rng = np.random.default_rng(7)
    k = 5
    X = rng.standard_normal((k, 3))
    Yb = X.copy()
    yo, var, Z = rng.standard_normal(3), np.array([0.5, 1.0, 2.0]), rng.standard_normal((3, k))
    Xa, w, d, K = orc.lwenkf_analyse(X, Yb, yo, var, Z, 1.0, 0.1, fn="cutoff", want_K=True)
    assert np.all(K[~np.eye(3, dtype=bool)] == 0.0)
    yo2 = yo.copy()
    yo2[2] += 5.0
    Xb, _, _, _ = orc.lwenkf_analyse(X, Yb, yo2, var, Z, 1.0, 0.1, fn="cutoff")
    np.testing.assert_array_equal(Xb[:, :2], Xa[:, :2])
    assert np.all(Xb[:, 2] != Xa[:, 2])
    Yp = Yb - Yb.mean(0)
    diag = (Yp ** 2).mean(0) + var
    assert abs(d["condition_number"] - diag.max() / diag.min()) < 1e-12 * d["condition_number"]


def test_large_R_leaves_only_the_inflation(orc):
    rng = np.random.default_rng(11)
    k, n, p, rho = 7, 30, 10, 1.21
    X = rng.standard_normal((k, n))
    Yb = X[:, :p].copy()
    Xa, _, _, _ = orc.lwenkf_analyse(X, Yb, rng.standard_normal(p), np.full(p, 1e30),
                                     rng.standard_normal((p, k)), rho, 0.3)
    m = X.mean(0)
    np.testing.assert_allclose(Xa, m + math.sqrt(rho) * (X - m), rtol=0, atol=1e-12)


def test_all_underflowing_likelihood_is_refused(orc):
    X = np.array([[0.0], [1.0]])
    with pytest.raises(RuntimeError):
        orc.lwenkf_analyse(X, X.copy(), [1e6], [1e-6], np.zeros((1, 2)), 1.0, 1.0,
                           weighting="likelihood")
