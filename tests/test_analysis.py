"""Madras-Sokal integrated autocorrelation time against processes whose
tau_int is known in closed form: AR(1) x_t = phi x_{t-1} + e_t has
rho(t) = phi^t and tau_int = (1 + phi) / (2 (1 - phi)); white noise has
tau_int = 1/2."""
import numpy as np
import pytest

from lceqcd_amd.analysis import autocorrelation, mean_error, tau_int


def ar1(phi, n, seed):
    rng = np.random.default_rng(seed)
    e = rng.normal(size=n)
    x = np.empty(n)
    x[0] = e[0] / np.sqrt(1 - phi * phi)
    for t in range(1, n):
        x[t] = phi * x[t - 1] + e[t]
    return x


@pytest.mark.parametrize("phi", [0.0, 0.5, 0.9, 0.98])
def test_ar1_tau(phi):
    x = ar1(phi, 400000, 1)
    tau, dtau, w = tau_int(x)
    exact = (1 + phi) / (2 * (1 - phi))
    assert abs(tau - exact) < 4 * dtau + 0.02 * exact, (tau, dtau, exact)


def test_autocorrelation_of_ar1():
    rho = autocorrelation(ar1(0.8, 200000, 2))
    np.testing.assert_allclose(rho[:6], 0.8 ** np.arange(6), atol=0.01)


def test_mean_error_covers_the_truth():
    # the mean of an AR(1) chain is 0; its error must account for tau
    hits = 0
    for s in range(40):
        m, e = mean_error(ar1(0.9, 20000, 100 + s))
        hits += abs(m) < 2 * e
    assert hits >= 32  # ~95% coverage expected
