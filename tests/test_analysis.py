"""Madras-Sokal integrated autocorrelation time against processes whose
tau_int is known in closed form: AR(1) x_t = phi x_{t-1} + e_t has
rho(t) = phi^t and tau_int = (1 + phi) / (2 (1 - phi)); white noise has
tau_int = 1/2."""
import numpy as np
import pytest

from lceqcd_amd.analysis import autocorrelation, cosh_mass, mean_error, screening_mass, tau_int


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


def _cosh_corr(L, m, A=2.0, B=5.0):
    z = np.arange(L)
    return A * np.cosh(m * (z - L / 2)) + B


@pytest.mark.parametrize("L,m", [(16, 0.7), (32, 0.25), (64, 1.3)])
def test_cosh_mass_exact(L, m):
    c = _cosh_corr(L, m)
    for z in range(1, L // 2 - 2):
        assert abs(cosh_mass(c, z) - m) < 1e-9 * max(1.0, m)
    assert np.isnan(cosh_mass(c, 0))


def test_screening_mass_jackknife_covers_the_truth():
    rng = np.random.default_rng(4)
    hits = 0
    for s in range(40):
        c = _cosh_corr(16, 0.7)
        samples = c + rng.normal(scale=0.05, size=(400, 16))
        m, e = screening_mass(samples, z=1, nblk=20)
        assert e > 0
        hits += abs(m - 0.7) < 2 * e
    assert hits >= 32
