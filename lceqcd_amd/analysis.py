"""Time-series analysis of Monte Carlo histories (host numpy).

tau_int: integrated autocorrelation time by the automatic windowing of
Madras and Sokal (J. Stat. Phys. 50 (1988) 109): the smallest window W with
W >= c tau_int(W), and the error estimate
dtau = tau sqrt(2 (2 W + 1) / n).  Effectively independent samples in n
sweeps: n / (2 tau_int).

screening_mass: the mass of a periodic z-correlator (the scalarcorr,
polyakov and correlator outputs) from the cosh form
C(z) = A cosh(m (z - L/2)) + B, with the constant B removed by differences,
with a delete-one-block jackknife error."""
import numpy as np


def autocorrelation(x):
    """Normalised autocorrelation rho(t), t = 0..n-1 (FFT, zero padded)."""
    x = np.asarray(x, dtype=np.float64)
    n = len(x)
    d = x - x.mean()
    f = np.fft.rfft(d, 2 * n)
    c = np.fft.irfft(f * np.conj(f))[:n] / np.arange(n, 0, -1)
    return c / c[0] if c[0] > 0 else np.zeros(n)


def tau_int(x, c=6.0):
    """(tau_int, error, window) of the series x."""
    rho = autocorrelation(x)
    n = len(rho)
    tau = 0.5
    for w in range(1, n):
        tau += rho[w]
        if w >= c * tau:
            return tau, tau * np.sqrt(2.0 * (2 * w + 1) / n), w
    return tau, tau * np.sqrt(2.0 * (2 * n - 1) / n), n - 1


def mean_error(x, c=6.0):
    """Mean and its error including the autocorrelation (2 tau_int / n)."""
    x = np.asarray(x, dtype=np.float64)
    tau, _, _ = tau_int(x, c)
    return x.mean(), np.sqrt(2.0 * tau * x.var(ddof=1) / len(x))


def cosh_mass(c, z):
    """Effective mass at separation z of a periodic correlator c[0..L-1]
    of the form A cosh(m (z - L/2)) + B.  The constant B (the disconnected
    part) cancels in the differences d(z) = c(z) - c(z+1), which satisfy
    d(z-1) + d(z+1) = 2 cosh(m) d(z) exactly, so
    m = arccosh((d(z-1) + d(z+1)) / (2 d(z))); nan where undefined."""
    c = np.asarray(c, dtype=np.float64)
    d = c[:-1] - c[1:]
    if z < 1 or z + 1 >= len(d) or d[z] == 0.0:
        return np.nan
    r = (d[z - 1] + d[z + 1]) / (2.0 * d[z])
    return float(np.arccosh(r)) if r >= 1.0 else np.nan


def screening_mass(samples, z=1, nblk=20):
    """Screening mass and its jackknife error from per-measurement
    correlators samples[k, 0..L-1] (consecutive measurements grouped into
    nblk blocks, so the error includes autocorrelations shorter than a
    block)."""
    x = np.asarray(samples, dtype=np.float64)
    nb = min(nblk, len(x))
    blocks = x[: (len(x) // nb) * nb].reshape(nb, -1, x.shape[1]).mean(axis=1)
    full = cosh_mass(blocks.mean(axis=0), z)
    jk = np.array([cosh_mass(np.delete(blocks, i, axis=0).mean(axis=0), z) for i in range(nb)])
    err = np.sqrt((nb - 1) / nb * np.sum((jk - jk.mean()) ** 2))
    return full, float(err)
