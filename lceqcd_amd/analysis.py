"""Time-series analysis of Monte Carlo histories (host numpy).

tau_int: integrated autocorrelation time by the automatic windowing of
Madras and Sokal (J. Stat. Phys. 50 (1988) 109): the smallest window W with
W >= c tau_int(W), and the error estimate
dtau = tau sqrt(2 (2 W + 1) / n).  Effectively independent samples in n
sweeps: n / (2 tau_int)."""
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
