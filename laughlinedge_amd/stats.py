"""Host-side Monte-Carlo statistics: integrated autocorrelation time and effective
sample size (Madras & Sokal 1988 automatic windowing), batch-mean error bars.

The reference leaves error bars to external post-processing of repeated runs
(main.cu:18-20, 43-51); these helpers turn the engine's per-walker traces and
per-block moments into tau_int, ESS and sigma directly.
"""
import numpy as np


def autocorr(series):
    """Normalised autocorrelation rho(t) of a (T, W) array: T steps of W independent
    chains, averaging the autocovariance over chains (FFT, biased estimator)."""
    x = np.asarray(series, dtype=np.float64)
    if x.ndim == 1:
        x = x[:, None]
    T = x.shape[0]
    x = x - x.mean()
    nfft = 1 << int(np.ceil(np.log2(2 * T)))
    f = np.fft.rfft(x, n=nfft, axis=0)
    acov = np.fft.irfft(f * np.conj(f), n=nfft, axis=0)[:T].mean(axis=1) / T
    if acov[0] <= 0:
        return np.ones(1)
    return acov / acov[0]


def tau_int(series, c=6.0):
    """Madras-Sokal windowed tau_int = 1/2 + sum_{t=1}^{M} rho(t), smallest M >= c tau(M)."""
    rho = autocorr(series)
    tau = 0.5
    for M in range(1, len(rho)):
        tau += rho[M]
        if M >= c * tau:
            return max(tau, 0.5)
    return max(tau, 0.5)


def ess(series, c=6.0):
    x = np.asarray(series)
    n = x.size
    return n / (2.0 * tau_int(x, c))


def batch_sigma(values):
    """standard error of the mean of equal-size batch means (axis 0 = batches)"""
    v = np.asarray(values, dtype=np.float64)
    nb = v.shape[0]
    if nb < 2:
        return np.full(v.shape[1:], np.nan)
    return v.std(axis=0, ddof=1) / np.sqrt(nb)
