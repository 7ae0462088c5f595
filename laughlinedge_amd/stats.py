"""Host-side Monte-Carlo statistics: integrated autocorrelation time and effective
sample size (Madras & Sokal 1988 automatic windowing), batch-mean error bars, and the
delete-one-block jackknife of generalized eigenvalues (bias-corrected, with standard errors).

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


def jackknife_eigenvalues(blocks, solve):
    """Delete-one-block jackknife of the generalized eigenvalues of a mean pencil.

    blocks: sequence of (S_b, H_b), the means of n equal-size blocks of samples.
    solve(S, H) -> ascending eigenvalues of H v = E S v.
    The eigenvalues of the full mean pencil carry a bias of order 1/samples (the lowest
    Rayleigh-Ritz value of a noisy pencil sits below the true one).  With E the full-mean
    eigenvalues and E_(b) those of the pencil without block b, the corrected estimate
    n E - (n - 1) mean_b E_(b) removes that leading term, and
    sqrt((n - 1) / n * sum_b (E_(b) - mean_b E_(b))^2) is its standard error.
    Returns (corrected, sigma, E).
    """
    n = len(blocks)
    if n < 2:
        raise ValueError("the jackknife needs at least two blocks")
    S = sum(b[0] for b in blocks) / n
    H = sum(b[1] for b in blocks) / n
    full = np.asarray(solve(S, H), dtype=np.float64)
    loo = np.array([solve((n * S - Sb) / (n - 1), (n * H - Hb) / (n - 1)) for Sb, Hb in blocks],
                   dtype=np.float64)
    mean_loo = loo.mean(axis=0)
    corrected = n * full - (n - 1) * mean_loo
    sigma = np.sqrt((n - 1) / n * ((loo - mean_loo) ** 2).sum(axis=0))
    return corrected, sigma, full
