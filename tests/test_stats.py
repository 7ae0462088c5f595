"""Host statistics helpers (no GPU): the delete-one-block jackknife of eigenvalues."""
import numpy as np
import scipy.linalg

from laughlinedge_amd.stats import batch_sigma, jackknife_eigenvalues


def _blocks_1x1(x):
    return [(np.ones((1, 1), complex), np.full((1, 1), v, complex)) for v in x]


def test_jackknife_linear_statistic_is_the_batch_mean():
    """For E = H/S with S = 1 the eigenvalue is linear in the block means: no correction, and
    the jackknife error equals the batch-mean standard error."""
    x = np.random.default_rng(1).standard_normal(32)
    corr, sig, full = jackknife_eigenvalues(_blocks_1x1(x), lambda S, H: np.array([(H / S)[0, 0].real]))
    assert np.allclose(full, x.mean()) and np.allclose(corr, x.mean())
    assert np.allclose(sig, batch_sigma(x[:, None]))


def test_jackknife_removes_the_bias_of_a_square():
    """For E = (mean)^2 the corrected estimate is mean^2 - s^2 / n, the unbiased estimator of
    mu^2 (s^2 the sample variance)."""
    x = 3.0 + np.random.default_rng(2).standard_normal(16)
    n = x.size
    corr, _, full = jackknife_eigenvalues(_blocks_1x1(x), lambda S, H: np.array([H[0, 0].real ** 2]))
    assert np.allclose(full, x.mean() ** 2)
    assert np.allclose(corr, x.mean() ** 2 - x.var(ddof=1) / n)


def test_jackknife_shrinks_the_rayleigh_ritz_bias():
    """Noisy Hermitian pencils around a known (S0, H0): averaged over repetitions, the lowest
    eigenvalue of the mean pencil lies below the true one, and the jackknife estimate has a
    bias several times smaller.  Its sigma matches the spread of the full-mean eigenvalue."""
    rng = np.random.default_rng(3)
    d, nb, reps, noise = 4, 16, 300, 0.6
    A = rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d))
    S0 = A @ A.conj().T / d + np.eye(d)
    B = rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d))
    H0 = (B + B.conj().T) / 2 + 3 * S0
    true = scipy.linalg.eigh(H0, S0, eigvals_only=True)

    def solve(S, H):
        return scipy.linalg.eigh((H + H.conj().T) / 2, (S + S.conj().T) / 2, eigvals_only=True)

    def herm(scale):
        Z = rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d))
        return scale * (Z + Z.conj().T) / 2

    fulls, corrs, sigs = [], [], []
    for _ in range(reps):
        blocks = []
        for _ in range(nb):
            # correlated noise in S and H, as when both come from the same samples
            E = herm(noise)
            blocks.append((S0 + E, H0 + 3 * E + herm(noise)))
        c, s, f = jackknife_eigenvalues(blocks, solve)
        fulls.append(f[0]); corrs.append(c[0]); sigs.append(s[0])
    bias_full = np.mean(fulls) - true[0]
    bias_corr = np.mean(corrs) - true[0]
    assert bias_full < -3 * np.std(fulls) / np.sqrt(reps), bias_full   # well resolved
    assert abs(bias_corr) < 0.3 * abs(bias_full), (bias_full, bias_corr)
    assert 0.7 < np.mean(sigs) / np.std(fulls) < 1.4, (np.mean(sigs), np.std(fulls))


def test_bench_spectra_comparison():
    """bench.compare_spectra: |E - E_ref| / sqrt(s^2 + s_ref^2) over every eigenvalue, sectors
    with a missing spectrum skipped, the lowest eigenvalue of each sector reported."""
    import bench
    n = len(bench.SECTORS)
    spec = [np.array([1.0, 2.0]) for _ in range(n)]
    ref = [np.array([1.3, 2.0]) for _ in range(n)]
    sig = [np.array([0.3, 1.0]) for _ in range(n)]
    ref_sig = [np.array([0.4, 1.0]) for _ in range(n)]
    spec[1] = None
    out = bench.compare_spectra(spec, sig, ref, ref_sig)
    assert out["eigenvalues"] == 2 * (n - 1)
    assert np.isclose(out["max_dev_in_combined_sigma"], 0.3 / 0.5)
    assert out["frac_within_1_sigma"] == 1.0
    assert str(bench.SECTORS[1]) not in out["lowest_per_sector_dev"]
    assert out["lowest_per_sector_dev"][str(bench.SECTORS[0])] == 0.6
