"""ctypes front-end of liboracle.so (the FP64 CPU oracle; test infrastructure only)."""
import ctypes
import os
import subprocess

import numpy as np

from .partitions import Basis

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        _LIB = ctypes.CDLL(path)
        _LIB.or_block_size.restype = ctypes.c_int
    return _LIB


class _Sys(ctypes.Structure):
    _fields_ = [("N", ctypes.c_int), ("m", ctypes.c_int), ("w", ctypes.c_double),
                ("eps", ctypes.c_double), ("strength", ctypes.c_double),
                ("expo", ctypes.c_int), ("R", ctypes.c_double)]


class _Basis(ctypes.Structure):
    _fields_ = [("D", ctypes.c_int), ("K", ctypes.c_int),
                ("fo", ctypes.POINTER(ctypes.c_int)), ("fk", ctypes.POINTER(ctypes.c_int)),
                ("fr", ctypes.POINTER(ctypes.c_int)), ("nsec", ctypes.c_int),
                ("so", ctypes.POINTER(ctypes.c_int))]


class _Mixture(ctypes.Structure):
    _fields_ = [("kmix", ctypes.c_int), ("c0", ctypes.c_double),
                ("a", ctypes.c_double * 32), ("b", ctypes.c_double * 32)]


def mixture_coefficients(L, nk, rho=0.5):
    """Mixture w = rho + sum_k c_k (|p_k|^2 / n_k)^(L / k), k = 1..L, with
    c_k = (1 - rho) / (L Gamma(L / k + 1)) (each term has mean (1 - rho) / L when |p_k|^2 / n_k
    is a unit exponential), as (kmix, c0, a, b) with term_k = exp(a_k ln|p_k|^2 + b_k)."""
    from math import lgamma, log
    a = [L / k for k in range(1, L + 1)]
    b = [log((1.0 - rho) / L) - lgamma(a[k - 1] + 1.0) - a[k - 1] * log(nk[k - 1]) for k in range(1, L + 1)]
    return L, float(rho), a, b


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Oracle:
    """System + basis bound to the oracle library."""

    def __init__(self, N, m, sectors, w=0.0, eps=0.0, strength=0.0, expo=0):
        self.N, self.m = int(N), int(m)
        self.R = float(np.sqrt(2.0 * m * N))
        self.basis = Basis(N, sectors)
        self.sys = _Sys(self.N, self.m, float(w), float(eps), float(strength), int(expo), self.R)
        fo, fk, fr = self.basis.factor_arrays()
        self._fo = (ctypes.c_int * len(fo))(*fo)
        self._fk = (ctypes.c_int * max(1, len(fk)))(*fk)
        self._fr = (ctypes.c_int * max(1, len(fr)))(*fr)
        so = self.basis.sector_offsets
        self._so = (ctypes.c_int * len(so))(*so)
        self.cb = _Basis(self.basis.D, self.basis.K, self._fo, self._fk, self._fr,
                         len(self.basis.sectors), self._so)
        self.dims = self.basis.sector_dims()
        self.nblock = sum(d * d for d in self.dims)

    def split_blocks(self, flat):
        """interleaved per-sector blocks -> list of complex d x d arrays"""
        out, off = [], 0
        for d in self.dims:
            blk = flat[2 * off:2 * (off + d * d)].reshape(d, d, 2)
            out.append(blk[..., 0] + 1j * blk[..., 1])
            off += d * d
        return out

    def accumulate_fixed(self, pos):
        """pos: (n, N, 2) float64 -> (S_blocks, H_blocks, Vsum) sums (not means)"""
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        n = pos.shape[0]
        S = np.zeros(2 * self.nblock)
        H = np.zeros(2 * self.nblock)
        v = ctypes.c_double()
        lib().or_accumulate_fixed(ctypes.byref(self.sys), ctypes.byref(self.cb), _dp(pos),
                                  ctypes.c_long(n), _dp(S), _dp(H), ctypes.byref(v))
        return self.split_blocks(S), self.split_blocks(H), v.value

    def accumulate_weighted(self, pos, w, s1, s2):
        """sum_s w_s conj(p_a) p_b over samples for a in sector s1, b in sector s2"""
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        w = np.ascontiguousarray(w, dtype=np.float64)
        d1, d2 = self.dims[s1], self.dims[s2]
        out = np.zeros(2 * d1 * d2)
        lib().or_accumulate_weighted(ctypes.byref(self.sys), ctypes.byref(self.cb), _dp(pos),
                                     ctypes.c_long(pos.shape[0]), _dp(w), ctypes.c_int(s1),
                                     ctypes.c_int(s2), _dp(out))
        return (out[0::2] + 1j * out[1::2]).reshape(d1, d2)

    def features(self, pos):
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        pl = np.zeros(2 * self.basis.D)
        v = ctypes.c_double()
        lib().or_features(ctypes.byref(self.sys), ctypes.byref(self.cb), _dp(pos), _dp(pl),
                          ctypes.byref(v))
        return pl[0::2] + 1j * pl[1::2], v.value

    def plain_run(self, walkers, burnin, samples, step, seed, thin=1, threads=0, want_pos=False):
        S = np.zeros(2 * self.nblock)
        H = np.zeros(2 * self.nblock)
        acc = ctypes.c_double()
        r2 = ctypes.c_double()
        fp = np.zeros((walkers, self.N, 2)) if want_pos else None
        lib().or_plain_run(ctypes.byref(self.sys), ctypes.byref(self.cb), ctypes.c_int(walkers),
                           ctypes.c_int(burnin), ctypes.c_int(samples), ctypes.c_int(thin),
                           ctypes.c_double(step), ctypes.c_uint64(seed), ctypes.c_int(threads),
                           _dp(S), _dp(H), ctypes.byref(acc),
                           _dp(fp) if want_pos else None, ctypes.byref(r2))
        n = walkers * samples
        Sb = [x / n for x in self.split_blocks(S)]
        Hb = [x / n for x in self.split_blocks(H)]
        return Sb, Hb, acc.value, r2.value, fp

    def mixture_weights(self, mix, pos):
        """w(z) of each configuration (pos: n x N x 2) for mix = (kmix, c0, a, b)"""
        kmix, c0, a, b = mix
        mx = _Mixture(int(kmix), float(c0), (ctypes.c_double * 32)(*a), (ctypes.c_double * 32)(*b))
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        w = np.zeros(pos.shape[0])
        if lib().or_mixture_weights(ctypes.byref(self.sys), ctypes.byref(mx), _dp(pos),
                                    ctypes.c_long(pos.shape[0]), _dp(w)) != 0:
            raise ValueError("mixture: kmix must be in 1..31")
        return w

    def mixture_run(self, mix, walkers, burnin, samples, step, seed, thin=1, threads=0, want_pos=False,
                    com_moves=False):
        """Chains on w |Psi_L|^2 (mix = (kmix, c0, a, b)); returns S, H blocks normalised by
        sum 1/w, that sum, the acceptance and optionally the final positions.  com_moves: the
        exact centre-of-mass random walk and rotation after every kept sample's sweeps."""
        kmix, c0, a, b = mix
        mx = _Mixture(int(kmix), float(c0), (ctypes.c_double * 32)(*a), (ctypes.c_double * 32)(*b))
        S = np.zeros(2 * self.nblock)
        H = np.zeros(2 * self.nblock)
        acc = ctypes.c_double()
        ws = ctypes.c_double()
        fp = np.zeros((walkers, self.N, 2)) if want_pos else None
        rc = lib().or_mixture_run(ctypes.byref(self.sys), ctypes.byref(self.cb), ctypes.byref(mx),
                                  ctypes.c_int(walkers), ctypes.c_int(burnin), ctypes.c_int(samples),
                                  ctypes.c_int(thin), ctypes.c_double(step), ctypes.c_uint64(seed),
                                  ctypes.c_int(threads), ctypes.c_int(1 if com_moves else 0), _dp(S), _dp(H),
                                  ctypes.byref(ws), ctypes.byref(acc),
                                  _dp(fp) if want_pos else None)
        if rc != 0:
            raise ValueError("mixture: kmix must be in 1..min(31, K)")
        Sb = [x / ws.value for x in self.split_blocks(S)]
        Hb = [x / ws.value for x in self.split_blocks(H)]
        return Sb, Hb, ws.value, acc.value, fp

    def plain_positions(self, walkers, burnin, samples, step, seed, thin=1):
        out = np.zeros((walkers, samples, self.N, 2))
        acc = ctypes.c_double()
        lib().or_plain_positions(ctypes.byref(self.sys), ctypes.c_int(walkers), ctypes.c_int(burnin),
                                 ctypes.c_int(samples), ctypes.c_int(thin), ctypes.c_double(step),
                                 ctypes.c_uint64(seed), _dp(out), ctypes.byref(acc))
        return out, acc.value

    def rowwise_sector(self, sec, walkers, burnin, samples, step, seed, threads=0, dl_exc=0, want_pos=False):
        d = self.dims[sec]
        H = np.zeros(2 * d * d); M = np.zeros(2 * d * d)
        VC = np.zeros(2 * d * d); VS = np.zeros(2 * d * d)
        acc = ctypes.c_double()
        fp = np.zeros((d * walkers, self.N, 2)) if want_pos else None
        lib().or_rowwise_sector(ctypes.byref(self.sys), ctypes.byref(self.cb), ctypes.c_int(sec),
                                ctypes.c_int(walkers), ctypes.c_int(burnin), ctypes.c_int(samples),
                                ctypes.c_double(step), ctypes.c_uint64(seed), ctypes.c_int(threads),
                                ctypes.c_int(dl_exc), _dp(H), _dp(M), _dp(VC), _dp(VS),
                                ctypes.byref(acc), _dp(fp) if want_pos else None)
        c = lambda a: (a[0::2] + 1j * a[1::2]).reshape(d, d)
        if want_pos:
            return c(H), c(M), c(VC), c(VS), acc.value, fp
        return c(H), c(M), c(VC), c(VS), acc.value


def harmonic_sums(pos, L):
    """per-sample sum_i cos(L theta_i), sum_i sin(L theta_i)"""
    th = np.arctan2(pos[..., 1], pos[..., 0])
    return np.cos(L * th).sum(-1), np.sin(L * th).sum(-1)


def potential(r2, w, eps, strength=0.0, expo=0, R=1.0):
    v = 0.5 * w * r2 + eps * r2 * r2
    if strength:
        v = v + strength * (np.sqrt(r2) / R) ** expo
    return v


def geneig(S, H, vectors=False):
    """FP64 Cholesky + Jacobi generalized eigensolve; returns evals (and S-normalised evecs)."""
    d = S.shape[0]
    Si = np.ascontiguousarray(np.stack([S.real, S.imag], -1).reshape(-1))
    Hi = np.ascontiguousarray(np.stack([H.real, H.imag], -1).reshape(-1))
    ev = np.zeros(d)
    vec = np.zeros(2 * d * d)
    rc = lib().or_geneig(ctypes.c_int(d), _dp(Si), _dp(Hi), _dp(ev), _dp(vec) if vectors else None)
    if rc != 0:
        raise np.linalg.LinAlgError("metric is not positive definite")
    if vectors:
        return ev, (vec[0::2] + 1j * vec[1::2]).reshape(d, d)
    return ev


def dsf_weights(H, M, VC, VS):
    dp1 = H.shape[0]
    pack = lambda A: np.ascontiguousarray(np.stack([A.real, A.imag], -1).reshape(-1))
    e = np.zeros(dp1 - 1)
    w = np.zeros(dp1 - 1)
    rc = lib().or_dsf_weights(ctypes.c_int(dp1), _dp(pack(H)), _dp(pack(M)), _dp(pack(VC)),
                              _dp(pack(VS)), _dp(e), _dp(w))
    if rc != 0:
        raise np.linalg.LinAlgError("metric is not positive definite")
    return e, w


def philox(ctr, key):
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    lib().or_philox4x32_10(c, k, o)
    return list(o)


_TIME_FN = ctypes.CFUNCTYPE(ctypes.c_double, ctypes.c_void_p, ctypes.c_double)


def evolve(M, H, U, rhos, F, end_time, save_step, eps_abs=1e-10, eps_rel=1e-8):
    """FP64 CPU RK-CK54 restatement of the reference's time evolution (one perturbation)."""
    pack = lambda A: np.ascontiguousarray(np.stack([np.asarray(A).real, np.asarray(A).imag], -1).reshape(-1))
    D = M.shape[0]
    nt = int(np.ceil(end_time / save_step))
    nr = len(rhos)
    out = np.zeros(2 * nt * max(1, nr))
    nrm = np.zeros(nt)
    R = np.concatenate([pack(r) for r in rhos]) if rhos else np.zeros(2)
    cb = _TIME_FN(lambda ctx, t: float(F(t)))
    steps = ctypes.c_long()
    rc = lib().or_evolve(ctypes.c_int(D), _dp(pack(M)), _dp(pack(H)), _dp(pack(U)), cb, None,
                         ctypes.c_double(end_time), ctypes.c_double(save_step), ctypes.c_double(eps_abs),
                         ctypes.c_double(eps_rel), ctypes.c_int(nr), _dp(R), _dp(out), _dp(nrm),
                         ctypes.byref(steps))
    if rc != 0:
        raise np.linalg.LinAlgError("singular metric")
    return (out[0::2] + 1j * out[1::2])[: nt * nr].reshape(nt, nr), nrm, steps.value
