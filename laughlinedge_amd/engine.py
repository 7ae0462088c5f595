"""ctypes binding of liblaughlinedge.so (the C ABI declared in include/laughlinedge.h).

This is the host-side mirror used by the tests and the benchmark; the drop-in C++
interface (edgeMC::spectrum_compute and friends) sits in include/edge_montecarlo.h and
calls the same C ABI.  Nothing here computes: every call goes to the HIP library, and
loading fails loudly when the library is missing.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblaughlinedge.so")

LE_OK = 0
LE_ERR_ARG = 1
LE_ERR_CAPACITY = 2
LE_ERR_HIP = 3
LE_ERR_NOT_PD = 4
LE_ERR_BISECTION = 5
LE_ERR_STATE = 6
LE_ERR_NOCONV = 7
PREC_FP32 = 0
PREC_FP64 = 1
PREC_BF16X3 = 2
PREC_BF16X6 = 3

EXPORTED = [
    "le_mc_params_default", "le_last_error", "le_engine_create", "le_engine_destroy",
    "le_engine_get_info", "le_sector_dims", "le_basis_state", "le_init_walkers",
    "le_tune_step", "le_burnin", "le_sample", "le_run", "le_get_stats", "le_n_blocks_done",
    "le_get_moments", "le_get_trace", "le_get_positions", "le_set_positions",
    "le_accumulate_positions", "le_reset_moments", "le_geneig", "le_engine_spectrum",
    "le_engine_spectra", "le_basis_describe", "le_get_block", "le_set_weight_callback",
    "le_evolve", "le_engine_create_occupation", "le_get_occupation", "le_train_params_default",
    "le_train_proposal", "le_set_proposal", "le_get_proposal_params", "le_set_proposal_params",
    "le_proposal_loss_grad", "le_set_thin", "le_removal_weights", "le_step_search",
    "le_set_mixture", "le_get_mixture",
]
PROPOSAL_PLAIN = 0
PROPOSAL_LEARNED = 1
EST_PLAIN = 0
EST_ROWWISE = 1
EST_MIXTURE = 2
N_MLP_PARAMS = 4546


class TrainParams(ctypes.Structure):
    _fields_ = [("steps", ctypes.c_int), ("lr", ctypes.c_double), ("batch", ctypes.c_int),
                ("n_configs", ctypes.c_int), ("snapshot_sweeps", ctypes.c_int), ("seed", ctypes.c_uint64),
                ("sigma0", ctypes.c_double), ("table_size", ctypes.c_int)]


class TrainReport(ctypes.Structure):
    _fields_ = [("loss_first", ctypes.c_double), ("loss_last", ctypes.c_double),
                ("accept_first", ctypes.c_double), ("accept_last", ctypes.c_double), ("ms", ctypes.c_double)]

TIME_FN = ctypes.CFUNCTYPE(ctypes.c_double, ctypes.c_void_p, ctypes.c_double)

W_ONE = 0
W_V = 1


def W_COS(L):
    return 0x100 | int(L)


def W_SIN(L):
    return 0x200 | int(L)


def W_USER(k):
    return 0x10000 + int(k)


class LEError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"laughlinedge error {code}: {msg}")
        self.code = code


class SystemParams(ctypes.Structure):
    _fields_ = [("particles_number", ctypes.c_int), ("inverse_filling", ctypes.c_int),
                ("conf_w", ctypes.c_double), ("conf_eps", ctypes.c_double),
                ("conf_strength", ctypes.c_double), ("conf_exponent", ctypes.c_int)]


class MCParams(ctypes.Structure):
    _fields_ = [("walkers", ctypes.c_int), ("burnin_sweeps", ctypes.c_int),
                ("samples", ctypes.c_int), ("thin", ctypes.c_int), ("step", ctypes.c_double),
                ("tune_sweeps", ctypes.c_int), ("target_acceptance", ctypes.c_double),
                ("acceptance_tol", ctypes.c_double), ("seed", ctypes.c_uint64),
                ("precision", ctypes.c_int), ("lanes_per_walker", ctypes.c_int),
                ("samples_per_launch", ctypes.c_int), ("n_blocks", ctypes.c_int),
                ("trace_walkers", ctypes.c_int), ("graphs", ctypes.c_int), ("symmetry_moves", ctypes.c_int),
                ("sweep_registers", ctypes.c_int), ("estimator", ctypes.c_int), ("gram_xcd_spread", ctypes.c_int),
                ("mixture_rho", ctypes.c_double), ("fused_step", ctypes.c_int), ("fused_rows", ctypes.c_int),
                ("gram_tile_cols", ctypes.c_int)]


class Observables(ctypes.Structure):
    _fields_ = [("n_harmonics", ctypes.c_int), ("harmonics", ctypes.POINTER(ctypes.c_int)),
                ("n_user_weights", ctypes.c_int)]


class EngineInfo(ctypes.Structure):
    _fields_ = [("D", ctypes.c_int), ("K", ctypes.c_int), ("n_sectors", ctypes.c_int),
                ("longest_partition", ctypes.c_int), ("block_elems", ctypes.c_long),
                ("lanes_per_walker", ctypes.c_int), ("samples_per_launch", ctypes.c_int),
                ("n_tiles", ctypes.c_int), ("x_ld", ctypes.c_long), ("sweep_in_registers", ctypes.c_int),
                ("gram_mfma_flop_per_row", ctypes.c_double), ("gram_xcd_spread", ctypes.c_int),
                ("chains", ctypes.c_long), ("gram_tile_cols", ctypes.c_int)]


class RunStats(ctypes.Structure):
    _fields_ = [("step", ctypes.c_double), ("acceptance", ctypes.c_double),
                ("sweeps", ctypes.c_longlong), ("samples", ctypes.c_longlong),
                ("ms_sweep", ctypes.c_double), ("ms_features", ctypes.c_double),
                ("ms_gram", ctypes.c_double), ("ms_sample_total", ctypes.c_double),
                ("gram_launches", ctypes.c_longlong), ("sweep_launches", ctypes.c_longlong),
                ("ms_sampler_graph", ctypes.c_double), ("ms_gram_launch_min", ctypes.c_double),
                ("ms_step", ctypes.c_double), ("step_launches", ctypes.c_longlong)]


_lib = None


def lib():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run python -m laughlinedge_amd.build")
        L = ctypes.CDLL(LIB_PATH)
        L.le_last_error.restype = ctypes.c_char_p
        L.le_n_blocks_done.restype = ctypes.c_int
        _lib = L
    return _lib


def _check(rc):
    if rc != LE_OK:
        raise LEError(rc, lib().le_last_error().decode())


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


ACCEPT_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_double, ctypes.POINTER(ctypes.c_double))


def step_search(measure, step0, target, tol):
    """The step search of edgeMC::optimal_step (le_step_search, no GPU call of its own) over a
    Python acceptance function measure(step) -> acceptance, or -> (error code, acceptance).
    Returns (step, number of measurements); raises LEError with the library's code."""
    def cb(_ctx, s, out):
        r = measure(s)
        code, a = (r if isinstance(r, tuple) else (LE_OK, r))
        out[0] = float(a)
        return int(code)
    fn = ACCEPT_FN(cb)
    st = ctypes.c_double(0.0)
    nm = ctypes.c_int(0)
    rc = lib().le_step_search(fn, None, ctypes.c_double(step0), ctypes.c_double(target), ctypes.c_double(tol),
                              ctypes.byref(st), ctypes.byref(nm))
    if rc != LE_OK:
        err = LEError(rc, lib().le_last_error().decode())
        err.measurements = nm.value
        raise err
    return st.value, nm.value


def default_mc():
    p = MCParams()
    lib().le_mc_params_default(ctypes.byref(p))
    return p


def _cplx(flat):
    return flat[0::2] + 1j * flat[1::2]


class Engine:
    """One engine = one system, one list of sectors, one set of walkers on the GPU."""

    def __init__(self, N, m, sectors, w=0.0, eps=0.0, strength=0.0, exponent=0, harmonics=(),
                 n_user_weights=0, **mc):
        self.sys = SystemParams(int(N), int(m), float(w), float(eps), float(strength), int(exponent))
        self.mc = default_mc()
        for k, v in mc.items():
            if not hasattr(self.mc, k):
                raise TypeError(f"unknown Monte-Carlo parameter {k}")
            setattr(self.mc, k, v)
        self.sectors = [int(s) for s in sectors]
        arr = (ctypes.c_int * len(self.sectors))(*self.sectors)
        self._h = ctypes.c_void_p()
        self.harmonics = [int(x) for x in harmonics]
        self._harr = (ctypes.c_int * max(1, len(self.harmonics)))(*self.harmonics)
        obs = Observables(len(self.harmonics), self._harr, int(n_user_weights))
        _check(lib().le_engine_create(ctypes.byref(self.sys), ctypes.byref(self.mc), arr,
                                      len(self.sectors), ctypes.byref(obs), ctypes.byref(self._h)))
        self._info = EngineInfo()
        _check(lib().le_engine_get_info(self._h, ctypes.byref(self._info)))
        ns = self.info.n_sectors
        dims = (ctypes.c_int * ns)()
        offs = (ctypes.c_int * (ns + 1))()
        _check(lib().le_sector_dims(self._h, dims, offs))
        self.dims = list(dims)
        self.offsets = list(offs)
        self.N, self.m = int(N), int(m)
        self.R = float(np.sqrt(2.0 * m * N))

    @property
    def info(self):
        """engine shape and the lane mapping in use (the autotuner may change it at init)"""
        if self._h:
            _check(lib().le_engine_get_info(self._h, ctypes.byref(self._info)))
        return self._info

    def close(self):
        if self._h:
            lib().le_engine_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # basis --------------------------------------------------------------
    def state(self, t):
        dl = ctypes.c_int()
        pairs = (ctypes.c_int * 16)()
        n = ctypes.c_int()
        _check(lib().le_basis_state(self._h, int(t), ctypes.byref(dl), pairs, ctypes.byref(n)))
        return dl.value, [(pairs[2 * q], pairs[2 * q + 1]) for q in range(n.value)]

    # sampling ------------------------------------------------------------
    def init_walkers(self):
        _check(lib().le_init_walkers(self._h))

    def tune_step(self):
        s = ctypes.c_double()
        _check(lib().le_tune_step(self._h, ctypes.byref(s)))
        return s.value

    def burnin(self, sweeps):
        _check(lib().le_burnin(self._h, int(sweeps)))

    def set_mixture(self, nk, rho):
        """Mixture density of a LE_EST_MIXTURE engine from its normalisers n_k (k = 1..L)."""
        v = np.ascontiguousarray(nk, dtype=np.float64)
        _check(lib().le_set_mixture(self._h, v.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                    ctypes.c_int(len(v)), ctypes.c_double(rho)))

    def get_mixture(self):
        """(L, c0, a, b) of the density in use (L = 0 when not set)."""
        L = ctypes.c_int()
        c0 = ctypes.c_double()
        a = np.zeros(32)
        b = np.zeros(32)
        _check(lib().le_get_mixture(self._h, ctypes.byref(L), ctypes.byref(c0),
                                    a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                    b.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return L.value, c0.value, a[:L.value].copy(), b[:L.value].copy()

    def sample(self, samples):
        _check(lib().le_sample(self._h, int(samples)))

    def run(self):
        _check(lib().le_run(self._h))

    def stats(self):
        st = RunStats()
        _check(lib().le_get_stats(self._h, ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in RunStats._fields_}

    def n_blocks(self):
        return lib().le_n_blocks_done(self._h)

    def moments(self, block=-1):
        """list over sectors of (S, H) complex d x d arrays"""
        be = self.info.block_elems
        S = np.zeros(2 * be)
        H = np.zeros(2 * be)
        _check(lib().le_get_moments(self._h, int(block), _dp(S), _dp(H)))
        out, off = [], 0
        Sc, Hc = _cplx(S), _cplx(H)
        for d in self.dims:
            out.append((Sc[off:off + d * d].reshape(d, d), Hc[off:off + d * d].reshape(d, d)))
            off += d * d
        return out

    def block(self, weight, s1, s2, block=-1):
        """d_s1 x d_s2 complex block <conj(p_a) w p_b> of one accumulated weight"""
        d1, d2 = self.dims[s1], self.dims[s2]
        out = np.zeros(2 * d1 * d2)
        _check(lib().le_get_block(self._h, int(block), int(weight), int(s1), int(s2), _dp(out)))
        return _cplx(out).reshape(d1, d2)

    def set_weight_callback(self, fn_ptr, ctx=None):
        """fn_ptr: address of a C function with the le_weight_fn signature"""
        self._wfn = fn_ptr
        _check(lib().le_set_weight_callback(self._h, ctypes.c_void_p(fn_ptr), ctypes.c_void_p(ctx)))

    def train_proposal(self, **kw):
        tp = TrainParams()
        lib().le_train_params_default(ctypes.byref(tp))
        for k, v in kw.items():
            if not hasattr(tp, k):
                raise TypeError(f"unknown training parameter {k}")
            setattr(tp, k, v)
        rep = TrainReport()
        _check(lib().le_train_proposal(self._h, ctypes.byref(tp), ctypes.byref(rep)))
        return {k: getattr(rep, k) for k, _ in TrainReport._fields_}

    def set_proposal(self, kind):
        _check(lib().le_set_proposal(self._h, int(kind)))

    def proposal_params(self):
        out = np.zeros(N_MLP_PARAMS, dtype=np.float32)
        _check(lib().le_get_proposal_params(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return out

    def set_proposal_params(self, p, sigma0, table_size=12):
        p = np.ascontiguousarray(p, dtype=np.float32)
        _check(lib().le_set_proposal_params(self._h, p.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                            ctypes.c_double(sigma0), int(table_size)))

    def proposal_loss_grad(self, configs, cfg, part, noise, sigma0):
        configs = np.ascontiguousarray(configs, dtype=np.float32)
        cfg = np.ascontiguousarray(cfg, dtype=np.int32)
        part = np.ascontiguousarray(part, dtype=np.int32)
        noise = np.ascontiguousarray(noise, dtype=np.float32)
        loss = ctypes.c_double()
        grad = np.zeros(N_MLP_PARAMS, dtype=np.float32)
        fp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        ip = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
        _check(lib().le_proposal_loss_grad(self._h, fp(configs), configs.shape[0], ip(cfg), ip(part), fp(noise),
                                           len(cfg), ctypes.c_double(sigma0), ctypes.byref(loss), fp(grad)))
        return loss.value, grad

    def trace(self):
        n = ctypes.c_long()
        _check(lib().le_get_trace(self._h, None, 0, ctypes.byref(n)))
        out = np.zeros((max(1, n.value), 9), dtype=np.float32)
        _check(lib().le_get_trace(self._h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                  n.value, ctypes.byref(n)))
        return out[:n.value]

    def positions(self):
        out = np.zeros((self.info.chains, self.N, 2))
        _check(lib().le_get_positions(self._h, _dp(out)))
        return out

    def set_positions(self, pos):
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        assert pos.shape == (self.info.chains, self.N, 2)
        _check(lib().le_set_positions(self._h, _dp(pos)))

    def accumulate_positions(self, pos):
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, self.N, 2)
        _check(lib().le_accumulate_positions(self._h, _dp(pos), ctypes.c_long(pos.shape[0])))

    def reset_moments(self):
        _check(lib().le_reset_moments(self._h))

    def set_thin(self, thin):
        _check(lib().le_set_thin(self._h, int(thin)))
        self.mc.thin = int(thin)

    def spectra(self, block=-1):
        ev = np.zeros(max(1, self.info.D))
        st = (ctypes.c_int * self.info.n_sectors)()
        _check(lib().le_engine_spectra(self._h, int(block), _dp(ev), st))
        out = []
        for s, d in enumerate(self.dims):
            o = self.offsets[s]
            out.append(ev[o:o + d].copy() if st[s] == 0 else None)
        return out

    def removal_weights(self, sector, block=-1):
        """Particle-removal spectral weights of the (N+1)-particle Laughlin state into this
        engine's sector: (orbital l, ascending energies E_j, weights summing to 1)."""
        d = self.dims[sector]
        ev = np.zeros(d)
        w = np.zeros(d)
        orb = ctypes.c_int()
        _check(lib().le_removal_weights(self._h, int(sector), int(block), ctypes.byref(orb), _dp(ev), _dp(w)))
        return orb.value, ev, w


def basis_describe(N, sectors):
    """Product basis (C++ enumerator) without touching the GPU: list of (dL, pairs)."""
    sec = (ctypes.c_int * len(sectors))(*[int(x) for x in sectors])
    D = ctypes.c_int()
    K = ctypes.c_int()
    _check(lib().le_basis_describe(int(N), sec, len(sectors), 0, ctypes.byref(D), ctypes.byref(K),
                                   None, None, None))
    n = D.value
    dls = (ctypes.c_int * max(1, n))()
    pairs = (ctypes.c_int * max(1, 16 * n))()
    npairs = (ctypes.c_int * max(1, n))()
    _check(lib().le_basis_describe(int(N), sec, len(sectors), n, ctypes.byref(D), ctypes.byref(K),
                                   dls, pairs, npairs))
    return [(dls[t], [(pairs[16 * t + 2 * q], pairs[16 * t + 2 * q + 1]) for q in range(npairs[t])])
            for t in range(n)]


def geneig(S, H, vectors=False):
    """FP64 generalized Hermitian eigensolve on the GPU."""
    d = S.shape[0]
    pack = lambda A: np.ascontiguousarray(np.stack([A.real, A.imag], -1).reshape(-1))
    ev = np.zeros(d)
    vec = np.zeros(2 * d * d) if vectors else None
    _check(lib().le_geneig(ctypes.c_int(d), _dp(pack(S)), _dp(pack(H)), _dp(ev),
                           _dp(vec) if vectors else None))
    if vectors:
        return ev, _cplx(vec).reshape(d, d)
    return ev


def _pack(A):
    return np.ascontiguousarray(np.stack([np.asarray(A).real, np.asarray(A).imag], -1).reshape(-1))


def evolve(M, H, Us, rhos, F, end_time, save_step, block_dims, eps_abs=1e-10, eps_rel=1e-8):
    """FP64 GPU time evolution: returns (rho[u][k][l] complex, norm[u][k], steps)."""
    D = M.shape[0]
    Us = list(Us)
    rhos = list(rhos)
    nt = int(np.ceil(end_time / save_step))
    out = np.zeros(2 * len(Us) * nt * max(1, len(rhos)))
    nrm = np.zeros(len(Us) * nt)
    cb = TIME_FN(lambda ctx, t: float(F(t)))
    dims = (ctypes.c_int * len(block_dims))(*[int(x) for x in block_dims])
    U = np.concatenate([_pack(u) for u in Us])
    R = np.concatenate([_pack(r) for r in rhos]) if rhos else np.zeros(2)
    steps = ctypes.c_long()
    _check(lib().le_evolve(ctypes.c_int(D), len(block_dims), dims, _dp(_pack(M)), _dp(_pack(H)), len(Us), _dp(U),
                           len(rhos), _dp(R), cb, None, ctypes.c_double(end_time), ctypes.c_double(save_step),
                           ctypes.c_double(eps_abs), ctypes.c_double(eps_rel), _dp(out), _dp(nrm),
                           ctypes.byref(steps)))
    rho = _cplx(out)[: len(Us) * nt * len(rhos)].reshape(len(Us), nt, len(rhos))
    return rho, nrm.reshape(len(Us), nt), steps.value


class OccupationEngine(Engine):
    """Two-sided chain for <L| a^dagger_l a_l |R>, all pairs of one sector at once."""

    def __init__(self, N, m, dL, lcut, **mc):
        self.sys = SystemParams(int(N), int(m), 0.0, 0.0, 0.0, 0)
        self.mc = default_mc()
        for k, v in mc.items():
            if not hasattr(self.mc, k):
                raise TypeError(f"unknown Monte-Carlo parameter {k}")
            setattr(self.mc, k, v)
        self._h = ctypes.c_void_p()
        _check(lib().le_engine_create_occupation(ctypes.byref(self.sys), ctypes.byref(self.mc), int(dL), int(lcut),
                                                 ctypes.byref(self._h)))
        self._info = EngineInfo()
        _check(lib().le_engine_get_info(self._h, ctypes.byref(self._info)))
        dims = (ctypes.c_int * 2)()
        offs = (ctypes.c_int * 3)()
        _check(lib().le_sector_dims(self._h, dims, offs))
        self.dims = list(dims)
        self.offsets = list(offs)
        self.N, self.m, self.lcut = int(N), int(m), int(lcut)
        self.R = float(np.sqrt(2.0 * m * N))

    def positions(self):
        out = np.zeros((self.mc.walkers, self.N + 1, 2))
        _check(lib().le_get_positions(self._h, _dp(out)))
        return out

    def accumulate_positions(self, pos):
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, self.N + 1, 2)
        _check(lib().le_accumulate_positions(self._h, _dp(pos), ctypes.c_long(pos.shape[0])))

    def occupation(self, block=-1):
        d = self.dims[0]
        out = np.zeros(2 * d * d * self.lcut)
        _check(lib().le_get_occupation(self._h, int(block), _dp(out)))
        return _cplx(out).reshape(d, d, self.lcut)
