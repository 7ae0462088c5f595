"""C ABI checks that need no GPU: the library loads, exports every declared symbol, the
C++ basis enumerator matches the oracle's, and argument errors come back as codes."""
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\**\s+(le_[a-z_0-9]+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    from laughlinedge_amd import lib
    from laughlinedge_amd.engine import EXPORTED
    names = _declared("laughlinedge.h")
    assert len(names) >= 20
    L = lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(EXPORTED) <= set(names)


@pytest.mark.parametrize("N,sectors", [(10, [0, 1, 2, 3, 4, 5, 6]), (3, [4, 5, 6]),
                                       (200, list(range(1, 17))), (400, [20, 0, 7])])
def test_product_basis_matches_oracle(N, sectors):
    from laughlinedge_amd.engine import basis_describe
    from oracle.partitions import Basis
    got = basis_describe(N, sectors)
    ref = Basis(N, sectors)
    assert [(dl, list(map(tuple, p))) for dl, p in got] == [(dl, list(map(tuple, p))) for dl, p in ref.states]


def test_argument_errors_are_codes():
    from laughlinedge_amd.engine import basis_describe, LEError, LE_ERR_ARG
    with pytest.raises(LEError) as ei:
        basis_describe(10, [-1])
    assert ei.value.code == LE_ERR_ARG


def test_mc_defaults_follow_reference():
    # em_parameters.cu:9-31: 1000 burn-in, 50000 samples, step 0.8, 100 tuning sweeps,
    # target 0.5 +- 0.01, 32 x 128 walkers
    from laughlinedge_amd.engine import default_mc
    p = default_mc()
    assert (p.burnin_sweeps, p.samples, p.tune_sweeps, p.walkers) == (1000, 50000, 100, 4096)
    assert abs(p.step - 0.8) < 1e-12 and abs(p.target_acceptance - 0.5) < 1e-12
    assert abs(p.acceptance_tol - 0.01) < 1e-12



def test_basis_matches_golden_listing():
    """C++ basis enumerator (no GPU) against the committed N = 4 listing"""
    import os
    import numpy as np
    from laughlinedge_amd.engine import basis_describe
    g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "exact_n4_m3.npz"))
    for dl in range(5):
        got = basis_describe(4, [dl])
        want = g[f"pairs{dl}"]
        assert len(got) == want.shape[0]
        for (d, pairs), row in zip(got, want):
            assert d == dl
            flat = [x for pr in pairs for x in pr]
            assert flat == [int(x) for x in row[:len(flat)]] and not any(row[len(flat):])
