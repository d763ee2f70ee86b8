"""Float problems solved alone at the sizes where the packed float filter is the default
(n >= 28): answers against the oracle's FP64 brute force, over value families that stress
the filter's quantization — wide dynamic range (one entry many orders above the rest),
extreme scales, near-ties, float32-rounded input, sign-definite matrices whose optimum sits
at an end of the energy range — and every tie width."""
import numpy as np
import pytest

import oracle
from qubo_lib_amd import engine

pytestmark = pytest.mark.gpu


def flush(*a):
    print(*a, flush=True)


def family(rng, n, k):
    G = rng.standard_normal((n, n))
    if k == 0:
        return G
    if k == 1:
        return G.astype(np.float32).astype(np.float64)
    if k == 2:          # near-ties: quarter units plus a 1e-9 perturbation
        return np.round(G * 4) / 4 + rng.standard_normal((n, n)) * 1e-9
    if k == 3:          # one entry six orders above the rest
        G[3, 3] = -1e6
        return G
    if k == 4:          # a coupling six orders above the rest, of the sign that frustrates it
        G[1, 2] = G[2, 1] = 5e5
        G[1, 1] = G[2, 2] = -4e5
        return G
    if k == 5:
        return G * 1e-30
    if k == 6:
        return G * 1e30
    if k == 7:          # all entries >= 0: the empty state is the optimum (E = 0)
        return np.abs(G)
    if k == 8:          # all entries <= 0: the full state is the optimum
        return -np.abs(G)
    if k == 9:          # a few distinct values: many exactly equal energies in FP64 (kept
        return rng.integers(-3, 4, (n, n)) * 0.125      # on the float path below)
    raise ValueError(k)


@pytest.mark.parametrize("k", range(10))
def test_packed_filter_alone_vs_oracle(gpu_ctx, k):
    rng = np.random.default_rng(5100 + k)
    for trial in range(3):
        n = 28 + trial % 2
        Q = family(rng, n, k)
        B = [None, 0, n][trial]
        flush("float solo", k, trial, n, B)
        p = engine.prepare(Q, lattice=False)
        assert p.kind == "f64", (k, trial)
        got = engine.solve_matrix(p, tie_B=B)
        assert got.info["walk"] == "float_pair16", (k, trial, got.info["walk"])
        E_ref, s_ref = oracle.brute_f64(Q, B)
        tol = 1e-13 * np.abs(Q).sum() + 1e-300
        assert got.state == s_ref and abs(got.E - E_ref) <= tol, (k, trial, n, B, got.E, E_ref)
        if k in (7, 9) or trial == 0:
            alt = engine.solve_matrix(p, tie_B=B, float_filter="int32")
            assert (alt.E, alt.state) == (got.E, got.state), (k, trial, n, B)
