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



def exact_energy(Q, state):
    """x^T Q x of the state (x[0] the most significant bit) in exact rational arithmetic."""
    from fractions import Fraction
    n = Q.shape[0]
    on = [i for i in range(n) if (state >> (n - 1 - i)) & 1]
    return float(sum((Fraction(float(Q[i, j])) for i in on for j in on), Fraction(0)))


def test_tie_heavy_float_inputs(gpu_ctx):
    """Ties in real arithmetic that FP64 keeps only to rounding: MaxCut QUBOs scaled by 0.1
    and by 1/3 (not on a power-of-two lattice, so they stay on the float path), and a
    float32 TSP (every tour six ways).  Which of the tied states is smallest in FP64 depends
    on the order of the sums (the brute force sums along its Gray walk, the kernels from each
    block's base), so the answer is checked as a state whose exact energy is the minimum to
    rounding: E against the FP64 brute force at n = 25-28, and the two filters, which walk
    the same blocks the same way, equal to each other at every size."""
    from qubo_lib_amd.problems.generators import maxcut_gnp
    from qubo_lib_amd.problems.tsp import TSP
    cases = []
    for n, seed in ((28, 1), (34, 2), (36, 3)):
        Q = maxcut_gnp(n, 0.5, seed=seed).astype(np.float64)
        cases += [(f"maxcut{n}x0.1", Q * 0.1), (f"maxcut{n}/3", Q / 3)]
    cases.append(("tsp5", np.asarray(TSP.init_random_instance(seed=42, n_cities=5).build_Q_matrix())))
    for name, Q in cases:
        p = engine.prepare(Q)
        assert p.kind == "f64", name
        runs = [engine.solve_matrix(p, float_filter=f) for f in ("packed", "int32")]
        flush("ties", name, [r.info["walk"] for r in runs])
        assert [r.info["walk"] for r in runs] == ["float_pair16", "float"], name
        assert (runs[0].E, runs[0].state) == (runs[1].E, runs[1].state), name
        tol = 1e-12 * np.abs(Q).sum()
        got = runs[0]
        assert abs(exact_energy(Q, got.state) - got.E) <= tol, name
        if p.n <= 28:
            E_ref, _ = oracle.brute_f64(Q.astype(np.float64), None)
            assert abs(got.E - E_ref) <= tol, (name, got.E, E_ref)
