"""The walk qbx_prep picks for an integer problem, against the rule restated here in numpy:
the packed pair walk when the energy bound sum |d| + sum_{a<b} |W_ab| is at most 15359;
else the rebased pair walk when its guard constant D (the inner couplings, the pair bit's
and the four group bits' couplings to the inner bits) leaves room for 1.5 x the inner
fields' root mean square S over the walked states; else the int32 walk.  Problems are
scaled so that D + 1.5 S straddles the bound; every answer is checked against the host
build of the same walk."""
import numpy as np
import pytest

from qubo_lib_amd import engine

pytestmark = pytest.mark.gpu

PAIR_BOUND = 15359
GROUP_BITS = 4


def expected_path(Q, M, L):
    n = Q.shape[0]
    Qi = Q.astype(np.int64)
    # bit a <-> variable n-1-a
    Qb = Qi[::-1, ::-1]
    W = Qb + Qb.T
    np.fill_diagonal(W, 0)
    d = np.diag(Qb)
    bound = int(np.abs(d).sum() + np.abs(np.triu(W, 1)).sum())
    if L > M and bound <= PAIR_BOUND:
        return "pair16", None
    if bound > 2 ** 31 - 2 or L <= M:
        return "int32", None
    D = int(np.abs(np.triu(W[:M, :M], 1)).sum()) + int(np.abs(W[M, :M]).sum())
    for i in range(GROUP_BITS):
        if M + 1 + i < L:
            D += int(np.abs(W[M + 1 + i, :M]).sum())
    S = 0.0
    for k in range(M):
        w = W[k, M:n].astype(np.float64)
        mean = float(d[k]) + 0.5 * w.sum()
        var = 0.25 * (w * w).sum()
        S += np.sqrt(mean * mean + var)
    margin = PAIR_BOUND - (D + 1.5 * S)
    return ("pair16_rebased" if D < PAIR_BOUND and margin >= 0 else "int32"), margin


def test_prep_path_choice_matches_rule(gpu_ctx):
    rng = np.random.default_rng(3100)
    seen = set()
    for trial in range(60):
        n = int(rng.integers(22, 31))
        M, L = [(7, n - 12), (6, n - 11), (4, n - 13)][trial % 3]
        scale = float(rng.choice([20, 60, 120, 200, 300, 500]))
        U = np.rint(rng.standard_normal((n, n)) * scale).astype(np.int64)
        Q = np.triu(U) + np.triu(U, 1).T
        want, margin = expected_path(Q, M, L)
        if margin is not None and abs(margin) < 1e-6 * PAIR_BOUND:
            continue     # on the boundary: either choice is right
        got = engine.solve_matrix(Q, walk_bits=L, inner_bits=M)
        assert got.info["walk"] == want, (trial, n, M, L, scale, margin, got.info["walk"])
        host = engine.solve_matrix(Q, device="cpu")
        assert (got.E, got.state) == (host.E, host.state), (trial, n, M, L, want)
        seen.add(want)
    assert seen == {"pair16", "pair16_rebased", "int32"}, seen
