"""Problem generators, Solution / StateVectors contracts, ReducedProblem algebra (CPU)."""
import json
from pathlib import Path

import numpy as np
import pytest

from qubo_lib_amd import Solution, StateVectors
from qubo_lib_amd.abc import ProblemABC
from qubo_lib_amd.problems import TSP, Just_Q, MaxCut, RandomQUBO, ReducedProblem
from qubo_lib_amd.problems import generators as gen

GOLDEN = json.loads((Path(__file__).resolve().parent / "golden" / "notebook.json").read_text())


def energy(Q, x):
    x = np.asarray(x, dtype=np.float64)
    return float(x @ np.asarray(Q, dtype=np.float64) @ x)


def test_tsp_cities_and_Q_blocks_match_notebook():
    t = TSP.init_random_instance(n_cities=7, seed=42)
    assert np.array_equal(t.cities, np.array(GOLDEN["tsp7_seed42_cities"]["cities"], dtype=np.float32))
    Q = t.build_Q_matrix()
    g = GOLDEN["tsp7_seed42_Q_round"]
    assert list(Q.shape) == g["shape"] and Q.dtype == np.float32
    assert np.array_equal(np.round(Q[:10, :10]), np.array(g["Q_round_10x10"]))
    g5 = GOLDEN["tsp5_cities_Q_round"]
    Q5 = TSP.init_from_cities(np.array(g5["cities"])).build_Q_matrix()
    assert list(Q5.shape) == g5["shape"]
    assert np.array_equal(np.round(Q5[:10, :10]), np.array(g5["Q_round_10x10"]))


def test_tsp_notebook_solutions_have_their_energy():
    # The notebook prints energies the reference accumulated in float32 (qbf.py:214-217;
    # cupy GEMM/row adds for the 6-city run): the exact energy of its own best_X differs
    # from the printed value by 2.8e-3 for TSP(6) (about 50 float32 ulps at 882).
    for key, nc, tol in (("tsp5_seed42", 5, 1e-3), ("tsp6_seed42", 6, 5e-3)):
        g = GOLDEN[key]
        Q = TSP.init_random_instance(n_cities=nc, seed=42).build_Q_matrix()
        x = [int(c) for c in g["best_X"]]
        assert energy(Q, x) == pytest.approx(g["best_E"], abs=tol)
        # a valid tour: one city per time and one time per city
        path = TSP.decode_x_to_path(np.array(x))
        assert sorted(path.tolist()) == list(range(nc))


def test_random_qubo_notebook_solutions_have_their_energy():
    g = GOLDEN["random37_seed42_f32"]
    Q = RandomQUBO.init_random_instance(seed=42, N=37).build_Q_matrix()
    assert Q.dtype == np.float32
    # printed by the reference's CPU QBF, whose tables (diag_BB, ABBA) are float32 sums
    # (qbf.py:251, 264-289): its energy is 0.026 below the exact energy of its best_X
    assert energy(Q, [int(c) for c in g["best_X"]]) == pytest.approx(g["best_E"], abs=0.05)
    # the int16 golden is exact: integer energies have no rounding
    g = GOLDEN["random44_seed42_int16"]
    Q = RandomQUBO.init_random_instance(N=44, seed=42, dtype=np.int16).build_Q_matrix()
    assert Q.dtype == np.int16 and Q.min() >= -100 and Q.max() <= 99
    x = np.array([int(c) for c in g["best_X"]], dtype=np.int64)
    assert int(x @ Q.astype(np.int64) @ x) == g["best_E"]


def test_random_qubo_is_deterministic_and_matches_pcg64():
    a = RandomQUBO.init_random_instance(seed=3, N=9, low=-5, high=5, dtype=np.int32).build_Q_matrix()
    b = np.random.Generator(np.random.PCG64(3)).integers(-5, 5, (9, 9), dtype=np.int32)
    assert np.array_equal(a, b)
    f = RandomQUBO.init_random_instance(seed=3, N=4).build_Q_matrix()
    rng = np.random.Generator(np.random.PCG64(3))
    assert np.array_equal(f, -100 + rng.random((4, 4), dtype=np.float32) * 200)


def test_reduced_problem_algebra():
    rng = np.random.default_rng(2)
    Q = rng.integers(-9, 10, (9, 9)).astype(np.int64)
    base = Just_Q(Q)
    for prefix in ([1, 0, 1], [0, 0, 0], [1, 1, 1]):
        rp = ReducedProblem(base, np.array(prefix))
        Qr = rp.build_Q_matrix()
        for t in range(1 << 6):
            rest = [(t >> (5 - i)) & 1 for i in range(6)]
            x = np.array(prefix + rest)
            y = np.array(rest)
            assert x @ Q @ x == rp.prefix_E + y @ Qr @ y


def test_maxcut_random_gnp_convention():
    m = MaxCut.random_gnp(12, 0.5, seed=4)
    Q = m.build_Q_matrix()
    for s in range(0, 1 << 12, 97):
        x = np.array([(s >> (11 - i)) & 1 for i in range(12)])
        cut = sum(w for i, j, w in m.edges if x[i] != x[j])
        assert x @ Q @ x == -cut


def test_generators_reproducible():
    assert np.array_equal(gen.symmetric_int_qubo(40, 0), gen.symmetric_int_qubo(40, 0))
    Q = gen.symmetric_int_qubo(40, 1)
    assert Q.dtype == np.int16 and np.array_equal(Q, Q.T) and Q.min() >= -127 and Q.max() <= 127
    probs, kinds = gen.batch_mix(64, seed=0)
    assert len(probs) == 64 and kinds[0] == "maxcut" and kinds[1] == "tsp"
    ns = [p.N for p in probs]
    assert all(16 <= n <= 32 for n, k in zip(ns, kinds) if k == "maxcut")
    assert all(n in (16, 25) for n, k in zip(ns, kinds) if k == "tsp")
    T = gen.tsp_int(4, seed=1)
    assert T.dtype == np.int32 and T.shape == (16, 16)
    reqs = gen.serving_requests(50, 1000.0, seed=2)
    assert all(20 <= q.shape[0] <= 36 for _, q, _ in reqs)
    assert all(b[0] > a[0] for a, b in zip(reqs, reqs[1:]))


def test_int_tsp_valid_tours_are_minimal():
    """Integer TSP with penalty A = n_c * max(D): the optimum is a tour whose energy is
    its path length minus n_c * A (each of the n_c assignment rows contributes -A)."""
    import itertools
    Q = gen.tsp_int(4, seed=9)
    import oracle
    E, s = oracle.brute_i64(Q.astype(np.int64))
    x = np.array([(s >> (15 - i)) & 1 for i in range(16)])
    path = TSP.decode_x_to_path(x)
    assert sorted(path.tolist()) == [0, 1, 2, 3]
    # no valid tour is cheaper
    best_tour = min(int(np.array([(1 if TSP.encode_path_to_x(list(p))[i] else 0) for i in range(16)])
                        @ Q.astype(np.int64) @ TSP.encode_path_to_x(list(p))) for p in itertools.permutations(range(4)))
    assert E == best_tour


def test_state_vectors_packing():
    wide = np.array([[1, 0, 1, 1, 0, 0, 0, 0, 1, 1]], dtype=np.int32)
    sv = StateVectors(wide)
    assert sv.bitwise_vectors.dtype == np.uint8
    assert sv.bitwise_vectors.tolist() == [[0b00001101, 0b00000011]]
    assert np.array_equal(sv.get_wide_vectors(), wide.astype(np.uint8))
    p16 = StateVectors.from_packed_vectors(np.array([[0b1000000001]], dtype=np.uint16), 10)
    assert p16.get_wide_vectors().tolist() == [[1, 0, 0, 0, 0, 0, 0, 0, 0, 1]]
    with pytest.raises(TypeError):
        StateVectors.from_packed_vectors(np.array([[1.0]]), 1)


def test_solution_json_round_trip_and_str():
    s = Solution(3, 2, 1, "{}", "{}")
    s.best_E = np.float64(-1.5)
    s.best_X = np.array([1, 0, 1], dtype=np.int32)
    s.state_vectors = StateVectors(s.best_X.reshape(1, 3))
    s.energies = np.array([-1.5])
    s.time_to_solve_ms = 1.25
    s.solver_class = "QBF"
    s.metrics = [{"energy_arr": np.array([0.0, -1.5])}]
    t = Solution.init_from_json(s.to_json())
    assert np.array_equal(t.best_X, s.best_X) and t.best_X.dtype == np.int32
    assert t.best_E == s.best_E and np.array_equal(t.metrics[0]["energy_arr"], [0.0, -1.5])
    assert np.array_equal(t.state_vectors.get_wide_vectors(), [[1, 0, 1]])
    assert "best_E=-1.500" in str(t) and "best_X='101'" in str(t)
    s.error = "boom"
    assert str(s) == "Error: boom"


def test_ising_offset():
    Q = np.array([[1.0, 2.0], [3.0, 4.0]])
    assert ProblemABC.get_Q_Ising_offset(np, Q) == pytest.approx(0.25 * 10 + 0.25 * 5)
