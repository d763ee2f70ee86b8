"""The drop-in solvers on the GPU: QBF / QB_F1 / Parallel / solve_batch return the
reference's Solution fields with the exact answers."""
import json
from pathlib import Path

import numpy as np
import pytest

import oracle
from qubo_lib_amd import Solution
from qubo_lib_amd.batch import solve_batch
from qubo_lib_amd.problems import TSP, Just_Q, MaxCut, RandomQUBO, ReducedProblem
from qubo_lib_amd.solvers import QB_F1, QBF, Parallel

pytestmark = pytest.mark.gpu
NB = json.loads((Path(__file__).resolve().parent / "golden" / "notebook.json").read_text())


def test_qbf_solution_fields_tsp5():
    prob = TSP.init_random_instance(n_cities=5, seed=42)
    sol = QBF().solve(prob, device="cuda")
    g = NB["tsp5_seed42"]
    assert isinstance(sol, Solution)
    assert sol.N == 25 and sol.samples_cnt == 1 and sol.solver_class == "QBF"
    assert sol.epochs_cnt == 2 ** 5                       # A = N - min(N, 20)
    assert "".join(map(str, sol.best_X)) == g["best_X"]
    assert sol.best_E == pytest.approx(g["best_E"], abs=1e-3)
    assert sol.best_X.dtype == np.int32
    assert np.array_equal(sol.state_vectors.get_wide_vectors()[0], sol.best_X)
    assert sol.energies.tolist() == [sol.best_E]
    Q = prob.build_Q_matrix()
    assert sol.ising_offset == pytest.approx(0.25 * Q.sum() + 0.25 * np.trace(Q))
    assert sol.time_to_solve_ms > 0 and sol.kernel_ms > 0
    labels = [l for _, l in sol.timer]
    assert "build_Q" in labels and "Kernel run time" in labels
    assert "best_X='" + g["best_X"] + "'" in str(sol)
    assert Solution.init_from_json(sol.to_json()).best_E == sol.best_E


def test_qbf_maxcut_known_answer():
    sol = QBF().solve(MaxCut("test"), device="cuda")
    assert sol.best_E == -19


def test_qbf_tie_rule_follows_A_B_kwargs():
    Q = np.diag([-1, 0, 1, 1, 1, 1, 1, 1]).astype(np.int32)
    prob = Just_Q(Q)
    assert "".join(map(str, QBF().solve(prob, B=0, device="cuda").best_X)) == "11000000"
    assert "".join(map(str, QBF().solve(prob, A=2, device="cuda").best_X)) == "11000000"
    assert "".join(map(str, QBF().solve(prob, device="cuda").best_X)) == "10000000"


def test_qbf_reduced_problem_prefix_E():
    base = RandomQUBO.init_random_instance(seed=3, N=14, dtype=np.int32)
    rp = ReducedProblem(base, np.array([1, 0, 1]))
    sol = QBF().solve(rp, device="cuda")
    assert sol.prefix_E == rp.prefix_E
    assert sol.best_E == oracle.brute_i64(rp.build_Q_matrix())[0]


def test_qbf_energy_arr_metric():
    prob = RandomQUBO.init_random_instance(seed=4, N=12, dtype=np.int32)
    sol = QBF().solve(prob, B=8, use_metrics="all", device="cuda")
    arr = sol.metrics[0]["energy_arr"]
    Q = prob.build_Q_matrix().astype(np.int64)
    assert arr.shape == (16,)
    for i in range(16):
        g = i ^ (i >> 1)
        states = [(g << 8) | t for t in range(256)]
        assert arr[i] == min(oracle.energy(Q, s) for s in states)
    assert sol.metrics[0]["time_arr"].shape == (16,)


def test_qbf1_contract():
    prob = RandomQUBO.init_random_instance(N=30, seed=42, dtype=np.int16)
    sol = QB_F1().solve(prob)
    assert sol.solver_class == "QB_F1" and sol.epochs_cnt == 1 and isinstance(sol.best_E, int)
    ref = oracle.brute_i64(prob.build_Q_matrix(), 18)
    assert sol.best_E == ref[0]
    assert int("".join(map(str, sol.best_X)), 2) == ref[1]


def test_parallel_matches_qbf():
    prob = RandomQUBO.init_random_instance(seed=9, N=22, dtype=np.int32)
    whole = QBF().solve(prob, device="cuda")
    for splits in (1, 2, 8):
        sol = Parallel(QBF, splits, ["cuda"]).solve(prob)
        assert sol.best_E == whole.best_E
        assert np.array_equal(sol.best_X, whole.best_X)
        assert sol.epochs_cnt == splits and len(sol.worker_logs) == splits
    sol = Parallel(QB_F1, 4, ["cuda:0"]).solve(RandomQUBO.init_random_instance(seed=9, N=20, dtype=np.int16))
    assert sol.solver_class == "QB_F1"


def test_solve_batch_matches_qbf():
    probs = [RandomQUBO.init_random_instance(seed=s, N=10 + s, dtype=np.int32) for s in range(8)]
    probs += [TSP.init_random_instance(n_cities=4, seed=s) for s in range(3)]
    sols = solve_batch(probs)
    for p, s in zip(probs, sols):
        one = QBF().solve(p, device="cuda")
        assert s.best_E == one.best_E and np.array_equal(s.best_X, one.best_X)
