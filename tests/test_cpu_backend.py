"""device='cpu' (the reference QBF's default device): libqbx_cpu.so, the GPU kernels' lane
walk compiled for the host with OpenMP over lanes, checked against the oracle; and the
Parallel merge's tie key on the same backend.  Runs without a GPU."""
import json
from pathlib import Path

import numpy as np
import pytest

import oracle
from helpers import KINDS, random_instance
from qubo_lib_amd import _native, engine
from qubo_lib_amd.problems import TSP, Just_Q, MaxCut, RandomQUBO, ReducedProblem
from qubo_lib_amd.problems import generators as gen
from qubo_lib_amd.solvers import QBF, Parallel

GOLDEN = Path(__file__).resolve().parent / "golden"


def bits(x):
    return "".join(map(str, np.asarray(x).tolist()))


def test_random_integer_problems_vs_brute_force():
    rng = np.random.default_rng(70)
    for t in range(300):
        n = int(rng.integers(1, 17))
        Q = random_instance(rng, n, KINDS[t % len(KINDS)])
        B = int(rng.integers(0, n + 1))
        r = engine.solve_matrix(Q, device="cpu", tie_B=B)
        assert (r.E, r.state) == oracle.brute_i64(Q, B), (n, B, KINDS[t % len(KINDS)])
        assert r.info["walk"] == "int32"


def test_every_tiling_gives_the_same_answer():
    rng = np.random.default_rng(71)
    Q = rng.integers(-3, 4, (18, 18))          # many exact ties
    for B in (0, 5, 18):
        ref = oracle.brute_i64(Q, B)
        for M in (1, 4, 6, 7):
            for L in range(M, min(18, M + 13) + 1, 2):
                r = engine.solve_matrix(Q, device="cpu", tie_B=B, inner_bits=M, walk_bits=L)
                assert (r.E, r.state) == ref, (B, M, L)
                assert r.info["inner_bits"] == M and r.info["walk_bits"] == L


def test_random_float_problems_vs_fp64_brute_force():
    rng = np.random.default_rng(72)
    for t in range(150):
        n = int(rng.integers(1, 17))
        Q = rng.standard_normal((n, n)) * 10.0 ** rng.integers(-3, 4)
        B = int(rng.integers(0, n + 1))
        r = engine.solve_matrix(Q, device="cpu", tie_B=B)
        E, s = oracle.brute_f64(Q, B)
        assert r.state == s and abs(r.E - E) <= 1e-9 * max(1.0, abs(E)), (n, B)


def test_float_exact_ties():
    # quarter-integer entries: every energy is exact, so ties are exact and follow the rank
    rng = np.random.default_rng(73)
    for t in range(40):
        n = int(rng.integers(2, 15))
        Q = rng.integers(-4, 5, (n, n)) / 4.0 + 0.125 * (t % 2)
        p = engine.prepare(Q, lattice=False)
        assert p.kind == "f64"
        B = int(rng.integers(0, n + 1))
        r = engine.solve_matrix(p, device="cpu", tie_B=B)
        assert (r.E, r.state) == oracle.brute_f64(Q, B)
        # the same matrix as the exact integer problem 8 Q: identical answer
        z = engine.solve_matrix(Q, device="cpu", tie_B=B)
        assert z.kind == "i32" and (z.E, z.state) == (r.E, r.state)


def test_reference_cpu_configuration_n24():
    """The reference's CPU configuration (n = 24 int32, B = 20, A = 4) through the drop-in
    solver's default device, against the brute-force goldens."""
    gold = json.loads((GOLDEN / "oracle_goldens.json").read_text())
    for seed in (0, 1, 2):
        Q = gen.symmetric_int_qubo(24, seed, -100, 100, np.int32)
        sol = QBF().solve(Just_Q(Q))
        g = gold[f"sym_int32_n24_s{seed}"]
        assert sol.best_E == g["E"] and bits(sol.best_X) == g["state_bits"]
        assert sol.epochs_cnt == 16 and sol.run_info["walk"] == "int32"


def test_known_answers():
    assert QBF().solve(MaxCut("test")).best_E == -19
    g = json.loads((GOLDEN / "notebook.json").read_text())["tsp5_seed42"]
    sol = QBF().solve(TSP.init_random_instance(n_cities=5, seed=42), device="cpu")
    assert bits(sol.best_X) == g["best_X"]
    # the notebook prints QBF's float32 sums; the solution's FP64 energy is within them
    assert abs(sol.best_E - g["best_E"]) < 0.05


def test_tie_rule_kwargs_and_reduced_problem():
    prob = Just_Q(np.diag([-1, 0, 1, 1, 1, 1, 1, 1]).astype(np.int32))
    assert bits(QBF().solve(prob, B=0).best_X) == "11000000"
    assert bits(QBF().solve(prob, A=2).best_X) == "11000000"
    assert bits(QBF().solve(prob).best_X) == "10000000"
    base = RandomQUBO.init_random_instance(seed=3, N=14, dtype=np.int32)
    rp = ReducedProblem(base, np.array([1, 0, 1]))
    sol = QBF().solve(rp)
    assert sol.prefix_E == rp.prefix_E
    assert sol.best_E == oracle.brute_i64(rp.build_Q_matrix())[0]


def test_range_errors():
    cpu = _native.get_backend("cpu")
    with pytest.raises(OverflowError):
        cpu.solve_i32(np.full((4, 4), 2 ** 29, dtype=np.int32))
    with pytest.raises(ValueError):
        cpu.solve_i32(np.zeros((57, 57), dtype=np.int32))
    with pytest.raises(ValueError):
        cpu.solve_i32(np.zeros((30, 30), dtype=np.int32), walk_bits=25, inner_bits=4)   # L > M + 16
    with pytest.raises(OverflowError):
        cpu.solve_f64(np.full((3, 3), 1e301))
    # integer Q beyond the int32 bound: the engine takes the exact FP64 walk instead
    Q = np.diag([2 ** 40, -(2 ** 40) + 3, 5]).astype(np.int64)
    r = engine.solve_matrix(Q, device="cpu")
    assert (r.E, r.state) == oracle.brute_i64(Q) and r.kind == "f64"


def test_solve_part_needs_a_gpu():
    with pytest.raises(ValueError, match="cuda"):
        engine.solve_part(np.eye(4, dtype=np.int32), 0, 2, device="cpu")


# ------------------------------------------------------------------ Parallel tie key
def tie_heavy(rng, n, free=3):
    """Minimisers all have x0 = 1 (a large negative Q_00) and x1..x_free are free (zero
    rows and columns): every minimum is an exact tie over 2^free states, spread inside the
    reference prefix when n > 20, where the rank order of a task with odd-parity fixed bits
    runs backwards."""
    Q = rng.integers(-5, 6, (n, n))
    Q[1:free + 1, :] = 0
    Q[:, 1:free + 1] = 0
    Q[0, 0] = -1000
    return Q


def parallel_cases():
    rng = np.random.default_rng(74)
    cases = [Just_Q(gen.tsp_int(nc, seed=s)) for nc in (4, 5) for s in range(3)]
    cases += [TSP.init_random_instance(seed=s, n_cities=nc) for nc in (4, 5) for s in range(2)]
    cases += [MaxCut.random_gnp(int(n), 0.5, seed=int(n)) for n in (12, 17, 22)]
    cases += [Just_Q(tie_heavy(rng, int(n))) for n in (21, 22, 23, 24)]
    return cases


@pytest.mark.parametrize("case", range(17))
def test_parallel_best_x_equals_qbf(case):
    prob = parallel_cases()[case]
    whole = QBF().solve(prob, device="cpu")
    for k in (1, 2, 3):
        sol = Parallel(QBF, 2 ** k, ["cpu", "cpu"]).solve(prob)
        assert bits(sol.best_X) == bits(whole.best_X), (case, k)
        assert sol.best_E == whole.best_E and sol.epochs_cnt == 2 ** k


def test_parallel_tie_key_with_explicit_widths():
    rng = np.random.default_rng(75)
    for t in range(30):
        n = int(rng.integers(4, 19))
        Q = tie_heavy(rng, n, free=min(3, n - 1))
        kw = {"B": int(rng.integers(0, n + 1))} if t % 2 else {"A": int(rng.integers(0, n + 1))}
        whole = QBF().solve(Just_Q(Q), device="cpu", **kw)
        for k in range(0, min(n, 4) + 1):
            sol = Parallel(QBF, 2 ** k, ["cpu"]).solve(Just_Q(Q), **kw)
            assert bits(sol.best_X) == bits(whole.best_X), (n, kw, k)


def test_power_of_two_lattice_floats_take_the_integer_walks():
    """Float Q whose entries are multiples of u = 2^-k (MaxCut weights / 4, half-integer
    RandomQUBOs): solved as the integer problem Q / u, energies u * E_int -- the float
    path's exact answer, state and ties included."""
    rng = np.random.default_rng(77)
    cases = [gen.maxcut_gnp(18, 0.5, seed=5) / 4.0, gen.symmetric_int_qubo(17, 6).astype(np.float64) / 2.0,
             rng.integers(-8, 9, (15, 15)) / 8.0, gen.tsp_int(4, seed=7) / 16.0]
    for Q in cases:
        p = engine.prepare(Q)
        assert p.kind == "i32" and p.unit < 1.0
        z = engine.solve_matrix(p, device="cpu")
        f = engine.solve_matrix(engine.prepare(Q, lattice=False), device="cpu")
        E, s = oracle.brute_f64(np.asarray(Q, dtype=np.float64))
        assert (z.E, z.state) == (f.E, f.state) == (E, s) and isinstance(z.E, float)
        _, mz, _ = engine.solve_matrix_minima(p, device="cpu", tie_B=10)
        _, mf, _ = engine.solve_matrix_minima(engine.prepare(Q, lattice=False), device="cpu", tie_B=10)
        assert np.array_equal(mz, mf)
