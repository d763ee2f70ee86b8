"""The host backend (device='cpu') and the GPU give the same answers, and Parallel over
GPU tasks merges by the full problem's tie key (best_X equal to QBF's bit for bit)."""
import json
from pathlib import Path

import numpy as np
import pytest

from qubo_lib_amd import engine
from qubo_lib_amd.problems import Just_Q, RandomQUBO, TSP
from qubo_lib_amd.problems import generators as gen
from qubo_lib_amd.solvers import QB_F1, QBF, Parallel
from test_cpu_backend import bits, parallel_cases, tie_heavy

pytestmark = pytest.mark.gpu
GOLD = json.loads((Path(__file__).resolve().parent / "golden" / "oracle_goldens.json").read_text())


def test_cpu_and_gpu_agree_on_the_reference_cpu_configuration(gpu_ctx):
    for seed in (0, 1, 2):
        Q = gen.symmetric_int_qubo(24, seed, -100, 100, np.int32)
        c = QBF().solve(Just_Q(Q), device="cpu")
        g = QBF().solve(Just_Q(Q), device="cuda")
        assert c.best_E == g.best_E == GOLD[f"sym_int32_n24_s{seed}"]["E"]
        assert bits(c.best_X) == bits(g.best_X) == GOLD[f"sym_int32_n24_s{seed}"]["state_bits"]


def test_cpu_and_gpu_agree_on_float_and_integer_problems(gpu_ctx):
    probs = [gen.symmetric_normal_qubo(n, s) for n, s in ((20, 1), (24, 2), (26, 3))]
    probs += [TSP.init_random_instance(n_cities=5, seed=42).build_Q_matrix(),
              gen.maxcut_gnp(24, 0.5, seed=4), gen.symmetric_int_qubo(27, 5)]
    for Q in probs:
        c = engine.solve_matrix(Q, device="cpu")
        g = engine.solve_matrix(Q, device=0)
        # same inner-block width: the FP64 energies are summed in the same order, bit for bit
        assert (c.E, c.state) == (g.E, g.state), Q.shape


@pytest.mark.parametrize("case", range(17))
def test_parallel_on_gpu_equals_qbf(gpu_ctx, case):
    prob = parallel_cases()[case]
    whole = QBF().solve(prob, device="cuda")
    for k in (1, 2, 3):
        sol = Parallel(QBF, 2 ** k, ["cuda"]).solve(prob)
        assert bits(sol.best_X) == bits(whole.best_X), (case, k)
        assert sol.best_E == whole.best_E


def test_parallel_qbf1_tie_key(gpu_ctx):
    rng = np.random.default_rng(76)
    for n in (20, 22, 23):
        prob = Just_Q(tie_heavy(rng, n).astype(np.int16))
        whole = QB_F1().solve(prob)
        for k in (1, 2, 3):
            sol = Parallel(QB_F1, 2 ** k, ["cuda:0"]).solve(prob)
            assert bits(sol.best_X) == bits(whole.best_X) and sol.best_E == whole.best_E, (n, k)
    prob = RandomQUBO.init_random_instance(seed=9, N=22, dtype=np.int16)
    assert bits(Parallel(QB_F1, 4, ["cuda"]).solve(prob).best_X) == bits(QB_F1().solve(prob).best_X)


def test_lattice_floats_on_the_gpu(gpu_ctx):
    """Float Q on a power-of-two lattice (MaxCut weights / 4, halved int16) takes the integer
    walks on the GPU too, with the float path's exact answer."""
    for Q in (gen.maxcut_gnp(30, 0.5, seed=8) / 4.0, gen.symmetric_int_qubo(28, 9).astype(np.float64) / 2.0):
        p = engine.prepare(Q)
        z = engine.solve_matrix(p, device=0)
        f = engine.solve_matrix(engine.prepare(Q, lattice=False), device=0)
        c = engine.solve_matrix(p, device="cpu")
        assert p.kind == "i32" and z.info["walk"] in ("pair16", "pair16_rebased")
        assert (z.E, z.state) == (f.E, f.state) == (c.E, c.state)
