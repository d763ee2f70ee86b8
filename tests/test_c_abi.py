"""include/qbx.h serves plain C and C++ callers: it compiles as C99 and C++17, and a C99
program linked against libqbx.so (tests/c/qbx_c_check.c) gets the Python binding's answers."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
C_CHECK = ROOT / "tests" / "_build" / "qbx_c_check"


@pytest.mark.parametrize("cmd", [["gcc", "-std=c99", "-x", "c"], ["g++", "-std=c++17", "-x", "c++"]])
def test_header_compiles_for_c_and_cpp(cmd, tmp_path):
    src = tmp_path / "use.txt"
    src.write_text('#include "qbx.h"\nint f(qbx_ctx_t *c) { qbx_result_t r; qbx_opts_t o = {-1, -1, -1, 0};\n'
                   'return qbx_solve_i32(c, 0, 0, &o, &r); }\n')
    r = subprocess.run(cmd + ["-Wall", "-Werror", "-fsyntax-only", f"-I{ROOT / 'include'}", str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def _lcg_matrix(n=12, seed=12345):
    x, out = seed, []
    for _ in range(n * n):
        x = (x * 1103515245 + 12345) & 0xFFFFFFFF
        out.append(((x >> 16) % 41) - 20)
    return np.array(out, dtype=np.int32).reshape(n, n)


@pytest.mark.gpu
def test_c_program_matches_python_binding(gpu_ctx):
    assert C_CHECK.exists(), "build it with __graft_entry__.build()"
    r = subprocess.run([str(C_CHECK)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    lines = dict(line.split(" ", 1) for line in r.stdout.strip().splitlines())
    Q = _lcg_matrix()
    a = gpu_ctx.solve_i32(Q)
    b = gpu_ctx.solve_i32(-Q)
    assert lines["solve"].split() == [str(a.E_int), str(a.state)]
    assert lines["batch"].split() == [str(a.E_int), str(a.state), str(b.E_int), str(b.state)]
    assert lines["server"].split() == lines["batch"].split()
    c = gpu_ctx.solve_i32(_lcg_matrix(20, 777))
    assert lines["parts"].split() == [str(c.E_int), str(c.state)]
    t = gpu_ctx.random_tsp_batch([42], [5], 0, 100, np.float32)[0]
    E, s = lines["tsp"].split()
    assert (float(E), int(s)) == (t.E, t.state)
    f = gpu_ctx.solve_f64(_lcg_matrix(20, 777).astype(np.float64) / 4)
    E, s = lines["float"].split()
    assert (float(E), int(s)) == (f.E, f.state)
