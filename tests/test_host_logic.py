"""Host-side logic that runs without a GPU: arithmetic-path selection, argument checks,
the reference's A/B split rules."""
import numpy as np
import pytest

from qubo_lib_amd import engine
from qubo_lib_amd.problems import RandomQUBO
from qubo_lib_amd.solvers import QB_F1, QBF, Parallel
from qubo_lib_amd.solvers.qbf import adjust_AB_to_fit_memory, define_chunk_size


def test_path_selection():
    p = engine.prepare(np.array([[1, 2], [3, 4]], dtype=np.int16))
    assert p.kind == "i32" and p.integer and p.bound == 1 + 4 + 5
    assert p.matrix.tolist() == [[1, 5], [0, 4]]          # folded upper triangle
    big = np.array([[2 ** 40, 0], [0, 1]], dtype=np.int64)
    p = engine.prepare(big)
    assert p.kind == "f64" and p.integer
    # huge entries that cancel in Q_ij + Q_ji still fit the int32 path
    c = np.array([[0, 2 ** 40], [-(2 ** 40), 0]], dtype=np.int64)
    assert engine.prepare(c).kind == "i32"
    with pytest.raises(OverflowError):
        engine.prepare(np.array([[2 ** 62, 0], [0, 2 ** 62]], dtype=np.int64))
    # multiples of one power of two: the integer walk on Q / u, energies u * E_int
    f = engine.prepare(np.eye(3, dtype=np.float32) * 0.5)
    assert f.kind == "i32" and not f.integer and f.unit == 0.5 and f.matrix.tolist() == np.eye(3).tolist()
    assert engine.prepare(np.eye(3, dtype=np.float32) * 0.5, lattice=False).kind == "f64"
    assert engine.prepare(np.eye(3) * 0.1).kind == "f64"                       # 0.1: no short lattice
    assert engine.prepare(np.full((4, 4), 2.0 ** 40 + 0.5)).kind == "f64"     # bound beyond int32
    assert engine.lattice_bits(np.array([[0.75, 0.0], [-1.5, 2.0]])) == 2
    assert engine.lattice_bits(np.zeros((2, 2))) == 0
    # integer-valued floats take the exact int32 walk (energies stay floats)
    g = engine.prepare(np.eye(3, dtype=np.float32))
    assert g.kind == "i32" and not g.integer
    # integer-valued floats whose integer bound is 2^53 or more are still solvable: the
    # float path takes them (the integer paths would refuse them)
    h = engine.prepare(np.full((40, 40), 1e14))
    assert h.kind == "f64" and not h.integer and h.bound >= 2.0 ** 53
    assert engine.prepare(np.full((40, 40), 1e14) + 0.5).kind == "f64"
    with pytest.raises(ValueError):
        engine.prepare(np.array([[np.nan]]))
    with pytest.raises(ValueError):
        engine.prepare(np.zeros((2, 3)))
    with pytest.raises(ValueError):
        engine.prepare(np.zeros((57, 57)))
    assert engine.prepare(np.ones((2, 2), dtype=bool)).kind == "i32"


def test_reference_split_rules():
    assert define_chunk_size(30) == (10, 20)
    assert define_chunk_size(12) == (0, 12)
    assert define_chunk_size(30, A=4) == (4, 26)
    assert define_chunk_size(30, B=7) == (23, 7)
    # qbf.py:455-466: largest B <= N-2 with (A + B) * 2^B * sizeof <= mem
    A, B = adjust_AB_to_fit_memory(30, 1.0, 4)
    assert A + B == 30 and (A + B) * 2 ** B * 4 <= 2 ** 30 < (30) * 2 ** (B + 1) * 4
    assert define_chunk_size(30, use_mem_GB=1.0, size_of=4) == (A, B)


def test_solver_argument_errors():
    prob = RandomQUBO.init_random_instance(seed=1, N=6, dtype=np.int16)
    with pytest.raises(ValueError, match="obsolete"):
        QBF().solve(prob, use_gpu=True)
    with pytest.raises(ValueError, match="obsolete"):
        QBF().solve(prob, gpu_id=0)
    with pytest.raises(ValueError, match="device"):
        QBF().solve(prob, devices=["cuda"])
    with pytest.raises(ValueError, match="Unknown device"):
        QBF().solve(prob, device="tpu")
    with pytest.raises(ValueError, match="Unknown device"):
        QBF().solve(prob, device="cuda:x")
    with pytest.raises(ValueError, match="GPU only"):
        QB_F1().solve(prob, device="cpu")
    with pytest.raises(TypeError):
        QB_F1().solve(RandomQUBO.init_random_instance(seed=1, N=6, dtype=np.int32))


def test_parallel_validation():
    with pytest.raises(ValueError, match="power of 2"):
        Parallel(QBF, 3, ["cuda"])
    with pytest.raises(ValueError, match="positive"):
        Parallel(QBF, 0, ["cuda"])
    with pytest.raises(ValueError, match="empty"):
        Parallel(QBF, 2, [])
    with pytest.raises(ValueError, match="mix"):
        Parallel(QBF, 2, ["cpu", "cuda"])
    with pytest.raises(ValueError, match="GPU only"):
        Parallel(QB_F1, 2, ["cpu"])
    with pytest.raises(ValueError, match="Unknown device"):
        Parallel(QBF, 2, ["tpu"])
    assert Parallel(QBF, 2, ["cpu"] * 2)._devices == ["cpu", "cpu"]
    with pytest.raises(ValueError, match="Unsupported"):
        Parallel(object, 2, ["cuda"])
    p = Parallel(QBF, 4, ["cuda"] * 2)
    assert p._devices == ["cuda:0", "cuda:1"]


def test_state_bits():
    assert engine.state_bits(0b1011, 4).tolist() == [1, 0, 1, 1]
    assert engine.state_bits(1, 3).dtype == np.int32


def test_graft_entry_build_wiring(monkeypatch):
    """build() calls the native build and the oracle recipe, then imports the package
    (compilers replaced by stubs here; the real build runs outside the test suite)."""
    import __graft_entry__ as g
    import build_native
    import oracle
    calls = []
    monkeypatch.setattr(build_native, "build_all", lambda: calls.append("native") or [])
    monkeypatch.setattr(oracle, "build", lambda *a, **k: calls.append("oracle"))
    g.build()
    assert calls == ["native", "oracle"]


def test_tuning_table_files(tmp_path, monkeypatch):
    """The committed gfx950 table parses and is valid for every n; lookup honours the
    environment override and 'off'; malformed tables are rejected."""
    import json

    from qubo_lib_amd import _native
    monkeypatch.delenv("QUBO_LIB_AMD_TUNING", raising=False)
    path = _native.tuning_path("gfx950")
    assert path is not None and path.name == "gfx950.json"
    tables = _native.read_tuning(path)
    assert set(tables) == {(k, m) for k in ("i32", "f64") for m in ("single", "batch")}
    for (kind, mode), t in tables.items():
        assert len(t) == 57
        tuned = [n for n, v in enumerate(t) if v >= 0]
        assert tuned and all(1 <= t[n] <= n for n in tuned)
    assert _native.tuning_path("gfx000") is None
    monkeypatch.setenv("QUBO_LIB_AMD_TUNING", "off")
    assert _native.tuning_path("gfx950") is None
    bad = tmp_path / "bad.json"
    bad.write_text(json.dumps({"walk_bits": {"i32/single": [-1, 5]}}))
    monkeypatch.setenv("QUBO_LIB_AMD_TUNING", str(bad))
    assert _native.tuning_path("gfx950") == bad
    with pytest.raises(ValueError):
        _native.read_tuning(bad)
    bad.write_text(json.dumps({"walk_bits": {"i16/single": [-1]}}))
    with pytest.raises(ValueError):
        _native.read_tuning(bad)


def test_device_draw_selection():
    from qubo_lib_amd.batch import _drawable_on_device
    from qubo_lib_amd.problems import MaxCut
    r = [RandomQUBO.init_random_instance(seed=i, N=12, dtype=np.int16) for i in range(3)]
    assert _drawable_on_device(r)
    assert not _drawable_on_device(r + [RandomQUBO.init_random_instance(seed=9, N=12, dtype=np.int32)])
    assert not _drawable_on_device(r + [MaxCut.random_gnp(8, 0.5, seed=1)])
    big = [RandomQUBO.init_random_instance(seed=1, N=12, low=-2 ** 30, high=2 ** 30, dtype=np.int32)]
    assert not _drawable_on_device(big)            # int32 bound cannot be guaranteed
    assert _drawable_on_device([RandomQUBO.init_random_instance(seed=1, N=40, dtype=np.float32)])
    assert not _drawable_on_device([])


def test_prepare_fold_matches_exact_python_ints():
    """The vectorised int64 fold equals the fold in Python ints for every integer dtype."""
    rng = np.random.default_rng(30)
    for dtype, lo, hi in [(np.int8, -128, 127), (np.int16, -32768, 32767), (np.int32, -2 ** 31, 2 ** 31 - 1),
                          (np.uint16, 0, 65535), (np.int64, -2 ** 31, 2 ** 31), (np.int64, -2 ** 40, 2 ** 40)]:
        for n in (1, 5, 17):
            Q = rng.integers(lo, hi, (n, n), dtype=np.int64, endpoint=True).astype(dtype)
            Qo = Q.astype(object)
            up = np.triu(Qo, 1) + np.triu(Qo.T, 1)
            bound = int(sum(abs(v) for v in np.diag(Qo))) + int(sum(abs(v) for v in up.ravel()))
            if bound >= engine.EXACT_F64_BOUND:
                with pytest.raises(OverflowError):
                    engine.prepare(Q)
                continue
            p = engine.prepare(Q)
            assert p.bound == bound and p.integer
            ref = (up + np.diag(np.diag(Qo))).astype(np.int64)
            assert np.array_equal(p.matrix.astype(np.int64), ref)
