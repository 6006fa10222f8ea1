"""CPU checks of the C-ABI library and the host-side generators (no GPU)."""
import ctypes

import numpy as np
import pytest

import meerkat_amd as mk
from meerkat_amd import _lib, gen


def test_library_loads_and_exports_every_header_symbol():
    L = mk.lib()
    names = mk.header_symbols()
    assert len(names) > 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_every_exported_symbol_is_declared():
    import subprocess

    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T mk_" in line}
    assert exported == set(mk.header_symbols())


def test_version_and_error_text():
    assert mk.lib().mk_version() >= 100
    # a NULL handle is rejected without touching the GPU
    rc = mk.lib().mk_pool_stats(None, None, None)
    assert rc == _lib.MK_ERR_BAD_ARG
    assert "NULL" in _lib.last_error()
    rc = mk.lib().mk_graph_create(None, 10, 0, 0.7, None, ctypes.byref(ctypes.c_void_p()))
    assert rc == _lib.MK_ERR_BAD_ARG


def test_rmat_properties_and_determinism():
    n, s, d, w = gen.rmat(10, 16, seed=1, weighted=True)
    assert n == 1024
    k = gen.keys_of(s, d)
    assert (np.diff(k.astype(np.int64)) > 0).all()  # sorted, unique
    assert (s != d).all() and s.max() < n and d.max() < n
    assert w.min() >= 1 and w.max() <= 100
    n2, s2, d2, w2 = gen.rmat(10, 16, seed=1, weighted=True)
    assert np.array_equal(s, s2) and np.array_equal(d, d2) and np.array_equal(w, w2)
    _, s3, _, _ = gen.rmat(10, 16, seed=2)
    assert not np.array_equal(s, s3[: len(s)]) or len(s) != len(s3)
    # RMAT skew: vertex 0 is the hub (a = 0.57 on both halves)
    deg = np.bincount(s, minlength=n)
    assert deg.argmax() == 0
    # quadrant shares on the top bit: P(src top bit = 1) = c + d = 0.24
    n20, s20, _, _ = gen.rmat(14, 16, seed=4)
    share = (s20 >= n20 // 2).mean()
    assert 0.18 < share < 0.28


def test_rmat_fixture_digest():
    # the generator's output is part of the workload definition: pin it
    _, s, d, w = gen.rmat(8, 8, seed=1, weighted=True)
    digest = int((s.astype(np.uint64) * 1000003 + d.astype(np.uint64) * 7 + w).sum())
    assert len(s) == 1508
    assert digest == 105161139707


def test_undirected_rmat_is_symmetric():
    _, s, d, _ = gen.rmat(10, 8, seed=7, undirected=True)
    k = set(gen.keys_of(s, d).tolist())
    assert all(((int(b) << 32) | int(a)) in k for a, b in zip(s[:500], d[:500]))


def test_er_properties():
    n, s, d, _ = gen.erdos_renyi(5000, 40000, seed=5)
    assert (s != d).all() and len(s) <= 40000 and len(s) > 39000
    k = gen.keys_of(s, d)
    assert (np.diff(k.astype(np.int64)) > 0).all()


def test_update_batches():
    n, s, d, _ = gen.rmat(10, 8, seed=1)
    keys = gen.keys_of(s, d)
    bs, bd, bw = gen.insert_batch(keys, n, 2000, seed=2, weighted=True)
    bk = gen.keys_of(bs, bd)
    assert len(np.unique(bk)) == 2000
    assert not np.isin(bk, keys).any()
    assert (bs != bd).all() and bw.min() >= 1 and bw.max() <= 100
    ds, dd = gen.delete_batch(keys, 1500, seed=3)
    dk = gen.keys_of(ds, dd)
    assert len(np.unique(dk)) == 1500 and np.isin(dk, keys).all()
    with pytest.raises(mk.MeerkatError):
        gen.delete_batch(keys, len(keys) + 1, seed=1)


def test_tune_knobs_validate_without_gpu():
    from meerkat_amd import autotune

    before = autotune.knobs()
    assert before == {"waves_per_cu": 16, "ms_unroll": 8, "ms_waves_per_cu": 24,
                      "pr_waves_per_cu": 32, "orphan_repair": 0,
                      "persistent_steps": 1, "tail_enter": 128,
                      "coop_launch": 0, "mut_waves_per_cu": 32, "barrier_timeout_ms": 5000,
                      "pull_enter": 14, "pull_exit": 24, "sssp_delta": 0, "sssp_far_cap": 0}
    try:
        autotune.set_knob("waves_per_cu", 32)
        assert autotune.get_knob("waves_per_cu") == 32
        for name, bad in [("waves_per_cu", 0), ("waves_per_cu", 65), ("ms_unroll", 6),
                          ("ms_waves_per_cu", 0), ("ms_waves_per_cu", 65), ("pr_waves_per_cu", 0),
                          ("orphan_repair", 2), ("persistent_steps", 3), ("tail_enter", 1025),
                          ("pull_enter", -1), ("pull_exit", 0), ("pull_exit", 4097),
                          ("no_such_knob", 1)]:
            with pytest.raises(_lib.MeerkatError) as e:
                autotune.set_knob(name, bad)
            assert e.value.code == _lib.MK_ERR_BAD_ARG
        assert autotune.knobs()["waves_per_cu"] == 32
    finally:
        for k, v in before.items():
            autotune.set_knob(k, v)


def test_profile_levels_validate_without_gpu():
    from meerkat_amd import profiling

    for on in (0, 1, 2, 0):
        _lib.call("mk_profile_enable", on)
    with pytest.raises(_lib.MeerkatError) as e:
        _lib.call("mk_profile_enable", 3)
    assert e.value.code == _lib.MK_ERR_BAD_ARG
    profiling.enable(True, loops_only=True)
    profiling.enable(False)
    profiling.reset()
    assert profiling.totals() == {} and profiling.steps() == []
