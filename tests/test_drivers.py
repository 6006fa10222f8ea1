"""The command-line drivers (meerkat_amd/bin/mk_*), built over the C++
classes of include/meerkat/: argument handling on the CPU; on the GPU, runs
over text graph and batch files (the reference drivers' formats) whose
final results (--dump) must equal the oracle's on the updated graph.

The SSSP driver draws weights with the C library's srand/rand exactly as the
reference's src/SSSP.cu:17-19,50-57,110-116 does; the test draws the same
sequence through ctypes from the same C library."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import oracle
from meerkat_amd import gen

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "meerkat_amd", "bin")


def prog(name):
    path = os.path.join(BIN, name)
    if not os.path.exists(path):
        pytest.fail(f"{path} is not built (make -C meerkat_amd/drivers)")
    return path


def run(args, timeout=120):
    r = subprocess.run(args, capture_output=True, text=True, timeout=timeout)
    return r.returncode, r.stdout, r.stderr


@pytest.mark.parametrize("name,nargs", [("mk_bfs", 4), ("mk_sssp", 5), ("mk_pagerank", 6),
                                        ("mk_tc", 4)])
def test_usage_and_missing_file(name, nargs, tmp_path):
    rc, _, err = run([prog(name)])
    assert rc == 1 and "usage" in err
    # a missing graph file is reported before any device work
    args = [str(tmp_path / "absent.txt"), "0", str(tmp_path / "u"), "1", "1", "1"][:nargs]
    rc, _, err = run([prog(name)] + args)
    assert rc == 2 and "cannot open" in err


def _libc_weights(seed, count):
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(ctypes.c_uint(seed))
    return np.array([1 + libc.rand() % 100 for _ in range(count)], np.uint32)


def _csv(out):
    vals = out.strip().splitlines()[0].split(",")
    return vals[0], [float(x) for x in vals[1:]]


@pytest.mark.gpu
@pytest.mark.parametrize("op", ["0", "1"])
def test_bfs_driver(gpu, tmp_path, op):
    n, s, d, _ = gen.rmat(12, 8, seed=1)
    keys = gen.keys_of(s, d)
    f = tmp_path / "g.txt"
    gen.write_edge_file(f, n, s, d)
    pre = str(tmp_path / "upd")
    for i in range(3):
        if op == "0":
            bs, bd, _ = gen.insert_batch(keys, n, 500, seed=10 + i)
            keys = np.unique(np.r_[keys, gen.keys_of(bs, bd)])
        else:
            bs, bd = gen.delete_batch(keys, 500, seed=20 + i)
            keys = np.setdiff1d(keys, gen.keys_of(bs, bd))
        gen.write_batch_file(pre, i, bs, bd)
    dump = tmp_path / "dist.bin"
    rc, out, err = run([prog("mk_bfs"), str(f), op, pre, "3", f"--dump={dump}"])
    assert rc == 0, err
    name, times = _csv(out)
    assert name == str(f) and len(times) == 4 and all(t >= 0 for t in times)
    s2, d2 = gen.split_keys(keys)
    got = np.fromfile(dump, np.uint32)
    assert np.array_equal(got, oracle.bfs(oracle.csr(n, s2, d2), 0))


@pytest.mark.gpu
@pytest.mark.parametrize("op", ["0", "1"])
def test_sssp_driver_weights_follow_srand(gpu, tmp_path, op):
    n, s, d, _ = gen.rmat(12, 8, seed=2)
    seed = 1234
    f = tmp_path / "g.txt"
    gen.write_edge_file(f, n, s, d)
    w_all = _libc_weights(seed, len(s) + 3 * 400)
    w = w_all[: len(s)]
    keys = gen.keys_of(s, d)
    wmap = dict(zip(keys.tolist(), w.tolist()))
    pre = str(tmp_path / "upd")
    pos = len(s)
    for i in range(3):
        if op == "0":
            bs, bd, _ = gen.insert_batch(keys, n, 400, seed=30 + i)
            for k, x in zip(gen.keys_of(bs, bd).tolist(), w_all[pos:pos + 400].tolist()):
                wmap[k] = x
            keys = np.unique(np.r_[keys, gen.keys_of(bs, bd)])
        else:
            bs, bd = gen.delete_batch(keys, 400, seed=40 + i)
            for k in gen.keys_of(bs, bd).tolist():
                del wmap[k]
            keys = np.setdiff1d(keys, gen.keys_of(bs, bd))
        pos += 400
        gen.write_batch_file(pre, i, bs, bd)
    dump = tmp_path / "dist.bin"
    rc, out, err = run([prog("mk_sssp"), str(f), op, pre, "3", str(seed), f"--dump={dump}"])
    assert rc == 0, err
    assert len(_csv(out)[1]) == 4
    s2, d2 = gen.split_keys(keys)
    w2 = np.array([wmap[k] for k in keys.tolist()], np.uint32)
    got = np.fromfile(dump, np.uint32)
    assert np.array_equal(got, oracle.dijkstra(oracle.csr(n, s2, d2, w2), 0))


@pytest.mark.gpu
def test_pagerank_driver_dedups_batches(gpu, tmp_path):
    n, s, d, _ = gen.erdos_renyi(4000, 24000, seed=3)
    keys = gen.keys_of(s, d)
    f = tmp_path / "g.txt"
    gen.write_edge_file(f, n, s, d)
    pre = str(tmp_path / "upd")
    for i in range(2):
        bs, bd, _ = gen.insert_batch(keys, n, 1000, seed=50 + i)
        keys = np.unique(np.r_[keys, gen.keys_of(bs, bd)])
        # every batch edge twice: the driver de-duplicates, as the reference
        gen.write_batch_file(pre, i, np.r_[bs, bs], np.r_[bd, bd])
    dump = tmp_path / "pr.bin"
    rc, out, err = run([prog("mk_pagerank"), str(f), "0", pre, "2", "0.85", "1e-7",
                        f"--dump={dump}"])
    assert rc == 0, err
    assert len(_csv(out)[1]) == 3
    s2, d2 = gen.split_keys(keys)
    g = oracle.csr(n, s2, d2)
    ref, _, delta = oracle.pagerank(oracle.transpose(g), g.degrees(), 0.85, 1e-14, 5000)
    got = np.fromfile(dump, np.float32)
    assert np.abs(got.astype(np.float64) - ref).sum() <= 1e-6


@pytest.mark.gpu
@pytest.mark.parametrize("op", ["0", "1"])
def test_tc_driver_counts_and_components(gpu, tmp_path, op):
    n, s, d, _ = gen.rmat(11, 8, seed=4, undirected=True)
    keys = gen.keys_of(s, d)
    f = tmp_path / "g.txt"
    gen.write_edge_file(f, n, s, d)
    pre = str(tmp_path / "upd")
    for i in range(2):
        if op == "0":
            bs, bd, _ = gen.insert_batch(keys, n, 300, seed=60 + i)
        else:
            bs, bd = gen.delete_batch(keys, 300, seed=70 + i)
        S, D = np.r_[bs, bd], np.r_[bd, bs]
        # both directions of every undirected edge, once
        k = np.unique(gen.keys_of(S, D))
        S, D = gen.split_keys(k)
        keys = np.unique(np.r_[keys, k]) if op == "0" else np.setdiff1d(keys, k)
        gen.write_batch_file(pre, i, S, D)
    dump = tmp_path / "tc.bin"
    rc, out, err = run([prog("mk_tc"), str(f), op, pre, "2", f"--dump={dump}"])
    assert rc == 0, err
    lines = out.strip().splitlines()
    assert len(lines) == 2 and lines[1].startswith("cc,")
    s2, d2 = gen.split_keys(keys)
    g = oracle.csr(n, s2, d2)
    raw = np.fromfile(dump, np.uint32)
    count = int(raw[:2].view(np.uint64)[0])
    assert count == oracle.tc(g)
    assert np.array_equal(raw[2:], oracle.cc(g)[0])
