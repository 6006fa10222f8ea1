"""CPU checks of the rocprofv3 CSV summaries in meerkat_amd.profiling:
marker-delimited regions select exactly the dispatches between a marker
pair, and the per-kernel statistics add up."""
import csv

import pytest

from meerkat_amd import profiling

MARK = "void at::native::vectorized_elementwise_kernel<4, at::native::FillFunctor<short>, x>"


def write_trace(path, names):
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Dispatch_Id", "Kernel_Name", "Start_Timestamp",
                                          "End_Timestamp"])
        w.writeheader()
        t = 1000
        # written out of order: selection must follow Dispatch_Id
        rows = []
        for i, (name, dur) in enumerate(names, start=1):
            rows.append({"Dispatch_Id": i, "Kernel_Name": name, "Start_Timestamp": t,
                         "End_Timestamp": t + dur})
            t += dur + 10
        for r in reversed(rows):
            w.writerow(r)


def test_trace_region_selection(tmp_path):
    p = tmp_path / "k_kernel_trace.csv"
    write_trace(p, [("setup", 5000), (MARK, 1), ("k_a", 2000), ("k_b", 1000), ("k_a", 4000),
                    (MARK, 1), ("k_a", 9999), (MARK, 1), ("copy", 300), (MARK, 1)])
    stats = {name: (n, tot, avg) for name, n, tot, avg, mn, mx in
             profiling.trace_stats(str(p), region=0)}
    assert set(stats) == {"k_a", "k_b"}
    assert stats["k_a"][0] == 2 and stats["k_a"][2] == pytest.approx(3.0)  # us
    assert stats["k_b"] == (1, pytest.approx(0.001), pytest.approx(1.0))
    r1 = profiling.trace_stats(str(p), region=1)
    assert [x[0] for x in r1] == ["copy"]
    with pytest.raises(ValueError):
        profiling.trace_stats(str(p), region=2)
    # no region: every dispatch
    assert sum(x[1] for x in profiling.trace_stats(str(p))) == 10


def test_pmc_region_means(tmp_path):
    p = tmp_path / "c_counter_collection.csv"
    rows = [(1, MARK, 1.0), (2, "k_a", 10.0), (3, "k_a", 30.0), (4, MARK, 1.0), (5, "k_a", 99.0)]
    with open(p, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Dispatch_Id", "Kernel_Name", "Counter_Name",
                                          "Counter_Value"])
        w.writeheader()
        for i, k, v in rows:
            w.writerow({"Dispatch_Id": i, "Kernel_Name": k, "Counter_Name": "FETCH_SIZE",
                        "Counter_Value": v})
            w.writerow({"Dispatch_Id": i, "Kernel_Name": k, "Counter_Name": "OTHER",
                        "Counter_Value": 0})
    got = profiling.pmc_per_kernel(str(p), "FETCH_SIZE", region=0)
    assert got == {"k_a": (2, 20.0)}


@pytest.mark.gpu
def test_step_trace_of_the_frontier_loop(gpu):
    import meerkat_amd as mk
    from meerkat_amd import autotune, gen

    before = autotune.get_knob("persistent_steps")
    autotune.set_knob("persistent_steps", 1)
    try:
        n, s, d, w = gen.rmat(12, 8, seed=5, weighted=True)
        pool = mk.SlabPool(1 << 26)
        g = mk.DynamicGraph(n, pool, 0.7, gen.degree_hints(n, s), weighted=True)
        g.InsertEdges(s, d, w)
        sp = mk.SSSP(g, 0)
        profiling.reset()
        profiling.enable(True)
        sp.Static()
        profiling.enable(False)
        st = profiling.steps()
        assert st and all(x["kernel"] == "k_expand_relax" for x in st)
        assert [x["step"] for x in st] == list(range(len(st)))
        assert st[0]["frontier"] + st[0]["hub_items"] >= 1  # the source
        assert all(x["us"] > 0 for x in st)
        profiling.reset()
        assert profiling.steps() == []
        sp.close()
    finally:
        profiling.enable(False)
        autotune.set_knob("persistent_steps", before)
