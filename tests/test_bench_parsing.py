"""CPU checks of bench.py's rocprofv3 csv parsing: both kernel forms of each
direction are recognised, and the per-kernel and calibration medians come
out of a counter csv as the traffic line uses them."""
import csv
import importlib.util
import os

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_kernel_side_names_both_forms():
    b = _bench()
    assert b._kernel_side("void (anonymous namespace)::a2b_fwd_chunk<float, 4>((anonymous namespace)::ChunkArgs)") == "forward"
    assert b._kernel_side("void (anonymous namespace)::a2b_fwd_chain<float, 4, 4>((anonymous namespace)::ChunkArgs)") == "forward"
    assert b._kernel_side("void (anonymous namespace)::a2b_bwd_chunk<float, 4>((anonymous namespace)::ChunkArgs)") == "backward"
    assert b._kernel_side("void (anonymous namespace)::a2b_bwd_chain<double, 2, 3>((anonymous namespace)::ChunkArgs)") == "backward"
    assert b._kernel_side("(anonymous namespace)::a2b_zero_words(unsigned int*, unsigned long)") is None


def test_pmc_medians_per_kernel_and_calibration(tmp_path):
    b = _bench()
    rows = [(1, "a2b_fwd_chain<float, 4, 4>", 10.0), (2, "a2b_bwd_chunk<float, 4>", 20.0),
            (3, "a2b_fwd_chain<float, 4, 4>", 30.0), (4, "a2b_bwd_chunk<float, 4>", 40.0),
            (5, "a2b_fwd_chain<float, 4, 4>", 20.0)]
    rows += [(6 + k, "elementwise_kernel", 100.0 + k) for k in range(5)]
    d = tmp_path / "run"
    d.mkdir()
    with open(d / "p_counter_collection.csv", "w", newline="") as fh:
        w = csv.writer(fh)
        w.writerow(["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"])
        for i, k, v in rows:
            w.writerow([i, k, "FETCH_SIZE", v])
            w.writerow([i, k, "OTHER", -1])
    med, calib = b._pmc_medians(str(tmp_path), "FETCH_SIZE")
    assert med["a2b_fwd_chain<float, 4, 4>"] == 20.0 and med["a2b_bwd_chunk<float, 4>"] == 30.0
    assert calib == 102.0
