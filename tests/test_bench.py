"""bench.py's output contract: one JSON line with the headline metric, the
roofline entry and the CPU baseline; one GPU only."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BENCH = os.path.join(ROOT, "bench.py")


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, BENCH, "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in out["error"]


def test_committed_counters_match_the_byte_model():
    """The HBM bytes per launch measured with FETCH_SIZE/WRITE_SIZE (the
    committed summary) agree with the algorithmic bytes of each kernel
    class: the passes move no more than they must.  Within 5%: the x pass's
    row pairs each write half of their output lines (two workgroups fill a
    line; its WRITE_SIZE measures about 5% above the bytes written), and the
    y passes' non-temporal state accesses move some partial lines (3-4%
    above the bytes they need; the first pass, which has none, 0.3%)."""
    sys.path.insert(0, ROOT)
    import bench
    from gturbo_amd import workloads
    w = workloads()["euler4096_rk4"]
    n = w.params.n
    m = n // 3  # the 2/3 rectangle: kx, ky in [-N/3, N/3]
    cb = bench.compact_bytes(n, (m, m, m), 4)
    for k in ("x_pass", "y_stage"):
        assert bench.read_traffic(k) is not None, k
    for k, v in cb.items():
        t = bench.read_traffic(k)
        if t is None:
            continue
        assert abs(t / v - 1.0) < (0.005 if k == "y_start" else 0.05), (k, t, v)
    # one RK4 step: y_start + 3 y_stage + y_final + 4 x passes
    per_step = cb["y_start"] + 3 * cb["y_stage"] + cb["y_final"] + 4 * cb["x_pass"]
    assert 250 < per_step / n ** 2 < 265


@pytest.mark.parametrize("rko,per_gp", [(4, 468.0), (2, 216.0)])
def test_fixed_byte_model_per_step(rko, per_gp):
    """The fixed byte model distributed over the kernel classes adds up to
    58.5A (RK4) and 27A (RK2) per step, A = 8 N^2 bytes."""
    sys.path.insert(0, ROOT)
    import bench
    n = 1024
    ab = bench.model_bytes(n, rko)
    if rko == 4:
        per_step = ab["y_start"] + 3 * ab["y_stage"] + ab["y_final"] + 4 * ab["x_pass"]
    else:
        per_step = ab["y_start"] + ab["y_stage"] + ab["y_final"] + 2 * ab["x_pass"]
    assert per_step / n ** 2 == pytest.approx(per_gp, rel=1e-12)


@pytest.mark.gpu
def test_bench_line_has_the_contract_keys():
    r = subprocess.run([sys.executable, BENCH, "--n", "256", "--steps", "3", "--warmup", "1",
                        "--rk2-steps", "2", "--ens-steps", "2"],
                       capture_output=True, text=True, timeout=600, cwd=ROOT)
    print(r.stderr[-3000:])
    assert r.returncode == 0
    lines = [ln for ln in r.stdout.strip().splitlines() if ln.strip()]
    assert len(lines) == 1
    d = json.loads(lines[0])
    for k in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step",
              "higher_is_better", "scaling", "vs_baseline", "dtype", "data", "config",
              "roofline", "cpu_baseline"):
        assert k in d, k
    assert d["n_gpus"] == 1 and d["steps"] == 3 and d["warmup"] == 1
    assert d["value"] > 0 and d["dtype"] == "f64" and d["higher_is_better"] is True
    assert "workload" in d["config"]
    roof = d["roofline"]
    assert roof["bound"] == "hbm" and roof["unit"] == "GB/s" and roof["peak"] == 8000.0
    assert abs(roof["frac"] - roof["achieved"] / roof["peak"]) < 1e-12
    assert roof["traffic"] is None  # the counter summary is of the 4096^2 workload only
    assert roof["frac_measured"] is None and roof["achieved"] <= roof["peak"]
    assert roof["model_frac"] == pytest.approx(roof["model_achieved"] / roof["peak"])
    for k, v in d["kernels"].items():
        assert v.get("achieved_GBs", 0.0) <= 8000.0, k
    cb = d["cpu_baseline"]
    assert cb["value"] > 0 and cb["kind"] == "port" and cb["cores"] >= 1 and cb["sample"]
    assert cb["one_core"]["cores"] == 1 and cb["one_core"]["value"] > 0
    assert cb["cpu_model"] and cb["nproc"] >= 1
    e = cb["euler256"]
    assert e["all_cores"]["value"] > 0 and e["one_core"]["cores"] == 1
    for leg in ("rk2", "ensemble", "adjoint"):
        assert d[leg]["finite"] is True, leg
    for leg in ("euler256", "sqg1024"):
        assert d["configs"][leg]["finite"] is True, leg
