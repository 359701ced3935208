"""bench.py's output contract: one JSON line with the metric, the roofline of the dominant kernel and the
configuration; more than one GPU is refused."""
import json
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def test_more_than_one_gpu_is_refused():
    res = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                         timeout=120)
    assert res.returncode == 2
    assert "one GPU only" in res.stderr
    assert res.stdout.strip() == ""


@pytest.mark.gpu
def test_bench_line_keeps_the_contract():
    print("case bench contract: bench.py --steps 2 --warmup 1", flush=True)
    res = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--steps", "2", "--warmup", "1",
                          "--no-cpu-baseline", "--serve-requests", "0"], capture_output=True, text=True, timeout=240)
    assert res.returncode == 0, res.stderr[-2000:]
    lines = [line for line in res.stdout.splitlines() if line.strip()]
    assert len(lines) == 1
    b = json.loads(lines[0])
    for key in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better", "scaling",
                "vs_baseline", "dtype", "data", "config", "roofline"):
        assert key in b, key
    assert b["n_gpus"] == 1 and b["steps"] == 2 and b["warmup"] == 1 and b["higher_is_better"] is True
    assert b["value"] > 0 and b["ms_per_step"] > 0 and b["unit"] == "slices/s"
    assert "workload" in b["config"]
    r = b["roofline"]
    assert r["bound"] in ("hbm", "mfma") and r["unit"] == "GB/s" and r["peak"] == 8000.0
    assert r["achieved"] > 0 and abs(r["frac"] - r["achieved"] / r["peak"]) < 1e-9
    assert b["two_streams"]["ms_per_step"] > 0 and b["two_streams"]["slices_per_s"] > 0


def test_counter_summaries_describe_the_current_kernels():
    """bench.py reads the HBM traffic of the dominant kernel from the `traffic:` lines of profiles/*.md; each line
    records a hash of the kernel sources it was measured with. The default variants must resolve against the sources
    as they are (a kernel change needs its counter passes run again)."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("bench_module", ROOT / "bench.py")
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    for op, slices, variant in (("insert", 10000, "slab"), ("insert", 10000, "one-wave"),
                                ("extract", 4096, "block-copy gather"), ("extract", 4096, "LDS tiles")):
        nbytes, source = bench.measured_traffic(op, slices, variant)
        assert nbytes is not None, f"{op} {variant}: {source}"
        assert nbytes > 0
