"""bench.py on the GPU at small sizes: every mode prints one JSON line with the
fields the measurement contract names (metric, value, unit, steps, timing,
dtype, config, the dominant kernel's roofline), the headline mode its CPU
baseline, and the numbers agree with each other (value = info bits / time,
roofline.frac = achieved / peak)."""
import json
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
pytestmark = pytest.mark.gpu

REQUIRED = ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better", "scaling",
            "vs_baseline", "dtype", "data", "config", "roofline")
ROOFLINE = ("bound", "achieved", "peak", "unit", "frac", "traffic")


def bench(*args):
    print("bench.py", *args, flush=True)
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), *args], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr[-2000:]
    lines = [ln for ln in r.stdout.splitlines() if ln.strip()]
    assert len(lines) == 1, r.stdout
    return json.loads(lines[0])


def check_line(d, steps):
    for k in REQUIRED:
        assert k in d, k
    assert d["n_gpus"] == 1 and d["steps"] == steps and d["higher_is_better"] is True
    assert d["value"] > 0 and d["ms_per_step"] > 0
    rl = d["roofline"]
    for k in ROOFLINE:
        assert k in rl, k
    assert rl["bound"] in ("hbm", "mfma") and rl["peak"] == 8000.0 and rl["unit"] == "GB/s"
    assert abs(rl["frac"] - rl["achieved"] / rl["peak"]) < 1e-3
    # achieved = algorithmic bytes / the timed step
    assert abs(rl["achieved"] - rl["algorithmic_bytes"] / (d["ms_per_step"] * 1e-3) / 1e9) <= 0.01 * rl["achieved"]


def test_headline_mode_small(gpu_available):
    frames = 1 << 16
    d = bench("--frames", str(frames), "--steps", "3", "--warmup", "1", "--cpu-baseline-words", "32",
              "--cpu-single-words", "8", "--cpu-runs", "1")
    check_line(d, 3)
    c = d["config"]
    assert c["frames"] == frames
    learned, plain = c["learned"], c["plain"]
    # value is the learned frame queue's info throughput: K = 2000 bits per frame
    assert abs(d["value"] - frames * 2000 / (d["ms_per_step"] * 1e-3) / 1e9) <= 1e-3 * d["value"] + 1e-3
    assert learned["avg_iterations"] <= 20 and plain["avg_iterations"] == 20
    assert 0 <= learned["fer"] <= 1 and 0 <= plain["fer"] <= 1
    cb = d["cpu_baseline"]
    assert cb["value"] > 0 and cb["kind"] in ("port", "reference") and cb["cores"] >= 1


def test_plain_mode_lists_all_variants(gpu_available):
    d = bench("--mode", "plain", "--frames", "16384", "--steps", "3", "--warmup", "1", "--no-cpu-baseline")
    check_line(d, 3)
    v = d["config"]["variants"]
    assert sorted(v) == sorted(["MS", "OMS", "OMS_v2", "NMS", "2NMS"])
    for row in v.values():
        for it in ("it20", "it10"):
            assert row[it]["ms_per_decode"] > 0 and 0 <= row[it]["fer"] <= 1
    # OMS and OMS_v2 decode the same frames identically (the v2 input is the word layout of the same frames)
    assert v["OMS"]["it20"]["fer"] == v["OMS_v2"]["it20"]["fer"]
    assert d["config"]["pcie_inclusive"]["decisions_equal_device_path"] is True


def test_ira_and_mixed_modes_small(gpu_available):
    d = bench("--mode", "ira", "--frames", "1024", "--steps", "2", "--warmup", "1", "--no-cpu-baseline")
    check_line(d, 2)
    legs = d["config"]["legs"]
    assert set(legs) == {"learned_pool", "oms_pool", "learned_pool_input_order", "oms50_fixed"}
    assert legs["oms50_fixed"]["avg_iterations"] == 50
    # the queue order changes no frame's result: same counts in either order
    a, b = legs["learned_pool"], legs["learned_pool_input_order"]
    assert a["order"] == "least reliable first" and b["order"] == "input"
    assert (a["avg_iterations"], a["bit_errors"], a["frame_errors"]) == \
        (b["avg_iterations"], b["bit_errors"], b["frame_errors"])
    d = bench("--mode", "mixed", "--frames", "8192", "--steps", "2", "--warmup", "1", "--no-cpu-baseline")
    check_line(d, 2)
    codes = d["config"]["codes"]
    assert [c["code"] for c in codes] == ["qc_12x24_z81", "qc_4x24_z96"]
    assert sum(c["frames"] for c in codes) == 8192
    assert all(c["avg_iterations"] <= 30 for c in codes)
