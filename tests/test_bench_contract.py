"""bench.py's contract on the CPU: one GPU only (any other count exits with
a message), and the result line is strict JSON (no NaN tokens)."""
import importlib.util
import json
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr and r.stdout == ""


def test_result_line_is_strict_json():
    b = _bench()
    line = {"value": 1.5, "p50": float("nan"), "nested": {"a": [np.float64("inf"), 2, np.int32(3)],
                                                           "b": (np.nan, "x")}}
    out = json.dumps(b._finite(line), allow_nan=False)
    back = json.loads(out)
    assert back == {"value": 1.5, "p50": None, "nested": {"a": [None, 2, 3], "b": [None, "x"]}}
