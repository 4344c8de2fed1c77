"""bench.py host logic that runs without a GPU: the --gpus guard and the
traffic figures it reads from the committed profile summary."""
from __future__ import annotations

import importlib.util
import json
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def _bench():
    spec = importlib.util.spec_from_file_location("bench_mod", ROOT / "bench.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_traffic_read_from_profile_summary():
    b = _bench()
    for wl in ("rotor512", "ot1m"):
        t = b.measured_traffic(wl)
        assert t is not None and t > b.FORCE_BYTES_PER_PARTICLE[2] * 262144 * 0.5
    assert b.measured_traffic("no-such-workload") is None


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU" in json.loads(r.stdout.strip().splitlines()[-1])["error"]
