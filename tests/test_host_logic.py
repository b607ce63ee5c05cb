"""Host-side logic that needs no GPU: the tuner's candidate set and cache, the
workload mixture, and bench.py's one-GPU contract."""
import json
import os
import subprocess
import sys

import pytest

import vxsgm
from vxsgm import tuner

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("D", [1, 16, 33, 64, 65, 128, 129, 200, 256])
def test_tuner_candidates_cover_d(D):
    c = tuner.candidates(D)
    assert len(c) == (2 if D > 128 else 3), f"no layouts for D={D}"
    for L, K in c:
        assert L * K >= D and L * K in (64, 128, 256) and 2 <= K <= 8
    s = tuner.schedules(D)
    assert len(s) == len(c) ** 2 and all(a in c and b in c for a, b in s)


def test_tuner_cache_roundtrip(tmp_path):
    p = tmp_path / "t.json"
    t = tuner.Tuner(str(p))
    t.table["batch:64x48:D16:P8"] = {"schedule": "32x2/8x8", "ms": {"32x2/8x8": 1.0}}
    t.save()
    t2 = tuner.Tuner(str(p))
    assert t2.lookup("batch", 64, 48, 16, 8) == ((32, 2), (8, 8))
    assert t2.lookup("plain", 64, 48, 16, 8) is None
    assert json.loads(p.read_text())["batch:64x48:D16:P8"]["schedule"] == "32x2/8x8"


def test_mixture_is_seeded_and_in_range():
    a, b = vxsgm.mixture(64), vxsgm.mixture(64)
    assert a == b and len(a) == 64
    assert {(w, h) for w, h, _, _ in a} <= set(vxsgm.MIX_SIZES)
    assert {d for _, _, d, _ in a} <= set(vxsgm.MIX_DISPARITIES)
    pxd = sum(w * h * d for w, h, d, _ in a)
    assert 4e9 < pxd < 13e9


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
