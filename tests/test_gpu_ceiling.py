"""The gather-rate microbenchmark that bench.py reports as the practical
ceiling of the push expansion and the PageRank pull."""
import json
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "cggraph_amd", "gather_ceiling")


def test_ceiling_is_built():
    assert os.access(EXE, os.X_OK), "make -C cggraph_amd builds cggraph_amd/gather_ceiling"


@pytest.mark.gpu
def test_ceiling_reports_every_pattern():
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    c = json.loads(r.stdout.strip().splitlines()[-1])
    keys = ["copy_GBps"] + [f"{op}_{kind}_v{s}_G_per_s" for s in (24, 25)
                            for kind in ("uniform", "skewed") for op in ("gather", "relax")]
    for k in keys:
        assert c[k] > 0, k
    # degree-skewed addresses reuse cache lines, uniform ones do not
    for s in (24, 25):
        assert c[f"gather_skewed_v{s}_G_per_s"] > c[f"gather_uniform_v{s}_G_per_s"]
