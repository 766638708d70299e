"""The C++ driver surface (include/cgmi.hpp: TwoLevelPartition, CGgraphEngine,
CG_co_execution) compiled into tests/cpp/engine_test and checked against the
golden toy vectors."""
import json
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CPP = os.path.join(HERE, "cpp")
BIN = os.path.join(CPP, "engine_test")
GOLD = json.load(open(os.path.join(HERE, "golden", "toy_golden.json")))


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not on PATH")
def test_cpp_header_compiles():
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-fsyntax-only", "engine_test.cpp"],
                       cwd=CPP, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("toy", ["toy_0", "toy_1"])
@pytest.mark.parametrize("alg", ["BFS", "SSSP"])
def test_cpp_engine_matches_golden(toy, alg):
    assert os.path.exists(BIN), "tests/cpp/engine_test is not built (run __graft_entry__.build())"
    gold = GOLD["graphs"][toy][alg.lower()]
    for root, want in gold.items():
        print(f"cpp driver {toy} {alg} root={root}", flush=True)
        r = subprocess.run([BIN, toy, alg, root, "5"], capture_output=True, text=True, timeout=60)
        assert r.returncode == 0, r.stdout + r.stderr
        lines = r.stdout.strip().splitlines()
        assert lines[0].startswith("[Total Run Time]:")
        assert [int(x) for x in lines[1].split()] == want
