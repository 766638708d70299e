"""ThreadSanitizer over the restated MULTI_CORE engine (oracle/multicore.c)
and the engine's host team (cggraph_amd/csrc/host_team.cpp).

The reference's CPU path relies on benign races (plain reads beside CAS /
atomicMin, P/host_algorithm.hpp:176-181; a volatile CAS loop, :23-39; busy
flags, P/CGgraph_engine.cuh:866-874).  The restatement uses __atomic
operations for each, so a TSan build (clang + libomp, with libomp's Archer tool
reporting OpenMP's synchronisation) must run BFS, SSSP and CC with no report
and the same values as the sequential oracle.  A planted plain-int race in the
driver shows that the same build does report races.  The host team (the
engine's LIGHT-iteration relaxation over OpenMP, with chunked hub edges) is
driven in its three modes from tests/host_team_tsan.cpp on a host-only graph.
CPU only.
"""
import os
import subprocess

import pytest

LLVM = "/opt/rocm/llvm"
HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
ORACLE = os.path.join(ROOT, "oracle")
CSRC = os.path.join(ROOT, "cggraph_amd", "csrc")
ARCHER = os.path.join(LLVM, "lib", "libarcher.so")


@pytest.fixture(scope="module")
def tsan_binary(tmp_path_factory):
    clang = os.path.join(LLVM, "bin", "clang")
    if not (os.path.exists(clang) and os.path.exists(ARCHER)):
        pytest.skip("clang with libomp's Archer tool is not installed")
    out = str(tmp_path_factory.mktemp("tsan") / "tsan_check")
    srcs = [os.path.join(ORACLE, f) for f in ("tsan_check.c", "oracle.c", "multicore.c")]
    cmd = [clang, "-O1", "-g", "-std=c11", "-fopenmp", "-fsanitize=thread", "-o", out, *srcs,
           f"-L{LLVM}/lib", f"-Wl,-rpath,{LLVM}/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    return out


@pytest.fixture(scope="module")
def host_team_binary(tmp_path_factory):
    clang = os.path.join(LLVM, "bin", "clang")
    if not (os.path.exists(clang) and os.path.exists(ARCHER)):
        pytest.skip("clang with libomp's Archer tool is not installed")
    d = tmp_path_factory.mktemp("tsan_team")
    obj, out = str(d / "oracle.o"), str(d / "host_team_tsan")
    flags = ["-O1", "-g", "-fopenmp", "-fsanitize=thread"]
    r = subprocess.run([clang, *flags, "-std=c11", "-c", os.path.join(ORACLE, "oracle.c"), "-o", obj],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    # host code only: HIP's headers for the graph type, its runtime only to link
    cmd = [clang + "++", *flags, "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", f"-I{CSRC}",
           f"-I{ORACLE}", os.path.join(HERE, "host_team_tsan.cpp"), os.path.join(CSRC, "host_team.cpp"), obj,
           "-o", out, f"-L{LLVM}/lib", f"-Wl,-rpath,{LLVM}/lib", "-L/opt/rocm/lib", "-Wl,-rpath,/opt/rocm/lib",
           "-lamdhip64"]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    return out


def run(binary, *args):
    env = dict(os.environ, OMP_TOOL_LIBRARIES=ARCHER, OMP_NUM_THREADS="8",
               TSAN_OPTIONS="ignore_noninstrumented_modules=1 halt_on_error=0")
    return subprocess.run([binary, *args], capture_output=True, text=True, timeout=600, env=env)


def test_multicore_engine_is_race_free(tsan_binary):
    r = run(tsan_binary)
    assert "ThreadSanitizer" not in r.stderr, r.stderr[-4000:]
    assert r.returncode == 0, (r.stdout, r.stderr[-2000:])
    lines = [ln for ln in r.stdout.splitlines() if ln.startswith("alg ")]
    assert len(lines) == 3 and all(ln.endswith("mismatches 0") for ln in lines), r.stdout


def test_planted_race_is_reported(tsan_binary):
    r = run(tsan_binary, "plant")
    assert "WARNING: ThreadSanitizer: data race" in r.stderr


def test_host_team_is_race_free(host_team_binary):
    """List (three-stage stealing with donation), bitmap (the engine without a
    GPU) and tail (team against a head-claiming agent thread) modes."""
    r = run(host_team_binary)
    assert "ThreadSanitizer" not in r.stderr, r.stderr[-4000:]
    assert r.returncode == 0, (r.stdout, r.stderr[-2000:])
    lines = [ln for ln in r.stdout.splitlines() if ln.startswith("alg ")]
    assert len(lines) == 9 and all(ln.endswith("mismatches 0") for ln in lines), r.stdout
    f = {(ln.split()[1], ln.split()[3]): dict(zip(ln.split()[4::2], map(int, ln.split()[5::2]))) for ln in lines}
    # the hub (16 000 edges) was donated in the list and bitmap modes
    assert all(f[(a, m)]["donated"] >= 1 for a in "012" for m in ("list", "bitmap"))
    # CC's first frontier is every vertex: both ends of the queue took entries
    assert f[("2", "tail")]["agent"] > 0 and f[("2", "tail")]["team"] > 0
