"""psf::AbstractSolver host semantics (include/psf/solver.hpp) against the
reference's AbstractSolver (Solver/abstractsolver.cpp:178-302, 379-397,
psfsolver.cpp:23-30, 62-67, 226-233): the particle ring of addParticle,
clearParticles, the rebuild calls and the benchmark.csv layout.  Compiled with
g++ on the host; no GPU call."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def run(tmp_path_factory):
    gxx = shutil.which("g++")
    if not gxx:
        pytest.skip("g++ not found")
    d = tmp_path_factory.mktemp("solver")
    exe = str(d / "solver_semantics")
    subprocess.run([gxx, "-std=c++17", "-O1", "-Wall", "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "solver_semantics.cpp"), "-o", exe], check=True)
    csv = str(d / "benchmark.csv")
    out = subprocess.run([exe, csv], capture_output=True, text=True, check=True).stdout
    lines = {}
    for ln in out.splitlines():
        k, _, v = ln.partition(" ")
        lines.setdefault(k, []).append(v)
    return lines, open(csv).read()


def _floats(s):
    return [float(x) for x in s.split()]


def test_particle_ring(run):
    lines, _ = run
    assert _floats(lines["cleared"][0]) == [0.0] * 12
    added = _floats(lines["added"][0])
    assert added[0:4] == pytest.approx([0.1, 0.2, 0.3, 5.0])
    assert added[4:8] == pytest.approx([0.0, 0.1, 0.0, 2.0])
    assert added[8:12] == pytest.approx([0.0, 0.0, 0.1, 3.0])
    assert lines["pointer"][0] == "0"
    after = _floats(lines["after5"][0])
    assert after[0:4] == pytest.approx([-0.1, 0.0, 0.0, 6.0])
    assert lines["simtime"][0] == "5"


def test_rebuilds(run):
    lines, _ = run
    assert lines["rebuild"][0] == "1 1 1 0 K 12"
    assert lines["rebuild"][1] == "2 2 2 1 simtime 0"
    assert lines["particles"][0] == "2"


def test_benchmark_csv(run):
    lines, csv = run
    rows = csv.split("\n")
    assert rows[0] == "2" and rows[1] == "12" and rows[2] == "4,5,6"
    assert len(rows[3].split(",")) == 3 and rows[4] == ""
    frames = [r for r in rows[5:] if r.startswith("1.500000,")]
    assert len(frames) == 4          # frames 1..4 are recorded; the 5th call ends the benchmark
    assert frames[0] == "1.500000,2.250000"
    assert rows[-1] == "2.000000,3.000000"   # the sums over maxFramesBenchmark = 3
    assert lines["benchmarking"][0] == "0"
