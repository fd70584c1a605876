"""Host sizing of the neighbour-search grid (vfd_amd/csrc/grid.h; ParticleSearch::ComputeCellInformation,
ParticleSearch.cu:57-93), on the CPU.

The header is compiled with g++ under the undefined-behaviour sanitizer into a small driver that reads integer cell
bounds and prints the grid. Ordinary bounds must give the oracle's grid bit for bit (the same float arithmetic);
bounds of a runaway particle (saturated float -> int conversions, a grid beyond the cell limit) must be refused without
any integer overflow, where the sizing used to wrap a 32-bit capacity to a small value."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = r"""
#include <climits>
#include <cstdio>
#include "grid.h"
int main() {
    int lo[3], hi[3];
    float R;
    unsigned long long before;
    while (std::scanf("%d %d %d %d %d %d %a %llu", &lo[0], &lo[1], &lo[2], &hi[0], &hi[1], &hi[2], &R, &before) == 8) {
        vfd::SceneGrid g;
        const unsigned long long cells = vfd::size_scene_grid(lo, hi, R, before, g);
        std::printf("%llu %a %a %a %a %a %a %u %u %u %u %u %u %u %llu\n", cells, g.gridMin[0], g.gridMin[1],
                    g.gridMin[2], g.gridDelta[0], g.gridDelta[1], g.gridDelta[2], g.dims[0], g.dims[1], g.dims[2],
                    g.meta[0], g.meta[1], g.meta[2], g.cellCount,
                    (unsigned long long)vfd::grid_capacity(cells + before));
    }
    return 0;
}
"""


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    d = tmp_path_factory.mktemp("grid")
    src = d / "grid_driver.cpp"
    src.write_text(DRIVER)
    exe = d / "grid_driver"
    subprocess.check_call(["g++", "-std=c++17", "-O1", "-ffp-contract=off", "-fsanitize=undefined",
                           "-fno-sanitize-recover=all", "-I", os.path.join(ROOT, "vfd_amd", "csrc"), str(src),
                           "-o", str(exe)])

    def run(rows):
        text = "".join("%d %d %d %d %d %d %s %d\n" % (*lo, *hi, float(np.float32(R)).hex(), before)
                       for lo, hi, R, before in rows)
        p = subprocess.run([str(exe)], input=text, capture_output=True, text=True)
        assert p.returncode == 0 and not p.stderr, p.stderr  # the sanitizer reports any overflow on stderr
        out = []
        for line in p.stdout.splitlines():
            f = line.split()
            out.append(dict(cells=int(f[0]), gridMin=np.array([float.fromhex(v) for v in f[1:4]], np.float32),
                            gridDelta=np.array([float.fromhex(v) for v in f[4:7]], np.float32),
                            dims=[int(v) for v in f[7:10]], meta=[int(v) for v in f[10:13]], cellCount=int(f[13]),
                            capacity=int(f[14])))
        return out
    return run


def test_grid_matches_the_oracle(oracle_lib, driver):
    rng = np.random.default_rng(11)
    rows, expect = [], []
    for R, span in [(0.016, (2.0, 1.0, 1.0)), (0.1, (2.0, 2.0, 2.0)), (0.04, (0.3, 5.0, 0.01))]:
        x = (rng.uniform(-0.5, 0.5, size=(5000, 3)) * span).astype(np.float32)
        Rf = np.float32(R)
        c = np.floor(x / Rf).astype(np.int64)
        rows.append((c.min(0), c.max(0), Rf, 0))
        expect.append(O.search_grid(x, Rf))
    for got, ref in zip(driver(rows), expect):
        assert np.array_equal(got["gridMin"], ref["gridMin"])
        assert np.array_equal(got["gridDelta"], ref["gridDelta"])
        assert got["dims"] == list(ref["dims"]) and got["meta"] == list(ref["meta"])
        assert got["cellCount"] == ref["cellCount"] == got["cells"]
        assert got["capacity"] == got["cells"] + got["cells"] // 4


def test_runaway_bounds_are_refused_without_overflow(driver):
    LIM = 1 << 30
    R = np.float32(0.016)
    rows = [
        ((-2 ** 31, -2 ** 31, -2 ** 31), (2 ** 31 - 1, 2 ** 31 - 1, 2 ** 31 - 1), R, 0),  # saturated conversions
        ((0, 0, 0), (2 ** 31 - 1, 0, 0), R, 0),                                           # hi + 1 would overflow int
        ((0, 0, 0), (1527, 1527, 1519), R, 0),  # 3.6e9 cells: the window where the 32-bit capacity used to wrap
        ((0, 0, 0), (100, 100, 100), R, LIM - 1000),                                      # earlier scenes fill it
        ((5, 0, 0), (4, 0, 0), R, 0),                                                      # empty bounds
        ((0, 0, 0), (10, 10, 10), np.float32(np.inf), 0),
    ]
    for got in driver(rows):
        assert got["cells"] == 0 and got["cellCount"] == 0 and got["dims"] == [0, 0, 0]
        assert got["capacity"] <= LIM


def _dims(lo, hi, R):
    """dims of one axis in the reference's float arithmetic (ParticleSearch.cu:66-72)."""
    R = np.float32(R)
    size = np.float32(np.float32(hi + 1) * R) - np.float32(np.float32(lo) * R)
    return int(np.ceil(np.float32(size / R))) + 4


def test_largest_accepted_grid_and_capacity_clamp(driver):
    LIM = 1 << 30
    R = np.float32(0.016)
    # the largest cube of bounds whose grid is at most 128 meta blocks (1024 cells) per axis: 2^30 cells exactly,
    # accepted, with the capacity clamped to the limit
    n = max(k for k in range(1000, 1030) if _dims(0, k, R) <= 1024)
    assert _dims(0, n, R) > 1016
    got = driver([((0, 0, 0), (n, n, n), R, 0)])[0]
    assert got["meta"] == [128] * 3
    assert got["cells"] == LIM and got["capacity"] == LIM
    # one meta block more along x: refused
    got = driver([((0, 0, 0), (n + 8, n, n), R, 0)])[0]
    assert got["cells"] == 0
