"""The work-unit chunk schedule shared by the host planner and the kernels
(render_launch.h): every sample of a pixel lies in exactly one chunk, fixed
chunking is the plain ceil split, and the guided schedule keeps full chunks
while more than two remain, then halves the rest down to one sample."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = "/opt/rocm/bin/hipcc"


@pytest.fixture(scope="module")
def schedules(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    exe = str(tmp_path_factory.mktemp("chunks") / "chunks")
    csrc = os.path.join(ROOT, "droplet-render_amd", "csrc")
    subprocess.run([HIPCC, "-O1", "-std=c++17", "-I" + os.path.join(ROOT, "include"), "-I" + csrc,
                    os.path.join(ROOT, "tests", "cpp", "chunks.cpp"), "-o", exe], check=True,
                   stdout=subprocess.PIPE, stderr=subprocess.PIPE)
    out = subprocess.run([exe], check=True, stdout=subprocess.PIPE, text=True).stdout
    rows = []
    for line in out.strip().split("\n"):
        v = [int(x) for x in line.split()]
        rows.append((v[0], v[1], bool(v[2]), v[3], v[4], v[5:]))
    return rows


def test_chunks_partition_the_samples(schedules):
    for s, c, guided, nfull, n, starts in schedules:
        assert len(starts) == n + 1
        assert starts[0] == 0 and starts[-1] == s, (s, c, guided, starts)
        sizes = [b - a for a, b in zip(starts, starts[1:])]
        assert all(z >= 1 for z in sizes), (s, c, guided, sizes)


def test_fixed_chunks_are_the_ceil_split(schedules):
    for s, c, guided, nfull, n, starts in schedules:
        if guided:
            continue
        assert n == nfull == -(-s // c)
        assert starts == [min(k * c, s) for k in range(n + 1)]


def test_guided_tail_halves_down_to_one(schedules):
    for s, c, guided, nfull, n, starts in schedules:
        if not guided:
            continue
        sizes = [b - a for a, b in zip(starts, starts[1:])]
        assert sizes[:nfull] == [c] * nfull
        rem = s - nfull * c
        assert rem <= 2 * c and (nfull == 0 or rem > c), (s, c, nfull, rem)
        tail = sizes[nfull:]
        want = []
        while rem:
            z = (rem + 1) // 2
            want.append(z)
            rem -= z
        assert tail == want, (s, c, tail, want)
        assert tail[-1] == 1


def test_invariant_division_is_exact(pkg):
    """The kernels map a work unit to its chunk and pixel with a
    multiply-high division by the band's pixel count and by the row width
    (UDiv, render_launch.h); its quotients equal n / d for every divisor
    shape (1, powers of two and their neighbours, the largest values, the C1
    band and row sizes, random ones) over dividends spread across the whole
    32-bit range, exhaustively near 0 and at the top."""
    import ctypes as C
    import numpy as np
    L = pkg._lib.lib()
    rng = np.random.default_rng(11)
    ds = [1, 2, 3, 5, 7, 10, 641, 1023, 1024, 1025, 65535, 65536, 65537, 73728, 589824, 1 << 24,
          (1 << 31) - 1, 1 << 31, (1 << 31) + 1, 0xFFFFFFFE, 0xFFFFFFFF]
    ds += [int(x) for x in rng.integers(1, 1 << 32, 40, dtype=np.uint64)]
    ds += [int(x) for x in rng.integers(1, 1 << 16, 20)]
    bad = C.c_uint64()
    for d in ds:
        assert L.drk_probe_udiv(d, 0, 1, 1 << 18, C.byref(bad)) == 0 and bad.value == 0, d
        # strided over all of [0, 2^32): an odd step visits dividends of every magnitude
        n0 = int(rng.integers(0, 1 << 32))
        assert L.drk_probe_udiv(d, n0, 2654435761, 1 << 20, C.byref(bad)) == 0 and bad.value == 0, d
        # multiples of d and their neighbours
        assert L.drk_probe_udiv(d, d - 1, d, min(1 << 16, (1 << 32) // d), C.byref(bad)) == 0 and bad.value == 0, d
    assert L.drk_probe_udiv(0, 0, 1, 1, C.byref(bad)) != 0
