"""The trilinear lookup's division by the leaf extent, as the kernels compute
it when every leaf of a scene has the same extent e: through y = RN(1/e) and
one Markstein correction, q = RN(a*y), r = fma(-e, q, a), q' = fma(r, y, q).
q' must be the correctly rounded quotient RN(a/e) for every numerator the
kernels send this way (|a| >= 2^-100; smaller ones, zero included, take the
division), so that the lookups stay bit-exact with the division the reference
writes (src/kernel.cpp:98). Checked here on the CPU: every float
numerator for extents that are not powers of two, and random numerators over
the whole range for 64 random extents; the GPU test of the lookup
(test_trilinear_bit_exact) runs the same arithmetic in the kernels' code.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("markstein") / "markstein_check")
    subprocess.run(["gcc", "-O2", "-mfma", "-ffp-contract=off", "-o", exe,
                    os.path.join(ROOT, "tests", "cpp", "markstein_check.c"), "-lm", "-lpthread"], check=True)
    return exe


@pytest.mark.parametrize("extent", ["0.0190", "0.3109712", "3.7e-5"])
def test_every_numerator(checker, extent):
    r = subprocess.run([checker, "all", extent], stdout=subprocess.PIPE, text=True)
    assert r.returncode == 0, r.stdout


def test_random_extents(checker):
    r = subprocess.run([checker, "random", "2000000", "7"], stdout=subprocess.PIPE, text=True)
    assert r.returncode == 0, r.stdout


def test_tiny_numerators_need_the_division(checker):
    """Below 2^-100 the correction does go wrong (its residual underflows),
    which is why the kernels divide there; the checker sees those cases."""
    r = subprocess.run([checker, "tiny", "0.0190"], stdout=subprocess.PIPE, text=True)
    assert r.returncode == 0, r.stdout
