"""The device's division by a generation-time constant (``fmr_divc`` in
``formura_amd/csrc/formura_hip_rt.h``) equals IEEE division bit for bit.

The step kernels divide by constants (``/(dx*dx)`` in every bundled program)
through a host-rounded reciprocal and two fused corrections; the oracle's C
divides.  The same header is compiled for the host (x86 FMA is correctly
rounded, as the device's is) and checked on random, near-midpoint and special
dividends for several divisors.  On the GPU the parity tests then compare the
kernels themselves with the oracle.

``fmr_divc1`` (one correction) is emitted only for divisors with
|b*RN(1/b) - 1| <= 2^-54, the condition under which q0 = RN(a*y) is faithful
and Markstein's theorem applies; the checker tests it on those divisors.  Its
remainder is formed at scale 2^120, which keeps it exact down to subnormal
dividends and quotients: the checker runs it without a range test on tiny and
subnormal dividends (near rounding midpoints too) for every |a| <= 2^900, and
requires the unscaled remainder to fail some of those cases.  A
failure of one correction elsewhere needs a quotient within about 2^-52 ulp of
a rounding midpoint, which random and near-midpoint dividends do not reach: the
condition rests on the proof in the header, not on this test.
"""
import os
import subprocess
from fractions import Fraction

import numpy as np

import pytest

from formura_amd import build

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("divc") / "divc_check")
    cmd = [build.HIPCC, "-O2", "-mfma", "-ffp-contract=off", "-I", build.CSRC,
           os.path.join(HERE, "native", "divc_check.cpp"), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, "host build of the division check failed: %s" % r.stderr[-1500:]
    return out


@pytest.mark.parametrize("seed", ["0x9E3779B97F4A7C15", "12345"])
def test_divc_matches_ieee_division(checker, seed):
    r = subprocess.run([checker, "2000000", seed], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:]
    assert "mismatches 0" in r.stdout


def test_emitted_divisions():
    """A non-power-of-two constant divisor becomes fmr_divc with its rounded
    reciprocal; a power-of-two one an exact product; host Setup keeps '/'.
    The tile kernels' point function divides with the branch-free forms and
    holds the guarded ones for the waves that need them."""
    from formura_amd.emit_hip import generate
    _, src, _ = generate(os.path.join(build.PROG_DIR, "grayscott2d.fmr"),
                         os.path.join(build.PROG_DIR, "gs_64.yaml"), "gs_64")
    # d_xx and d_yy of u and of v: the fused sequence (remainder at scale
    # 2^120), and the guarded form for the waves that hold a dividend above
    # 2^900 (one comparison per division)
    assert src.count("fmr_divc1_fast(") == 4 and src.count("fmr_divc1(") == 4
    assert src.count(", 0.0001, 10000.0") == 8 and src.count("fmr_divc1_out_of_range(") == 4
    assert "fmr_divc_out_of_range(" not in src and "fmr_divc1_redo(" not in src
    _, fsrc, _ = generate(os.path.join(build.PROG_DIR, "grayscott2d.fmr"),
                          os.path.join(build.PROG_DIR, "gs_64.yaml"), "gs_64", "flat")
    assert fsrc.count("fmr_divc1(") == 4 and "fmr_divc1s(" not in fsrc
    setup = src[src.index("static void Formura_Setup"):]
    assert "fmr_divc" not in setup[:setup.index("\n}\n")]
    _, hsrc, _ = generate(os.path.join(build.PROG_DIR, "heat3d.fmr"),
                          os.path.join(build.PROG_DIR, "heat3d_16.yaml"), "heat3d_16")
    kernels = hsrc[:hsrc.index("static void Formura_Setup")]
    assert "fmr_divc" not in kernels and "*256.0" in kernels


def test_one_step_condition():
    """The exact condition for one correction, on divisors either side of it."""
    from formura_amd.emit import _divc_one_step
    assert _divc_one_step(1e-4) and _divc_one_step(0.01) and _divc_one_step(3.0)
    assert not _divc_one_step(1e-5) and not _divc_one_step(9.999999999999999e-5)


# --- the per-level range test of the tile kernels (div "level") -----------------

def _step_nodes(name):
    import os
    from formura_amd import build as B
    from formura_amd.emit_hip import generate
    fmr, yml, _ = B.PROGRAMS[name]
    _, _, mm = generate(os.path.join(B.PROG_DIR, fmr), os.path.join(B.PROG_DIR, yml), name, "flat")
    return mm.step


def test_level_division_bound_grayscott_and_planted_dividends():
    """Gray-Scott's dividends (a + b) - 2c: XMAX = 2^897 (4 XMAX <= 2^899).  A
    dividend with a non-integer factor, or a product of loads, has no bound."""
    import copy
    from formura_amd.emit import level_division_bound
    nodes = _step_nodes("gs_64")
    assert [level_division_bound(n) for _, n in sorted(nodes.items())] == [2.0 ** 897] * 2
    node = copy.deepcopy(sorted(nodes.items())[0][1])
    # the literal 2.0 of a 2*c term of a dividend becomes 2.5: not an integer combination
    twos = [mm.inst[2] for mm in node.micro if mm.inst[0] == "Binop" and mm.inst[1] == "*"
            and node.micro[mm.inst[2]].inst[0] == "Imm" and float(Fraction(node.micro[mm.inst[2]].inst[1])) == 2.0
            and node.micro[mm.inst[3]].inst[0] == "LoadCursorStatic"]
    assert twos
    m = node.micro[twos[0]]
    m.inst = ("Imm", Fraction(5, 2)) + tuple(m.inst[2:])
    assert level_division_bound(node) is None
    node = copy.deepcopy(sorted(nodes.items())[0][1])
    # a dividend term c*2 -> c*c (a product of loads)
    muls = [j for j, mm in enumerate(node.micro) if mm.inst[0] == "Binop" and mm.inst[1] == "*"
            and node.micro[mm.inst[2]].inst[0] == "Imm" and node.micro[mm.inst[3]].inst[0] == "LoadCursorStatic"]
    j = muls[0]
    a, b = node.micro[j].inst[2], node.micro[j].inst[3]
    node.micro[j].inst = ("Binop", "*", b, b)
    assert level_division_bound(node) is None


def test_level_values_give_dividends_in_range():
    """Values that pass fmr_level_ok (0, or 2^-848 <= |x| <= 2^897) give second
    differences (a + b) - 2c in double that are 0 or within [2^-900, 2^899]:
    exact rational check on values near both ends of the range and on mixtures."""
    rng = np.random.default_rng(5)
    lo, hi = 2.0 ** -848, 2.0 ** 897
    pools = [lo * (1 + rng.random(4000)), lo * 2.0 ** rng.integers(0, 60, 4000) * (1 + rng.random(4000)),
             hi * rng.random(4000), np.zeros(100), 1.0 + rng.random(1000)]
    vals = np.concatenate(pools)
    vals *= np.where(rng.random(vals.size) < 0.5, -1.0, 1.0)
    vals = vals[(np.abs(vals) >= lo) & (np.abs(vals) <= hi) | (vals == 0)]
    n = 200000
    a, b, c = (vals[rng.integers(0, vals.size, n)] for _ in range(3))
    d = (a + b) - 2.0 * c
    nz = d[d != 0]
    assert np.all(np.abs(nz) >= 2.0 ** -900) and np.all(np.abs(nz) <= 2.0 ** 899) and np.all(np.isfinite(d))
    # every nonzero one is a multiple of 2^-900 (checked exactly on a sample)
    for x in nz[:2000]:
        assert (Fraction(float(x)) / Fraction(1, 2 ** 900)).denominator == 1
