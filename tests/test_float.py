"""Programs on FP32 state (``float`` fields and values).

The generated C gives every micro value its declared type (Templates.hs:459-467)
and C's usual arithmetic conversions apply: float + float is a float sum,
float * double (a literal) is a double product rounded to float on assignment.
``mm_interp.step_typed`` restates that with numpy and is pinned here by a hand
transcription of the step of ``float2d.fmr`` in the same types; the GPU
libraries (flat and tile2d, a 32 x 48 grid) must give it bit for bit.
"""
import os

import numpy as np
import pytest

import mm_interp
from conftest import hip_lib
from formura_amd import build as B
from formura_amd.emit_hip import generate


def program():
    return generate(os.path.join(B.PROG_DIR, "float2d.fmr"), os.path.join(B.PROG_DIR, "float2d_32.yaml"),
                    "float2d", "flat")[2]


def hand_step(q):
    """q + r*(q[i+1,j] + q[i-1,j] + q[i,j+1] + q[i,j-1] - 4*q) with float loads
    and sums, the double literals 4 and r = 1/10 (products in double, rounded
    to float), in the drifting frame: output i reads input i - 1 at the centre."""
    f32 = np.float32
    c = np.roll(q, (1, 1), axis=(0, 1))          # q[i-1, j-1]: the centre
    xp = np.roll(q, (0, 1), axis=(0, 1))         # q[i, j-1]
    xm = np.roll(q, (2, 1), axis=(0, 1))         # q[i-2, j-1]
    yp = np.roll(q, (1, 0), axis=(0, 1))
    ym = np.roll(q, (1, 2), axis=(0, 1))
    s = ((xp + xm) + yp) + ym                    # float sums
    four_c = (4.0 * c.astype(np.float64)).astype(f32)
    d = s - four_c
    rd = (0.1 * d.astype(np.float64)).astype(f32)
    return (c + rd).astype(f32)


def test_typed_interpreter_matches_hand_transcription():
    mm = program()
    q = np.random.default_rng(4).random((32, 48)).astype(np.float32)
    a, b = {"q": q}, q
    for _ in range(5):
        a = mm_interp.step_typed(mm.step, 1, a)
        b = hand_step(b)
        assert a["q"].dtype == np.float32
        np.testing.assert_array_equal(a["q"], b)
    # and the double-precision reading of the same graph differs (the types matter)
    d = mm_interp.step(mm.step, 1, {"q": q.astype(np.float64)})["q"]
    assert not np.array_equal(d.astype(np.float32), hand_step(q))


@pytest.mark.gpu
@pytest.mark.parametrize("name,steps", [("float2d_32", 7), ("float2d_32_tile", 10)])
def test_float_library_bit_exact(name, steps):
    print("%s: %d steps" % (name, steps), flush=True)
    mm = program()
    h = hip_lib(name)
    h.init()
    q = {"q": h.field("q").copy()}
    assert q["q"].dtype == np.float32 and q["q"].shape == (32, 48)
    rng = np.random.default_rng(9)
    q = {"q": (q["q"] + rng.random(q["q"].shape).astype(np.float32)).astype(np.float32)}
    h.set_state(q)
    h.forward(steps)
    for _ in range(steps):
        q = mm_interp.step_typed(mm.step, 1, q)
    np.testing.assert_array_equal(h.field("q"), q["q"])


def div_program():
    return generate(os.path.join(B.PROG_DIR, "floatdiv2d.fmr"), os.path.join(B.PROG_DIR, "float2d_32.yaml"),
                    "floatdiv2d", "flat")[2]


def hand_div_step(q):
    """s/3 + (q - q[i,j-1])/10 - q/a with s = (q[i+1,j] + q[i-1,j]) + q[i,j+1]:
    float sums and differences; each quotient by a literal (3, 10, and a =
    7/10, a Rational constant) taken in double and rounded to float."""
    f32, f64 = np.float32, np.float64
    c = np.roll(q, (1, 1), axis=(0, 1))          # q[i-1, j-1]: the centre
    xp = np.roll(q, (0, 1), axis=(0, 1))
    xm = np.roll(q, (2, 1), axis=(0, 1))
    yp = np.roll(q, (1, 0), axis=(0, 1))
    ym = np.roll(q, (1, 2), axis=(0, 1))
    s = (xp + xm) + yp
    t1 = (s.astype(f64) / 3.0).astype(f32)
    t2 = ((c - ym).astype(f64) / 10.0).astype(f32)
    t3 = (c.astype(f64) / (7.0 / 10.0)).astype(f32)
    return ((t1 + t2) - t3).astype(f32)


def test_typed_interpreter_divides_like_c():
    mm = div_program()
    q = np.random.default_rng(6).random((32, 48)).astype(np.float32)
    a, b = {"q": q}, q
    for _ in range(5):
        a = mm_interp.step_typed(mm.step, 1, a)
        b = hand_div_step(b)
        np.testing.assert_array_equal(a["q"], b)
    # the product with the float reciprocal is a different number in many cells
    bad = (q.astype(np.float32) * np.float32(1.0 / 3.0)).astype(np.float32)
    good = (q.astype(np.float64) / 3.0).astype(np.float32)
    assert np.count_nonzero(bad != good) > 100


@pytest.mark.gpu
@pytest.mark.parametrize("name,steps", [("floatdiv2d_32", 6), ("floatdiv2d_32_tile", 9)])
def test_float_division_bit_exact(name, steps):
    print("%s: %d steps" % (name, steps), flush=True)
    mm = div_program()
    h = hip_lib(name)
    h.init()
    q = {"q": np.random.default_rng(10).random((32, 48)).astype(np.float32)}
    h.set_state(q)
    h.forward(steps)
    for _ in range(steps):
        q = mm_interp.step_typed(mm.step, 1, q)
    np.testing.assert_array_equal(h.field("q"), q["q"])


def kahan_program():
    return generate(os.path.join(B.PROG_DIR, "heat3d_kahan.fmr"), os.path.join(B.PROG_DIR, "heat3d_kahan_16.yaml"),
                    "heat3d_kahan", "flat")[2]


def hand_kahan_step(q, c):
    """heat3d_kahan.fmr by hand in float32: second differences from the centre,
    the increment r*(lx + ly + lz) with r = 1/8 (a product in double, exact, then
    float), and the Kahan update y = dq - c, t = q + y, c' = (t - q) - y, q' = t,
    in the drifting frame (output i reads input i - 1 at the centre)."""
    f32 = np.float32
    ax = (0, 1, 2)
    cen = np.roll(q, (1, 1, 1), ax)

    def nb(axis, d):
        s = [1, 1, 1]
        s[axis] = 1 + d
        return np.roll(q, s, ax)
    lx = (nb(0, -1) - cen) + (nb(0, 1) - cen)
    ly = (nb(1, -1) - cen) + (nb(1, 1) - cen)
    lz = (nb(2, -1) - cen) + (nb(2, 1) - cen)
    dq = (0.125 * ((lx + ly) + lz).astype(np.float64)).astype(f32)
    y = dq - np.roll(c, (1, 1, 1), ax)
    t = cen + y
    return t.astype(f32), ((t - cen) - y).astype(f32)


def test_kahan_heat_interpreter_matches_hand_transcription_and_beats_plain_fp32():
    """The typed interpreter runs the Kahan program as the hand transcription does,
    bit for bit; over 400 steps from the same state the compensated FP32 run
    stays closer to the FP64 oracle (heat3d_16, the same stencil in FP64) than
    plain FP32 accumulation (q' = q + dq in float) does."""
    from conftest import layout_of, oracle_lib
    mm = kahan_program()
    q0 = np.random.default_rng(12).random((16, 16, 16))
    st = {"q": q0.astype(np.float32), "c": np.zeros((16, 16, 16), np.float32)}
    hq, hc = st["q"], st["c"]
    for _ in range(5):
        st = mm_interp.step_typed(mm.step, 1, st)
        hq, hc = hand_kahan_step(hq, hc)
        np.testing.assert_array_equal(st["q"], hq)
        np.testing.assert_array_equal(st["c"], hc)
    o = oracle_lib("heat3d_16", layout_of("heat3d_16"))
    o.init()
    o.set_state({"q": q0})
    n = 400
    o.forward(n)
    ref = o.field("q")
    kq, kc = q0.astype(np.float32), np.zeros((16, 16, 16), np.float32)
    pq = q0.astype(np.float32)
    for _ in range(n):
        kq, kc = hand_kahan_step(kq, kc)
        pq, _ = hand_kahan_step(pq, np.zeros_like(pq))    # c = 0 every step: plain accumulation
    ek = np.sqrt(np.sum((kq - ref) ** 2) / np.sum(ref ** 2))
    ep = np.sqrt(np.sum((pq - ref) ** 2) / np.sum(ref ** 2))
    print("relative L2 vs FP64 after %d steps: Kahan %.2e, plain %.2e" % (n, ek, ep))
    assert ek < 1e-7 and ek < ep


@pytest.mark.gpu
@pytest.mark.parametrize("name,steps,resident", [("heat3d_kahan_16", 9, False), ("heat3d_kahan_16_stream", 11, True),
                                                  ("heat3d_kahan_16_skew", 11, True)])
def test_kahan_heat_library_bit_exact(name, steps, resident):
    print("%s: %d steps" % (name, steps), flush=True)
    mm = kahan_program()
    h = hip_lib(name)
    h.init()
    q0 = np.random.default_rng(13).random((16, 16, 16))
    st = {"q": q0.astype(np.float32), "c": np.zeros((16, 16, 16), np.float32)}
    h.set_state(st)
    if resident:
        h.set_resident(True)
        h.upload()
        h.forward_steps(steps)
        h.download()
        h.set_resident(False)
    else:
        h.forward(steps)
    for _ in range(steps):
        st = mm_interp.step_typed(mm.step, 1, st)
    for f in ("q", "c"):
        assert h.field(f).dtype == np.float32
        np.testing.assert_array_equal(h.field(f), st[f])
