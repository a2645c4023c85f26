"""Adjoint (reverse-mode) kernels behind torch.autograd, with binomial checkpointing.

* the adjoint library's forward at the program's own parameters is bit-exact
  with the oracle;
* gradients w.r.t. the initial fields and the parameters F, k, Du, Dv match an
  independent FP64 torch CPU autograd reference of the Gray-Scott step (and of
  the heat step) to 1e-10;
* gradients match FP64 central differences (eps = 1e-6) along seeded random
  directions to 1e-6;
* the snapshot count changes the schedule, not the gradient (bit-identical).
"""
import numpy as np
import pytest
import torch

from conftest import layout_of, oracle_lib
from formura_amd import adjoint as A
from formura_amd import batch as B
from formura_amd import build, inputs

pytestmark = pytest.mark.gpu
DEV = "cuda"
THETA = [0.04, 0.06, 2e-5, 1e-5]


def _stepper(name):
    from conftest import current_lib
    current_lib(name)
    return A.Stepper(name)


def torch_gs_step(u, v, F, k, Du, Dv, dt=1.0, dx=0.01):
    """Gray-Scott step in the drifting frame: output i holds position i-1."""
    c = lambda x: torch.roll(x, (1, 1), (0, 1))
    xp = lambda x: torch.roll(x, (0, 1), (0, 1))
    xm = lambda x: torch.roll(x, (2, 1), (0, 1))
    yp = lambda x: torch.roll(x, (1, 0), (0, 1))
    ym = lambda x: torch.roll(x, (1, 2), (0, 1))
    lap = lambda x: ((xp(x) + xm(x)) - 2 * c(x)) / (dx * dx) + ((yp(x) + ym(x)) - 2 * c(x)) / (dx * dx)
    uc, vc = c(u), c(v)
    return (uc + dt * ((Du * lap(u) - uc * vc * vc) + F * (1 - uc)),
            vc + dt * ((Dv * lap(v) + uc * vc * vc) - (F + k) * vc))


def _gs_problem(n=64, steps=30, seed=3):
    st = B.grayscott_init(n, seed)
    x0 = np.stack([st["u"], st["v"]])
    s = _stepper("gs_%d_adj" % n)
    tgt = A.run(s, torch.tensor(x0, device=DEV), torch.tensor([0.035, 0.065, 2e-5, 1e-5], dtype=torch.float64), steps, 2)
    return s, x0, tgt[1].detach()


def _gpu_loss_and_grads(s, x0, theta, vt, steps, snaps, overlap="one"):
    x = torch.tensor(x0, device=DEV, requires_grad=True)
    p = torch.tensor(theta, dtype=torch.float64, requires_grad=True)
    out = A.run(s, x, p, steps, snaps, overlap=overlap)
    loss = 0.5 * ((out[1] - vt) ** 2).sum()
    loss.backward()
    return loss.item(), x.grad.cpu().numpy(), p.grad.cpu().numpy()


def test_adjoint_library_forward_is_bit_exact():
    s = _stepper("gs_64_adj")
    o = oracle_lib("gs_64", layout_of("gs_64"))
    o.init()
    st = B.grayscott_init(64, 9)
    o.set_state(st)
    o.forward(57)
    x = torch.tensor(np.stack([st["u"], st["v"]]), device=DEV)
    out = A.run(s, x, torch.tensor(THETA, dtype=torch.float64), 57, 0).cpu().numpy()
    np.testing.assert_array_equal(out[0], o.field("u"))
    np.testing.assert_array_equal(out[1], o.field("v"))


def test_gradients_match_torch_cpu_reference():
    steps = 30
    s, x0, vt = _gs_problem(steps=steps)
    loss, gx, gp = _gpu_loss_and_grads(s, x0, THETA, vt, steps, 3)
    u = torch.tensor(x0[0], requires_grad=True)
    v = torch.tensor(x0[1], requires_grad=True)
    th = torch.tensor(THETA, dtype=torch.float64, requires_grad=True)
    a, b = u, v
    for _ in range(steps):
        a, b = torch_gs_step(a, b, th[0], th[1], th[2], th[3])
    ref = 0.5 * ((b - vt.cpu()) ** 2).sum()
    ref.backward()
    assert abs(loss - ref.item()) <= 1e-12 * abs(ref.item())
    for got, want in ((gx[0], u.grad.numpy()), (gx[1], v.grad.numpy()), (gp, th.grad.numpy())):
        err = np.linalg.norm(got - want) / np.linalg.norm(want)
        assert err < 1e-10, err


def test_gradients_match_central_differences():
    steps = 30
    s, x0, vt = _gs_problem(steps=steps)
    loss, gx, gp = _gpu_loss_and_grads(s, x0, THETA, vt, steps, 4)
    rng = np.random.default_rng(11)
    scale = np.array(THETA)
    eps = 1e-6

    def L(x, th):
        with torch.no_grad():
            out = A.run(s, torch.tensor(x, device=DEV), torch.tensor(th, dtype=torch.float64), steps, 0)
            return 0.5 * float(((out[1] - vt) ** 2).sum())

    for it in range(8):
        dx = rng.standard_normal(x0.shape)
        dp = rng.standard_normal(4) * scale
        fd = (L(x0 + eps * dx, np.array(THETA) + eps * dp) - L(x0 - eps * dx, np.array(THETA) - eps * dp)) / (2 * eps)
        an = float((gx * dx).sum() + (gp * dp).sum())
        assert abs(fd - an) <= 1e-6 * max(abs(an), 1e-30), (it, fd, an)


def test_snapshot_count_does_not_change_gradients():
    steps = 30
    s, x0, vt = _gs_problem(steps=steps)
    res = [_gpu_loss_and_grads(s, x0, THETA, vt, steps, snaps) for snaps in (0, 2, 32)]
    for r in res[1:]:
        np.testing.assert_array_equal(r[1], res[0][1])
        np.testing.assert_array_equal(r[2], res[0][2])
    assert s.last_stats["forward_steps"] == A.recompute_steps(steps, 32, s.levels)
    assert s.last_stats["adjoint_steps"] == steps


@pytest.mark.parametrize("mode", ["one", "two"])
def test_native_reversal_matches_python_executor(mode):
    """The library's executor of the schedule (one stream, or the forward work
    on a second stream ordered by events) gives the same gradients bit for bit
    as the reference executor (one library call per operation)."""
    steps = 45
    s, x0, vt = _gs_problem(steps=steps)
    a = _gpu_loss_and_grads(s, x0, THETA, vt, steps, 3, overlap=mode)
    b = _gpu_loss_and_grads(s, x0, THETA, vt, steps, 3, overlap="python")
    np.testing.assert_array_equal(a[1], b[1])
    np.testing.assert_array_equal(a[2], b[2])


def test_heat_adjoint_matches_torch_cpu_reference():
    s = _stepper("heat3d_16_adj")
    n, steps = 16, 12
    q0 = inputs.uniform(5, (n, n, n))
    tgt = torch.tensor(inputs.uniform(6, (n, n, n)))
    x = torch.tensor(q0[None], device=DEV, requires_grad=True)
    out = A.run(s, x, torch.zeros(0, dtype=torch.float64), steps, 2)
    loss = 0.5 * ((out[0] - tgt.to(DEV)) ** 2).sum()
    loss.backward()
    q = torch.tensor(q0, requires_grad=True)
    a = q
    dx = 1.0 / n
    for _ in range(steps):
        c = torch.roll(a, (1, 1, 1), (0, 1, 2))
        lap = 0
        for ax in range(3):
            sp = [1, 1, 1]
            sm = [1, 1, 1]
            sp[ax] = 0
            sm[ax] = 2
            lap = lap + ((torch.roll(a, sp, (0, 1, 2)) + torch.roll(a, sm, (0, 1, 2))) - 2 * c) / dx ** 2
        a = c + (dx * dx / 8) * lap
    ref = 0.5 * ((a - tgt) ** 2).sum()
    ref.backward()
    # the heat step is linear with no parameters: no checkpoint, no recomputation
    assert s.state_free and s.last_stats["forward_steps"] == 0
    g = x.grad[0].cpu().numpy()
    assert np.linalg.norm(g - q.grad.numpy()) / np.linalg.norm(q.grad.numpy()) < 1e-12


def test_fp32_adjoint_close_to_fp64():
    steps = 30
    s, x0, vt = _gs_problem(steps=steps)
    _, gx, gp = _gpu_loss_and_grads(s, x0, THETA, vt, steps, 3)
    s32 = _stepper("gs_64_adj_f32")
    x = torch.tensor(x0, device=DEV, dtype=torch.float32, requires_grad=True)
    p = torch.tensor(THETA, dtype=torch.float64, requires_grad=True)
    out = A.run(s32, x, p, steps, 3)
    loss = 0.5 * ((out[1] - vt.float()) ** 2).sum()
    loss.backward()
    g32 = x.grad.double().cpu().numpy()
    err = np.linalg.norm(g32 - gx) / np.linalg.norm(gx)
    assert err < 1e-3, err


@pytest.mark.parametrize("name", ["gs_64_adj", "gs_64_adj_f32"])
def test_fused_adjoint_steps_match_single_steps(name):
    """One fused call over k steps (inner states recomputed in the call) gives the
    adjoint of k single steps bit for bit; parameter gradients agree to rounding
    (their summation order differs)."""
    s = _stepper(name)
    st = B.grayscott_init(64, 4)
    x = torch.tensor(np.stack([st["u"], st["v"]]), device=DEV, dtype=s.dtype)
    p = torch.tensor(THETA, dtype=torch.float64, device=DEV)
    g = torch.Generator().manual_seed(2)
    lam = torch.rand(x.shape, generator=g, dtype=torch.float64).to(DEV, s.dtype)
    assert s.levels >= 2
    for k in range(1, s.levels + 1):
        xs, tmp = [x], torch.empty_like(x)
        for _ in range(k - 1):
            y = torch.empty_like(x)
            s.forward(xs[-1], y, tmp, 1, p)
            xs.append(y)
        pg1 = torch.zeros(s.slots, s.nparams, dtype=torch.float64, device=DEV)
        a, b = lam.clone(), torch.empty_like(lam)
        for xm in reversed(xs):
            s.adjoint(xm, a, b, p, pg1)
            a, b = b, a
        pg2 = torch.zeros_like(pg1)
        out = torch.empty_like(lam)
        s.adjoint_steps(x, s.scratch(DEV), lam, out, k, p, pg2)
        torch.cuda.synchronize()
        assert torch.equal(out, a), k
        torch.testing.assert_close(pg2.sum(0), pg1.sum(0), rtol=1e-12, atol=0)


def test_euler_hllc_adjoint_forward_bit_exact_and_central_differences():
    """The HLLC step's adjoint (reverse mode through the wave speeds, the star
    states and the flux selection, one work-item per point): its forward equals
    the oracle bit for bit, and the gradient of a linear functional of the state
    after 6 steps matches FP64 central differences (eps = 1e-6) along 8 seeded
    directions to 1e-6."""
    s = _stepper("euler3d_16_adj")
    fields = s.desc["fields"]
    st = inputs.euler_state(2, (16, 16, 16))
    x0 = np.stack([st[f] for f in fields])
    steps = 6
    o = oracle_lib("euler3d_16", layout_of("euler3d_16"))
    o.init()
    o.set_state(st)
    o.forward(steps)
    none = torch.zeros(0, dtype=torch.float64)
    y = A.run(s, torch.tensor(x0, device=DEV), none, steps, 0).cpu().numpy()
    for i, f in enumerate(fields):
        np.testing.assert_array_equal(y[i], o.field(f))
    rng = np.random.default_rng(77)
    w = torch.tensor(rng.standard_normal(x0.shape), device=DEV)
    x = torch.tensor(x0, device=DEV, requires_grad=True)
    (w * A.run(s, x, none, steps, 2)).sum().backward()
    g = x.grad.cpu().numpy()

    def L(xx):
        with torch.no_grad():
            return float((w * A.run(s, torch.tensor(xx, device=DEV), none, steps, 0)).sum())

    eps = 1e-6
    for _ in range(8):
        d = rng.standard_normal(x0.shape) * np.abs(x0).mean(axis=(1, 2, 3), keepdims=True) * 0.01
        fd = (L(x0 + eps * d) - L(x0 - eps * d)) / (2 * eps)
        an = float((g * d).sum())
        print("HLLC adjoint: directional derivative %.12g, central difference %.12g" % (an, fd), flush=True)
        assert abs(fd - an) <= 1e-6 * abs(an), (fd, an)


def test_heat_256_adjoint_transpose_identity():
    """The heat step is linear: after 100 steps at 256^3, <A^T lam, x> =
    <lam, A x> to rounding, for seeded random x and lam."""
    s = _stepper("heat3d_256_adj")
    n, steps = 256, 100
    none = torch.zeros(0, dtype=torch.float64)
    xq = torch.tensor(inputs.uniform(31, (n, n, n))[None] - 0.5, device=DEV)
    lam = torch.tensor(inputs.uniform(32, (n, n, n))[None] - 0.5, device=DEV)
    ax = A.run(s, xq, none, steps, 0)
    x = xq.clone().requires_grad_(True)
    (lam * A.run(s, x, none, steps, 4)).sum().backward()
    lhs = float((x.grad * xq).sum())
    rhs = float((lam * ax).sum())
    scale = float(torch.linalg.norm(lam) * torch.linalg.norm(ax))
    print("heat 256^3: <A^T lam, x> = %.15g, <lam, A x> = %.15g" % (lhs, rhs), flush=True)
    assert abs(lhs - rhs) <= 1e-12 * scale


def test_configuration_gradients_central_differences():
    """The adjoint configuration: Gray-Scott 1024^2, 2000 steps, 32 snapshots,
    loss 0.5 ||v_T - v_target||^2 with v_target from a seeded forward run at
    (F, k) = (0.035, 0.065); gradients w.r.t. the initial u, v and F, k, Du, Dv
    match FP64 central differences (eps = 1e-6) along 16 seeded directions to
    1e-6."""
    steps, snaps = 2000, 32
    s = _stepper("gs_1024_adj")
    st = B.grayscott_init(1024, 11)
    x0 = np.stack([st["u"], st["v"]])
    with torch.no_grad():
        vt = A.run(s, torch.tensor(x0, device=DEV), torch.tensor([0.035, 0.065, 2e-5, 1e-5], dtype=torch.float64),
                   steps, 0)[1].clone()
    loss, gx, gp = _gpu_loss_and_grads(s, x0, THETA, vt, steps, snaps)

    def L(x, th):
        with torch.no_grad():
            y = A.run(s, torch.tensor(x, device=DEV), torch.tensor(th, dtype=torch.float64), steps, 0)
            return 0.5 * float(((y[1] - vt) ** 2).sum())

    rng = np.random.default_rng(2024)
    eps, worst = 1e-6, 0.0
    for _ in range(16):
        dx = rng.standard_normal(x0.shape)
        dp = rng.standard_normal(4) * np.array(THETA)
        fd = (L(x0 + eps * dx, np.array(THETA) + eps * dp) - L(x0 - eps * dx, np.array(THETA) - eps * dp)) / (2 * eps)
        an = float((gx * dx).sum() + (gp * dp).sum())
        worst = max(worst, abs(fd - an) / abs(an))
    print("Gray-Scott 1024^2 x 2000 steps, 32 snapshots: worst relative FD error %.3g" % worst, flush=True)
    assert worst <= 1e-6
