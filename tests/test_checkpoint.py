"""Binomial checkpointing (CPU): the schedule is optimal and reverses correctly."""
import functools

import pytest
import torch

import formura_amd.adjoint as A


@functools.lru_cache(maxsize=None)
def brute(l, s):
    """Exhaustive minimum of forward steps: advance j, reverse the tail with s-1, the head with s."""
    if l <= 1:
        return 0
    if s == 0:
        return l * (l - 1) // 2
    return min(j + brute(l - j, s - 1) + brute(j, s) for j in range(1, l))


@pytest.mark.parametrize("s", [0, 1, 2, 3, 5])
def test_schedule_is_optimal(s):
    for l in range(1, 60):
        assert A.forward_steps_needed(l, s, stored=False) == brute(l, s) == A.optimal_cost(l, s), (l, s)
        # the forward sweep's kept checkpoints save exactly the first path's advances
        assert A.forward_steps_needed(l, s) == brute(l, s) - sum(A.first_descent(l, s)), (l, s)


def test_config_counts():
    # 2000 steps, 32 free snapshots (33 checkpoints with the initial state):
    # repetition number 3, 3*2000 - C(36, 2) recomputed steps
    assert A.repetitions(2000, 32) == 3
    assert A.forward_steps_needed(2000, 32, stored=False) == 3 * 2000 - 630


class MockStepper:
    """x -> M x with a fixed random matrix; adjoint: lam -> M^T lam.  CPU only."""

    def __init__(self, n=6, seed=0, levels=1):
        g = torch.Generator().manual_seed(seed)
        self.levels = levels
        self.M = torch.randn(n, n, dtype=torch.float64, generator=g) / n ** 0.5
        self.dtype = torch.float64
        self.nparams = 0
        self.slots = 1
        self.calls = []

    def forward(self, x, out, tmp, nsteps, p):
        y = x.clone()
        for _ in range(nsteps):
            y = self.M @ y
        out.copy_(y)
        self.calls.append(("F", nsteps))

    def adjoint(self, state, lam, lam_out, p, pgrad):
        lam_out.copy_(self.M.T @ lam)
        self.calls.append(("A", state.clone()))

    def scratch(self, device):
        return torch.empty(max(1, self.levels - 1), 6, dtype=torch.float64)

    def adjoint_steps(self, state, scratch, lam, lam_out, k, p, pgrad):
        assert 1 <= k <= self.levels
        inner = [state.clone()]
        for _ in range(k - 1):
            inner.append(self.M @ inner[-1])
        y = lam.clone()
        for x in reversed(inner):
            y = self.M.T @ y
            self.calls.append(("A", x))
        lam_out.copy_(y)


@pytest.mark.parametrize("levels", [1, 3, 4])
@pytest.mark.parametrize("steps,snaps", [(1, 3), (7, 0), (13, 2), (40, 3), (25, 32), (30, 1)])
def test_reverse_schedule_with_mock_stepper(steps, snaps, levels):
    st = MockStepper(levels=levels)
    x0 = torch.randn(6, dtype=torch.float64, requires_grad=True)
    y = A._Run.apply(x0, torch.zeros(0, dtype=torch.float64), st, steps, snaps)
    w = torch.randn(6, dtype=torch.float64)
    (y * w).sum().backward()
    expect = w.clone()
    for _ in range(steps):
        expect = st.M.T @ expect
    assert torch.allclose(x0.grad, expect, rtol=1e-12, atol=1e-14)
    # the adjoint saw the states x_{T-1}, ..., x_0 in this order
    seen = [c[1] for c in st.calls if c[0] == "A"]
    states = [x0.detach().clone()]
    for _ in range(steps - 1):
        states.append(st.M @ states[-1])
    assert len(seen) == steps
    for a, b in zip(seen, reversed(states)):
        assert torch.allclose(a, b, rtol=1e-13, atol=1e-15)
    assert st.last_stats["forward_steps"] == A.recompute_steps(steps, snaps, levels)
    assert st.last_stats["adjoint_steps"] == steps
    if levels == 1:
        assert A.recompute_steps(steps, snaps) == A.forward_steps_needed(steps, snaps)
    # forward calls of the backward pass = recomputation; the sweep itself ran nsteps
    fwd = [c[1] for c in st.calls if c[0] == "F"]
    assert sum(fwd) == steps + A.recompute_steps(steps, snaps, levels)


def test_units_lower_the_repetition_number():
    # 2000 steps, 32 snapshots: 500 units of 4 steps need repetition number 2
    assert A.repetitions(500, 32) == 2
    assert A.recompute_steps(2000, 32, 4) == 4 * A.forward_steps_needed(500, 32)
    assert A.recompute_steps(2000, 32, 4) < 0.6 * A.recompute_steps(2000, 32, 1)


def test_reversal_buffers_and_initial_state():
    """The adjoint configuration (2000 steps, units of 4, 32 snapshots) holds at
    most 34 states at once and uses 36 buffers (checkpoint buffers, kept
    snapshots included, are recycled once the units after them are reversed;
    two spare work buffers let recomputation run ahead); no operation writes
    buffer 0, the caller's initial state."""
    ops, nbufs = A.reversal_ops(2000, 4, 32)
    assert nbufs <= 36
    assert all(not (kind == 0 and b == 0) for kind, a, b, k in ops)
    # every buffer an operation reads holds the state of the position it expects
    pos = {0: 0}
    first = A.first_descent(500, 32)
    p = 0
    for i, j in enumerate(first):
        p += j
        pos[1 + i] = p
    for kind, a, b, k in ops:
        assert a in pos, "a read of a buffer that holds no state"
        if kind == 0:
            pos[b] = pos[a] + k // 4
    assert len([o for o in ops if o[0] == 1]) == 500
