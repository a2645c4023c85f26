"""The skewed stream3d pipeline (spec key ``skew``), replayed on the CPU.

In trip t, the input plane t + 1 enters its slot first, then level j (1..K)
computes plane p = t - (j - 1) from planes p - 2s .. p of level j - 1; a level's
plane p lives in slot (p - base) mod R of its ring, R = 2s + 2, base = x0 - 2Ks.
The replay follows the slot arithmetic of the generated kernel
(``STREAM3D_SKEW`` in formura_amd/emit_hip.py) and checks, for every trip, that
each read finds the plane it needs in its slot, that no slot read in a trip is
written in the same trip (one barrier per trip is enough), and that every
output plane of the chunk is produced exactly once.
"""
import pytest

from formura_amd.emit_hip import STREAM3D_SKEW


def _slot(rs, c, R):
    """The kernel's slot of plane (t - base) + c: rs + ((c % R) + R) % R, then
    one conditional subtraction."""
    q = rs + ((c % R) + R) % R
    return q - R if q >= R else q


def replay(K, s, CX, x0):
    R = 2 * s + 2
    base = x0 - 2 * K * s
    ring = [[None] * R for _ in range(K)]
    ring[0][0] = base
    out = []
    for t in range(base, x0 + CX + K - 1):
        rs = (t - base) % R
        reads, writes = set(), []
        pre = None
        if t + 1 < x0 + CX:  # the kernel stores input plane t + 1 before the trip's levels run
            pre = (0, _slot(rs, 1, R), t + 1)
            ring[0][pre[1]] = t + 1
        for j in range(1, K + 1):
            p = t - (j - 1)
            if not (x0 - 2 * (K - j) * s <= p < x0 + CX):
                continue
            for d in range(2 * s + 1):
                sl = _slot(rs, d - 2 * s - (j - 1), R)
                assert ring[j - 1][sl] == p - 2 * s + d, (t, j, d, ring[j - 1][sl], p - 2 * s + d)
                reads.add((j - 1, sl))
            ws = _slot(rs, -(j - 1), R)
            if j < K:
                writes.append((j, ws, p))
            else:
                out.append(p)
        if pre:
            writes.append(pre)
        for lvl, sl, p in writes:
            assert (lvl, sl) not in reads, ("write over a slot read in the same trip", t, lvl, sl)
            ring[lvl][sl] = p
    assert out == list(range(x0, x0 + CX))


@pytest.mark.parametrize("K,s,CX,x0", [(1, 1, 8, 0), (2, 1, 8, 8), (2, 1, 64, 448), (3, 1, 16, 16), (4, 1, 8, 24),
                                       (5, 1, 4, 0), (2, 2, 8, 16), (3, 2, 5, 10), (1, 3, 6, 6)])
def test_skewed_ring_bookkeeping(K, s, CX, x0):
    replay(K, s, CX, x0)


def test_kernel_uses_the_replayed_slot_arithmetic():
    """The kernel text computes slots as the replay does (a constant offset
    reduced mod R, then one conditional subtraction)."""
    assert "((d - 2 * SLEEVE - (j - 1)) %% R + R) %% R" in STREAM3D_SKEW
    assert "((-(j - 1)) %% R + R) %% R" in STREAM3D_SKEW
    assert "R = %(R)d" in STREAM3D_SKEW


def test_a_single_conditional_add_fails_for_deep_pipelines():
    """Planted mistake: the slot as rs + c with one conditional add of R (c in
    [-(K - 1) - 2s, 0]) reads a wrong slot once K - 1 + 2s > R."""
    K, s, R = 4, 1, 4
    rs, c = 0, -(K - 1) - 2 * s
    naive = rs + c + R if rs + c < 0 else rs + c
    assert naive < 0 and _slot(rs, c, R) == (rs + c) % R
