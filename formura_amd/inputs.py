"""Seeded synthetic inputs for the bundled programs.

All random fields come from SplitMix64 (Steele, Lea and Flood, "Fast splittable
pseudorandom number generators", OOPSLA 2014): state += 0x9E3779B97F4A7C15,
then the xor-shift-multiply finaliser; a uniform double in [0,1) is the top 53
bits times 2^-53.  Values are drawn in C row-major order of the field, so the
same seed gives the same field on every platform.
"""
from __future__ import annotations

import numpy as np

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, n: int, start: int = 0) -> np.ndarray:
    """The ``n`` outputs ``start .. start+n-1`` of SplitMix64 seeded with ``seed``."""
    with np.errstate(over="ignore"):
        k = np.arange(start + 1, start + n + 1, dtype=np.uint64)
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + k * _GAMMA
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed: int, shape, start: int = 0) -> np.ndarray:
    """U[0,1) doubles in row-major order."""
    n = int(np.prod(shape))
    z = splitmix64(seed, n, start)
    return ((z >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)).reshape(shape)


def heat_field(seed: int, shape) -> dict:
    """Heat equation input: q i.i.d. U[0,1)."""
    return {"q": uniform(seed, shape)}


def grayscott_perturb(u: np.ndarray, v: np.ndarray, seed: int) -> dict:
    """Seeded form of the reference driver's ``perturbate`` (examples/pearson2/main.c:13-22):
    where ``u != 1`` add ``0.01*r`` to u and to v, r ~ U[0,1) drawn u-then-v per cell
    in row-major order."""
    mask = u != 1.0
    r = uniform(seed, (u.size, 2)).reshape(u.shape + (2,))
    u2 = u.copy()
    v2 = v.copy()
    u2[mask] += 0.01 * r[..., 0][mask]
    v2[mask] += 0.01 * r[..., 1][mask]
    return {"u": u2, "v": v2}


EULER_GAMMA = 1.4


def euler_primitive(seed: int, shape):
    """Seeded compressible-flow state: rho = 1 + 0.1 U[-1,1), p = 1 + 0.1 U[-1,1),
    u, v, w ~ N(0, 0.1^2) (Box-Muller on independent SplitMix64 streams)."""
    U = [uniform(seed * 16 + k, shape) for k in range(6)]
    rho = 1.0 + 0.1 * (2.0 * U[0] - 1.0)
    p = 1.0 + 0.1 * (2.0 * U[1] - 1.0)
    r1 = np.sqrt(-2.0 * np.log1p(-U[2]))
    u = 0.1 * r1 * np.cos(2.0 * np.pi * U[3])
    v = 0.1 * r1 * np.sin(2.0 * np.pi * U[3])
    w = 0.1 * np.sqrt(-2.0 * np.log1p(-U[4])) * np.cos(2.0 * np.pi * U[5])
    return rho, u, v, w, p


def euler_state(seed: int, shape) -> dict:
    """Conserved fields rho, rhou, rhov, rhow (momentum), e (total energy) of
    ``euler_primitive``: the field names of the Euler program."""
    rho, u, v, w, p = euler_primitive(seed, shape)
    e = p / (EULER_GAMMA - 1.0) + 0.5 * rho * (u * u + v * v + w * w)
    return {"rho": rho, "rhou": rho * u, "rhov": rho * v, "rhow": rho * w, "e": e}


def euler_cfl_dt(seed: int, shape, dx: float, cfl: float = 0.4) -> float:
    """dt = CFL * dx / max(|u| + |v| + |w| + 3c) over the initial state (unsplit 3-D bound)."""
    rho, u, v, w, p = euler_primitive(seed, shape)
    c = np.sqrt(EULER_GAMMA * p / rho)
    return float(cfl * dx / np.max(np.abs(u) + np.abs(v) + np.abs(w) + 3.0 * c))


def conserved_state(seed: int, shape, fields, gamma: float) -> dict:
    """A seeded, physically valid conserved state for an Euler or MHD program,
    keyed by the program's own field names: rho = 1 + 0.1 U[-1,1), p = 1 + 0.1
    U[-1,1), velocity components 0.1 U[-1,1), magnetic components (B_x, B_y,
    B_z) 0.3 U[-1,1), psi 0.01 U[-1,1); momenta rho*V (rhou, rhov, rhow), total
    energy e = p/(gamma-1) + rho |V|^2 / 2 + |B|^2 / 2 (the B fields present).
    Each quantity draws from its own SplitMix64 stream in row-major order."""
    def draw(k):
        return 2.0 * uniform(seed * 32 + k, shape) - 1.0
    rho = 1.0 + 0.1 * draw(0)
    p = 1.0 + 0.1 * draw(1)
    vel = {c: 0.1 * draw(2 + i) for i, c in enumerate("uvw")}
    B = {c: 0.3 * draw(5 + i) for i, c in enumerate("xyz") if "B_" + c in fields}
    out = {}
    kin = sum(v * v for v in vel.values())
    mag = sum(b * b for b in B.values()) if B else 0.0
    e = p / (gamma - 1.0) + 0.5 * rho * kin + 0.5 * mag
    for f in fields:
        if f == "rho":
            out[f] = rho
        elif f in ("rhou", "rhov", "rhow"):
            out[f] = rho * vel[f[-1]]
        elif f == "e":
            out[f] = e
        elif f.startswith("B_"):
            out[f] = B[f[-1]]
        elif f == "psi":
            out[f] = 0.01 * draw(8)
        else:
            raise ValueError("conserved_state: unknown field %r" % f)
    return out
