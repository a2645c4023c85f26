"""Parity of the HIP libraries with the oracle, through the generated C API.

Both sides export the API of the reference's generated C, so each test seeds
``formura_data`` identically in the oracle library (a transcription of what
formura v2.3.2 emits) and in the HIP library, calls ``Formura_Forward`` the
same number of times, and compares the state and the ``Formura_Navi`` structs.
The bar is bit-exact: both sides round every operation as written (no FMA
contraction; host libm only in the init graph).
"""
import numpy as np
import pytest

from conftest import hip_lib, layout_of, oracle_lib
from formura_amd import inputs

pytestmark = pytest.mark.gpu


def _run_pair(name, oracle_name, seed_state, steps, resident=False):
    lay = layout_of(oracle_name)
    o = oracle_lib(oracle_name, lay)
    h = hip_lib(name)
    for lib in (o, h):
        lib.init()
    st = seed_state(o)
    o.set_state(st)
    h.set_state(st)
    if resident:
        h.set_resident(True)
        h.upload()
        h.forward_steps(steps)
        h.download()
        h.set_resident(False)
    else:
        h.forward(steps)
    o.forward(steps)
    return o, h


def _assert_same(o, h):
    assert o.navi_dict() == h.navi_dict()
    for f in o.fields:
        a, b = o.field(f), h.field(f)
        if not np.array_equal(a, b):
            d = np.abs(a - b)
            i = np.unravel_index(np.argmax(d), d.shape)
            raise AssertionError("field %s differs: max |diff| %g at %s (oracle %r, hip %r)"
                                 % (f, d.max(), i, a[i], b[i]))


@pytest.mark.parametrize("name,steps", [("heat3d_16", 21), ("heat3d_32", 7), ("heat3d_64", 100)])
def test_heat3d_bit_exact(name, steps):
    o, h = _run_pair(name, name, lambda lib: inputs.heat_field(11, lib.grid), steps)
    _assert_same(o, h)
    assert h.navi.offset_x == (-steps) % h.grid[0]


@pytest.mark.parametrize("name,forwards", [("heat3d_16_tb", 5), ("heat3d_64_tb", 25)])
def test_heat3d_temporal_blocking_bit_exact(name, forwards):
    o, h = _run_pair(name, name, lambda lib: inputs.heat_field(5, lib.grid), forwards)
    _assert_same(o, h)


def test_heat3d_init_graph_matches():
    lay = layout_of("heat3d_32")
    o = oracle_lib("heat3d_32", lay)
    h = hip_lib("heat3d_32")
    o.init()
    h.init()
    np.testing.assert_array_equal(o.field("q"), h.field("q"))
    assert o.navi_dict() == h.navi_dict()
    o.forward(3)
    h.forward(3)
    _assert_same(o, h)


def _gs_seed(lib):
    return inputs.grayscott_perturb(lib.field("u").copy(), lib.field("v").copy(), 7)


@pytest.mark.parametrize("name,steps", [("gs_64", 300), ("gs_2048", 40)])
def test_grayscott_bit_exact(name, steps):
    o, h = _run_pair(name, name, _gs_seed, steps)
    _assert_same(o, h)


def _gs_extreme_seed(lib):
    """Gray-Scott state whose Laplacian numerators leave the range where the
    branch-free constant division is correctly rounded (|a| < 2^-900: bands of
    v = 1e-300 and of subnormal v = 1e-310 next to each other, which stay tiny
    as v decays; |a| > 2^900: a band of u = 1e275; exact zeros inside the
    bands), among ordinary values: the waves that hold such a point are
    evaluated again with guarded divisions.  v is exactly 0 around the huge
    band until v's front reaches it (after 11 steps of this 64^2 run), so the
    reaction term u*v*v stays finite for the steps run."""
    st = _gs_seed(lib)
    u, v = st["u"], st["v"]
    assert u.shape[0] == 64 and not v[:26].any() and not v[39:].any()
    v[8:11, :] = 1e-300
    v[11:13, :] = 1e-310
    u[55:58, ::3] = 1e275
    return {"u": u, "v": v}


@pytest.mark.parametrize("name,oracle,steps,resident", [
    ("gs_64", "gs_64", 9, False),               # one step per launch (tile2d, 1 level)
    ("gs_64", "gs_64", 10, True),               # 4 steps per launch
    ("gs_64_tile_k5_inplace", "gs_64", 10, True),
    ("gs_64_tile_spec", "gs_64", 10, True),     # such waves evaluate the whole point again
    ("gs_64_tile_guard", "gs_64", 10, True),    # a per-lane branch in each division
    ("gs_64_tile_colmap", "gs_64", 10, True),
    ("gs_64_tile_vpair", "gs_64", 10, True),
    ("gs_64_tile_level", "gs_64", 10, True),    # per-level range test, workgroups with such values guarded
    ("gs_64_tile_level_flat", "gs_64", 9, True),
    ("gs_64_flat", "gs_64", 7, False),          # the guarded divisions of the flat kernels
])
def test_grayscott_out_of_range_dividends(name, oracle, steps, resident):
    """Bit for bit (signed zeros included) with the oracle's IEEE division where
    the dividends are tiny, subnormal, huge or zero."""
    print("%s: %d steps, resident %s" % (name, steps, resident), flush=True)
    o, h = _run_pair(name, oracle, _gs_extreme_seed, steps, resident=resident)
    assert o.navi_dict() == h.navi_dict()
    u, v = o.field("u"), o.field("v")
    assert np.all(np.isfinite(u)) and np.all(np.isfinite(v)) and (np.abs(u) > 1e270).any()
    assert ((v != 0) & (np.abs(v) < 1e-300)).any() and ((v != 0) & (np.abs(v) < 1e-307)).any()
    for f in o.fields:
        a, b = o.field(f), h.field(f)
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), f


@pytest.mark.parametrize("name,steps", [("gs_64", 50), ("gs_2048", 43)])
def test_resident_multi_step_launches(name, steps):
    """Formura_Forward_Steps runs several steps per launch (temporal blocking in LDS)."""
    o, h = _run_pair(name, name, _gs_seed, steps, resident=True)
    _assert_same(o, h)


@pytest.mark.parametrize("name,oracle,steps,resident", [
    ("heat3d_32_stream", "heat3d_32", 11, False),     # stream3d, 1 step per launch
    ("heat3d_32_stream", "heat3d_32", 23, True),      # stream3d, 2 steps per launch + remainder
    ("heat3d_64_tb_stream", "heat3d_64_tb", 7, True),  # 4 steps per Forward, 3 per launch
    ("heat3d_32_skew", "heat3d_32", 11, False),       # skewed stream3d: one barrier per plane
    ("heat3d_32_skew", "heat3d_32", 23, True),
    ("heat3d_64_tb_skew", "heat3d_64_tb", 7, True),   # skewed, 3 levels in flight
    ("gs_64_flat", "gs_64", 33, False),
    ("gs_64_tile_k8", "gs_64", 61, True),             # 8 steps per launch, halo 16 of a 64 grid
    ("gs_64_tile_k5_inplace", "gs_64", 47, True),     # one LDS buffer per field, levels overwritten in place
    ("gs_64_wave_k3", "gs_64", 47, True),             # register windows along x, DPP lane shifts, no LDS
    ("gs_64_tile_t128", "gs_64", 29, True),           # 128 work-items per workgroup
    ("heat3d_32_stream512", "heat3d_32", 9, True),    # stream3d with 512 work-items per workgroup
    ("gs_64_tile_92k", "gs_64", 25, True),            # 92 KiB of LDS in one workgroup (gfx950 allows 160)
    ("gs_2048_wave", "gs_2048", 31, True),            # the same at full size (strips overlap at the last column)
    ("gs_64_tile_pair", "gs_64", 39, True),           # two adjacent points per work-item, 4 steps per launch
    ("gs_2048_tile_pair", "gs_2048", 41, True),       # the same at full size, 32 x 32 tiles
    ("gs_64_tile_spec", "gs_64", 37, True),           # divisions: whole point again for a wave out of range
    ("gs_64_tile_guard", "gs_64", 37, True),          # divisions: a per-lane branch each
    ("gs_64_tile_colmap", "gs_64", 43, True),         # a fixed column per work-item
    ("gs_64_tile_k5_inplace_colmap", "gs_64", 45, True),
    ("gs_64_tile_vpair", "gs_64", 39, True),          # two rows of one column per work-item
    ("gs_64_tile_level", "gs_64", 45, True),          # divisions untested where a level's values passed
    ("gs_64_tile_level_flat", "gs_64", 41, True),
    ("gs_2048_level", "gs_2048", 43, True),
])
def test_forced_schedules_bit_exact(name, oracle, steps, resident):
    seed = (lambda lib: inputs.heat_field(13, lib.grid)) if name.startswith("heat") else _gs_seed
    o, h = _run_pair(name, oracle, seed, steps, resident=resident)
    _assert_same(o, h)


@pytest.mark.parametrize("name,oracle,steps,resident", [
    ("gs_72x120", "gs_72x120", 37, False),                  # default schedule (8 x 8 tiles)
    ("gs_72x120", "gs_72x120", 37, True),
    ("gs_72x120_tile", "gs_72x120", 41, True),              # 24 x 40 tiles, 3 steps per launch + remainder
    ("gs_72x120_wave", "gs_72x120", 29, True),              # register windows on a 72-row grid
    ("gs_72x120_tile_pair", "gs_72x120", 35, True),         # 24 x 40 tiles, two points per work-item
    ("gs_72x120_tile_colmap", "gs_72x120", 35, True),       # 24 x 40 tiles, a fixed column per work-item
    ("heat3d_24x40x56", "heat3d_24x40x56", 13, False),      # default (flat) schedule
    ("heat3d_24x40x56_stream", "heat3d_24x40x56", 15, True),  # 8 x 8 columns along a 24-plane axis
    ("heat3d_24x40x56_skew", "heat3d_24x40x56", 15, True),    # the skewed pipeline on the same grid
    ("heat3d_24x40x56_tb", "heat3d_24x40x56_tb", 5, False),  # 2 steps per Forward, blocks 7 x 11 x 6
    ("heat3d_24x40x56_tb_stream", "heat3d_24x40x56_tb", 6, True),
])
def test_grids_no_power_of_two_divides(name, oracle, steps, resident):
    """Unequal sides that no power of two above 8 divides (72 x 120, 24 x 40 x 56)."""
    print("%s: %d steps, resident %s" % (name, steps, resident), flush=True)
    seed = (lambda lib: inputs.heat_field(19, lib.grid)) if name.startswith("heat") else _gs_seed
    o, h = _run_pair(name, oracle, seed, steps, resident=resident)
    _assert_same(o, h)
    assert o.navi_dict()["offset_x"] == (-steps * o.layout["steps_per_forward"]) % o.grid[0]


@pytest.mark.parametrize("name,oracle,steps", [("heat3d_256", "heat3d_256", 6)])
def test_full_size_tuned_schedule_bit_exact(name, oracle, steps):
    o, h = _run_pair(name, oracle, lambda lib: inputs.heat_field(17, lib.grid), steps, resident=True)
    _assert_same(o, h)


def test_heat3d_512_properties():
    """512^3 (beyond the oracle's reach in a test): a constant field is invariant
    bit-exactly and the sum is conserved to rounding after many launches."""
    h = hip_lib("heat3d_512")
    h.init()
    q = inputs.uniform(21, h.grid)
    h.set_state({"q": q})
    s0 = float(q.sum())
    h.set_resident(True)
    h.upload()
    h.forward_steps(20)
    h.download()
    h.set_resident(False)
    assert abs(float(h.field("q").sum()) - s0) < 1e-11 * s0
    h.set_state({"q": np.full(h.grid, 0.7)})
    h.forward(3)
    assert np.all(h.field("q") == 0.7)
    assert h.navi.offset_x == (-23) % 512


def _euler_seed(lib):
    return inputs.euler_state(1, lib.grid)


@pytest.mark.parametrize("name,oracle,steps,resident", [
    ("euler3d_16", "euler3d_16", 10, False),                 # 15 flux-temporary kernels + store kernel
    ("euler3d_32", "euler3d_32", 5, True),
    ("euler3d_16_inline", "euler3d_16", 8, False),           # temporaries recomputed in one kernel
    ("euler3d_32_inline_stream", "euler3d_32", 6, True),     # ... and streamed through LDS
    ("euler3d_16_staged", "euler3d_16", 7, False),           # each face flux once, staged in LDS rings
    ("euler3d_32_staged", "euler3d_32", 6, True),            # ... 16x16 tiles, 512 work-items
])
def test_euler_hllc_bit_exact(name, oracle, steps, resident):
    o, h = _run_pair(name, oracle, _euler_seed, steps, resident=resident)
    _assert_same(o, h)


def test_euler_hllc_full_size():
    o, h = _run_pair("euler3d_256", "euler3d_256", _euler_seed, 2, resident=True)
    _assert_same(o, h)


def test_leaving_residency_keeps_device_progress():
    """Resident steps, then Set_Resident(0) without an explicit download, then
    plain Forward calls (each uploads formura_data): the run continues from the
    device state, as if all steps had been plain calls."""
    o = oracle_lib("gs_64", layout_of("gs_64"))
    h = hip_lib("gs_64")
    for lib in (o, h):
        lib.init()
    st = inputs.grayscott_perturb(o.field("u").copy(), o.field("v").copy(), 8)
    o.set_state(st)
    h.set_state(st)
    h.set_resident(True)
    h.upload()
    h.forward_steps(5)
    h.set_resident(False)
    h.forward(3)
    o.forward(8)
    _assert_same(o, h)
