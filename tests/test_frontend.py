"""Front-end: parsing, OM graph, sleeve, manifestation and the emitted header.

The bundled config programs are checked through a numpy interpreter of the
manifest-machine graphs against the oracle (the hand transcription of the
reference's generated C) -- bit-exact, so the front-end reproduces the
reference's arithmetic and placement without the HIP emitter in the loop.
The reference's own example programs are parsed when /root/reference exists.
"""
import os

import numpy as np
import pytest

from conftest import layout_of, oracle_lib
from formura_amd import build as B
from formura_amd import inputs
from formura_amd.config import ConfigException, decode_config
from formura_amd.emit import EmitError, emit_header
from formura_amd.emit_hip import LibEmitter, generate
from formura_amd.manifest import calc_range, gen_mm_program
from formura_amd.parser import ParseError, parse_program
from formura_amd.translate import TranslateError, gen_om_program
import mm_interp

REF = "/root/reference"


def _mm(name):
    fmr, yml, _ = B.PROGRAMS[name]
    return generate(os.path.join(B.PROG_DIR, fmr), os.path.join(B.PROG_DIR, yml), name)[2]


def _compile(text, yaml_text):
    nc = decode_config(yaml_text)
    return gen_mm_program(gen_om_program(parse_program(text, nc)))


def test_heat3d_structure():
    mm = _mm("heat3d_16")
    om = mm.om
    assert om.config.sleeve == 1
    assert list(om.state_signature) == ["q"]
    assert len(mm.step) == 1
    (k, node), = mm.step.items()
    cursors = sorted(m.inst[1] for m in node.micro if m.inst[0] == "LoadCursorStatic")
    assert cursors == sorted([(0, 0, 0), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)])
    assert calc_range(mm.step)[k].lower == -1 and calc_range(mm.step)[k].upper == 1


def test_grayscott_structure():
    mm = _mm("gs_64")
    assert list(mm.om.state_signature) == ["u", "v"]          # Data.Map key order
    names = [n.name for n in mm.step.values()]
    assert names == ["u_next", "v_next"]                       # one nest per field, ascending OM id


def test_state_struct_sorted_by_name():
    text = ("dimension :: 1\naxes :: x\nbegin function (z, a) = init()\n  double [] :: z = 0\n"
            "  double [] :: a = 1\nend function\nbegin function (z2, a2) = step(z, a)\n"
            "  z2 = z[i+1] + z[i-1]\n  a2 = a\nend function\n")
    mm = _compile(text, "length_per_node: [1.0]\ngrid_per_node: [8]\n")
    assert list(mm.om.state_signature) == ["a", "z"]


def test_sleeve_and_half_offsets():
    # a staggered flux: F at i+1/2 from q[i] and q[i+1]; div reads F[i+1/2] - F[i-1/2]
    text = ("dimension :: 1\naxes :: x\nXf = fun(U) U[+1/2]\nXb = fun(U) U[-1/2]\n"
            "begin function q = init()\n  double [] :: q = 0\nend function\n"
            "begin function q2 = step(q)\n  manifest :: F = (Xb(q) + Xf(q))/2\n"
            "  q2 = q - (Xf(F) - Xb(F))\nend function\n")
    mm = _compile(text, "length_per_node: [1.0]\ngrid_per_node: [16]\n")
    assert mm.om.config.sleeve == 1
    rng = calc_range(mm.step)
    temps = [k for k, n in mm.step.items() if n.name == "F"]
    stores = [k for k, n in mm.step.items() if n.name == "q_next"]
    assert len(temps) == 1 and len(stores) == 1
    assert (rng[temps[0]].lower, rng[temps[0]].upper) == (0, 1)
    assert (rng[stores[0]].lower, rng[stores[0]].upper) == (-1, 1)


def test_precedence_unary_minus_below_product():
    # -a*b parses as -(a*b) and ** binds tighter than * (Parser.hs:199-211)
    text = ("dimension :: 1\naxes :: x\nbegin function q = init()\n  double [] :: q = 0\nend function\n"
            "begin function q2 = step(q)\n  q2 = -q*2**3 + q\nend function\n")
    mm = _compile(text, "length_per_node: [1.0]\ngrid_per_node: [8]\n")
    node = list(mm.step.values())[0]
    ops = [m.inst[1] for m in node.micro if m.inst[0] in ("Uniop", "Binop")]
    assert ops == ["**", "*", "-", "+"]


def test_type_errors_are_reported():
    text = ("dimension :: 1\naxes :: x\nbegin function q = init()\n  double [] :: q = 0\nend function\n"
            "begin function q2 = step(q)\n  q2 = q + q[i+1/2]\nend function\n")
    with pytest.raises(TranslateError, match="common type"):
        _compile(text, "length_per_node: [1.0]\ngrid_per_node: [8]\n")


def test_parse_error_location():
    with pytest.raises(ParseError, match=r":3:"):
        parse_program("dimension :: 1\naxes :: x\nq = (1 +\n", decode_config("length_per_node: [1.0]\ngrid_per_node: [8]\n"))


def test_inconsistent_temporal_blocking_rejected():
    text = open(os.path.join(B.PROG_DIR, "heat3d.fmr")).read()
    with pytest.raises(ConfigException, match="Inconsistent config"):
        _compile(text, "length_per_node: [1.0,1.0,1.0]\ngrid_per_node: [16,16,16]\n"
                       "grid_per_block: [6,6,6]\ntemporal_blocking_interval: 2\n")


def test_mpi_shape_rejected_by_hip_backend():
    text = open(os.path.join(B.PROG_DIR, "heat3d.fmr")).read()
    mm = _compile(text, "length_per_node: [1.0,1.0,1.0]\ngrid_per_node: [16,16,16]\nmpi_shape: [1,1,1]\n")
    with pytest.raises(EmitError, match="mpi_shape"):
        LibEmitter(mm, "x")


def test_header_layout_matches_reference_order():
    h = emit_header(_mm("gs_64").om, with_ext=False).splitlines()
    assert h[:5] == ["#pragma once", "#include <stdio.h>", "#include <stdlib.h>", "#include <stdbool.h>",
                     "#include <math.h>"]
    assert "double u[64][64];" in h and h.index("double u[64][64];") < h.index("double v[64][64];")
    navi = h[h.index("typedef struct {", h.index("} Formura_Grid_Struct;")) + 1:h.index("} Formura_Navi;")]
    assert [l.split()[1].rstrip(";") for l in navi] == [
        "time_step", "lower_x", "lower_y", "upper_x", "upper_y", "space_interval_x", "space_interval_y",
        "my_rank", "offset_x", "offset_y", "length_x", "length_y", "total_grid_x", "total_grid_y"]
    assert h[-3:] == ["void Formura_Init(int *,char ***,Formura_Navi *);", "void Formura_Forward(Formura_Navi *);",
                      "void Formura_Finalize();"]


@pytest.mark.parametrize("name,seed,steps", [("heat3d_16", 3, 6), ("gs_64", 4, 25)])
def test_frontend_graph_matches_oracle_bit_exact(name, seed, steps):
    mm = _mm(name)
    o = oracle_lib(name, layout_of(name))
    o.init()
    if name.startswith("gs"):
        st = inputs.grayscott_perturb(o.field("u").copy(), o.field("v").copy(), seed)
    else:
        st = inputs.heat_field(seed, o.grid)
    o.set_state(st)
    fields = {k: v.copy() for k, v in st.items()}
    for _ in range(steps):
        fields = mm_interp.step(mm.step, mm.om.config.sleeve, fields)
    o.forward(steps)
    for f in fields:
        np.testing.assert_array_equal(fields[f], o.field(f))


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference checkout not present")
@pytest.mark.parametrize("fmr,yml,sleeve,manifest", [
    ("examples/diffusion1/diffusion1.fmr", "examples/diffusion1/diffusion1.yaml", 1, 1),
    ("examples/diffusion2/diffusion2.fmr", "examples/diffusion2/diffusion2.yaml", 1, 1),
    ("examples/diffusion3/diffusion3.fmr", "examples/diffusion3/diffusion3.yaml", 1, 1),
    ("examples/pearson2/pearson2.fmr", "examples/pearson2/pearson2.yaml", 1, 2),
    ("sample/3d_hllc/3d_hllc.fmr", "sample/3d_hllc/config.yaml", 2, 64),
])
def test_reference_programs_compile(fmr, yml, sleeve, manifest):
    from formura_amd.config import read_config
    nc = read_config(os.path.join(REF, yml))
    om = gen_om_program(parse_program(open(os.path.join(REF, fmr)).read(), nc, fmr))
    mm = gen_mm_program(om)
    assert om.config.sleeve == sleeve
    assert len(mm.step) == manifest


def test_euler_frontend_graph_matches_oracle_bit_exact():
    """The HLLC program: 24 manifest temporaries (S_L, S_R, S_M and five flux
    components per face direction) + 5 stores, through the numpy interpreter
    of the front-end's graphs, against the oracle."""
    name = "euler3d_16"
    mm = _mm(name)
    assert len([n for n in mm.step.values() if n.name.endswith("_next")]) == 5
    assert len(mm.step) == 29
    rng = calc_range(mm.step)
    assert all((r.lower, r.upper) == ((0, 1) if not mm.step[k].name.endswith("_next") else (-1, 1))
               for k, r in rng.items())
    o = oracle_lib(name, layout_of(name))
    o.init()
    st = inputs.euler_state(2, o.grid)
    o.set_state(st)
    fields = {k: v.copy() for k, v in st.items()}
    for _ in range(3):
        fields = mm_interp.step(mm.step, mm.om.config.sleeve, fields)
    o.forward(3)
    for f in fields:
        np.testing.assert_array_equal(fields[f], o.field(f))


FIRST_ORDER_LINE = "  #(U_xL, U_xR, U_yL, U_yR, U_zL, U_zR) = first_order_LR(U)   # first-order scheme\n"
MUSCL_LINE = "  (U_xL, U_xR, U_yL, U_yR, U_zL, U_zR) = MUSCL_Hancock(U)   # second-order scheme\n"
DT_LINE = "double :: dt = C_cfl*min3(dx,dy,dz)/max_signal_speed\n"


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference checkout not present")
def test_euler_program_is_the_reference_first_order_hllc(tmp_path):
    """The bundled Euler program computes what the reference's own
    sample/3d_hllc/3d_hllc.fmr computes with its first-order reconstruction
    (first_order_LR in place of MUSCL_Hancock, :210-211) and the same fixed dt:
    the same 24 temporaries with the same ranges and instruction counts, and
    the same numbers bit for bit through the interpreter."""
    text = open(os.path.join(REF, "sample/3d_hllc/3d_hllc.fmr")).read()
    assert FIRST_ORDER_LINE in text and MUSCL_LINE in text and DT_LINE in text
    ours = open(os.path.join(B.PROG_DIR, "euler3d.fmr")).read()
    dt = [ln for ln in ours.splitlines() if ln.startswith("double :: dt")][0]
    text = (text.replace(FIRST_ORDER_LINE, FIRST_ORDER_LINE.replace("#(", "(", 1))
            .replace(MUSCL_LINE, "  #" + MUSCL_LINE[2:]).replace(DT_LINE, dt + "\n"))
    p = tmp_path / "first_order.fmr"
    p.write_text(text)
    yml = os.path.join(B.PROG_DIR, "euler3d_16.yaml")
    ref = generate(str(p), yml, "ref_first_order", "flat")[2]
    mm = _mm("euler3d_16")
    assert list(ref.om.state_signature) == list(mm.om.state_signature)
    r_ref, r_mm = calc_range(ref.step), calc_range(mm.step)
    shape = lambda g, r: [(r[k].lower, r[k].upper, len(g[k].micro)) for k in sorted(g)]
    assert shape(ref.step, r_ref) == shape(mm.step, r_mm)
    st = inputs.euler_state(4, mm.om.config.grid_per_node)
    a = {k: v.copy() for k, v in st.items()}
    b = {k: v.copy() for k, v in st.items()}
    for _ in range(3):
        a = mm_interp.step(ref.step, ref.om.config.sleeve, a)
        b = mm_interp.step(mm.step, mm.om.config.sleeve, b)
    for f in a:
        np.testing.assert_array_equal(a[f], b[f], err_msg=f)


def test_inlined_temporaries_match_oracle_bit_exact():
    """Recomputing the manifest temporaries where they are read (the fused GPU
    schedules) gives the same numbers as materialising them."""
    from formura_amd.manifest import inline_temporaries
    name = "euler3d_16"
    mm = _mm(name)
    inl = inline_temporaries(mm.step)
    assert len(inl) == 5 and all(m.inst[0] != "LoadCursor" for n in inl.values() for m in n.micro)
    assert {k: (r.lower, r.upper) for k, r in calc_range(inl).items()} == \
        {k: (r.lower, r.upper) for k, r in calc_range(mm.step).items() if k in inl}
    o = oracle_lib(name, layout_of(name))
    o.init()
    st = inputs.euler_state(6, o.grid)
    o.set_state(st)
    fields = {k: v.copy() for k, v in st.items()}
    for _ in range(2):
        fields = mm_interp.step(inl, mm.om.config.sleeve, fields)
    o.forward(2)
    for f in fields:
        np.testing.assert_array_equal(fields[f], o.field(f))
