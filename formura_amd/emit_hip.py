"""HIP library emitter: kernels, host orchestration and the C ABI of one program.

The generated ``.hip`` file is self-contained apart from the runtime header
``csrc/formura_hip_rt.h``.  It defines ``formura_data`` and the reference API
(``Formura_Init``/``Formura_Forward``/``Formura_Finalize``/``to_pos_*``), and the
device-resident extension declared in ``include/formura_hip.h``.

Kernel schedules (chosen per program, see ``choose_schedule``):

``flat``     one work-item per output point, every load through the cache
             hierarchy with periodic wrap; manifest temporaries get their own
             kernels (any dimension, any program).
``stream3d`` 3-D programs without temporaries: a workgroup owns a
             (TY x TZ) column tile and streams along the slowest axis; input
             planes are staged in an LDS ring of ``2s+1`` planes per field with
             their halos, so each input value is read from HBM once per tile.
``tile2d``   2-D programs without temporaries: a workgroup stages a tile plus
             a halo of ``s*k`` in LDS and advances ``k`` steps there (temporal
             blocking), writing only the ``k``-step result.
"""
from __future__ import annotations

import json
from typing import Dict, List

from . import syntax as S
from .emit import (level_division_bound, LDS_LOAD_LINE, tile_load, mixed_precision, mixed_precision_graph, Body, EmitError, GraphPlan, c_elem_type, emit_header,
                   lib_description, navi_fields, show_double, state_fields, CTYPE)
from .manifest import MMNode, calc_range


def _idx_expr(dim, names):
    # row-major linear index with compile-time extents N1..Nd
    e = names[0]
    for k in range(1, dim):
        e = "(%s)*N%d+%s" % (e, k + 1, names[k])
    return e


def _temps_read_index(g) -> bool:
    return any(n.micro[-1].inst[0] != "Store" and any(m.inst[0] == "LoadIndex" for m in n.micro)
               for n in g.values())


def _index_inlined(g):
    """A hook graph with its temporaries inlined when one of them reads an index."""
    if not _temps_read_index(g):
        return g
    from .manifest import inline_temporaries
    return inline_temporaries(g)


class LibEmitter:
    def __init__(self, mm, name: str, schedule: str = "auto", options=None):
        self.mm = mm
        self.om = mm.om
        self.name = name
        self.dim = self.om.dimension
        self.cfg = self.om.config
        self.N = self.cfg.grid_per_node
        self.fields = state_fields(self.om)
        self.field_index = {n: i for i, (n, _) in enumerate(self.fields)}
        self.options = dict(options or {})
        if self.cfg.mpi_shape is not None:
            raise EmitError("mpi_shape is set: the HIP backend runs one rank on one GPU; omit mpi_shape")
        step_mm = mm.step
        spec = schedule if isinstance(schedule, dict) else self.cfg.extra.get("hip_schedule")
        if schedule == "auto" and spec is None:
            from .emit_hip import _tuned_spec
            spec = _tuned_spec(name)
        if schedule == "auto" and spec is None:
            # untuned programs with cheap manifest temporaries (limited
            # reconstructions, single-cursor wave speeds): recompute those where
            # read, one kernel per remaining level (the reference's MUSCL-Hancock
            # HLLC at 128^3: 2.3 -> 5.9 GLUPS, HLLD MHD 1.3 -> 2.8, bit-identical)
            from .manifest import cheap_temporaries
            if cheap_temporaries(mm.step):
                spec = {"kind": "flat", "inline": "cheap"}
        self.inlined = bool(isinstance(spec, dict) and spec.get("inline"))
        # "inline": "inner" recomputes only the temporaries no Store node reads
        # (wave speeds read by the flux temporaries) inside their readers;
        # the temporaries the update reads stay materialized
        self.inline_inner = self.inlined and spec.get("inline") in ("inner", "cheap")
        self.inline_policy = spec.get("inline") if self.inline_inner else None
        if self.inline_inner:
            self.inlined = False
        # mixed precision: FP32 arithmetic except branch decisions, square-root
        # arguments and the update chain (emit.mixed_precision), decided per
        # node before any inlining
        self.mixed = bool(isinstance(spec, dict) and spec.get("mixed"))
        # contract: an opt-in FP64 form that is not bit-exact -- FMA contraction
        # across the instruction list and division by a constant as the
        # product with its reciprocal (checked against the oracle to a stated
        # tolerance, not bit for bit)
        self.contract = bool(isinstance(spec, dict) and spec.get("contract"))
        self._prec = {k: mixed_precision(n) for k, n in mm.step.items()} if self.mixed else None
        storage = None
        if self.mixed and not self.inlined:
            # temporaries stay materialized: precision decided over the whole graph
            self._prec, storage = mixed_precision_graph(mm.step)
        step_mm = mm.step
        if _temps_read_index(mm.step) and self.cfg.blocking is not None:
            raise EmitError("index access inside a manifest temporary is supported without temporal blocking only")
        if _temps_read_index(mm.step) and not self.inlined:
            # a temporary's index reads take its own nest's loop index, which a
            # periodic materialization cannot hold: recompute it where it is read
            self.inlined, self.inline_inner, self.inline_policy = True, False, None
        if self.inline_inner:
            from .manifest import cheap_temporaries, inline_temporaries, inner_temporaries
            inner = inner_temporaries(mm.step) if self.inline_policy == "inner" else cheap_temporaries(mm.step)
            keep = frozenset(k for k in mm.step if mm.step[k].micro[-1].inst[0] != "Store" and k not in inner)
            orig_rng = calc_range(mm.step)
            step_mm = inline_temporaries(mm.step, keep=keep)
            new_rng = calc_range(step_mm)
            for k in step_mm:
                if (orig_rng[k].lower, orig_rng[k].upper) != (new_rng[k].lower, new_rng[k].upper):
                    raise EmitError("inlining changed the stencil range of %s" % step_mm[k].name)
            if self.mixed:
                # the precision of the remaining temporaries is decided over the
                # graph the kernels evaluate
                self._prec, storage = mixed_precision_graph(step_mm)
        elif self.inlined:
            from .manifest import inline_temporaries
            orig_rng = calc_range(mm.step)
            step_mm = inline_temporaries(mm.step, prec=self._prec)
            new_rng = calc_range(step_mm)
            for k in step_mm:
                if (orig_rng[k].lower, orig_rng[k].upper) != (new_rng[k].lower, new_rng[k].upper):
                    raise EmitError("inlining changed the stencil range of %s" % step_mm[k].name)
        self.step = GraphPlan(step_mm, self.cfg.sleeve, self.dim, "step", storage=storage)
        self.filter = (GraphPlan(_index_inlined(mm.filter), self.cfg.filter_sleeve, self.dim, "filter")
                       if mm.filter else None)
        self.first = (GraphPlan(_index_inlined(mm.first_step), self.cfg.sleeve0, self.dim, "first_step")
                      if mm.first_step else None)
        # index reads in a temporally blocked step: every sub-step of a Forward
        # reads i + offset + block_offset with the Navi of the Forward's start
        # (Templates.hs:373,481), which only a schedule that runs one Forward
        # per launch over unwrapped tile coordinates reproduces (tile2d, k = nt)
        self.tb_index = self.cfg.blocking is not None and any(
            m.inst[0] == "LoadIndex" for k in self.step.stores for m in self.step.mm[k].micro)
        self.schedule = choose_schedule(self, schedule)
        if self.tb_index and not (
                (isinstance(self.schedule, Tile2DSchedule) and not self.schedule.inplace
                 or isinstance(self.schedule, Stream3DSchedule))
                and self.schedule.k == self.cfg.steps_per_forward):
            raise EmitError("index access in a temporally blocked step needs the tile2d (ping-pong) or stream3d "
                            "schedule with k = temporal_blocking_interval (%d)" % self.cfg.steps_per_forward)

    # -- helpers -----------------------------------------------------------------
    def prec_of(self, node):
        if not self.mixed:
            return None
        if any(m.prec for m in node.micro):  # inlined: precision carried from the source nodes
            return {j: m.prec for j, m in enumerate(node.micro) if m.prec}
        return self._prec.get(node.omid) or mixed_precision(node)

    def wrap_fn(self, k, reach):
        n = self.N[k]
        if reach < n:
            return "fmr_wrap<%d>" % n
        return "fmr_mod<%d>" % n

    def axes_vars(self, base="i"):
        return ["%s%d" % (base, k + 1) for k in range(self.dim)]

    # -- flat schedule --------------------------------------------------------------
    def flat_kernels(self, plan: GraphPlan) -> List[str]:
        out = []
        dim = self.dim
        ivs = self.axes_vars()
        wrap = [self.wrap_fn(k, plan.max_reach + 2 * plan.sleeve) for k in range(dim)]
        tmp_params = "".join(", %s* __restrict__ t_%d" % (plan.temp_ctype(k), k) for k in plan.temps)
        tmp_cparams = "".join(", const %s* __restrict__ t_%d" % (plan.temp_ctype(k), k) for k in plan.temps)
        navi_params = "".join(", int off_%d" % (k + 1) for k in range(dim))

        def decompose():
            lines = ["const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;",
                     "if (t >= NTOT) return;"]
            rem = "t"
            for k in range(dim - 1, -1, -1):
                if k == 0:
                    lines.append("const int %s = (int)(%s);" % (ivs[k], rem))
                else:
                    lines.append("const int %s = (int)(%s %% N%d);" % (ivs[k], rem, k + 1))
                    rem = "(%s / N%d)" % (rem, k + 1)
            return lines

        def loaders(base):
            def pos(c):
                return [("%s(%s%+d)" % (wrap[k], ivs[k], c[k] + base[k])) if (c[k] + base[k]) else ivs[k]
                        for k in range(dim)]

            def load(field, c):
                fi = self.field_index[field]
                return "FLD%d(st_in)[%s]" % (fi, _idx_expr(dim, pos(c)))

            def tmp_load(omid, c):
                return "t_%d[%s]" % (omid, _idx_expr(dim, pos(c)))

            return load, tmp_load

        for gi, grp in enumerate(temp_groups(plan)):
            # the temporaries of one dependency level, as one kernel: inputs read
            # once, common instructions shared (same values as one kernel each)
            load, tmp_load = loaders([0] * dim)
            precs = [self.prec_of(plan.mm[k]) or {} for k in grp] if self.mixed else None
            node, roots, mprec = _merge_nodes([plan.mm[k] for k in grp], precs)
            body = Body(node, load, tmp_load, lambda ax: "0", lambda n, v: "", "a", prec=mprec, recip=self.contract, fastdiv=self.contract)
            body.lines += ["t_%d[t] = (%s)a%d;" % (k, plan.temp_ctype(k), roots[i]) for i, k in enumerate(grp)]
            deps = temp_group_inputs(plan, grp)
            outs_ = "".join(", %s* __restrict__ t_%d" % (plan.temp_ctype(k), k) for k in grp)
            params = "".join(", const %s* __restrict__ t_%d" % (plan.temp_ctype(d), d) for d in deps)
            out.append(
                "__global__ void __launch_bounds__(256) k_%s_tmpg%d(const char* __restrict__ st_in%s%s) {\n"
                % (plan.tag, gi, outs_, params)
                + "\n".join("  " + l for l in decompose() + body.lines) + "\n}\n")
        # one fused kernel for every Store node
        lines = decompose()
        for k in plan.stores:
            node = plan.mm[k]
            r = plan.rng[k]
            base = [r.offset - 2 * plan.sleeve] * dim
            load, tmp_load = loaders(base)

            def store(fname, v):
                fi = self.field_index[fname]
                return "FLD%d(st_out)[t] = %s;" % (fi, v)

            def index(ax):
                return "(%s+off_%d)" % (ivs[ax], ax + 1)

            body = Body(node, load, tmp_load, index, store, "s%d_" % k, prec=self.prec_of(node), recip=self.contract, fastdiv=self.contract)
            lines.append("{")
            lines += ["  " + l for l in body.lines]
            lines.append("}")
        out.append(
            "__global__ void __launch_bounds__(256) k_%s_store(const char* __restrict__ st_in, char* __restrict__ st_out%s%s) {\n"
            % (plan.tag, tmp_cparams, navi_params)
            + "\n".join("  " + l for l in lines) + "\n}\n")
        return out

    def flat_launch(self, plan: GraphPlan) -> str:
        """Host function that runs one application of ``plan`` (one step)."""
        dim = self.dim
        L = ["static void run_%s(const char* in, char* out, const int* offs, hipStream_t s) {" % plan.tag,
             "  const unsigned nb = (unsigned)((NTOT + 255) / 256);"]
        for gi, grp in enumerate(temp_groups(plan)):
            deps = temp_group_inputs(plan, grp)
            L.append("  hipLaunchKernelGGL(k_%s_tmpg%d, dim3(nb), dim3(256), 0, s, in%s%s);"
                     % (plan.tag, gi,
                        "".join(", (%s*)g_tmp_%s[%d]" % (plan.temp_ctype(k), plan.tag, plan.temps.index(k)) for k in grp),
                        "".join(", (const %s*)g_tmp_%s[%d]" % (plan.temp_ctype(d), plan.tag, plan.temps.index(d)) for d in deps)))
        args = "".join(", (const %s*)g_tmp_%s[%d]" % (plan.temp_ctype(d), plan.tag, plan.temps.index(d)) for d in plan.temps)
        offs = "".join(", offs[%d]" % k for k in range(dim))
        L.append("  hipLaunchKernelGGL(k_%s_store, dim3(nb), dim3(256), 0, s, in, out%s%s);" % (plan.tag, args, offs))
        L.append("  FMR_CHECK(hipGetLastError());")
        L.append("}")
        return "\n".join(L)

    # -- host Setup (init graph), mirroring mkKernel with sleeve 0 ------------------------
    def setup_function(self) -> str:
        dim = self.dim
        mm = self.mm.init
        rng = calc_range(mm)
        ivs = self.axes_vars()
        L = ["static void Formura_Setup(Formura_Navi n) {"]
        inst = self.om.grid_struct_instance_name
        for k in sorted(mm):
            node = mm[k]
            r = rng[k]
            ext = [n - r.size for n in self.N]
            is_tmp = not _is_void(node.typ)
            if is_tmp:
                ct = c_elem_type(node.typ)
                L.append("  static %s formura_mn%d%s;" % (ct, k, "".join("[%d]" % max(e, 1) for e in ext)))

            def load(field, c, r=r):
                return "%s.%s%s" % (inst, field, "".join("[%s%+d]" % (ivs[a], c[a] + r.offset) for a in range(dim)))

            def tmp_load(omid, c, r=r):
                ro = rng[omid].offset
                return "formura_mn%d%s" % (omid, "".join("[%s%+d]" % (ivs[a], c[a] + r.offset - ro) for a in range(dim)))

            def index(ax):
                return "(double)(%s+n.offset_%s+0)" % (ivs[ax], self.om.axes[ax])

            def store(fname, v, k=k):
                if is_tmp:
                    return "formura_mn%d%s = %s;" % (k, "".join("[%s]" % i for i in ivs), v)
                return "%s.%s%s = %s;" % (inst, fname, "".join("[%s]" % i for i in ivs), v)

            body = Body(node, load, tmp_load, index, store, "a", fold=False, divc=False)
            lines = list(body.lines)
            if is_tmp:
                lines.append(store(None, body.result))
            loop = "  " + "".join("for (int %s = 0; %s < %d; %s++) " % (ivs[a], ivs[a], ext[a], ivs[a]) for a in range(dim))
            L.append(loop + "{")
            L += ["    " + l for l in lines]
            L.append("  }")
        L.append("}")
        return "\n".join(L)

    # -- whole file ---------------------------------------------------------------
    def emit(self) -> str:
        om, dim = self.om, self.dim
        tname = om.grid_struct_type_name
        inst = om.grid_struct_instance_name
        desc = lib_description(om, self.name, self.schedule.name)
        L = ["// Generated by formura_amd from %s. Do not edit." % self.name,
             "#include \"formura_hip_rt.h\"",
             "#pragma clang fp contract(fast)" if self.contract else "// no FMA contraction: every operation rounded as written",
             "extern \"C\" {",
             "#include \"%s.h\"" % self.name,
             "}",
             "extern \"C\" { %s %s; }" % (tname, inst),
             "namespace {"]
        L += ["constexpr int N%d = %d;" % (k + 1, n) for k, n in enumerate(self.N)]
        L.append("constexpr long long NTOT = %s;" % "*".join("(long long)%d" % n for n in self.N))
        L.append("constexpr int SLEEVE = %d;" % self.cfg.sleeve)
        L.append("constexpr int STEPS_PER_FORWARD = %d;" % self.cfg.steps_per_forward)
        L.append("constexpr size_t STATE_BYTES = sizeof(%s);" % tname)
        for i, (n, ct) in enumerate(self.fields):
            L.append("#define FLD%d(st) ((%s*)((st) + offsetof(%s, %s)))" % (i, ct, tname, n))
        L.append("const char* const kDescription = R\"FMRJSON(%s)FMRJSON\";" % json.dumps(desc))
        plans = [p for p in (self.step, self.filter, self.first) if p is not None]
        for p in plans:
            L.append("void* g_tmp_%s[%d] = {0};" % (p.tag, max(1, len(p.temps))))
            L.append("const size_t g_tmp_bytes_%s[%d] = {%s};" % (
                p.tag, max(1, len(p.temps)),
                ", ".join("sizeof(%s) * (size_t)NTOT" % p.temp_ctype(k) for k in p.temps) or "0"))
        for p in plans:
            if p is self.step:
                L += self.schedule.kernels(self, p)
            else:
                L += self.flat_kernels(p)
        for p in plans:
            if p is self.step:
                L.append(self.schedule.launch(self, p))
            else:
                L.append(self.flat_launch(p))
        L.append(self.setup_function())
        L.append("}  // namespace")
        L.append(self.api(plans))
        return "\n".join(L) + "\n"

    def forward_body(self, shift, adv_filter) -> str:
        axes = self.om.axes
        if self.filter is None:
            lines = ["  const long long steps = forwards * STEPS_PER_FORWARD;",
                     "  fmr_advance(steps, *n);"]
            for a in axes:
                lines.append("  { const long long sh = ((long long)SLEEVE * steps) %% n->total_grid_%s;\n"
                             "    n->offset_%s = (int)((n->offset_%s - sh + n->total_grid_%s) %% n->total_grid_%s); }"
                             % (a, a, a, a, a))
            lines.append("  n->time_step += (int)steps;")
            return "\n".join(lines)
        return ("  for (long long f = 0; f < forwards; ++f) {\n"
                "    fmr_advance(STEPS_PER_FORWARD, *n);\n%s\n%s"
                "    n->time_step += STEPS_PER_FORWARD;\n  }"
                % ("\n".join("  " + l for l in shift(self.cfg.sleeve * self.cfg.steps_per_forward).split("\n")), adv_filter))

    def api(self, plans) -> str:
        om = self.om
        inst = om.grid_struct_instance_name
        L = []
        for a in om.axes:
            L.append("extern \"C\" double to_pos_%s(int i, Formura_Navi n) {\n"
                     "  return n.space_interval_%s*((i+n.offset_%s)%%n.total_grid_%s);\n}" % (a, a, a, a))
        shift = lambda s: "\n".join(
            "  n->offset_%s = (n->offset_%s - %d + n->total_grid_%s)%%n->total_grid_%s;" % (a, a, s, a, a)
            for a in om.axes)
        tmp_alloc = "\n".join(
            "  for (int k = 0; k < %d; ++k) g_tmp_%s[k] = fmr_rt::device_alloc(g_tmp_bytes_%s[k]);"
            % (len(p.temps), p.tag, p.tag) for p in plans if p.temps)
        tmp_free = "\n".join(
            "  for (int k = 0; k < %d; ++k) { fmr_rt::device_free(g_tmp_%s[k]); g_tmp_%s[k] = 0; }"
            % (len(p.temps), p.tag, p.tag) for p in plans if p.temps)
        first = ""
        if self.first:
            first = ("  { const int offs[] = {%s}; run_first_step(g_rt.cur, g_rt.nxt, offs, g_rt.stream); g_rt.swap(); }\n" % ", ".join("n->offset_%s" % a for a in om.axes) + shift(self.cfg.sleeve0)
                     + "\n  if (!g_rt.resident) fmr_rt::download(g_rt, &%s);" % inst)
        filt = ""
        if self.filter:
            filt = ("    if (n->time_step %% %d == 0) {\n      run_filter(g_rt.cur, g_rt.nxt, *n, g_rt.stream); g_rt.swap();\n%s\n    }"
                    % (self.cfg.filter_interval, "\n".join("  " + l for l in shift(self.cfg.filter_sleeve).split("\n"))))
        navi_init = "\n".join("  n->%s = %s;" % (f, v) for f, _, v in navi_fields(om))
        offs_list = ", ".join("n->offset_%s" % a for a in om.axes)
        total_list = ", ".join("n->total_grid_%s" % a for a in om.axes)
        adv_filter = ""
        if self.filter:
            adv_filter = ("    if (n->time_step %% %d == 0) {\n"
                          "      { const int offs[] = {%s}; run_filter(g_rt.cur, g_rt.nxt, offs, g_rt.stream); g_rt.swap(); ++g_rt.launches; }\n%s\n    }\n"
                          % (self.cfg.filter_interval, ", ".join("n->offset_%s" % a for a in om.axes),
                             "\n".join("  " + l for l in shift(self.cfg.filter_sleeve).split("\n"))))
        L.append("""
static fmr_rt::Runtime g_rt;

%(advance)s

// Advance by `forwards` calls of Formura_Forward on the device.  Without a
// filter hook the steps of all calls run back to back (in multi-step launches
// where the schedule has them); the Navi bookkeeping is the closed form of
// repeated Templates.hs:328,346 updates.
static void fmr_forward_device(Formura_Navi *n, long long forwards) {
  if (forwards <= 0) return;
%(forward_body)s
}

extern "C" void Formura_Init(int *argc, char ***argv, Formura_Navi *n) {
  (void)argc; (void)argv;
%(navi_init)s
  Formura_Setup(*n);
  fmr_rt::init(g_rt, STATE_BYTES);
  (void)fmr_rt::static_scratch();  // device address of the totals' static scratch, once per device
%(tmp_alloc)s
  fmr_rt::upload(g_rt, &%(inst)s);
%(first)s
}

extern "C" void Formura_Forward(Formura_Navi *n) {
  if (!g_rt.ready) fmr_rt::fatal("Formura_Forward called before Formura_Init");
  if (!g_rt.resident) fmr_rt::upload(g_rt, &%(inst)s);
  fmr_forward_device(n, 1);
  if (!g_rt.resident) fmr_rt::download(g_rt, &%(inst)s);
}

extern "C" void Formura_Forward_Steps(Formura_Navi *n, int k) {
  if (!g_rt.ready) fmr_rt::fatal("Formura_Forward_Steps called before Formura_Init");
  if (k < 0) fmr_rt::fatal("Formura_Forward_Steps: negative count");
  if (!g_rt.resident) fmr_rt::upload(g_rt, &%(inst)s);
  g_rt.launches = 0;
  fmr_rt::timer_start(g_rt);
  fmr_forward_device(n, k);
  fmr_rt::timer_stop(g_rt);
  if (!g_rt.resident) fmr_rt::download(g_rt, &%(inst)s);
}

extern "C" void Formura_Finalize() {
  if (!g_rt.ready) return;
  if (g_rt.resident) fmr_rt::download(g_rt, &%(inst)s);
%(tmp_free)s
  fmr_rt::shutdown(g_rt);
}

// Leaving residency downloads the device state into formura_data, so that the
// next Formura_Forward (which uploads formura_data) continues from it.
extern "C" void Formura_HIP_Set_Resident(int on) {
  const int now = on ? 1 : 0;
  if (g_rt.ready && g_rt.resident && !now) fmr_rt::download(g_rt, &%(inst)s);
  g_rt.resident = now;
}
extern "C" void Formura_HIP_Upload(void) { if (g_rt.ready) fmr_rt::upload(g_rt, &%(inst)s); }
extern "C" void Formura_HIP_Download(void) { if (g_rt.ready) fmr_rt::download(g_rt, &%(inst)s); }
extern "C" const char *Formura_HIP_Describe(void) { return kDescription; }
extern "C" void *Formura_HIP_State_Device(void) { return g_rt.cur; }
extern "C" void *Formura_HIP_Stream(void) { return (void*)g_rt.stream; }
extern "C" double Formura_HIP_Last_Elapsed_Ms(void) { return g_rt.last_ms; }
extern "C" long long Formura_HIP_Last_Launches(void) { return g_rt.launches; }

extern "C" void Formura_HIP_Totals(double *out) {
  if (!g_rt.ready) fmr_rt::fatal("Formura_HIP_Totals called before Formura_Init");
  double *res = fmr_rt::lazy_scratch(g_rt, fmr_rt::TOTALS_BLOCKS + %(nf)d) + fmr_rt::TOTALS_BLOCKS;
%(totals)s
  FMR_CHECK(hipMemcpyAsync(out, res, %(nf)d * sizeof(double), hipMemcpyDeviceToHost, g_rt.stream));
  FMR_CHECK(hipStreamSynchronize(g_rt.stream));
}

// Allocation-free totals: nfields fields of `cells` values each (element size
// elem_bytes: 8 FP64, 4 FP32), back to back at device address `state` (a state
// image such as Formura_HIP_State_Device()), summed as Formura_HIP_Totals does
// into out[f] (device memory), enqueued on `stream` (0: the null stream).
// The partial sums live in the file-scope device array of the runtime; no
// memory is allocated or freed.
extern "C" void Formura_HIP_Totals_Static(const void *state, int nfields, long long cells, int elem_bytes,
                                          double *out, void *stream) {
  if (nfields < 1 || nfields > FMR_TOTALS_STATIC_MAX_FIELDS)
    fmr_rt::fatal("Formura_HIP_Totals_Static: nfields must be 1..64");
  if (cells < 0 || cells > FMR_TOTALS_STATIC_MAX_CELLS)
    fmr_rt::fatal("Formura_HIP_Totals_Static: cells must be 0..2^40");
  if (elem_bytes != 4 && elem_bytes != 8) fmr_rt::fatal("Formura_HIP_Totals_Static: elem_bytes must be 4 or 8");
  if (!state || !out) fmr_rt::fatal("Formura_HIP_Totals_Static: null pointer");
  double *part = fmr_rt::static_scratch();
  hipStream_t s = (hipStream_t)stream;
  for (int f = 0; f < nfields; ++f) {
    if (elem_bytes == 8)
      fmr_rt::field_total_with(part, (const double *)state + (long long)f * cells, cells, out + f, s);
    else
      fmr_rt::field_total_with(part, (const float *)state + (long long)f * cells, cells, out + f, s);
  }
}

extern "C" long long Formura_HIP_Positivity_Violations(void) { return fmr_rt::violations(g_rt); }
""" % dict(nf=len(self.fields), totals="\n".join(
    "  fmr_rt::field_total(g_rt, FLD%d(g_rt.cur), %s, res + %d);" % (i, "*".join("%dLL" % n for n in self.N), i)
    for i in range(len(self.fields))), advance=self.schedule.advance(self), forward_body=self.forward_body(shift, adv_filter),
           navi_init=navi_init, tmp_alloc=tmp_alloc, tmp_free=tmp_free, first=first, inst=inst))
        return "\n".join(L)


TEMP_GROUP_MICRO = 20000  # instructions of one merged temporary kernel (before sharing): one kernel per level in practice


def _indent(text: str, pre: str) -> str:
    return "\n".join(pre + l if l.strip() else l for l in text.split("\n"))


def temp_groups(plan):
    """Manifest temporaries in launch order, grouped: by dependency level (a
    temporary's level is one more than the highest level it reads; state-only
    ones are level 1), each level cut in id order into groups of at most
    TEMP_GROUP_MICRO instructions.  A group reads only lower levels
    (``temp_group_inputs``)."""
    level = {}
    for k in plan.temps:
        deps = [m.inst[2] for m in plan.mm[k].micro if m.inst[0] == "LoadCursor"]
        level[k] = 1 + max([level[d] for d in deps] + [0])
    out = []
    for lv in sorted(set(level.values())):
        cur, size = [], 0
        for k in plan.temps:
            if level[k] != lv:
                continue
            n = len(plan.mm[k].micro)
            if cur and size + n > TEMP_GROUP_MICRO:
                out.append(cur)
                cur, size = [], 0
            cur.append(k)
            size += n
        if cur:
            out.append(cur)
    return out


def temp_group_inputs(plan, grp):
    """The temporaries a group may read: every one of a lower level."""
    level = {}
    for k in plan.temps:
        deps = [m.inst[2] for m in plan.mm[k].micro if m.inst[0] == "LoadCursor"]
        level[k] = 1 + max([level[d] for d in deps] + [0])
    return [d for d in plan.temps if level[d] < level[grp[0]]]


def _is_void(t):
    return isinstance(t, S.ElemType) and t.name == "void"


# --- schedules ---------------------------------------------------------------------

def _offsets_at(axes, base="n0", t="t"):
    """Navi offsets after t steps of sleeve SLEEVE, as C expressions."""
    return ["(int)((((long long)%s.offset_%s - (long long)SLEEVE * %s) %% N%d + N%d) %% N%d)" % (base, a, t, k + 1, k + 1, k + 1)
            for k, a in enumerate(axes)]


class FlatSchedule:
    """One launch per step (plus one per manifest temporary)."""

    name = "flat"

    def kernels(self, em, plan):
        return em.flat_kernels(plan)

    def launch(self, em, plan):
        return em.flat_launch(plan)

    def advance(self, em):
        offs = _offsets_at(em.om.axes)
        return ("static void fmr_advance(long long steps, const Formura_Navi &n0) {\n"
                "  for (long long t = 0; t < steps; ++t) {\n"
                "    const int offs[] = {%s};\n"
                "    run_step(g_rt.cur, g_rt.nxt, offs, g_rt.stream); g_rt.swap();\n"
                "    g_rt.launches += %d;\n  }\n}" % (", ".join(offs), len(temp_groups(em.step)) + 1))


class Tile2DSchedule:
    """2-D, no temporaries: K steps per launch in LDS (temporal blocking).

    A workgroup owns a TX x TY tile of the output.  With the drifting frame an
    output index i depends only on input indices [i-2s, i], so K steps need the
    tile plus a low-side halo of H = 2*K*s rows and columns.  The halo tile is
    loaded once, level j = 1..K computes local indices [2*j*s, TX+H) x
    [2*j*s, TY+H) from level j-1 (ping-pong LDS buffers), and level K is
    written to HBM.  Every point is evaluated with the same instruction list as
    the one-step kernel, so the result is identical to K single steps.
    """

    name = "tile2d"

    def __init__(self, tx, ty, k, inplace=False, threads=256, pair=False, div="fix", colmap=False, vpair=False):
        self.tx, self.ty, self.k = tx, ty, k
        # div: how divisions by constants are guarded against dividends outside
        # the range of the fused sequence.  "spec": no branch per division
        # (fmr_divc1s); a wave with such a dividend evaluates the point again
        # with the guarded forms (one wave-uniform branch per point).  "fix":
        # each division redone with the guarded form where its wave needs it
        # (a wave-uniform branch per division).  "guard": a per-lane branch in
        # each division (fmr_divc1).  In every form the point's loads are
        # issued first.
        # "level": the values each level writes (and the loaded tile) are tested
        # against fmr_level_ok; where no wave of the workgroup found one outside
        # it, the next level divides with the fused sequences and no test (the
        # generator proved every dividend an integer combination of such values,
        # emit.level_division_bound), else with the guarded form.
        if div not in ("spec", "fix", "guard", "level"):
            raise EmitError("div must be spec, fix, guard or level")
        self.div = div
        # colmap: a level's work-items form P = NTH / W row groups of W columns
        # (W = the level's width); each owns one column and steps down rows
        # P apart, so a point's LDS offset advances by a constant instead of
        # being decomposed from a flat index
        self.colmap = bool(colmap)
        # vpair: each work-item evaluates two vertically adjacent points (rows
        # r, r+1 of one column): the column values they share are loaded once
        self.vpair = bool(vpair)

        # work-items per workgroup (a level's points are dealt round-robin)
        self.threads = int(threads)
        # pair: each work-item evaluates two adjacent points of a row; a level's
        # work-items form a fixed 2-D grid (P pairs per row, NTH/P rows per
        # pass), so a point's LDS offset advances by a constant per pass
        # instead of being decomposed from a flat index
        self.pair = bool(pair)
        # in place: one LDS buffer per field; a level's values go to registers
        # and overwrite the previous level after a barrier (half the LDS, so
        # more workgroups per CU, for one more barrier per level)
        self.inplace = bool(inplace)

    def kernels(self, em, plan):
        out = []
        s = plan.sleeve
        TX, TY = self.tx, self.ty
        fields = em.fields
        # point function: reads level-(j-1) LDS buffers around local offset o
        loads_used = {}
        body_lines = []
        params = ", ".join("const %s* __restrict__ lds%d" % (ct, i) for i, (n, ct) in enumerate(fields))
        outs = ", ".join("%s& o%d" % (ct, i) for i, (n, ct) in enumerate(fields))
        tbi = getattr(em, "tb_index", False)

        def bodies(spec_ok, fast=False):
            loads, lines, used = [], [], False
            for k in plan.stores:
                node = plan.mm[k]
                r = plan.rng[k]
                base = r.offset - 2 * s

                def load(field, c, base=base):
                    fi = em.field_index[field]
                    dr, dc = c[0] + base, c[1] + base
                    return "lds%d[o + (%d)*RY + (%d)]" % (fi, dr, dc)

                def store(fname, v):
                    return "o%d = %s;" % (em.field_index[fname], v)

                def index(ax):
                    # unwrapped tile coordinate + 2s(nt - j) + the Forward's offset
                    return "(%s + ib%d)" % ("gx" if ax == 0 else "gy", ax)

                b = Body(node, load, lambda *a: None, index if tbi else (lambda ax: None), store, "s%d_" % k,
                         prec=em.prec_of(node), recip=em.contract, fastdiv=em.contract, spec_ok=spec_ok,
                         divfix=self.div == "fix" and spec_ok is None, divfast=fast)
                used |= b.spec_used
                # the node's loads (names unique to the node) before every node's arithmetic
                loads += [l for l in b.lines if LDS_LOAD_LINE.search(l)]
                lines.append("{")
                lines += ["  " + l for l in b.lines if not LDS_LOAD_LINE.search(l)]
                lines.append("}")
            return loads + lines, used

        level = self.div == "level"
        xmax = None
        if level:
            if self.inplace or self.pair or self.vpair or tbi or em.contract or any(em.prec_of(plan.mm[k])
                                                                               for k in plan.stores):
                raise EmitError("div level needs the ping-pong tile2d (flat or column mapping) of the plain path")
            bounds = [level_division_bound(plan.mm[k]) for k in plan.stores]
            if any(b is None for b in bounds):
                raise EmitError("div level: a dividend is not an integer combination of loads and literals")
            xmax = min(bounds)
        body_lines, used = bodies("bad_" if self.div == "spec" else None, fast=level)
        if self.div == "spec" and used:
            exact, _ = bodies(None)
            body_lines = (["unsigned long long bad_ = 0;  // lanes with a dividend out of range"] + body_lines
                          + ["if (__builtin_expect(bad_ != 0, 0)) {  // the wave again, with guarded divisions"]
                          + ["  " + l for l in exact] + ["}"])
        ipar = ", const int gx, const int gy, const int ib0, const int ib1" if tbi else ""
        out.append("template <int RY>\n__device__ __forceinline__ void pt_step(%s, const int o%s, %s) {\n%s\n}\n"
                   % (params, ipar, outs, "\n".join("  " + l for l in body_lines)))
        if level:
            guarded, _ = bodies(None)
            out.append("template <int RY>\n__device__ __forceinline__ void pt_step_g(%s, const int o%s, %s) {\n%s\n}\n"
                       % (params, ipar, outs, "\n".join("  " + l for l in guarded)))
        wrap1 = em.wrap_fn(0, 2 * self.k * s + 1)
        wrap2 = em.wrap_fn(1, 2 * self.k * s + 1)
        nb = 1 if self.inplace else 2
        lds = "\n".join("  __shared__ %s B%d[%d][RX * RY];" % (ct, i, nb) for i, (n, ct) in enumerate(fields))
        lcheck = ("lb0 |= !fmr_level_ok(%%s, %s);" % show_double(xmax)) if level else None
        load_tile = tile_load("RX * RY", ["const int r = idx / RY, c = idx - r * RY;",
                                          "const long long g = (long long)%s(x0 - H + r) * N2 + %s(y0 - H + c);"
                                          % (wrap1, wrap2)],
                              [("B%d[0][idx]" % i, "FLD%d(st_in)[g]" % i) for i in range(len(fields))],
                              fields[0][1], "ld", threads=self.threads, check=lcheck)
        if level:
            load_tile = ("  bool lb0 = false;  // a loaded value outside fmr_level_ok\n" + load_tile +
                         "\n  { const unsigned long long m = __builtin_amdgcn_ballot_w64(lb0);\n"
                         "    if ((threadIdx.x & 63) == 0) wbad[0][threadIdx.x >> 6] = m != 0ull; }")
            lds += "\n  __shared__ int wbad[K][%d];  // wave w found a value of level j outside fmr_level_ok" \
                % (self.threads // 64)
        srcs = ", ".join("B%d[src]" % i for i in range(len(fields)))
        if tbi:
            srcs += ", o, x0 + r - H, y0 + c - H, ib0, ib1"
        else:
            srcs += ", o"
        decl_o = " ".join("%s o%d;" % (ct, i) for i, (n, ct) in enumerate(fields))
        refs_o = ", ".join("o%d" % i for i in range(len(fields)))
        st_lds = " ".join("B%d[dst][o] = o%d;" % (i, i) for i in range(len(fields)))
        st_glb = " ".join("FLD%d(st_out)[g] = o%d;" % (i, i) for i in range(len(fields)))
        if self.pair and not self.inplace and not tbi:
            decl_p = " ".join("%s p%d;" % (ct, i) for i, (n, ct) in enumerate(fields))
            refs_p = ", ".join("p%d" % i for i in range(len(fields)))
            srcs0 = ", ".join("B%d[src]" % i for i in range(len(fields)))
            st_lds2 = " ".join("B%d[dst][o] = o%d; B%d[dst][o + 1] = p%d;" % (i, i, i, i) for i in range(len(fields)))
            st_glb2 = " ".join("FLD%d(st_out)[g] = o%d; FLD%d(st_out)[g + 1] = p%d;" % (i, i, i, i)
                               for i in range(len(fields)))
            out.append(TILE2D_PAIR % dict(TX=TX, TY=TY, NTH=self.threads, lds=lds, load_tile=load_tile,
                                          decl_o=decl_o + " " + decl_p, srcs=srcs0, refs_o=refs_o, refs_p=refs_p,
                                          st_lds=st_lds2, st_glb=st_glb2))
            return out
        ibdecl = ("\n    // index reads: the value at unwrapped tile coordinate g of level j of the\n"
                  "    // Forward reads g + 2s(K - j) + the offset at the Forward's start\n"
                  "    const int ib0 = 2 * SLEEVE * (K - j) + offT0, ib1 = 2 * SLEEVE * (K - j) + offT1;"
                  if tbi else "")
        if self.inplace:
            if self.colmap:
                out.append(TILE_ROWS_FN)
            levels = (LEVELS_INPLACE_COLS if self.colmap else LEVELS_INPLACE) % dict(
                NTH=self.threads, decl_o=decl_o, refs_o=refs_o,
                srcs=", ".join("B%d[0]" % i for i in range(len(fields))),
                rdecl=" ".join("%s res%d[NI];" % (ct, i) for i, (n, ct) in enumerate(fields)),
                rset=" ".join("res%d[it] = o%d;" % (i, i) for i in range(len(fields))),
                rput=" ".join("B%d[0][o] = res%d[it];" % (i, i) for i in range(len(fields))),
                st_glb=st_glb)
        elif self.vpair:
            decl_p = " ".join("%s p%d;" % (ct, i) for i, (n, ct) in enumerate(fields))
            refs_p = ", ".join("p%d" % i for i in range(len(fields)))
            srcs0 = ", ".join("B%d[src]" % i for i in range(len(fields)))
            levels = LEVELS_PINGPONG % dict(ibdecl=ibdecl, loop=TILE2D_VPAIR_LOOP % dict(
                NTH=self.threads, decl_o=decl_o + " " + decl_p, srcs=srcs0, refs_o=refs_o, refs_p=refs_p,
                st_lds=st_lds + " " + " ".join("B%d[dst][o + RY] = p%d;" % (i, i) for i in range(len(fields))),
                st_glb=st_glb + " " + " ".join("FLD%d(st_out)[g + N2] = p%d;" % (i, i) for i in range(len(fields)))))
        else:
            loop = TILE2D_COLS_LOOP if self.colmap else TILE2D_FLAT_LOOP
            tmpl = LEVELS_PINGPONG
            if level:
                # the level's loop twice, the branch on the workgroup-uniform flag outside it
                loop = ("    if (safe) {\n" + loop + "\n    } else {\n"
                        + loop.replace("pt_step<RY>(%(srcs)s, %(refs_o)s);", "pt_step_g<RY>(%(srcs)s, %(refs_o)s);")
                        + "\n    }")
                st_lds = st_lds + " lb |= " + " | ".join("!fmr_level_ok(o%d, %s)" % (i, show_double(xmax))
                                                        for i in range(len(fields))) + ";"
                tmpl = LEVELS_PINGPONG_LEVEL.replace("NW_", str(self.threads // 64))
            levels = tmpl % dict(ibdecl=ibdecl, loop=loop % dict(NTH=self.threads, decl_o=decl_o, srcs=srcs,
                                                                 refs_o=refs_o, st_lds=st_lds, st_glb=st_glb))
        head = TILE2D_HEAD % dict(NTH=self.threads, TX=TX, TY=TY, lds=lds,
                                  offp=", const int offT0, const int offT1" if tbi else "")
        out.append(head + TILE2D_ONE % dict(load_tile=load_tile, levels=levels))
        return out

    def launch(self, em, plan):
        nt = (em.N[0] // self.tx) * (em.N[1] // self.ty)
        if getattr(em, "tb_index", False):
            # one Forward per launch; offsets of the Navi at the Forward's start
            return ("static void run_step_tile(int k, const char* in, char* out, hipStream_t s, int o0, int o1) {\n"
                    "  if (k != %d) fmr_rt::fatal(\"bad step count\");\n"
                    "  hipLaunchKernelGGL(k_step_tile<%d>, dim3(%d), dim3(%d), 0, s, in, out, o0, o1);\n"
                    "  FMR_CHECK(hipGetLastError());\n}\n" % (self.k, self.k, nt, self.threads))
        return ("static void run_step_tile(int k, const char* in, char* out, hipStream_t s) {\n"
                "  switch (k) {\n%s  default: fmr_rt::fatal(\"bad step count\");\n  }\n"
                "  FMR_CHECK(hipGetLastError());\n}\n"
                % "".join("  case %d: hipLaunchKernelGGL(k_step_tile<%d>, dim3(%d), dim3(%d), 0, s, in, out); break;\n"
                          % (k, k, nt, self.threads) for k in range(1, self.k + 1)))

    def advance(self, em):
        if getattr(em, "tb_index", False):
            offs = _offsets_at(em.om.axes, "n0", "done")
            return ("static void fmr_advance(long long steps, const Formura_Navi &n0) {\n"
                    "  if (steps %% %d) fmr_rt::fatal(\"a temporally blocked step with index reads advances whole Forwards\");\n"
                    "  for (long long done = 0; done < steps; done += %d) {\n"
                    "    run_step_tile(%d, g_rt.cur, g_rt.nxt, g_rt.stream, %s); g_rt.swap();\n"
                    "    g_rt.launches += 1;\n  }\n}" % (self.k, self.k, self.k, ", ".join(offs)))
        return ("static void fmr_advance(long long steps, const Formura_Navi &n0) {\n"
                "  (void)n0;\n"
                "  while (steps > 0) {\n"
                "    const int k = steps >= %d ? %d : (int)steps;\n"
                "    run_step_tile(k, g_rt.cur, g_rt.nxt, g_rt.stream); g_rt.swap();\n"
                "    g_rt.launches += 1; steps -= k;\n  }\n}" % (self.k, self.k))


TILE2D_FLAT_LOOP = """    for (int idx = threadIdx.x; idx < HH * W; idx += %(NTH)d) {
      const int rr = idx / W;
      const int r = lo + rr, c = lo + (idx - rr * W);
      const int o = r * RY + c;
      %(decl_o)s
      pt_step<RY>(%(srcs)s, %(refs_o)s);
      if (j < K) {
        %(st_lds)s
      } else {
        const long long g = (long long)(x0 + r - H) * N2 + (y0 + c - H);
        %(st_glb)s
      }
    }"""

# work-item pairs of rows: pair (rp, c) is the points (lo + 2 rp, c) and
# (lo + 2 rp + 1, c); a level's height RX - lo is even (TX is)
TILE2D_VPAIR_LOOP = """    static_assert((RX - 2 * SLEEVE) %% 2 == 0, "a level's rows must pair up");
    for (int idx = threadIdx.x; idx < (HH / 2) * W; idx += %(NTH)d) {
      const int rp = idx / W;
      const int r = lo + 2 * rp, c = lo + (idx - rp * W);
      const int o = r * RY + c;
      %(decl_o)s
      pt_step<RY>(%(srcs)s, o, %(refs_o)s);
      pt_step<RY>(%(srcs)s, o + RY, %(refs_p)s);
      if (j < K) {
        %(st_lds)s
      } else {
        const long long g = (long long)(x0 + r - H) * N2 + (y0 + c - H);
        %(st_glb)s
      }
    }"""

# work-item t < P*W owns column lo + t %% W of rows lo + t / W + n*P (P = NTH / W
# row groups; the level's width W is at most NTH)
TILE2D_COLS_LOOP = """    static_assert(RY - 2 * SLEEVE <= %(NTH)d, "a level's columns must fit the workgroup");
    const int P = %(NTH)d / W;
    (void)HH;
    if ((int)threadIdx.x < P * W) {
      const int rg = (int)threadIdx.x / W;
      const int c = lo + ((int)threadIdx.x - rg * W);
      int o = (lo + rg) * RY + c;
      for (int r = lo + rg; r < RX; r += P, o += P * RY) {
        %(decl_o)s
        pt_step<RY>(%(srcs)s, %(refs_o)s);
        if (j < K) {
          %(st_lds)s
        } else {
          const long long g = (long long)(x0 + r - H) * N2 + (y0 + c - H);
          %(st_glb)s
        }
      }
    }"""


TILE_ROWS_FN = """
// most rows a work-item evaluates in one level of the column mapping
__host__ __device__ constexpr int fmr_tile_rows(int RX, int RY, int S, int K, int NTH) {
  int m = 0;
  for (int j = 1; j <= K; ++j) {
    const int lo = 2 * j * S, P = NTH / (RY - lo), n = (RX - lo + P - 1) / P;
    m = n > m ? n : m;
  }
  return m;
}
"""

TILE2D_HEAD = """
template <int K>
__global__ void __launch_bounds__(%(NTH)d) k_step_tile(const char* __restrict__ st_in, char* __restrict__ st_out%(offp)s) {
  constexpr int TX = %(TX)d, TY = %(TY)d, H = 2 * K * SLEEVE, RX = TX + H, RY = TY + H;
  constexpr int NTX = N1 / TX, NTY = N2 / TY, NT = NTX * NTY;
%(lds)s
"""

# one tile per workgroup
TILE2D_ONE = """  // XCD-aware tile order: workgroup ids are dealt round-robin to the 8 XCDs, so
  // give each XCD a contiguous band of tiles (shared halos stay in its L2).
  const int b = blockIdx.x;
  const int tile = (NT %% 8 == 0) ? ((b %% 8) * (NT / 8) + b / 8) : b;
  const int x0 = (tile / NTY) * TX, y0 = (tile %% NTY) * TY;
%(load_tile)s
  __syncthreads();
%(levels)s
}
"""

# K levels, ping-pong LDS buffers (level j reads buffer (j-1)&1, writes j&1)
LEVELS_PINGPONG = """#pragma unroll
  for (int j = 1; j <= K; ++j) {
    const int src = (j - 1) & 1, dst = j & 1;
    const int lo = 2 * j * SLEEVE, W = RY - lo, HH = RX - lo;%(ibdecl)s
%(loop)s
    if (j < K) __syncthreads();
  }"""

# the same with the per-level range test (div "level"): level j divides without
# tests when no wave found a value of level j-1 (or of the loaded tile) outside
# fmr_level_ok
LEVELS_PINGPONG_LEVEL = """#pragma unroll
  for (int j = 1; j <= K; ++j) {
    const int src = (j - 1) & 1, dst = j & 1;
    const int lo = 2 * j * SLEEVE, W = RY - lo, HH = RX - lo;%(ibdecl)s
    int nbad = 0;
#pragma unroll
    for (int w = 0; w < NW_; ++w) nbad |= wbad[j - 1][w];
    const bool safe = __builtin_amdgcn_readfirstlane(nbad) == 0;  // the same in every lane
    bool lb = false;
%(loop)s
    if (j < K) {
      const unsigned long long m = __builtin_amdgcn_ballot_w64(lb);
      if ((threadIdx.x & 63) == 0) wbad[j][threadIdx.x >> 6] = m != 0ull;
      __syncthreads();
    }
  }"""

# K levels in one LDS buffer: a level's values go to registers and overwrite
# the previous level after a barrier
LEVELS_INPLACE = """  // points of the largest level per thread; level j covers [2js, RX) x [2js, RY)
  constexpr int NI = ((RX - 2 * SLEEVE) * (RY - 2 * SLEEVE) + %(NTH)d - 1) / %(NTH)d;
#pragma unroll
  for (int j = 1; j <= K; ++j) {
    const int lo = 2 * j * SLEEVE, W = RY - lo, HH = RX - lo;
    %(rdecl)s
#pragma unroll
    for (int it = 0; it < NI; ++it) {
      const int idx = threadIdx.x + it * %(NTH)d;
      if (idx < HH * W) {
        const int rr = idx / W;
        const int r = lo + rr, c = lo + (idx - rr * W);
        const int o = r * RY + c;
        %(decl_o)s
        pt_step<RY>(%(srcs)s, o, %(refs_o)s);
        if (j < K) {
          %(rset)s
        } else {
          const long long g = (long long)(x0 + r - H) * N2 + (y0 + c - H);
          %(st_glb)s
        }
      }
    }
    if (j < K) {
      __syncthreads();  // every read of level j-1 is done
#pragma unroll
      for (int it = 0; it < NI; ++it) {
        const int idx = threadIdx.x + it * %(NTH)d;
        if (idx < HH * W) {
          const int rr = idx / W;
          const int o = (lo + rr) * RY + lo + (idx - rr * W);
          %(rput)s
        }
      }
      __syncthreads();
    }
  }"""

LEVELS_INPLACE_COLS = """  static_assert(RY - 2 * SLEEVE <= %(NTH)d, "a level's columns must fit the workgroup");
  constexpr int NI = fmr_tile_rows(RX, RY, SLEEVE, K, %(NTH)d);
#pragma unroll
  for (int j = 1; j <= K; ++j) {
    // level j covers [lo, RX) x [lo, RY): P = NTH / W row groups of W columns;
    // work-item t < P*W owns column lo + t %% W of rows lo + t / W + it*P
    const int lo = 2 * j * SLEEVE, W = RY - lo, P = %(NTH)d / W, NJ = (RX - lo + P - 1) / P;
    const int t = threadIdx.x;
    const bool on = t < P * W;
    const int rg = on ? t / W : 0;
    const int c = lo + (t - rg * W);
    %(rdecl)s
#pragma unroll
    for (int it = 0; it < NJ; ++it) {
      const int r = lo + rg + it * P;
      if (on && r < RX) {
        const int o = r * RY + c;
        %(decl_o)s
        pt_step<RY>(%(srcs)s, o, %(refs_o)s);
        if (j < K) {
          %(rset)s
        } else {
          const long long g = (long long)(x0 + r - H) * N2 + (y0 + c - H);
          %(st_glb)s
        }
      }
    }
    if (j < K) {
      __syncthreads();  // every read of level j-1 is done
#pragma unroll
      for (int it = 0; it < NJ; ++it) {
        const int r = lo + rg + it * P;
        if (on && r < RX) {
          const int o = r * RY + c;
          %(rput)s
        }
      }
      __syncthreads();
    }
  }"""


TILE2D_PAIR = """
template <int K>
__global__ void __launch_bounds__(%(NTH)d) k_step_tile(const char* __restrict__ st_in, char* __restrict__ st_out) {
  constexpr int TX = %(TX)d, TY = %(TY)d, H = 2 * K * SLEEVE, RX = TX + H, RY = TY + H;
  constexpr int NTX = N1 / TX, NTY = N2 / TY, NT = NTX * NTY;
%(lds)s
  // XCD-aware tile order: workgroup ids are dealt round-robin to the 8 XCDs, so
  // give each XCD a contiguous band of tiles (shared halos stay in its L2).
  const int b = blockIdx.x;
  const int tile = (NT %% 8 == 0) ? ((b %% 8) * (NT / 8) + b / 8) : b;
  const int x0 = (tile / NTY) * TX, y0 = (tile %% NTY) * TY;
%(load_tile)s
  __syncthreads();
#pragma unroll
  for (int j = 1; j <= K; ++j) {
    const int src = (j - 1) & 1, dst = j & 1;
    // level j covers rows and columns [lo, RX) x [lo, RY): W = RY - lo columns
    // (even: TY and lo are), P = W/2 pairs per row, R rows per pass
    constexpr int lo0 = 2 * SLEEVE;
    static_assert((RY - lo0) %% 2 == 0 && (RY - lo0) / 2 <= %(NTH)d, "a row's pairs must fit the workgroup");
    const int lo = lo0 * j, W = RY - lo, HH = RX - lo, P = W / 2, R = %(NTH)d / P;
    const int t = threadIdx.x;
    if (t < R * P) {
      const int pr = t / P, pc = t - pr * P;
      const int c = lo + 2 * pc;
      int o = (lo + pr) * RY + c;
      for (int r = lo + pr; r < RX; r += R, o += R * RY) {
        %(decl_o)s
        pt_step<RY>(%(srcs)s, o, %(refs_o)s);
        pt_step<RY>(%(srcs)s, o + 1, %(refs_p)s);
        if (j < K) {
          %(st_lds)s
        } else {
          const long long g = (long long)(x0 + r - H) * N2 + (y0 + c - H);
          %(st_glb)s
        }
      }
    }
    (void)HH;
    if (j < K) __syncthreads();
  }
}
"""


class Wave2DSchedule:
    """2-D, no temporaries: one wavefront streams a 64-column strip along x
    and keeps every level in registers (no LDS, no barriers).

    Lane l of a wave holds column c0 - H + l (H = 2*K*s), so the strip's
    outputs are lanes H..63 and the lower lanes are its low-side halo.  The
    wave walks the rows r0 - H .. r0 + R - 1 of its chunk.  With the drifting
    frame an output (r, c) reads its input at rows r-2s..r and columns
    c-2s..c, so all K levels of row r can be computed in the iteration that
    loads input row r: level j reads level j-1's rows r-2s..r from a short
    register window, and a column offset -d is the same window shifted by d
    lanes (wave_shr:1 DPP moves, applied once when a row enters the window).
    Loads and stores are whole rows of 64 consecutive values.  Every point is
    evaluated with the same instruction list as the one-step kernel, so the
    result is identical to K single steps.  Halo lanes and warm-up rows
    compute values that are never stored."""

    name = "wave2d"

    def __init__(self, rows, k, unroll=1):
        self.rows, self.k = rows, k
        # rows per loop trip: the rows of one trip at the same level are
        # independent, so unrolling gives the scheduler two chains to interleave
        self.unroll = int(unroll)

    def kernels(self, em, plan):
        s = plan.sleeve
        K, R = self.k, self.rows
        H = 2 * K * s
        fields = em.fields
        nf = len(fields)
        # which (field, row offset, column offset) each level reads
        uses = {}
        for k in plan.stores:
            r = plan.rng[k]
            base = r.offset - 2 * s
            for m in plan.mm[k].micro:
                if m.inst[0] == "LoadCursorStatic":
                    fi = em.field_index[m.inst[2]]
                    dr, dc = m.inst[1][0] + base, m.inst[1][1] + base
                    if not (-2 * s <= dr <= 0 and -2 * s <= dc <= 0):
                        raise EmitError("wave2d: load offset (%d,%d) outside [-2s,0]" % (dr, dc))
                    uses.setdefault((fi, -dc), set()).add(-dr)
        depth = {key: max(v) + 1 for key, v in uses.items()}
        shifts = {fi: sorted(d for (f, d) in uses if f == fi) for fi in range(nf)}
        out = []
        shr = []
        for ct in sorted(set(ct for _, ct in fields)):
            if ct == "double":
                shr.append("""__device__ __forceinline__ double fmr_shr1(double x) {
  // lane l receives lane l-1 (lane 0 receives 0): one wave_shr:1 DPP move per 32-bit half
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, 0x138, 0xf, 0xf, true);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), 0x138, 0xf, 0xf, true);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}""")
            elif ct == "float":
                shr.append("""__device__ __forceinline__ float fmr_shr1(float x) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x138, 0xf, 0xf, true));
}""")
        out.append("\n".join(shr) + "\n")

        def wname(lvl, fi, d, row):
            return "w%d_%d_%d_%d" % (lvl, fi, d, row)

        def level_body(lvl_in, prefix):
            """Instructions computing level lvl_in+1 from the windows of level lvl_in."""
            lines = []
            for k in plan.stores:
                node = plan.mm[k]
                r = plan.rng[k]
                base = r.offset - 2 * s

                def load(field, c, base=base):
                    fi = em.field_index[field]
                    return wname(lvl_in, fi, -(c[1] + base), -(c[0] + base))

                def store(fname, v):
                    return "n%d = %s;" % (em.field_index[fname], v)

                b = Body(node, load, lambda *a: None, lambda ax: None, store, "%s%d_" % (prefix, k),
                         prec=em.prec_of(node), recip=em.contract, fastdiv=em.contract)
                lines.append("{")
                lines += ["  " + l for l in b.lines]
                lines.append("}")
            return lines

        def push(lvl, src_of):
            """Rotate level lvl's windows and enter the new row (n<f>) with its shifts."""
            lines = []
            for fi in range(nf):
                ct = fields[fi][1]
                prev = src_of(fi)
                for d in range(0, max(shifts[fi] or [0]) + 1):
                    cur = "sh%d_%d" % (fi, d)
                    lines.append("const %s %s = %s;" % (ct, cur, prev if d == 0 else "fmr_shr1(%s)" % prev))
                    prev = cur
                    if d not in shifts[fi]:
                        continue
                    dep = depth[(fi, d)]
                    for row in range(dep - 1, 0, -1):
                        lines.append("%s = %s;" % (wname(lvl, fi, d, row), wname(lvl, fi, d, row - 1)))
                    lines.append("%s = %s;" % (wname(lvl, fi, d, 0), cur))
            return lines

        decl = []
        for lvl in range(K):
            for (fi, d), dep in sorted(depth.items()):
                ct = fields[fi][1]
                decl.append(" ".join("%s %s = 0;" % (ct, wname(lvl, fi, d, row)) for row in range(dep)))
        body = []
        # level 0: the input row (prefetched one iteration ahead)
        body.append("{")
        body += ["  " + l for l in push(0, lambda fi: "pf%d" % fi)]
        body.append("}")
        body.append("if (i + 1 < R) {")
        body.append("  const long long g = (long long)fmr_wrap<N1>(r0 + i + 1) * N2 + col;")
        body += ["  pf%d = FLD%d(st_in)[g];" % (fi, fi) for fi in range(nf)]
        body.append("}")
        # levels 1..K of the launch's template K (<= the schedule's K): level j
        # enters its window unless it is the last, which is stored
        for j in range(1, K + 1):
            body.append("if constexpr (%d <= K) {  // level %d" % (j, j))
            body.append("  " + " ".join("%s n%d;" % (ct, fi) for fi, (n, ct) in enumerate(fields)))
            body += ["  " + l for l in level_body(j - 1, "l%d_" % j)]
            if j < K:
                body.append("  if constexpr (%d < K) {" % j)
                body += ["    " + l for l in push(j, lambda fi: "n%d" % fi)]
                body.append("  }")
            body.append("  if constexpr (%d == K) {" % j)
            body.append("    if (i >= 0 && lane >= H) {")
            body.append("      const long long g = (long long)(r0 + i) * N2 + col;")
            body += ["      FLD%d(st_out)[g] = n%d;" % (fi, fi) for fi in range(nf)]
            body.append("    }")
            body.append("  }")
            body.append("}")
        out.append("""
template <int K>
__global__ void __launch_bounds__(256) k_step_wave(const char* __restrict__ st_in, char* __restrict__ st_out) {
  constexpr int H = 2 * K * SLEEVE, WOUT = 64 - H, R = %(R)d;
  constexpr int NSTRIP = (N2 + WOUT - 1) / WOUT, NCHUNK = N1 / R;
  const int lane = threadIdx.x & 63;
  const int wv = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (wv >= NSTRIP * NCHUNK) return;  // whole waves only
  const int strip = wv %% NSTRIP, chunk = wv / NSTRIP;
  // the last strip is moved back to end at column N2 (it recomputes columns of its
  // neighbour, which both store with identical values)
  const int c0 = strip * WOUT + WOUT <= N2 ? strip * WOUT : N2 - WOUT;
  const int col = fmr_wrap<N2>(c0 - H + lane);
  const int r0 = chunk * R;
  %(decl)s
  %(pfdecl)s
  {
    const long long g = (long long)fmr_wrap<N1>(r0 - H) * N2 + col;
    %(pfload)s
  }
#pragma unroll %(unroll)d
  for (int i = -H; i < R; ++i) {
%(body)s
  }
}
""" % dict(R=R, unroll=self.unroll, decl="\n  ".join(decl),
           pfdecl=" ".join("%s pf%d;" % (ct, fi) for fi, (n, ct) in enumerate(fields)),
           pfload=" ".join("pf%d = FLD%d(st_in)[g];" % (fi, fi) for fi in range(nf)),
           body="\n".join("    " + l for l in body)))
        return out

    def nwaves(self, em, k):
        wout = 64 - 2 * k * em.cfg.sleeve
        return ((em.N[1] + wout - 1) // wout) * (em.N[0] // self.rows)

    def launch(self, em, plan):
        return ("static void run_step_wave(int k, const char* in, char* out, hipStream_t s) {\n"
                "  switch (k) {\n%s  default: fmr_rt::fatal(\"bad step count\");\n  }\n"
                "  FMR_CHECK(hipGetLastError());\n}\n"
                % "".join("  case %d: hipLaunchKernelGGL(k_step_wave<%d>, dim3(%d), dim3(256), 0, s, in, out); break;\n"
                          % (k, k, (self.nwaves(em, k) + 3) // 4) for k in range(1, self.k + 1)))

    def advance(self, em):
        return ("static void fmr_advance(long long steps, const Formura_Navi &n0) {\n"
                "  (void)n0;\n"
                "  while (steps > 0) {\n"
                "    const int k = steps >= %d ? %d : (int)steps;\n"
                "    run_step_wave(k, g_rt.cur, g_rt.nxt, g_rt.stream); g_rt.swap();\n"
                "    g_rt.launches += 1; steps -= k;\n  }\n}" % (self.k, self.k))


def _merge_nodes(nodes, precs=None):
    """One micro DAG evaluating several nodes at the same point: identical
    instructions are shared (so a group of flux components computes its wave
    speeds once); ``precs``: each node's per-micro precision (mixed form), part
    of the sharing key.  Returns (node-like object, index of each node's value,
    the merged per-micro precision or None)."""
    from .manifest import Micro
    merged, memo, roots = [], {}, []
    mprec = {} if precs is not None else None
    for ni, node in enumerate(nodes):
        pr = precs[ni] if precs is not None else None
        remap = {}
        for j, m in enumerate(node.micro):
            inst = m.inst
            k = inst[0]
            if k == "Uniop":
                inst = (k, inst[1], remap[inst[2]])
            elif k == "Binop":
                inst = (k, inst[1], remap[inst[2]], remap[inst[3]])
            elif k == "Triop":
                inst = (k, inst[1], remap[inst[2]], remap[inst[3]], remap[inst[4]])
            elif k not in ("Imm", "LoadCursorStatic", "LoadCursor"):
                raise EmitError("cannot merge the instruction %s" % k)
            p = pr.get(j) if pr else None
            key = (inst, m.typ, m.src if k == "Imm" else None, p)
            if key not in memo:
                memo[key] = len(merged)
                if p is not None:
                    mprec[len(merged)] = p
                merged.append(Micro(inst, m.typ, src=m.src))
            remap[j] = memo[key]
        roots.append(remap[len(node.micro) - 1])

    class _Node:
        pass
    out = _Node()
    out.micro, out.omid, out.name = merged, None, "group"
    return out, roots, mprec


class Staged3DSchedule:
    """3-D, one level of manifest temporaries read only the state (flux
    components of a finite-volume update): stream along x, one step per
    launch, every temporary evaluated once per point into LDS.

    A workgroup owns a TY x TZ column tile and a chunk of CX planes.  The
    temporaries are grouped by the set of offsets at which the update reads
    them (an x-flux at faces x-1 and x, a y-flux at faces y-1 and y, ...).  At
    output plane x the workgroup loads the newest state plane into a ring of
    DS planes (with the y/z halo every read needs), evaluates each group's
    newest plane over the group's region (the tile plus the group's y/z read
    offsets) into a ring of DG planes, one merged instruction DAG per point,
    and then evaluates the update of plane x from the rings.  A face flux is
    computed once instead of once per neighbouring cell.  Ring depths, halos
    and the warm-up planes before the chunk follow from the offsets, so every
    value is the one the flat schedule materializes: bit-identical."""

    name = "staged3d"

    def __init__(self, ty, tz, cx, threads=256):
        self.ty, self.tz, self.cx = ty, tz, cx
        self.k = 1
        self.threads = int(threads)

    @staticmethod
    def analyse(em, ty, tz):
        plan = em.step
        mm, s = plan.mm, plan.sleeve
        if em.dim != 3 or not plan.temps:
            raise EmitError("staged3d needs a 3-D program with manifest temporaries")
        O = {t: set() for t in plan.temps}
        store_state = set()
        for k in plan.stores:
            base = plan.rng[k].offset - 2 * s
            for m in mm[k].micro:
                if m.inst[0] == "LoadCursor":
                    O[m.inst[2]].add(tuple(c + base for c in m.inst[1]))
                elif m.inst[0] == "LoadCursorStatic":
                    store_state.add(tuple(c + base for c in m.inst[1]))
                elif m.inst[0] == "LoadIndex":
                    raise EmitError("staged3d: index reads are not supported")
        groups = {}
        for t in plan.temps:
            for m in mm[t].micro:
                if m.inst[0] in ("LoadCursor", "LoadIndex"):
                    raise EmitError("staged3d: temporaries may read only the state")
            if not O[t]:
                continue
            groups.setdefault(frozenset(O[t]), []).append(t)
        G = []
        for key, ts in sorted(groups.items(), key=lambda kv: min(kv[1])):
            offs = list(key)
            D = sorted(set(tuple(m.inst[1]) for t in ts for m in mm[t].micro if m.inst[0] == "LoadCursorStatic"))
            g = {"temps": ts, "oxmin": min(o[0] for o in offs), "oxmax": max(o[0] for o in offs),
                 "ymin": min(o[1] for o in offs), "ymax": max(o[1] for o in offs),
                 "zmin": min(o[2] for o in offs), "zmax": max(o[2] for o in offs), "D": D}
            g["DG"] = g["oxmax"] - g["oxmin"] + 1
            g["RY"] = ty + g["ymax"] - g["ymin"]
            g["RZ"] = tz + g["zmax"] - g["zmin"]
            G.append(g)
        planes, ys, zs = set(), [], []
        for g in G:
            for d in g["D"]:
                planes.add(g["oxmax"] + d[0])
                ys += [g["ymin"] + d[1], ty - 1 + g["ymax"] + d[1]]
                zs += [g["zmin"] + d[2], tz - 1 + g["zmax"] + d[2]]
        for o in store_state:
            planes.add(o[0])
            ys += [o[1], ty - 1 + o[1]]
            zs += [o[2], tz - 1 + o[2]]
        st = {"Smin": min(planes), "Smax": max(planes), "SYmin": min(ys), "SYmax": max(ys),
              "SZmin": min(zs), "SZmax": max(zs)}
        st["DS"] = st["Smax"] - st["Smin"] + 1
        st["SRY"] = st["SYmax"] - st["SYmin"] + 1
        st["SRZ"] = st["SZmax"] - st["SZmin"] + 1
        st["W"] = max(g["DG"] for g in G) - 1 + st["DS"] - 1
        return G, st

    @staticmethod
    def lds_bytes(em, ty, tz):
        G, st = Staged3DSchedule.analyse(em, ty, tz)
        bpv = {n: (8 if ct == "double" else 4) for n, ct in em.fields}
        state = sum(bpv.values()) * st["DS"] * st["SRY"] * st["SRZ"]
        tmp = sum(g["DG"] * g["RY"] * g["RZ"] * (8 if em.step.temp_ctype(t) == "double" else 4)
                  for g in G for t in g["temps"])
        return state + tmp

    def kernels(self, em, plan):
        ty, tz, cx, nth = self.ty, self.tz, self.cx, self.threads
        G, st = self.analyse(em, ty, tz)
        mm = plan.mm
        fields = em.fields
        nf = len(fields)
        tctype = {t: plan.temp_ctype(t) for t in plan.temps}
        L = []
        L += ["  __shared__ %s S%d[%d][%d];" % (ct, fi, st["DS"], st["SRY"] * st["SRZ"])
              for fi, (n, ct) in enumerate(fields)]
        for g in G:
            L += ["  __shared__ %s T%d[%d][%d];" % (tctype[t], t, g["DG"], g["RY"] * g["RZ"]) for t in g["temps"]]
        lds = "\n".join(L)
        # group evaluation: one loop over the concatenated regions of every group,
        # so that all work-items are busy whatever the size of each region
        gbodies = ["{"]
        starts, tot = [], 0
        for g in G:
            starts.append(tot)
            tot += g["RY"] * g["RZ"]
        gbodies.append("  for (int gidx = threadIdx.x; gidx < %d; gidx += %d) {" % (tot, nth))
        for gi, g in enumerate(G):
            precs = [em.prec_of(mm[t]) or {} for t in g["temps"]] if em.mixed else None
            node, roots, mprec = _merge_nodes([mm[t] for t in g["temps"]], precs)
            dxs = sorted(set(d[0] for d in g["D"]))

            def load(field, c, gi=gi):
                fi = em.field_index[field]
                return "S%d[sl_%d_%s][so + (%d)*%d + (%d)]" % (fi, gi, str(c[0]).replace("-", "m"), c[1], st["SRZ"], c[2])

            b = Body(node, load, lambda *a: None, lambda ax: None, lambda n, v: "", "g%d_" % gi, prec=mprec,
                     recip=em.contract, fastdiv=em.contract)
            kw = "if" if gi == 0 else "} else if"
            # planes of this group computed before iteration x0 - (DG - 1) are never read
            lines = ["%s (gidx < %d) {  // group %d: temporaries %s, read at x offsets %d..%d"
                     % (kw, starts[gi] + g["RY"] * g["RZ"], gi, g["temps"], g["oxmin"], g["oxmax"]),
                     "  if (x >= x0 - %d) {" % (g["DG"] - 1),
                     "    const int P = x + (%d);" % g["oxmax"],
                     "    const int gsl = (P - xs + 64 * %d) %% %d;" % (g["DG"], g["DG"])]
            lines += ["    const int sl_%d_%s = (P + (%d) - xs + 64 * DS) %% DS;" % (gi, str(d).replace("-", "m"), d) for d in dxs]
            lines.append("    const int idx = gidx - %d;" % starts[gi])
            lines.append("    const int r = idx / %d, c = idx - r * %d;" % (g["RZ"], g["RZ"]))
            lines.append("    const int so = (r + (%d)) * SRZ + (c + (%d));" % (g["ymin"] - st["SYmin"], g["zmin"] - st["SZmin"]))
            lines += ["    " + l for l in b.lines]
            lines += ["    T%d[gsl][idx] = (%s)g%d_%d;" % (t, tctype[t], gi, roots[i]) for i, t in enumerate(g["temps"])]
            lines.append("  }")
            gbodies += ["    " + l for l in lines]
        gbodies += ["    }", "  }", "}"]
        # the update
        gof = {t: (gi, g) for gi, g in enumerate(G) for t in g["temps"]}
        sbody = []
        xslots = set()
        for k in plan.stores:
            node = mm[k]
            base = plan.rng[k].offset - 2 * plan.sleeve

            def sload(field, c, base=base):
                fi = em.field_index[field]
                o = [c[a] + base for a in range(3)]
                xslots.add(("s", o[0]))
                return "S%d[ss_%s][(r + (%d)) * SRZ + (c + (%d))]" % (fi, str(o[0]).replace("-", "m"),
                                                                      o[1] - st["SYmin"], o[2] - st["SZmin"])

            def tload(omid, c, base=base):
                gi, g = gof[omid]
                o = [c[a] + base for a in range(3)]
                xslots.add(("t%d" % gi, o[0]))
                return "T%d[ts_%d_%s][(r + (%d)) * %d + (c + (%d))]" % (omid, gi, str(o[0]).replace("-", "m"),
                                                                     o[1] - g["ymin"], g["RZ"], o[2] - g["zmin"])

            def store(fname, v):
                return "FLD%d(st_out)[gout] = %s;" % (em.field_index[fname], v)

            b = Body(node, sload, tload, lambda ax: None, store, "u%d_" % k, prec=em.prec_of(node), recip=em.contract, fastdiv=em.contract)
            sbody += ["{"] + ["  " + l for l in b.lines] + ["}"]
        slot_lines = []
        for kind, o in sorted(xslots):
            if kind == "s":
                slot_lines.append("const int ss_%s = (x + (%d) - xs + 64 * DS) %% DS;" % (str(o).replace("-", "m"), o))
            else:
                gi = int(kind[1:])
                dg = G[gi]["DG"]
                slot_lines.append("const int ts_%d_%s = (x + (%d) - xs + 64 * %d) %% %d;"
                                  % (gi, str(o).replace("-", "m"), o, dg, dg))
        reach = 1 + max(st["W"] - st["Smin"], abs(st["Smax"]), -st["SYmin"], st["SYmax"], -st["SZmin"], st["SZmax"])
        wrap = [em.wrap_fn(a, reach) for a in range(3)]
        return ["""
__global__ void __launch_bounds__(%(NTH)d) k_step_staged(const char* __restrict__ st_in, char* __restrict__ st_out) {
  constexpr int TY = %(TY)d, TZ = %(TZ)d, CX = %(CX)d;
  constexpr int NTY = N2 / TY, NTZ = N3 / TZ, NT = NTY * NTZ * (N1 / CX);
  constexpr int DS = %(DS)d, SRY = %(SRY)d, SRZ = %(SRZ)d;
%(lds)s
  const int b = blockIdx.x;
  // consecutive workgroups go to different XCDs: neighbouring tiles share an L2
  const int tile = (NT %% 8 == 0) ? ((b %% 8) * (NT / 8) + b / 8) : b;
  const int zt = tile %% NTZ, yt = (tile / NTZ) %% NTY, xc = tile / (NTZ * NTY);
  const int x0 = xc * CX, y0 = yt * TY, z0 = zt * TZ;
  const int xs = x0 - %(W)d;  // warm-up: the rings fill before the first output plane
  for (int x = xs; x < x0 + CX; ++x) {
    __syncthreads();  // every read of the ring slots reused below is done
    {  // state plane x + Smax into its slot
      const int P = x + (%(SMAX)d);
      const int sl = (P - xs + 64 * DS) %% DS;
      const long long gp = (long long)%(w0)s(P) * N2;
      for (int idx = threadIdx.x; idx < SRY * SRZ; idx += %(NTH)d) {
        const int r = idx / SRZ, c = idx - r * SRZ;
        const long long g = (gp + %(w1)s(y0 + (%(SYMIN)d) + r)) * N3 + %(w2)s(z0 + (%(SZMIN)d) + c);
%(sload)s
      }
    }
    __syncthreads();
%(groups)s
    __syncthreads();
    if (x >= x0) {
%(slots)s
      for (int idx = threadIdx.x; idx < TY * TZ; idx += %(NTH)d) {
        const int r = idx / TZ, c = idx - r * TZ;
        const long long gout = ((long long)x * N2 + (y0 + r)) * N3 + (z0 + c);
%(sbody)s
      }
    }
  }
}
""" % dict(NTH=nth, TY=ty, TZ=tz, CX=cx, DS=st["DS"], SRY=st["SRY"], SRZ=st["SRZ"], lds=lds, W=st["W"],
           SMAX=st["Smax"], SYMIN=st["SYmin"], SZMIN=st["SZmin"], w0=wrap[0], w1=wrap[1], w2=wrap[2],
           sload="\n".join("        S%d[sl][idx] = FLD%d(st_in)[g];" % (fi, fi) for fi in range(nf)),
           groups="\n".join("    " + l for l in gbodies),
           slots="\n".join("      " + l for l in slot_lines),
           sbody="\n".join("        " + l for l in sbody))]

    def launch(self, em, plan):
        nb = (em.N[0] // self.cx) * (em.N[1] // self.ty) * (em.N[2] // self.tz)
        return ("static void run_step_staged(const char* in, char* out, hipStream_t s) {\n"
                "  hipLaunchKernelGGL(k_step_staged, dim3(%d), dim3(%d), 0, s, in, out);\n"
                "  FMR_CHECK(hipGetLastError());\n}\n" % (nb, self.threads))

    def advance(self, em):
        return ("static void fmr_advance(long long steps, const Formura_Navi &n0) {\n"
                "  (void)n0;\n"
                "  for (long long t = 0; t < steps; ++t) {\n"
                "    run_step_staged(g_rt.cur, g_rt.nxt, g_rt.stream); g_rt.swap();\n"
                "    g_rt.launches += 1;\n  }\n}")


def _staged3d_ok(em, ty, tz, cx, threads=256):
    plan = em.step
    if em.dim != 3 or not plan.temps or em.inlined:
        return False
    if em.N[0] % cx or em.N[1] % ty or em.N[2] % tz:
        return False
    try:
        G, st = Staged3DSchedule.analyse(em, ty, tz)
        lds = Staged3DSchedule.lds_bytes(em, ty, tz)
    except EmitError:
        return False
    return lds <= LDS_PER_WORKGROUP and st["W"] < min(em.N) and cx >= 1


class Stream3DSchedule:
    """3-D, no temporaries: stream along the slowest axis with K levels in LDS.

    A workgroup owns a TY x TZ column tile and a chunk of CX planes along x.
    Input planes enter an LDS ring of 2s+1 planes (with a low-side halo of
    2*K*s in y and z); level j = 1..K computes its plane t from planes
    t-2s..t of level j-1 as soon as plane t of level j-1 exists, so K steps
    are pipelined along x and only level K is written to HBM.  With the
    drifting frame every dependency points to lower indices, so the pipeline
    starts 2*K*s planes before the chunk.
    """

    name = "stream3d"

    def __init__(self, ty, tz, cx, k, threads=256, skew=False):
        self.ty, self.tz, self.cx, self.k = ty, tz, cx, k
        # work-items per workgroup
        self.threads = int(threads)
        # skew: level j computes plane t - (j - 1) in the trip where level 1
        # computes plane t, so that every level reads only planes finished in
        # earlier trips: one barrier per trip instead of K + 1, for rings of
        # 2s + 2 planes (one more)
        self.skew = bool(skew)

    def kernels(self, em, plan):
        s = plan.sleeve
        R = 2 * s + 1
        if self.skew:
            if getattr(em, "tb_index", False):
                raise EmitError("stream3d: skew does not take index reads")
            R = 2 * s + 2
        fields = em.fields
        params = ", ".join("const %s* __restrict__ lds%d" % (ct, i) for i, (n, ct) in enumerate(fields))
        outs = ", ".join("%s& o%d" % (ct, i) for i, (n, ct) in enumerate(fields))
        body_lines = []
        tbi = getattr(em, "tb_index", False)
        for k in plan.stores:
            node = plan.mm[k]
            base = plan.rng[k].offset - 2 * s

            def load(field, c, base=base):
                fi = em.field_index[field]
                dx, dy, dz = c[0] + base, c[1] + base, c[2] + base
                return "lds%d[pl[%d] + o + (%d)*RZ + (%d)]" % (fi, dx + 2 * s, dy, dz)

            def store(fname, v):
                return "o%d = %s;" % (em.field_index[fname], v)

            def index(ax):
                # unwrapped coordinate + 2s(nt - j) + the offset at the Forward's start
                return "(g%d + ib%d)" % (ax, ax)

            b = Body(node, load, lambda *a: None, index if tbi else (lambda ax: None), store, "s%d_" % k,
                     prec=em.prec_of(node), recip=em.contract, fastdiv=em.contract)
            body_lines.append("{")
            body_lines += ["  " + l for l in b.lines]
            body_lines.append("}")
        ipar = ", const int g0, const int g1, const int g2, const int ib0, const int ib1, const int ib2" if tbi else ""
        out = ["template <int RZ>\n__device__ __forceinline__ void pt_step3(%s, const int* pl, const int o%s, %s) {\n%s\n}\n"
               % (params, ipar, outs, "\n".join("  " + l for l in body_lines))]
        nf = len(fields)
        lds = "\n".join("  __shared__ %s B%d[K][R * PLANE];" % (ct, i) for i, (n, ct) in enumerate(fields))
        wrap = [em.wrap_fn(a, 2 * self.k * s + 1) for a in range(3)]
        # software pipeline: plane t+1 is loaded into registers (pf*) while plane t
        # is computed, and written into its ring slot at the top of the next trip
        pf_decl = " ".join("%s pf%d[PIT];" % (ct, i) for i, (n, ct) in enumerate(fields))
        pf_load = "\n".join("          pf%d[it] = FLD%d(st_in)[g];" % (i, i) for i in range(nf))
        pf_store = "\n".join("        B%d[0][slot * PLANE + idx] = pf%d[it];" % (i, i) for i in range(nf))
        srcs = ", ".join("B%d[j - 1]" % i for i in range(nf))
        decl_o = " ".join("%s o%d;" % (ct, i) for i, (n, ct) in enumerate(fields))
        refs_o = ", ".join("o%d" % i for i in range(nf))
        st_lds = " ".join("B%d[j][slot * PLANE + o] = o%d;" % (i, i) for i in range(nf))
        st_glb = " ".join("FLD%d(st_out)[g] = o%d;" % (i, i) for i in range(nf))
        if self.skew:
            st_lds_k = " ".join("B%d[j][ws * PLANE + o] = o%d;" % (i, i) for i in range(nf))
            out.append(STREAM3D_SKEW % dict(TY=self.ty, TZ=self.tz, CX=self.cx, R=R, NTH=self.threads, lds=lds,
                                            w0=wrap[0], w1=wrap[1], w2=wrap[2], pf_decl=pf_decl, pf_load=pf_load,
                                            pf_store=pf_store, srcs=srcs, decl_o=decl_o, refs_o=refs_o,
                                            st_lds=st_lds_k, st_glb=st_glb))
            return out
        out.append("""
template <int K>
__global__ void __launch_bounds__(%(NTH)d) k_step_stream(const char* __restrict__ st_in, char* __restrict__ st_out%(offp)s) {
  constexpr int TY = %(TY)d, TZ = %(TZ)d, CX = %(CX)d, R = %(R)d;
  constexpr int H = 2 * K * SLEEVE, RY = TY + H, RZ = TZ + H, PLANE = RY * RZ;
  constexpr int NTY = N2 / TY, NTZ = N3 / TZ, NCX = N1 / CX, NT = NTY * NTZ * NCX;
  // levels 0..K-1 live in LDS (level K goes to HBM)
%(lds)s
  const int b = blockIdx.x;
  const int tile = (NT %% 8 == 0) ? ((b %% 8) * (NT / 8) + b / 8) : b;
  const int cz = tile %% NTZ, cy = (tile / NTZ) %% NTY, cxi = tile / (NTZ * NTY);
  const int x0 = cxi * CX, y0 = cy * TY, z0 = cz * TZ;
  constexpr int PIT = (PLANE + %(NTH)d - 1) / %(NTH)d;
  %(pf_decl)s
  auto prefetch = [&](int t) {
    const long long gx = (long long)%(w0)s(t) * N2;
#pragma unroll
    for (int it = 0; it < PIT; ++it) {
      const int idx = threadIdx.x + it * %(NTH)d;
      if (idx < PLANE) {
        const int ly = idx / RZ, lz = idx - ly * RZ;
        const long long g = (gx + %(w1)s(y0 - H + ly)) * N3 + %(w2)s(z0 - H + lz);
%(pf_load)s
      }
    }
  };
  prefetch(x0 - H);
  // ring slot of plane t, carried along (no division in the plane loop)
  int rs = (((x0 - H) %% R) + R) %% R;
  for (int t = x0 - H; t < x0 + CX; ++t) {
    const int slot = rs;
    // every read of this ring slot's previous plane ended before the last barrier
#pragma unroll
    for (int it = 0; it < PIT; ++it) {
      const int idx = threadIdx.x + it * %(NTH)d;
      if (idx < PLANE) {
%(pf_store)s
      }
    }
    __syncthreads();
    if (t + 1 < x0 + CX) prefetch(t + 1);
#pragma unroll
    for (int j = 1; j <= K; ++j) {
      if (t >= x0 - 2 * (K - j) * SLEEVE) {
        int pl[R];
#pragma unroll
        for (int d = 0; d < R; ++d) {  // slot of plane t - 2s + d
          const int q = rs + d - 2 * SLEEVE;
          pl[d] = (q < 0 ? q + R : q) * PLANE;
        }
        const int lo = 2 * j * SLEEVE, W = RZ - lo, HH = RY - lo;%(ibdecl)s
        for (int idx = threadIdx.x; idx < HH * W; idx += %(NTH)d) {
          const int rr = idx / W;
          const int ly = lo + rr, lz = lo + (idx - rr * W);
          const int o = ly * RZ + lz;
          %(decl_o)s
          pt_step3<RZ>(%(srcs)s, pl, o%(ipass)s, %(refs_o)s);
          if (j < K) {
            %(st_lds)s
          } else {
            const long long g = ((long long)t * N2 + (y0 + ly - H)) * N3 + (z0 + lz - H);
            %(st_glb)s
          }
        }
      }
      __syncthreads();
    }
    rs = rs + 1 == R ? 0 : rs + 1;
  }
}
""" % dict(TY=self.ty, TZ=self.tz, CX=self.cx, R=R, NTH=self.threads, lds=lds, w0=wrap[0], w1=wrap[1], w2=wrap[2],
           pf_decl=pf_decl, pf_load=pf_load, pf_store=pf_store, srcs=srcs, decl_o=decl_o, refs_o=refs_o, st_lds=st_lds, st_glb=st_glb,
           offp=", const int offT0, const int offT1, const int offT2" if tbi else "",
           ibdecl=("\n        // index reads: level j of the Forward at unwrapped coordinate g reads\n"
                   "        // g + 2s(K - j) + the offset at the Forward's start\n"
                   "        const int ib0 = 2 * SLEEVE * (K - j) + offT0, ib1 = 2 * SLEEVE * (K - j) + offT1, "
                   "ib2 = 2 * SLEEVE * (K - j) + offT2;") if tbi else "",
           ipass=", t, y0 + ly - H, z0 + lz - H, ib0, ib1, ib2" if tbi else ""))
        return out

    def launch(self, em, plan):
        nt = (em.N[0] // self.cx) * (em.N[1] // self.ty) * (em.N[2] // self.tz)
        if getattr(em, "tb_index", False):
            return ("static void run_step_stream(int k, const char* in, char* out, hipStream_t s, int o0, int o1, int o2) {\n"
                    "  if (k != %d) fmr_rt::fatal(\"bad step count\");\n"
                    "  hipLaunchKernelGGL(k_step_stream<%d>, dim3(%d), dim3(%d), 0, s, in, out, o0, o1, o2);\n"
                    "  FMR_CHECK(hipGetLastError());\n}\n" % (self.k, self.k, nt, self.threads))
        return ("static void run_step_stream(int k, const char* in, char* out, hipStream_t s) {\n"
                "  switch (k) {\n%s  default: fmr_rt::fatal(\"bad step count\");\n  }\n"
                "  FMR_CHECK(hipGetLastError());\n}\n"
                % "".join("  case %d: hipLaunchKernelGGL(k_step_stream<%d>, dim3(%d), dim3(%d), 0, s, in, out); break;\n"
                          % (k, k, nt, self.threads) for k in range(1, self.k + 1)))

    def advance(self, em):
        if getattr(em, "tb_index", False):
            offs = _offsets_at(em.om.axes, "n0", "done")
            return ("static void fmr_advance(long long steps, const Formura_Navi &n0) {\n"
                    "  if (steps %% %d) fmr_rt::fatal(\"a temporally blocked step with index reads advances whole Forwards\");\n"
                    "  for (long long done = 0; done < steps; done += %d) {\n"
                    "    run_step_stream(%d, g_rt.cur, g_rt.nxt, g_rt.stream, %s); g_rt.swap();\n"
                    "    g_rt.launches += 1;\n  }\n}" % (self.k, self.k, self.k, ", ".join(offs)))
        return ("static void fmr_advance(long long steps, const Formura_Navi &n0) {\n"
                "  (void)n0;\n"
                "  while (steps > 0) {\n"
                "    const int k = steps >= %d ? %d : (int)steps;\n"
                "    run_step_stream(k, g_rt.cur, g_rt.nxt, g_rt.stream); g_rt.swap();\n"
                "    g_rt.launches += 1; steps -= k;\n  }\n}" % (self.k, self.k))


# stream3d with skewed levels (spec key skew): in trip t level j computes plane
# p = t - (j - 1) from planes p - 2s .. p of level j - 1, all finished in earlier
# trips; the input plane t + 1 enters its slot at the top of the trip.  Rings of
# R = 2s + 2 planes: a level's plane is written while the next level reads the
# 2s + 1 planes before it.  One barrier per trip.
STREAM3D_SKEW = """
template <int K>
__global__ void __launch_bounds__(%(NTH)d) k_step_stream(const char* __restrict__ st_in, char* __restrict__ st_out) {
  constexpr int TY = %(TY)d, TZ = %(TZ)d, CX = %(CX)d, R = %(R)d;
  constexpr int H = 2 * K * SLEEVE, RY = TY + H, RZ = TZ + H, PLANE = RY * RZ;
  constexpr int NTY = N2 / TY, NTZ = N3 / TZ, NCX = N1 / CX, NT = NTY * NTZ * NCX;
  // levels 0..K-1 live in LDS rings of R planes (level K goes to HBM)
%(lds)s
  const int b = blockIdx.x;
  const int tile = (NT %% 8 == 0) ? ((b %% 8) * (NT / 8) + b / 8) : b;
  const int cz = tile %% NTZ, cy = (tile / NTZ) %% NTY, cxi = tile / (NTZ * NTY);
  const int x0 = cxi * CX, y0 = cy * TY, z0 = cz * TZ;
  const int base = x0 - H;  // first plane of the pipeline: plane p lives in slot (p - base) %% R
  constexpr int PIT = (PLANE + %(NTH)d - 1) / %(NTH)d;
  %(pf_decl)s
  auto prefetch = [&](int t) {
    const long long gx = (long long)%(w0)s(t) * N2;
#pragma unroll
    for (int it = 0; it < PIT; ++it) {
      const int idx = threadIdx.x + it * %(NTH)d;
      if (idx < PLANE) {
        const int ly = idx / RZ, lz = idx - ly * RZ;
        const long long g = (gx + %(w1)s(y0 - H + ly)) * N3 + %(w2)s(z0 - H + lz);
%(pf_load)s
      }
    }
  };
  auto put0 = [&](int slot) {  // the prefetched input plane into its slot
#pragma unroll
    for (int it = 0; it < PIT; ++it) {
      const int idx = threadIdx.x + it * %(NTH)d;
      if (idx < PLANE) {
%(pf_store)s
      }
    }
  };
  prefetch(base);
  put0(0);
  prefetch(base + 1);
  __syncthreads();
  int rs = 0;  // (t - base) %% R
  for (int t = base; t < x0 + CX + K - 1; ++t) {
    if (t + 1 < x0 + CX) {
      // input plane t + 1 into the slot of plane t + 1 - R, whose last read (by
      // level 1, in the trip before) ended at the last barrier; level 1 reads
      // planes t - 2s .. t in this trip.  The next plane's loads are issued
      // before this trip's work so that they overlap it.
      put0(rs + 1 == R ? 0 : rs + 1);
      if (t + 2 < x0 + CX) prefetch(t + 2);
    }
#pragma unroll
    for (int j = 1; j <= K; ++j) {
      const int p = t - (j - 1);  // the plane level j computes in this trip
      if (p >= x0 - 2 * (K - j) * SLEEVE && p < x0 + CX) {
        int pl[2 * SLEEVE + 1];
#pragma unroll
        for (int d = 0; d <= 2 * SLEEVE; ++d) {  // slot of plane p - 2s + d
          const int q = rs + ((d - 2 * SLEEVE - (j - 1)) %% R + R) %% R;
          pl[d] = (q >= R ? q - R : q) * PLANE;
        }
        const int wq = rs + ((-(j - 1)) %% R + R) %% R, ws = wq >= R ? wq - R : wq;  // slot of plane p
        const int lo = 2 * j * SLEEVE, W = RZ - lo, HH = RY - lo;
        for (int idx = threadIdx.x; idx < HH * W; idx += %(NTH)d) {
          const int rr = idx / W;
          const int ly = lo + rr, lz = lo + (idx - rr * W);
          const int o = ly * RZ + lz;
          %(decl_o)s
          pt_step3<RZ>(%(srcs)s, pl, o, %(refs_o)s);
          if (j < K) {
            %(st_lds)s
          } else {
            const long long g = ((long long)p * N2 + (y0 + ly - H)) * N3 + (z0 + lz - H);
            %(st_glb)s
          }
        }
      }
    }
    __syncthreads();
    rs = rs + 1 == R ? 0 : rs + 1;
  }
}
"""


def _stream3d_ok(em, ty, tz, cx, k, skew=False):
    plan = em.step
    if em.dim != 3 or plan.temps:
        return False
    if any(m.inst[0] == "LoadIndex" for kk in plan.stores for m in plan.mm[kk].micro):
        # only the temporally blocked form, one Forward per launch (see LibEmitter)
        if not getattr(em, "tb_index", False) or k != em.cfg.steps_per_forward:
            return False
    if em.N[0] % cx or em.N[1] % ty or em.N[2] % tz or cx < 1:
        return False
    s = plan.sleeve
    if 2 * k * s >= min(em.N):
        return False
    if skew and getattr(em, "tb_index", False):
        return False
    plane = (ty + 2 * k * s) * (tz + 2 * k * s)
    lds = sum(8 if ct == "double" else 4 for _, ct in em.fields) * k * (2 * s + (2 if skew else 1)) * plane
    return lds <= LDS_PER_WORKGROUP


def _wave2d_ok(em, rows, k):
    plan = em.step
    if em.dim != 2 or plan.temps:
        return False
    if any(m.inst[0] == "LoadIndex" for kk in plan.stores for m in plan.mm[kk].micro):
        return False
    h = 2 * k * plan.sleeve
    return h < 48 and em.N[1] >= 64 and em.N[0] % rows == 0 and em.N[0] > h and em.N[1] > h


def _tile2d_ok(em, tx, ty, k, inplace=False):
    plan = em.step
    if em.dim != 2 or plan.temps:
        return False
    if any(m.inst[0] == "LoadIndex" for kk in plan.stores for m in plan.mm[kk].micro):
        # only the temporally blocked form, one Forward per launch (see LibEmitter)
        if not getattr(em, "tb_index", False) or k != em.cfg.steps_per_forward or inplace:
            return False
    if em.N[0] % tx or em.N[1] % ty:
        return False
    return _tile2d_lds(em, tx, ty, k, inplace) <= LDS_PER_WORKGROUP


# gfx950: 160 KiB of LDS per CU, all of it available to one workgroup
LDS_PER_WORKGROUP = 160 * 1024


def _tile2d_lds(em, tx, ty, k, inplace=False):
    s = em.step.sleeve
    return (sum(8 if ct == "double" else 4 for _, ct in em.fields) * (1 if inplace else 2)
            * (tx + 2 * k * s) * (ty + 2 * k * s))


def schedule_from_spec(em, spec: dict):
    """Build a schedule from an explicit spec ({"kind": ..., tile sizes, "k": ...})."""
    kind = spec.get("kind")
    if kind == "flat":
        return FlatSchedule()
    if kind == "tile2d":
        args = (int(spec["tx"]), int(spec["ty"]), int(spec["k"]))
        if not _tile2d_ok(em, *args, inplace=bool(spec.get("inplace", 0))):
            raise EmitError("tile2d %r does not apply to %s" % (spec, em.name))
        if spec.get("pair") and (int(spec["ty"]) % 2 or spec.get("inplace") or getattr(em, "tb_index", False)
                                 or (int(spec["ty"]) + 2 * (int(spec["k"]) - 1) * em.step.sleeve) // 2
                                 > int(spec.get("threads", 256))):
            raise EmitError("tile2d pairs need an even tile width, a row's pairs within the workgroup, "
                            "ping-pong buffers and no index reads")
        if spec.get("vpair") and (int(spec["tx"]) % 2 or spec.get("inplace") or spec.get("colmap") or spec.get("pair")
                                  or getattr(em, "tb_index", False)):
            raise EmitError("row pairs need an even tile height, ping-pong buffers and no index reads")
        if spec.get("colmap") and (int(spec["ty"]) + 2 * (int(spec["k"]) - 1) * em.step.sleeve
                                   > int(spec.get("threads", 256))):
            raise EmitError("the column mapping needs a level's width within the workgroup")
        return Tile2DSchedule(*args, inplace=bool(spec.get("inplace", 0)), threads=int(spec.get("threads", 256)),
                              pair=bool(spec.get("pair", 0)), div=spec.get("div", "fix"),
                              colmap=bool(spec.get("colmap", 0)), vpair=bool(spec.get("vpair", 0)))
    if kind == "wave2d":
        args = (int(spec["rows"]), int(spec["k"]))
        if not _wave2d_ok(em, *args):
            raise EmitError("wave2d %r does not apply to %s" % (spec, em.name))
        return Wave2DSchedule(*args, unroll=int(spec.get("unroll", 1)))
    if kind == "staged3d":
        args = (int(spec["ty"]), int(spec["tz"]), int(spec["cx"]))
        if not _staged3d_ok(em, *args):
            raise EmitError("staged3d %r does not apply to %s" % (spec, em.name))
        return Staged3DSchedule(*args, threads=int(spec.get("threads", 256)))
    if kind == "stream3d":
        args = (int(spec["ty"]), int(spec["tz"]), int(spec["cx"]), int(spec["k"]))
        skew = bool(spec.get("skew", 0))
        if not _stream3d_ok(em, *args, skew=skew):
            raise EmitError("stream3d %r does not apply to %s" % (spec, em.name))
        return Stream3DSchedule(*args, threads=int(spec.get("threads", 256)), skew=skew)
    raise EmitError("unknown schedule kind %r" % (kind,))


def candidate_specs(em) -> List[dict]:
    """Every schedule the autotuner tries for this program (valid ones only)."""
    out = [{"kind": "flat"}]
    if em.mm.step and any(n.micro[-1].inst[0] != "Store" for n in em.mm.step.values()):
        # programs with manifest temporaries: also try recomputing them where read
        from .manifest import cheap_temporaries, inline_temporaries
        out.append({"kind": "flat", "inline": 1})
        if cheap_temporaries(em.mm.step):
            out.append({"kind": "flat", "inline": "cheap"})
        if em.dim == 3:
            bpc = sum(8 if ct == "double" else 4 for _, ct in em.fields)
            s = em.cfg.sleeve
            for ty, tz, cx in ((8, 32, 16), (8, 32, 32), (4, 64, 16), (8, 16, 32), (16, 32, 16), (4, 32, 32),
                               (4, 16, 32), (8, 8, 32), (16, 32, 32), (8, 64, 32), (16, 16, 32), (32, 32, 32)):
                if (em.N[1] % ty == 0 and em.N[2] % tz == 0 and em.N[0] % cx == 0
                        and bpc * (2 * s + 1) * (ty + 2 * s) * (tz + 2 * s) <= LDS_PER_WORKGROUP
                        and (em.N[0] // cx) * (em.N[1] // ty) * (em.N[2] // tz) >= 128):
                    out.append({"kind": "stream3d", "ty": ty, "tz": tz, "cx": cx, "k": 1, "inline": 1})
                    if ty * tz >= 512:
                        out.append({"kind": "stream3d", "ty": ty, "tz": tz, "cx": cx, "k": 1, "inline": 1,
                                    "threads": 512})
            # ... or evaluating each of them once per point into LDS rings
            # (temporaries that only other temporaries read, such as the wave
            # speeds a flux reads, recomputed inside their readers)
            from .manifest import inner_temporaries
            extra = {"inline": "inner"} if inner_temporaries(em.mm.step) else {}
            em_s = LibEmitter(em.mm, em.name, dict({"kind": "flat"}, **extra)) if extra else em
            for ty, tz in ((8, 16), (16, 16), (8, 32), (16, 32), (4, 64), (8, 64)):
                for cx in (8, 16, 32):
                    if (_staged3d_ok(em_s, ty, tz, cx)
                            and (em.N[0] // cx) * (em.N[1] // ty) * (em.N[2] // tz) >= 256):
                        for nth in (256, 512):
                            if nth <= 2 * ty * tz:
                                spec = dict({"kind": "staged3d", "ty": ty, "tz": tz, "cx": cx}, **extra)
                                if nth != 256:
                                    spec["threads"] = nth
                                out.append(spec)
        return out
    if em.dim == 2:
        for tx, ty in ((16, 32), (32, 32), (16, 64), (32, 64), (64, 32), (64, 64), (8, 128), (16, 128)):
            for k in (1, 2, 3, 4, 6, 8):
                for nth in (256, 512, 1024):
                    if nth > tx * ty:
                        continue
                    for inplace in ((0,) if k == 1 else (0, 1)):
                        if _tile2d_ok(em, tx, ty, k, inplace):
                            spec = {"kind": "tile2d", "tx": tx, "ty": ty, "k": k}
                            if inplace:
                                spec["inplace"] = 1
                            if nth != 256:
                                spec["threads"] = nth
                            out.append(spec)
                            if not inplace and ty + 2 * (k - 1) * em.step.sleeve <= nth:
                                out.append(dict(spec, colmap=1))
                            if not inplace and tx % 2 == 0 and not getattr(em, "tb_index", False):
                                out.append(dict(spec, vpair=1))
                            if not inplace and ty % 2 == 0 and not getattr(em, "tb_index", False):
                                out.append(dict(spec, pair=1))
        for rows in (8, 16, 32, 64):
            for k in (1, 2, 3, 4):
                if _wave2d_ok(em, rows, k):
                    out.append({"kind": "wave2d", "rows": rows, "k": k})
                    out.append({"kind": "wave2d", "rows": rows, "k": k, "unroll": 2})
    if em.dim == 3:
        for ty, tz in ((16, 64), (8, 64), (32, 32), (16, 32), (8, 32), (4, 64), (4, 32), (8, 16), (4, 16)):
            for cx in (8, 16, 32, 64):
                for k in (1, 2, 3, 4):
                    nwg = (em.N[0] // cx) * (em.N[1] // ty) * (em.N[2] // tz) if _stream3d_ok(em, ty, tz, cx, k) else 0
                    if nwg >= 128 and (nwg <= 4096 or cx >= 32):
                        out.append({"kind": "stream3d", "ty": ty, "tz": tz, "cx": cx, "k": k})
                        if ty * tz >= 512 and k in (1, 2, 3) and cx >= 32:
                            out.append({"kind": "stream3d", "ty": ty, "tz": tz, "cx": cx, "k": k, "threads": 512})
        # the skewed pipeline (one barrier per plane) on large cross-sections
        # and long columns, where its extra ring plane still fits
        for ty, tz in ((32, 64), (64, 32), (16, 64), (32, 32), (16, 32)):
            for cx in (32, 64, 128, 256):
                for k in (1, 2, 3):
                    if not _stream3d_ok(em, ty, tz, cx, k, skew=True):
                        continue
                    if (em.N[0] // cx) * (em.N[1] // ty) * (em.N[2] // tz) >= 128:
                        out.append({"kind": "stream3d", "ty": ty, "tz": tz, "cx": cx, "k": k, "skew": 1,
                                    "threads": 1024 if ty * tz >= 1024 else 512})
    return out


def _tuned_spec(name: str):
    import os
    p = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning", "%s.json" % name)
    if os.path.exists(p):
        with open(p) as f:
            return json.load(f).get("best")
    return None


def choose_schedule(em, want):
    """``want``: a spec dict, "flat"/"tile2d"/"stream3d", or "auto" (tuning file, then defaults)."""
    if isinstance(want, dict):
        return schedule_from_spec(em, want)
    if want == "flat":
        return FlatSchedule()
    spec = em.cfg.extra.get("hip_schedule")
    if spec is None and want == "auto":
        spec = _tuned_spec(em.name)
    if spec is not None:
        try:
            return schedule_from_spec(em, spec)
        except EmitError:
            if want != "auto":
                raise
    if want in ("auto", "tile2d") and em.dim == 2:
        if getattr(em, "tb_index", False):
            nt = em.cfg.steps_per_forward
            for tx, ty in ((32, 32), (16, 32), (16, 16), (8, 8)):
                if _tile2d_ok(em, tx, ty, nt):
                    return Tile2DSchedule(tx, ty, nt)
        for tx, ty, k in ((32, 32, 4), (16, 16, 4), (32, 32, 1), (16, 16, 1), (8, 8, 1)):
            if _tile2d_ok(em, tx, ty, k):
                return Tile2DSchedule(tx, ty, k)
    if want in ("auto", "stream3d") and em.dim == 3 and getattr(em, "tb_index", False):
        nt = em.cfg.steps_per_forward
        for ty, tz, cx in ((8, 32, 16), (8, 16, 8), (4, 8, 8), (4, 4, 4)):
            if _stream3d_ok(em, ty, tz, cx, nt):
                return Stream3DSchedule(ty, tz, cx, nt)
    if want == "stream3d" and em.dim == 3:
        for c in ((16, 64, 32, 1), (8, 32, 16, 1), (4, 8, 8, 1)):
            if _stream3d_ok(em, *c):
                return Stream3DSchedule(*c)
    if want in ("tile2d", "stream3d"):
        raise EmitError("the %s schedule does not apply to this program" % want)
    return FlatSchedule()


# --- entry points --------------------------------------------------------------------

def generate(fmr_path: str, yaml_path: str, name: str, schedule: str = "auto"):
    """Front-end + emitters: returns (header text, hip source text, mm program)."""
    from .config import read_config
    from .parser import parse_program
    from .translate import gen_om_program
    from .manifest import gen_mm_program
    nc = read_config(yaml_path)
    with open(fmr_path) as f:
        prog = parse_program(f.read(), nc, fmr_path)
    om = gen_om_program(prog)
    mm = gen_mm_program(om)
    em = LibEmitter(mm, name, schedule)
    return emit_header(om), em.emit(), mm
