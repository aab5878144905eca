"""ctypes view of the oracle (oracle/liboracle.so) — TEST INFRASTRUCTURE.

Only tests/, the smoke check and the benchmark's CPU-baseline leg import this module; the
product package never does. See fxo.h for what each function restates.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


class Model(C.Structure):
    _fields_ = [("nnode", C.c_int32), ("xyz", C.c_void_p), ("nelem", C.c_int32), ("conn", C.c_void_p),
                ("group", C.c_void_p), ("region", C.c_void_p), ("nfix", C.c_int32), ("fix_node", C.c_void_p),
                ("fix_dof", C.c_void_p), ("ntie", C.c_int32), ("tie_retained", C.c_void_p),
                ("tie_constrained", C.c_void_p), ("tie_dof", C.c_void_p), ("tie_command", C.c_void_p)]


class Compiled(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("nnode", "nelem", "padded", "neq", "nparam", "ncolor")] + [
        ("node_order", C.POINTER(C.c_int32)), ("node_rank", C.POINTER(C.c_int32)),
        ("elem_order", C.POINTER(C.c_int32)), ("node_eq", C.POINTER(C.c_int32)),
        ("coords", C.POINTER(C.c_double)), ("node", C.POINTER(C.c_int32)), ("eq", C.POINTER(C.c_int32)),
        ("param", C.POINTER(C.c_int32)), ("param_group", C.POINTER(C.c_int32)),
        ("param_region", C.POINTER(C.c_int32)), ("color_start", C.POINTER(C.c_int32)),
        ("color_elements", C.POINTER(C.c_int32)), ("gather_start", C.POINTER(C.c_int64)),
        ("gather_list", C.POINTER(C.c_int64)), ("slots", C.POINTER(C.c_double))]


class Rayleigh(C.Structure):
    _fields_ = [("alphaM", C.c_double), ("betaK", C.c_double), ("betaKinit", C.c_double),
                ("betaKcomm", C.c_double)]


class Elastic(C.Structure):
    _fields_ = [("E", C.c_double), ("nu", C.c_double), ("rho", C.c_double), ("b", C.c_double * 3),
                ("mass_type", C.c_int32), ("damp", Rayleigh), ("element", C.c_int32)]


class Coeffs(C.Structure):
    _fields_ = [("stiffness", C.c_double), ("damping", C.c_double), ("mass", C.c_double), ("initial", C.c_int32)]


class PimyParams(C.Structure):
    _fields_ = [("load_stage", C.c_int32), ("nsurf", C.c_int32)] + [
        (n, C.c_double) for n in ("rho", "G", "K", "phi", "gamma_peak", "pref", "cohesion", "pdc", "residual_press")
    ] + [("size", C.c_double * 21), ("modulus", C.c_double * 21)]


class PimyElem(C.Structure):
    _fields_ = [("mat", PimyParams), ("b", C.c_double * 3), ("mass_type", C.c_int32), ("damp", Rayleigh),
                ("element", C.c_int32)]


PIMY_NSTATE = 190


def pimy_elems(materials, stage=0):
    """materials: dicts with rho G K cohesion gamma_peak phi pref pdc [nsurf gredu b mass_type rayleigh
    element ("stdBrick" or "SSPbrick": one material point, the element's point 0)]
    -> ctypes array of fxo_pimy_elem (setUpSurfaces by the oracle)."""
    arr = (PimyElem * max(1, len(materials)))()
    for i, m in enumerate(materials):
        inp = np.array([m["rho"], m["G"], m["K"], m["cohesion"], m["gamma_peak"], m["phi"], m["pref"], m["pdc"]])
        g = m.get("gredu")
        g = None if g is None else np.ascontiguousarray(g, np.float64)
        rc = lib().fxo_pimy_setup(inp.ctypes.data, int(m.get("nsurf", 20)), None if g is None else g.ctypes.data,
                                  C.byref(arr[i].mat))
        if rc != 0:
            raise ValueError(f"fxo_pimy_setup: {rc}")
        arr[i].mat.load_stage = stage
        arr[i].b = (C.c_double * 3)(*m.get("b", (0.0, 0.0, 0.0)))
        arr[i].mass_type = int(m.get("mass_type", 0))
        arr[i].damp = Rayleigh(*m.get("rayleigh", (0.0, 0.0, 0.0, 0.0)))
        arr[i].element = 1 if m.get("element", "stdBrick") == "SSPbrick" else 0
    return arr


class System(C.Structure):
    _fields_ = [("c", C.c_void_p), ("elastic", C.c_void_p), ("pimy", C.c_void_p), ("C", C.c_void_p), ("T", C.c_void_p),
                ("D", C.c_void_p), ("nthreads", C.c_int)]


class Series(C.Structure):
    _fields_ = [("values", C.c_void_p), ("count", C.c_int32), ("dt", C.c_double), ("factor", C.c_double),
                ("start", C.c_double), ("use_last", C.c_int32), ("time", C.c_void_p), ("loc", C.c_void_p)]


def series_value(values, dt, t, factor=1.0, start=0.0, use_last=False, fact=1.0):
    """PathSeries value at time t (uniform step), times the pattern factor."""
    sv = np.ascontiguousarray(values, np.float64)
    s = Series(sv.ctypes.data, len(sv), dt, factor, start, int(bool(use_last)), None, None)
    return lib().fxo_series_value(C.cast(C.pointer(s), C.c_void_p), fact, t)


class TimeArraySeries:
    """PathTimeSeries (values at given times) with its interval walk kept between calls, as
    GroundMotions keeps it per pattern (ground_motion.cpp:43-74, :133-139)."""

    def __init__(self, values, time, factor=1.0):
        self.values = np.ascontiguousarray(values, np.float64)
        self.time = np.ascontiguousarray(time, np.float64)
        assert len(self.values) == len(self.time) >= 1
        self.factor = factor
        self.loc = C.c_int32(0)

    def struct(self):
        return Series(self.values.ctypes.data, len(self.values), 0.0, self.factor, 0.0, 0,
                      self.time.ctypes.data, C.cast(C.pointer(self.loc), C.c_void_p))

    def __call__(self, t, fact=1.0):
        s = self.struct()
        return lib().fxo_series_value(C.cast(C.pointer(s), C.c_void_p), fact, t)


class NewmarkOpts(C.Structure):
    _fields_ = [("gamma", C.c_double), ("beta", C.c_double), ("dt", C.c_double), ("nsteps", C.c_int32),
                ("test", C.c_int32), ("tol", C.c_double), ("max_iter", C.c_int32), ("cg_rtol", C.c_double),
                ("cg_max_it", C.c_int32), ("lanes", C.c_int32), ("ground_accel", C.c_void_p),
                ("max_retries", C.c_int32), ("num_sublevels", C.c_int32), ("num_substeps", C.c_int32),
                ("ground_series", C.c_void_p), ("ground_factor", C.c_double), ("time", C.c_double),
                ("analysis", C.c_int32), ("algorithm", C.c_int32), ("modified_initial", C.c_int32),
                ("factor_once", C.c_int32), ("load_increment", C.c_double), ("load_min_increment", C.c_double),
                ("load_max_increment", C.c_double), ("load_num_iter", C.c_int32)]


class NewmarkStats(C.Structure):
    _fields_ = [("steps", C.c_int32), ("newton_iterations", C.c_int32), ("max_newton_iterations", C.c_int32),
                ("krylov_iterations", C.c_int32), ("last_test", C.c_double), ("commits", C.c_int32),
                ("failed_attempts", C.c_int32), ("retries", C.c_int32), ("substepped", C.c_int32),
                ("time", C.c_double)]


class SolveResult(C.Structure):
    _fields_ = [("converged", C.c_int32), ("iterations", C.c_int32), ("indefinite", C.c_int32),
                ("rhs_norm", C.c_double), ("initial_residual", C.c_double), ("final_residual", C.c_double)]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run __graft_entry__.build())")
        L = C.CDLL(path)
        L.fxo_lcg_symmetric.restype = C.c_double
        L.fxo_lcg_symmetric.argtypes = [C.POINTER(C.c_uint64)]
        L.fxo_compile.argtypes = [C.POINTER(Model), C.c_int, C.c_int, C.c_int, C.POINTER(Compiled)]
        L.fxo_compiled_free.argtypes = [C.POINTER(Compiled)]
        L.fxo_node_bandwidth.restype = C.c_int32
        L.fxo_node_bandwidth.argtypes = [C.POINTER(Model), i32p]
        L.fxo_brick_shape.argtypes = [f64p, C.c_int, f64p, C.POINTER(C.c_double)]
        L.fxo_elastic_tangent.argtypes = [C.c_double, C.c_double, f64p]
        L.fxo_brick_force_elastic.argtypes = [C.POINTER(Elastic), f64p, f64p, f64p, f64p, C.c_void_p]
        L.fxo_brick_mass.argtypes = [C.POINTER(Elastic), f64p, f64p]
        L.fxo_brick_lumped_mass.argtypes = [C.c_double, f64p, f64p]
        L.fxo_brick_operator_apply.argtypes = [C.POINTER(Elastic), C.POINTER(Coeffs), f64p, f64p, f64p, f64p]
        L.fxo_brick_operator_diagonal.argtypes = [C.POINTER(Elastic), C.POINTER(Coeffs), f64p, f64p, i32p, f64p]
        L.fxo_apply.argtypes = [C.POINTER(Compiled), C.c_void_p, C.POINTER(Coeffs), f64p, f64p, C.c_int]
        L.fxo_diagonal.argtypes = [C.POINTER(Compiled), C.c_void_p, C.POINTER(Coeffs), f64p, C.c_int]
        L.fxo_lumped_mass.argtypes = [C.POINTER(Compiled), C.c_void_p, f64p, f64p]
        L.fxo_body_load.argtypes = [C.POINTER(Compiled), C.c_void_p, f64p]
        L.fxo_dot.restype = C.c_double
        L.fxo_dot.argtypes = [C.c_int64, f64p, f64p, C.c_int]
        L.fxo_series_value.restype = C.c_double
        L.fxo_series_value.argtypes = [C.c_void_p, C.c_double, C.c_double]
        L.fxo_force.argtypes = [C.POINTER(Compiled), C.c_void_p, f64p, C.c_void_p, f64p, C.c_int]
        L.fxo_brick_max_omega_sq.restype = C.c_double
        L.fxo_brick_max_omega_sq.argtypes = [C.POINTER(Elastic), f64p]
        L.fxo_stable_dt.restype = C.c_double
        L.fxo_stable_dt.argtypes = [C.c_int, C.POINTER(Rayleigh), C.c_double]
        L.fxo_critical_dt.restype = C.c_double
        L.fxo_critical_dt.argtypes = [C.POINTER(Compiled), C.c_void_p, C.c_int]
        L.fxo_cd_run.argtypes = [C.POINTER(Compiled), C.c_void_p, f64p, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p,
                                 C.c_void_p, C.c_double, C.c_int32, f64p, f64p, f64p, f64p, C.c_int]
        L.fxo_explicit_run.argtypes = [C.POINTER(Compiled), C.c_void_p, f64p, C.c_void_p, C.c_void_p, C.c_int32,
                                       C.c_void_p, C.c_void_p, C.c_double, C.c_int32, C.c_int, C.c_int, C.c_void_p,
                                       f64p, f64p, f64p, f64p, C.c_int]
        L.fxo_pimy_setup.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.POINTER(PimyParams)]
        L.fxo_pimy_nstate.restype = C.c_int
        L.fxo_pimy_initial_state.argtypes = [C.POINTER(PimyParams), C.c_void_p]
        L.fxo_pimy_force.argtypes = [C.POINTER(Compiled), C.c_void_p, f64p, f64p, f64p, C.c_void_p, f64p, C.c_void_p,
                                     C.c_int]
        L.fxo_apply_gp.argtypes = [C.POINTER(Compiled), C.c_void_p, C.POINTER(Coeffs), f64p, C.c_void_p, f64p, f64p,
                                   C.c_int]
        L.fxo_diagonal_gp.argtypes = [C.POINTER(Compiled), C.c_void_p, C.POINTER(Coeffs), f64p, C.c_void_p, f64p,
                                      C.c_int]
        L.fxo_pimy_committed_tangent.argtypes = [C.POINTER(PimyParams), f64p, f64p]
        L.fxo_gmres_system.restype = C.c_int
        L.fxo_gmres_system.argtypes = [C.POINTER(System), C.POINTER(Coeffs), f64p, f64p, C.c_double, C.c_double,
                                       C.c_int32, C.c_int32, C.c_int, C.POINTER(SolveResult)]
        L.fxo_newmark.restype = C.c_int
        L.fxo_newmark.argtypes = [C.POINTER(System), C.POINTER(NewmarkOpts), C.c_int, f64p, f64p, f64p, f64p,
                                  C.POINTER(NewmarkStats)]
        L.fxo_mass_times_ground.argtypes = [C.POINTER(Compiled), C.c_void_p, C.c_void_p, C.c_int, f64p, C.c_int]
        L.fxo_cg.restype = C.c_int
        L.fxo_cg.argtypes = [C.POINTER(Compiled), C.c_void_p, C.POINTER(Coeffs), f64p, f64p, C.c_double,
                             C.c_double, C.c_int32, C.c_int, C.c_int, C.POINTER(SolveResult)]
        _LIB = L
    return _LIB


def ref():
    """The reference's own physics compiled from its headers (oracle/_ref), or None."""
    global _REF
    if _REF is None:
        path = os.path.join(HERE, "_ref", "libfxref.so")
        if not os.path.exists(path):
            return None
        R = C.CDLL(path)
        vp, i = C.c_void_p, C.c_int
        R.fxref_brick_force.argtypes = [vp, vp, i, vp, vp, vp, vp, vp, vp]
        R.fxref_brick_mass.argtypes = [vp, i, vp, vp]
        R.fxref_operator.argtypes = [vp, i, vp, vp, i, vp, vp, vp, i, vp]
        R.fxref_pimy_setup.argtypes = [vp, i, vp, vp]
        R.fxref_pimy_initial_state.argtypes = [vp, vp]
        R.fxref_pimy_update.argtypes = [vp, i, vp, vp, vp, vp, vp]
        _REF = R
    return _REF


def lcg_sequence(seed, n):
    s = C.c_uint64(seed)
    L = lib()
    return np.array([L.fxo_lcg_symmetric(C.byref(s)) for _ in range(n)])


def coeffs(stiffness=1.0, damping=0.0, mass=0.0, initial=False):
    return Coeffs(stiffness, damping, mass, int(initial))


def elastic_params(params):
    """params: list of dicts with E, nu, rho, b (3), mass_type, rayleigh (4), element ("stdBrick" or
    "SSPbrick", default stdBrick) -> ctypes array."""
    arr = (Elastic * max(1, len(params)))()
    for i, p in enumerate(params):
        b = p.get("b", (0.0, 0.0, 0.0))
        r = p.get("rayleigh", (0.0, 0.0, 0.0, 0.0))
        el = p.get("element", 0)
        el = {"stdBrick": 0, "SSPbrick": 1}.get(el, el)
        arr[i] = Elastic(p["E"], p["nu"], p.get("rho", 0.0), (C.c_double * 3)(*b), int(p.get("mass_type", 0)),
                         Rayleigh(*r), int(el))
    return arr


REFUSALS = {2: "equalDOF commands", 3: "the fix is silently lost", 4: "partial chain", 5: "has a larger tag",
            6: "other"}


class RefusedModel(RuntimeError):
    """fxo_compile refused the model; .code is the pattern (FXO_REFUSE_*)."""

    def __init__(self, code):
        super().__init__(f"fxo_compile refused the model: {REFUSALS.get(code, code)}")
        self.code = code


class OracleMesh:
    """A compiled model (fxo_compile) with numpy views of its arrays."""

    def __init__(self, model, renumber=True, simd=64, coloring=True):
        self._keep = []

        def ptr(a, dt):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=dt)
            self._keep.append(a)
            return a.ctypes.data

        m = Model()
        m.nnode = model["nnode"]
        m.xyz = ptr(model["xyz"], np.float64)
        m.nelem = model["nelem"]
        m.conn = ptr(model["conn"], np.int32)
        m.group = ptr(model.get("group"), np.int32)
        m.region = ptr(model.get("region"), np.int32)
        fix = model.get("fix", np.zeros((0, 2), np.int32))
        tie = model.get("tie", np.zeros((0, 3), np.int32))
        m.nfix = len(fix)
        m.fix_node = ptr(np.asarray(fix)[:, 0] if len(fix) else np.zeros(1), np.int32)
        m.fix_dof = ptr(np.asarray(fix)[:, 1] if len(fix) else np.zeros(1), np.int32)
        m.ntie = len(tie)
        m.tie_retained = ptr(np.asarray(tie)[:, 0] if len(tie) else np.zeros(1), np.int32)
        m.tie_constrained = ptr(np.asarray(tie)[:, 1] if len(tie) else np.zeros(1), np.int32)
        m.tie_dof = ptr(np.asarray(tie)[:, 2] if len(tie) else np.zeros(1), np.int32)
        cmd = model.get("tie_command")
        m.tie_command = ptr(cmd, np.int32) if cmd is not None and len(tie) else None
        self.model = m
        self.c = Compiled()
        rc = lib().fxo_compile(C.byref(m), int(renumber), int(simd), int(coloring), C.byref(self.c))
        if rc != 0:
            raise RefusedModel(rc)
        c = self.c
        N, P, E = c.nnode, c.padded, c.nelem
        cp = lambda p, n: np.ctypeslib.as_array(p, shape=(n,)).copy() if n > 0 else np.zeros(0)
        self.nnode, self.nelem, self.padded, self.neq, self.nparam = N, E, P, c.neq, c.nparam
        self.node_order = cp(c.node_order, N)
        self.node_rank = cp(c.node_rank, N)
        self.elem_order = cp(c.elem_order, E)
        self.node_eq = cp(c.node_eq, 3 * N)
        self.coords = cp(c.coords, 3 * N)
        self.node = cp(c.node, 8 * P)
        self.eq = cp(c.eq, 24 * P)
        self.param = cp(c.param, P)
        self.param_group = cp(c.param_group, c.nparam)
        self.param_region = cp(c.param_region, c.nparam)
        self.ncolor = c.ncolor
        self.color_start = cp(c.color_start, c.ncolor + 1) if coloring else None
        self.color_elements = cp(c.color_elements, E) if coloring else None

    def __del__(self):
        try:
            lib().fxo_compiled_free(C.byref(self.c))
        except Exception:
            pass

    def bandwidth(self):
        return lib().fxo_node_bandwidth(C.byref(self.model), self.node_rank.astype(np.int32))

    def apply(self, params, x, k=None, nthreads=0):
        y = np.zeros(self.neq)
        lib().fxo_apply(C.byref(self.c), C.cast(params, C.c_void_p), C.byref(k or coeffs()),
                        np.ascontiguousarray(x, np.float64), y, nthreads)
        return y

    def diagonal(self, params, k=None, nthreads=0):
        d = np.zeros(self.neq)
        lib().fxo_diagonal(C.byref(self.c), C.cast(params, C.c_void_p), C.byref(k or coeffs()), d, nthreads)
        return d

    def lumped_mass(self, params):
        m = np.zeros(self.neq)
        am = np.zeros(self.neq)
        lib().fxo_lumped_mass(C.byref(self.c), C.cast(params, C.c_void_p), m, am)
        return m, am

    def body_load(self, params):
        b = np.zeros(self.neq)
        lib().fxo_body_load(C.byref(self.c), C.cast(params, C.c_void_p), b)
        return b

    def force(self, params, u, v=None, nthreads=0):
        f = np.zeros(self.neq)
        vv = None if v is None else np.ascontiguousarray(v, np.float64)
        lib().fxo_force(C.byref(self.c), C.cast(params, C.c_void_p), np.ascontiguousarray(u, np.float64),
                        None if vv is None else vv.ctypes.data, f, nthreads)
        return f

    def critical_dt(self, params, scheme=0):
        return lib().fxo_critical_dt(C.byref(self.c), C.cast(params, C.c_void_p), scheme)

    def cd_run(self, params, m, cm, dt, nsteps, u, um, v, a, load_eq=None, load_hist=None, dash=None, nthreads=0):
        """nsteps CentralDifference steps; u, um, v, a (numpy) are updated in place."""
        cmv = None if cm is None else np.ascontiguousarray(cm, np.float64)
        dsv = None if dash is None else np.ascontiguousarray(dash, np.float64)
        le = np.ascontiguousarray(load_eq if load_eq is not None else np.zeros(0), np.int32)
        lh = np.ascontiguousarray(load_hist if load_hist is not None else np.zeros(0), np.float64)
        lib().fxo_cd_run(C.byref(self.c), C.cast(params, C.c_void_p), np.ascontiguousarray(m, np.float64),
                         None if cmv is None else cmv.ctypes.data, None if dsv is None else dsv.ctypes.data, len(le), le.ctypes.data, lh.ctypes.data, dt,
                         nsteps, u, um, v, a, nthreads)

    def explicit_run(self, params, m, cm, dt, nsteps, u, um, v, a, scheme=0, ground_dir=-1, ground_accel=None,
                     load_eq=None, load_hist=None, dash=None, nthreads=0):
        """nsteps explicit steps (scheme 0 CentralDifference, 1 ExplicitDifference), a uniform
        excitation in ground_dir with ground_accel[k] at each step's load time; in place."""
        cmv = None if cm is None else np.ascontiguousarray(cm, np.float64)
        dsv = None if dash is None else np.ascontiguousarray(dash, np.float64)
        le = np.ascontiguousarray(load_eq if load_eq is not None else np.zeros(0), np.int32)
        lh = np.ascontiguousarray(load_hist if load_hist is not None else np.zeros(0), np.float64)
        ga = None if ground_accel is None else np.ascontiguousarray(ground_accel, np.float64)
        lib().fxo_explicit_run(C.byref(self.c), C.cast(params, C.c_void_p), np.ascontiguousarray(m, np.float64),
                               None if cmv is None else cmv.ctypes.data, None if dsv is None else dsv.ctypes.data,
                               len(le), le.ctypes.data, lh.ctypes.data, dt, nsteps, scheme, ground_dir,
                               None if ga is None else ga.ctypes.data, u, um, v, a, nthreads)

    # ---- PressureIndependMultiYield batches: points e * 8 + gp, history [point, PIMY_NSTATE]
    def pimy_initial_state(self, elems):
        S = np.zeros((self.nelem * 8, PIMY_NSTATE))
        one = np.zeros(PIMY_NSTATE)
        for e in range(self.nelem):
            lib().fxo_pimy_initial_state(C.byref(elems[self.param[e]].mat), one.ctypes.data)
            S[8 * e:8 * e + 8] = one
        return S

    def pimy_force(self, elems, Cs, Ts, u, want_D=False, nthreads=0, v=None):
        """f(u) with the return map at every point (trial history written); v: the velocities of
        the Rayleigh damping force, or None."""
        f = np.zeros(self.neq)
        D = np.zeros((self.nelem * 8, 36)) if want_D else None
        vv = None if v is None else np.ascontiguousarray(v, np.float64)
        lib().fxo_pimy_force(C.byref(self.c), C.cast(elems, C.c_void_p), Cs, Ts, np.ascontiguousarray(u, np.float64),
                             None if vv is None else vv.ctypes.data, f, None if D is None else D.ctypes.data, nthreads)
        return (f, D) if want_D else f

    def apply_gp(self, elems, D, x, k=None, nthreads=0, committed=None):
        """K_eff x on the stored tangents D; committed: the committed histories (betaKcomm)."""
        y = np.zeros(self.neq)
        cm = None if committed is None else np.ascontiguousarray(committed)
        lib().fxo_apply_gp(C.byref(self.c), C.cast(elems, C.c_void_p), C.byref(k or coeffs()), np.ascontiguousarray(D),
                           None if cm is None else cm.ctypes.data, np.ascontiguousarray(x, np.float64), y, nthreads)
        return y

    def diagonal_gp(self, elems, D, k=None, nthreads=0, committed=None):
        d = np.zeros(self.neq)
        cm = None if committed is None else np.ascontiguousarray(committed)
        lib().fxo_diagonal_gp(C.byref(self.c), C.cast(elems, C.c_void_p), C.byref(k or coeffs()), np.ascontiguousarray(D),
                              None if cm is None else cm.ctypes.data, d, nthreads)
        return d

    def ground_direction(self, ground_dir):
        """ground_dir: None, a direction 0..2, or the influence vector of one direction (1 at its
        free equations) -> the direction, or -1."""
        if ground_dir is None:
            return -1
        if np.isscalar(ground_dir):
            return int(ground_dir)
        r = np.asarray(ground_dir)
        dirs = {d for d in range(3) for q in self.node_eq.reshape(-1, 3)[:, d] if q >= 0 and r[q] != 0.0}
        assert len(dirs) == 1, "the influence vector names one direction"
        return dirs.pop()

    def mass_times_ground(self, d, elastic=None, pimy=None, nthreads=0):
        """M r with r the rigid unit translation in direction d at every element DOF"""
        y = np.zeros(self.neq)
        lib().fxo_mass_times_ground(C.byref(self.c), None if elastic is None else C.cast(elastic, C.c_void_p),
                                    None if pimy is None else C.cast(pimy, C.c_void_p), d, y, nthreads)
        return y

    def newmark(self, u, v, a, du, dt, nsteps, elastic=None, pimy=None, pimy_state=None, gamma=0.5, beta=0.25,
                test=1, tol=1e-10, max_iter=25, cg_rtol=1e-12, cg_max_it=10000, ground_dir=None, ground_accel=None,
                lanes=64, nthreads=0, max_retries=0, num_sublevels=0, num_substeps=2, series=None, ground_factor=1.0,
                time=0.0, static=False, algorithm=0, modified_initial=False, factor_once=False, load_increment=None,
                load_min_increment=0.0, load_max_increment=0.0, load_num_iter=1):
        """nsteps steps of an implicit stage (implicit_stage.cpp) on u, v, a, du (numpy, in place):
        Newmark (static=False) or LoadControl (static=True) steps; algorithm 0 Newton, 1 Linear,
        2 ModifiedNewton. pimy_state: [C, T] point-major histories (swapped in place on commits).
        Returns (status, stats dict)."""
        s = System()
        s.c = C.cast(C.pointer(self.c), C.c_void_p)
        s.elastic = C.cast(elastic, C.c_void_p) if elastic is not None else None
        s.pimy = C.cast(pimy, C.c_void_p) if pimy is not None else None
        keep = []
        if pimy is not None:
            Cs, Ts = pimy_state
            D = np.zeros((self.nelem * 8, 36))
            keep += [Cs, Ts, D]
            s.C, s.T, s.D = Cs.ctypes.data, Ts.ctypes.data, D.ctypes.data
        s.nthreads = nthreads
        ga = None if ground_accel is None else np.ascontiguousarray(ground_accel, np.float64)
        ser = None
        if isinstance(series, TimeArraySeries):
            ser = series.struct()
            keep += [series, ser]
        elif series is not None:
            sv = np.ascontiguousarray(series["values"], np.float64)
            ser = Series(sv.ctypes.data, len(sv), series["dt"], series.get("factor", 1.0), series.get("start", 0.0),
                         int(bool(series.get("use_last", False))), None, None)
            keep += [sv, ser]
        o = NewmarkOpts(gamma, beta, dt, nsteps, test, tol, max_iter, cg_rtol, cg_max_it, lanes,
                        None if ga is None else ga.ctypes.data, max_retries, num_sublevels, num_substeps,
                        None if ser is None else C.cast(C.pointer(ser), C.c_void_p), ground_factor, time,
                        int(bool(static)), algorithm, int(bool(modified_initial)), int(bool(factor_once)),
                        dt if load_increment is None else load_increment, load_min_increment, load_max_increment,
                        load_num_iter)
        st = NewmarkStats()
        rc = lib().fxo_newmark(C.byref(s), C.byref(o), self.ground_direction(ground_dir), u, v, a, du, C.byref(st))
        if pimy is not None and s.C != pimy_state[0].ctypes.data:  # an odd number of commits
            tmp = pimy_state[0].copy()
            pimy_state[0][...] = pimy_state[1]
            pimy_state[1][...] = tmp
        return rc, {f: getattr(st, f) for f, _ in NewmarkStats._fields_}

    def gmres(self, params, b, x0=None, rtol=1e-8, atol=0.0, max_it=10000, restart=30, lanes=64, k=None,
              nthreads=0):
        """FGMRES(restart) with Jacobi on the elastic operator (gmres.cpp:32-123)."""
        s = System()
        s.c = C.cast(C.pointer(self.c), C.c_void_p)
        s.elastic = C.cast(params, C.c_void_p)
        s.nthreads = nthreads
        x = np.zeros(self.neq) if x0 is None else np.array(x0, dtype=np.float64)
        r = SolveResult()
        rc = lib().fxo_gmres_system(C.byref(s), C.byref(k or coeffs()), np.ascontiguousarray(b, np.float64), x, rtol,
                                    atol, max_it, restart, lanes, C.byref(r))
        if rc != 0:
            raise RuntimeError("Jacobi: zero diagonal entries")
        return x, {f: getattr(r, f) for f, _ in SolveResult._fields_}

    def cg(self, params, b, x0=None, rtol=1e-8, atol=0.0, max_it=10000, lanes=1, k=None, nthreads=0):
        x = np.zeros(self.neq) if x0 is None else np.array(x0, dtype=np.float64)
        r = SolveResult()
        rc = lib().fxo_cg(C.byref(self.c), C.cast(params, C.c_void_p), C.byref(k or coeffs()),
                          np.ascontiguousarray(b, np.float64), x, rtol, atol, max_it, lanes, nthreads, C.byref(r))
        if rc != 0:
            raise RuntimeError("Jacobi: zero diagonal entries")
        return x, {f: getattr(r, f) for f, _ in SolveResult._fields_}


def dot(x, y, lanes=1):
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    return lib().fxo_dot(len(x), x, y, lanes)
