"""Python host mirror of the C ABI in include/femorax_amd.h (ctypes over libfemorax_amd.so).

The names follow the reference's exec/solvers interfaces: Mesh (compile::Compiled), Batch
(InternalForce::Batch / InternalForce), apply_operator / operator_diagonal, cg
(ConjugateGradient + Jacobi). There is no fallback: if the HIP library is not built, importing
the bindings raises, and every compute call goes to the device.
"""
import ctypes as C
import weakref
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfemorax_amd.so")
_LIB = None

FX_OK = 0
FX_EINVAL = -1
ERRORS = {-1: "EINVAL", -2: "EDEVICE", -3: "ENOMEM", -4: "EZERO", -5: "EUNSUP", -6: "EFULL", -7: "EIO"}
FX_BATCH_FP32 = 1
FX_BATCH_GENERAL = 2
FX_BATCH_XCD_SLABS = 8
FX_BATCH_DYNAMIC = 32
FX_BATCH_GMRES = 64
FX_BATCH_NO_TUNE = 16
FX_BATCH_SLOTS = 2048
FX_BATCH_SSPBRICK = 4096
FX_OUT_PLAIN, FX_OUT_ATOMIC, FX_OUT_SLOTS = 0, 1, 2
FX_TANGENT_SPD, FX_TANGENT_SPSD = 0, 1


class FxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class ModelDesc(C.Structure):
    _fields_ = [("nnode", C.c_int32), ("xyz", C.c_void_p), ("nelem", C.c_int32), ("conn", C.c_void_p),
                ("group", C.c_void_p), ("region", C.c_void_p), ("nfix", C.c_int32), ("fix_node", C.c_void_p),
                ("fix_dof", C.c_void_p), ("ntie", C.c_int32), ("tie_retained", C.c_void_p),
                ("tie_constrained", C.c_void_p), ("tie_dof", C.c_void_p), ("tie_command", C.c_void_p)]


class MeshView(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("nnode", "nelem", "padded", "neq", "nparam", "ncolor")] + [
        (n, C.c_void_p) for n in ("node_order", "node_rank", "elem_order", "node_eq", "coords", "node", "eq",
                                  "param", "param_group", "param_region", "color_start", "color_elements")]


class Rayleigh(C.Structure):
    _fields_ = [("alphaM", C.c_double), ("betaK", C.c_double), ("betaKinit", C.c_double), ("betaKcomm", C.c_double)]


class ElasticParams(C.Structure):
    _fields_ = [("E", C.c_double), ("nu", C.c_double), ("rho", C.c_double), ("b", C.c_double * 3),
                ("mass_type", C.c_int32), ("damp", Rayleigh)]


class Coeffs(C.Structure):
    _fields_ = [("stiffness", C.c_double), ("damping", C.c_double), ("mass", C.c_double), ("initial", C.c_int32)]


class BatchDesc(C.Structure):
    _fields_ = [("count", C.c_int32), ("padded", C.c_int32), ("node", C.c_void_p), ("eq", C.c_void_p),
                ("param", C.c_void_p), ("nparam", C.c_int32), ("params", C.c_void_p), ("nnode", C.c_int32),
                ("coords", C.c_void_p), ("neq", C.c_int32)]


class BatchInfo(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("count", "neq", "nnode", "nblocks", "ncolors", "tied", "lattice",
                                         "tiles", "variant", "tangent_symmetry")] + [("device_bytes", C.c_uint64)]


class SolveOptions(C.Structure):
    _fields_ = [("relative_tolerance", C.c_double), ("absolute_tolerance", C.c_double),
                ("max_iterations", C.c_int32)]


class CdDesc(C.Structure):
    _fields_ = [("dt", C.c_double), ("nsteps", C.c_int32), ("dash", C.c_void_p), ("nload", C.c_int32),
                ("load_eq", C.c_void_p), ("load_hist", C.c_void_p), ("scheme", C.c_int32), ("ground_dir", C.c_int32),
                ("ground_accel", C.c_void_p), ("ground_series", C.c_void_p), ("ground_factor", C.c_double),
                ("time", C.c_double), ("recorder", C.c_void_p)]


class TimeSeries(C.Structure):
    _fields_ = [("values", C.c_void_p), ("count", C.c_int32), ("dt", C.c_double), ("factor", C.c_double),
                ("start", C.c_double), ("use_last", C.c_int32), ("time", C.c_void_p), ("loc", C.c_void_p)]


class NewmarkDesc(C.Structure):
    _fields_ = [("gamma", C.c_double), ("beta", C.c_double), ("dt", C.c_double), ("nsteps", C.c_int32),
                ("test", C.c_int32), ("tol", C.c_double), ("max_iter", C.c_int32), ("cg_rtol", C.c_double),
                ("cg_max_it", C.c_int32), ("ground_dir", C.c_int32), ("ground_accel", C.c_void_p),
                ("max_retries", C.c_int32), ("num_sublevels", C.c_int32), ("num_substeps", C.c_int32),
                ("ground_series", C.c_void_p), ("ground_factor", C.c_double), ("time", C.c_double),
                ("mixed", C.c_int32), ("inner_rtol", C.c_double), ("analysis", C.c_int32), ("algorithm", C.c_int32),
                ("modified_initial", C.c_int32), ("factor_once", C.c_int32), ("load_increment", C.c_double),
                ("load_min_increment", C.c_double), ("load_max_increment", C.c_double), ("load_num_iter", C.c_int32),
                ("recorder", C.c_void_p)]


class NewmarkResult(C.Structure):
    _fields_ = [("steps", C.c_int32), ("newton_iterations", C.c_int32), ("max_newton_iterations", C.c_int32),
                ("krylov_iterations", C.c_int32), ("failed", C.c_int32), ("last_test", C.c_double),
                ("commits", C.c_int32), ("failed_attempts", C.c_int32), ("retries", C.c_int32),
                ("substepped", C.c_int32), ("time", C.c_double)]


class PimyBatchDesc(C.Structure):
    _fields_ = [("count", C.c_int32), ("padded", C.c_int32), ("node", C.c_void_p), ("eq", C.c_void_p),
                ("param", C.c_void_p), ("nparam", C.c_int32), ("params", C.c_void_p), ("nnode", C.c_int32),
                ("coords", C.c_void_p), ("neq", C.c_int32), ("damp", C.c_void_p)]


class PimyMaterial(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("rho", "G", "K", "cohesion", "gamma_peak", "phi", "pref", "pdc")] + [
        ("nsurf", C.c_int32), ("gredu", C.c_void_p), ("b", C.c_double * 3), ("mass_type", C.c_int32)]


PIMY_NSTATE = 190
FX_EFULL = -6
FX_EIO = -7
FX_COLUMN_MAX_ELEMS = 512
FX_COLUMN_MIXED = 1
FX_COLUMN_NO_GRAPH = 2
FX_COLUMN_COMPACT = 4
FX_COLUMN_TUNE = 16
FX_BATCH_COMPACT = 128
FX_BATCH_TILE4 = 256
FX_BATCH_TILE8 = 512
FX_BATCH_TILE16 = 1024


class ColumnPoolDesc(C.Structure):
    _fields_ = [("max_elements", C.c_int64), ("max_columns", C.c_int32), ("max_materials", C.c_int32),
                ("max_record", C.c_int64), ("gamma", C.c_double), ("beta", C.c_double), ("dt", C.c_double),
                ("test", C.c_int32), ("tol", C.c_double), ("max_iter", C.c_int32), ("cg_rtol", C.c_double),
                ("cg_max_it", C.c_int32), ("flags", C.c_uint32), ("inner_rtol", C.c_double),
                ("max_retries", C.c_int32), ("num_sublevels", C.c_int32), ("num_substeps", C.c_int32)]


class ColumnDesc(C.Structure):
    _fields_ = [("nelem", C.c_int32), ("hx", C.c_double), ("hy", C.c_double), ("hz", C.c_double),
                ("nmat", C.c_int32), ("mats", C.c_void_p), ("mat_of", C.c_void_p), ("stage", C.c_int32),
                ("nsteps", C.c_int32), ("ground_dir", C.c_int32), ("ground_accel", C.c_void_p), ("record", C.c_int32)]


class ColumnResult(C.Structure):
    _fields_ = [("status", C.c_int32), ("steps", C.c_int32), ("newton_iterations", C.c_int32),
                ("max_newton_iterations", C.c_int32), ("krylov_iterations", C.c_int32), ("last_test", C.c_double),
                ("commits", C.c_int32), ("failed_attempts", C.c_int32), ("retries", C.c_int32),
                ("substepped", C.c_int32)]


def pimy_material(rho, G, K, cohesion, gamma_peak, phi, pref, pdc, nsurf=20, gredu=None, b=(0.0, 0.0, 0.0),
                  mass_type=0, keep=None):
    """fx_pimy_material; gredu: optional (gamma, G/Gmax) pairs (kept alive in `keep`)."""
    g = None
    if gredu is not None and len(gredu):
        arr = np.ascontiguousarray(gredu, np.float64)
        if keep is not None:
            keep.append(arr)
        g = arr.ctypes.data
    return PimyMaterial(rho, G, K, cohesion, gamma_peak, phi, pref, pdc, nsurf, g, (C.c_double * 3)(*b), mass_type)


def pimy_materials(dicts, keep=None):
    """fx_pimy_material list from dicts (rho G K cohesion gamma_peak phi pref pdc [nsurf gredu b
    mass_type])."""
    return [pimy_material(m["rho"], m["G"], m["K"], m["cohesion"], m["gamma_peak"], m["phi"], m["pref"], m["pdc"],
                          nsurf=m.get("nsurf", 20), gredu=m.get("gredu"), b=m.get("b", (0.0, 0.0, 0.0)),
                          mass_type=m.get("mass_type", 0), keep=keep) for m in dicts]


def pimy_surfaces(m):
    size, modulus = np.zeros(21), np.zeros(21)
    _check(lib().fx_pimy_surfaces(C.byref(m), size.ctypes.data, modulus.ctypes.data))
    return size, modulus


def pimy_init_state(ctx, mats, mat, state):
    arr = (PimyMaterial * len(mats))(*mats)
    _check(lib().fx_pimy_init_state(ctx.h, len(mats), arr, mat.n, mat.ptr, state.ptr), ctx)


def pimy_points(ctx, mats, load_stage, mat, committed, trial, eps, sig, D=None):
    arr = (PimyMaterial * len(mats))(*mats)
    st = np.ascontiguousarray(load_stage, np.int32)
    _check(lib().fx_pimy_points(ctx.h, len(mats), arr, st.ctypes.data, mat.n, mat.ptr, committed.ptr, trial.ptr,
                                eps.ptr, sig.ptr, D.ptr if D is not None else None), ctx)


class SolveResult(C.Structure):
    _fields_ = [("converged", C.c_int32), ("iterations", C.c_int32), ("indefinite", C.c_int32),
                ("outer_iterations", C.c_int32), ("rhs_norm", C.c_double), ("initial_residual", C.c_double),
                ("final_residual", C.c_double), ("true_residual", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


# (name, restype, argtypes) of every entry point in include/femorax_amd.h
vp, i32, i64, u32, u64, dbl = C.c_void_p, C.c_int32, C.c_int64, C.c_uint32, C.c_uint64, C.c_double
SIGNATURES = [
    ("fx_ctx_create", C.c_int, [C.c_int, u64, C.POINTER(vp)]),
    ("fx_ctx_destroy", None, [vp]),
    ("fx_last_error", C.c_char_p, [vp]),
    ("fx_sync", C.c_int, [vp]),
    ("fx_ctx_set_option", C.c_int, [vp, C.c_int, i64]),
    ("fx_device_memory", C.c_int, [vp, C.POINTER(u64), C.POINTER(u64)]),
    ("fx_budget_used", u64, [vp]),
    ("fx_alloc", C.c_int, [vp, C.c_size_t, C.POINTER(vp)]),
    ("fx_free", C.c_int, [vp, vp]),
    ("fx_upload", C.c_int, [vp, vp, vp, C.c_size_t]),
    ("fx_download", C.c_int, [vp, vp, vp, C.c_size_t]),
    ("fx_memset0", C.c_int, [vp, vp, C.c_size_t]),
    ("fx_compile_model", C.c_int, [C.POINTER(ModelDesc), C.c_int, C.c_int, C.c_int, C.POINTER(vp)]),
    ("fx_mesh_get", C.c_int, [vp, C.POINTER(MeshView)]),
    ("fx_mesh_destroy", None, [vp]),
    ("fx_batch_create", C.c_int, [vp, C.POINTER(BatchDesc), u32, C.POINTER(vp)]),
    ("fx_batch_destroy", None, [vp]),
    ("fx_batch_info_get", C.c_int, [vp, C.POINTER(BatchInfo)]),
    ("fx_batch_tune", C.c_int, [vp, C.c_int]),
    ("fx_operator_apply", C.c_int, [vp, C.POINTER(Coeffs), vp, vp]),
    ("fx_operator_apply_f32", C.c_int, [vp, C.POINTER(Coeffs), vp, vp]),
    ("fx_operator_diagonal", C.c_int, [vp, C.POINTER(Coeffs), vp]),
    ("fx_run_operator", C.c_int, [vp, C.c_int, C.POINTER(Coeffs), C.c_int, vp, vp, vp, i64, C.c_int]),
    ("fx_batch_set_nodal_mass", C.c_int, [vp, vp, vp]),
    ("fx_mass_times_ground", C.c_int, [vp, C.c_int, vp]),
    ("fx_operator_time", C.c_int, [vp, C.POINTER(Coeffs), C.c_int, vp, vp, C.c_int, C.POINTER(C.c_float)]),
    ("fx_copy", C.c_int, [vp, i64, vp, vp]),
    ("fx_dot", C.c_int, [vp, i64, vp, vp, C.POINTER(dbl)]),
    ("fx_dot_static", C.c_int, [vp, i64, vp, vp, vp]),
    ("fx_cg", C.c_int, [vp, C.POINTER(Coeffs), C.POINTER(SolveOptions), vp, vp, C.POINTER(SolveResult)]),
    ("fx_cg_many", C.c_int, [C.c_int32, C.POINTER(vp), C.POINTER(Coeffs), C.POINTER(SolveOptions), C.POINTER(vp),
                             C.POINTER(vp), C.POINTER(SolveResult)]),
    ("fx_cg_mixed", C.c_int, [vp, C.POINTER(Coeffs), C.POINTER(SolveOptions), dbl, vp, vp, C.POINTER(SolveResult)]),
    ("fx_gmres", C.c_int, [vp, C.POINTER(Coeffs), C.POINTER(SolveOptions), vp, vp, C.POINTER(SolveResult)]),
    ("fx_krylov", C.c_int, [vp, C.POINTER(Coeffs), C.POINTER(SolveOptions), vp, vp, C.POINTER(SolveResult)]),
    ("fx_lumped_mass", C.c_int, [vp, vp, vp]),
    ("fx_body_load", C.c_int, [vp, vp]),
    ("fx_force", C.c_int, [vp, vp, vp, vp]),
    ("fx_critical_dt", C.c_int, [vp, C.c_int, C.POINTER(dbl)]),
    ("fx_explicit_cd", C.c_int, [vp, C.POINTER(CdDesc), vp, vp, vp, vp]),
    ("fx_pimy_surfaces", C.c_int, [C.POINTER(PimyMaterial), vp, vp]),
    ("fx_pimy_batch_create", C.c_int, [vp, C.POINTER(PimyBatchDesc), u32, C.POINTER(vp)]),
    ("fx_set_material_stage", C.c_int, [vp, C.c_int32]),
    ("fx_newmark", C.c_int, [vp, C.POINTER(NewmarkDesc), vp, vp, vp, vp, C.POINTER(NewmarkResult)]),
    ("fx_store_tangents", C.c_int, [vp, vp, vp]),
    ("fx_commit_state", C.c_int, [vp]),
    ("fx_pimy_state", C.c_int, [vp, C.c_int, vp]),
    ("fx_pimy_init_state", C.c_int, [vp, C.c_int32, C.POINTER(PimyMaterial), i64, vp, vp]),
    ("fx_pimy_points", C.c_int, [vp, C.c_int32, C.POINTER(PimyMaterial), vp, i64, vp, vp, vp, vp, vp, vp]),
    ("fx_column_pool_create", C.c_int, [vp, C.POINTER(ColumnPoolDesc), C.POINTER(vp)]),
    ("fx_column_pool_destroy", None, [vp]),
    ("fx_column_add", C.c_int, [vp, C.POINTER(ColumnDesc), C.POINTER(i32)]),
    ("fx_column_run", C.c_int, [vp, i32, C.POINTER(i32)]),
    ("fx_column_status", C.c_int, [vp, i32, C.POINTER(ColumnResult)]),
    ("fx_column_finished", C.c_int, [vp, vp, i32, C.POINTER(i32)]),
    ("fx_column_fetch", C.c_int, [vp, i32, vp, vp, vp, vp]),
    ("fx_column_state", C.c_int, [vp, i32, vp]),
    ("fx_column_remove", C.c_int, [vp, i32]),
    ("fx_recorder_create", C.c_int, [vp, C.POINTER(vp)]),
    ("fx_recorder_add_node", C.c_int, [vp, C.c_char_p, i32, i32, dbl, i32, i64, vp, C.POINTER(i32)]),
    ("fx_recorder_activate", C.c_int, [vp, i32]),
    ("fx_recorder_record", C.c_int, [vp, dbl, i64, vp, vp, vp]),
    ("fx_recorder_finish", C.c_int, [vp]),
    ("fx_recorder_records", i64, [vp]),
    ("fx_recorder_destroy", None, [vp]),
]


def lib():
    """Load libfemorax_amd.so (raises if it is not built: there is no CPU fallback)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built: run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def _check(rc, ctx=None):
    if rc != FX_OK:
        msg = lib().fx_last_error(ctx.h if ctx is not None else None)
        raise FxError(rc, msg.decode() if msg else "")


class Context:
    """One device, one stream, an HBM budget (bytes; 0 = 90% of free memory)."""

    def __init__(self, device=0, budget=0):
        h = C.c_void_p()
        _check(lib().fx_ctx_create(device, budget, C.byref(h)))
        self.h = h
        self._children = weakref.WeakSet()  # batches and buffers: released before the context

    def adopt(self, child):
        self._children.add(child)

    def close(self):
        if self.h:
            for child in list(self._children):
                child.release()
            lib().fx_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        _check(lib().fx_sync(self.h), self)

    def set_graphs(self, enable):
        """CG iterations from captured hipGraphs, or launched one by one (default)."""
        _check(lib().fx_ctx_set_option(self.h, 1, int(bool(enable))), self)

    def set_stream(self, handle):
        """Queue the context's work on the caller's HIP stream (an integer hipStream_t, e.g.
        torch.cuda.Stream().cuda_stream), or on its own again with 0."""
        _check(lib().fx_ctx_set_option(self.h, 2, int(handle)), self)

    def set_stream_ordered(self, enable):
        """The drop-in operator calls return once queued (results ordered on the stream)."""
        _check(lib().fx_ctx_set_option(self.h, 3, int(bool(enable))), self)

    def memory(self):
        f, t = C.c_uint64(), C.c_uint64()
        _check(lib().fx_device_memory(self.h, C.byref(f), C.byref(t)), self)
        return f.value, t.value


class DeviceArray:
    """A device buffer owned by Python (tests and the benchmark)."""

    def __init__(self, ctx, n, dtype=np.float64):
        self.ctx, self.n, self.dtype = ctx, int(n), np.dtype(dtype)
        p = C.c_void_p()
        _check(lib().fx_alloc(ctx.h, max(1, self.n) * self.dtype.itemsize, C.byref(p)), ctx)
        self.ptr = p
        ctx.adopt(self)

    @classmethod
    def from_host(cls, ctx, a, dtype=None):
        a = np.ascontiguousarray(a, dtype=dtype or a.dtype)
        d = cls(ctx, a.size, a.dtype)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.size == self.n
        _check(lib().fx_upload(self.ctx.h, self.ptr, a.ctypes.data, a.nbytes), self.ctx)

    def get(self):
        out = np.empty(self.n, self.dtype)
        _check(lib().fx_download(self.ctx.h, out.ctypes.data, self.ptr, out.nbytes), self.ctx)
        return out

    def zero(self):
        _check(lib().fx_memset0(self.ctx.h, self.ptr, self.n * self.dtype.itemsize), self.ctx)

    def free(self):
        if self.ptr and self.ctx.h:
            lib().fx_free(self.ctx.h, self.ptr)
        self.ptr = None

    def release(self):
        self.free()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _arr(keep, a, dt):
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype=dt)
    keep.append(a)
    return a.ctypes.data


def model_desc(model, keep):
    m = ModelDesc()
    m.nnode = int(model["nnode"])
    m.xyz = _arr(keep, model["xyz"], np.float64)
    m.nelem = int(model["nelem"])
    m.conn = _arr(keep, model["conn"], np.int32)
    m.group = _arr(keep, model.get("group"), np.int32)
    m.region = _arr(keep, model.get("region"), np.int32)
    fix = np.asarray(model.get("fix", np.zeros((0, 2))), np.int32).reshape(-1, 2)
    tie = np.asarray(model.get("tie", np.zeros((0, 3))), np.int32).reshape(-1, 3)
    m.nfix = len(fix)
    m.fix_node = _arr(keep, fix[:, 0].copy() if len(fix) else np.zeros(1), np.int32)
    m.fix_dof = _arr(keep, fix[:, 1].copy() if len(fix) else np.zeros(1), np.int32)
    m.ntie = len(tie)
    m.tie_retained = _arr(keep, tie[:, 0].copy() if len(tie) else np.zeros(1), np.int32)
    m.tie_constrained = _arr(keep, tie[:, 1].copy() if len(tie) else np.zeros(1), np.int32)
    m.tie_dof = _arr(keep, tie[:, 2].copy() if len(tie) else np.zeros(1), np.int32)
    cmd = model.get("tie_command")
    m.tie_command = _arr(keep, cmd, np.int32) if cmd is not None and len(tie) else None
    return m


class Mesh:
    """fx_compile_model: the compile step's output as numpy arrays (host only, no GPU needed)."""

    def __init__(self, model, renumber=True, simd=64, coloring=True):
        keep = []
        desc = model_desc(model, keep)
        h = C.c_void_p()
        _check(lib().fx_compile_model(C.byref(desc), int(renumber), int(simd), int(coloring), C.byref(h)))
        v = MeshView()
        _check(lib().fx_mesh_get(h, C.byref(v)))
        N, P, E = v.nnode, v.padded, v.nelem

        def get(p, n, dt):
            if not p or n == 0:
                return np.zeros(0, dt)
            return np.ctypeslib.as_array(C.cast(p, C.POINTER(np.ctypeslib.as_ctypes_type(dt))), shape=(n,)).copy()

        self.nnode, self.nelem, self.padded, self.neq, self.nparam, self.ncolor = N, E, P, v.neq, v.nparam, v.ncolor
        self.node_order = get(v.node_order, N, np.int32)
        self.node_rank = get(v.node_rank, N, np.int32)
        self.elem_order = get(v.elem_order, E, np.int32)
        self.node_eq = get(v.node_eq, 3 * N, np.int32)
        self.coords = get(v.coords, 3 * N, np.float64)
        self.node = get(v.node, 8 * P, np.int32)
        self.eq = get(v.eq, 24 * P, np.int32)
        self.param = get(v.param, P, np.int32)
        self.param_group = get(v.param_group, v.nparam, np.int32)
        self.param_region = get(v.param_region, v.nparam, np.int32)
        self.color_start = get(v.color_start, v.ncolor + 1, np.int32) if v.color_start else None
        self.color_elements = get(v.color_elements, E, np.int32) if v.color_elements else None
        lib().fx_mesh_destroy(h)

    def params_for(self, model):
        """Per param set (group, region) -> elastic parameters from model['params'] by group and
        model.get('rayleigh') by region."""
        regions = model.get("rayleigh", [])
        out = []
        for g, r in zip(self.param_group, self.param_region):
            p = dict(model["params"][g])
            if r >= 0:
                p["rayleigh"] = tuple(regions[r])
            out.append(p)
        return out

    def to_equations(self, nodal):
        """Nodal values [nnode, 3] in model order -> equation space (fixed drop, tied add)."""
        b = np.zeros(self.neq)
        eqs = self.node_eq.reshape(-1, 3)[self.node_rank]
        for d in range(3):
            q = eqs[:, d]
            m = q >= 0
            np.add.at(b, q[m], nodal[m, d])
        return b

    def recorder_columns(self, nodes, dofs=(1, 2, 3)):
        """The equations a Node recorder of model nodes `nodes` and DOFs `dofs` (1-based) reads,
        node-major (-1 for a fixed DOF; a tied DOF reads its retained equation)."""
        eqs = self.node_eq.reshape(-1, 3)[self.node_rank[np.asarray(nodes, np.int64)]]
        return np.ascontiguousarray(eqs[:, [d - 1 for d in dofs]].ravel(), np.int32)

    def to_nodes(self, x):
        """Equation values -> [nnode, 3] in model order (fixed DOFs 0)."""
        eqs = self.node_eq.reshape(-1, 3)[self.node_rank]
        out = np.zeros((self.nnode, 3))
        m = eqs >= 0
        out[m] = x[eqs[m]]
        return out


def elastic_array(params):
    arr = (ElasticParams * max(1, len(params)))()
    for i, p in enumerate(params):
        b = p.get("b", (0.0, 0.0, 0.0))
        r = p.get("rayleigh", (0.0, 0.0, 0.0, 0.0))
        arr[i] = ElasticParams(p["E"], p["nu"], p.get("rho", 0.0), (C.c_double * 3)(*b), int(p.get("mass_type", 0)),
                               Rayleigh(*r))
    return arr


def coeffs(stiffness=1.0, damping=0.0, mass=0.0, initial=False):
    return Coeffs(stiffness, damping, mass, int(initial))


def cg_many(ctx, batches, bs, xs, rtol=1e-8, atol=0.0, max_it=10000, k=None):
    """fx_cg_many: one CG + Jacobi solve per batch, advanced together (batch i solves A_i x_i = b_i);
    the results as fx_cg gives them, in order."""
    n = len(batches)
    vp = C.c_void_p
    hs = (vp * max(1, n))(*[b.h.value for b in batches])
    rb = (vp * max(1, n))(*[b.ptr for b in bs])
    rx = (vp * max(1, n))(*[x.ptr for x in xs])
    res = (SolveResult * max(1, n))()
    o = SolveOptions(rtol, atol, max_it)
    _check(lib().fx_cg_many(n, hs, C.byref(k or coeffs()), C.byref(o), rb, rx, res), ctx)
    return [res[i].as_dict() for i in range(n)]


class Batch:
    """An element batch on the device (fx_batch_create): the operator, its diagonal, CG."""

    def __init__(self, ctx, mesh, params, fp32=False, general=False, xcd=False, tune=True, dynamic=False,
                 gmres=False, tile8=None, tile_depth=None, slots=False, element="stdBrick"):
        self.ctx = ctx
        self.neq = mesh.neq
        keep = []
        self._params = elastic_array(params)
        d = BatchDesc()
        d.count = mesh.nelem
        d.padded = mesh.padded
        d.node = _arr(keep, mesh.node, np.int32)
        d.eq = _arr(keep, mesh.eq, np.int32)
        d.param = _arr(keep, mesh.param, np.int32)
        d.nparam = len(params)
        d.params = C.cast(self._params, C.c_void_p)
        d.nnode = mesh.nnode
        d.coords = _arr(keep, mesh.coords, np.float64)
        d.neq = mesh.neq
        h = C.c_void_p()
        flags = (FX_BATCH_FP32 if fp32 else 0) | (FX_BATCH_GENERAL if general else 0) | \
            (FX_BATCH_XCD_SLABS if xcd else 0) | (0 if tune else FX_BATCH_NO_TUNE) | (FX_BATCH_DYNAMIC if dynamic else 0) | \
            (FX_BATCH_GMRES if gmres else 0) | (0 if tile8 is None else FX_BATCH_TILE8 if tile8 else FX_BATCH_TILE4) | \
            (FX_BATCH_SLOTS if slots else 0)
        if element not in ("stdBrick", "SSPbrick"):
            raise ValueError(f"element {element!r}: stdBrick or SSPbrick")
        if element == "SSPbrick":
            flags |= FX_BATCH_SSPBRICK
        if tile_depth is not None:  # cells per lattice tile in z: 4, 8 or 16
            flags |= {4: FX_BATCH_TILE4, 8: FX_BATCH_TILE8, 16: FX_BATCH_TILE16}[tile_depth]
        _check(lib().fx_batch_create(ctx.h, C.byref(d), flags, C.byref(h)), ctx)
        self.h = h
        ctx.adopt(self)

    @classmethod
    def pimy(cls, ctx, mesh, materials, compact=False, rayleigh=None, fp32=False, gmres=False, element="stdBrick"):
        """A stdBrick<PressureIndependMultiYield> batch (fx_pimy_batch_create), or SSPbrick's with
        element="SSPbrick"; materials: one fx_pimy_material per compiled parameter set; compact: the
        compacted return map; rayleigh: per parameter set (alphaM, betaK, betaKinit, betaKcomm), or
        None."""
        self = cls.__new__(cls)
        self.ctx = ctx
        self.neq = mesh.neq
        self.npts = 8 * mesh.nelem
        keep = []
        self._params = (PimyMaterial * len(materials))(*materials)
        d = PimyBatchDesc()
        d.count = mesh.nelem
        d.padded = mesh.padded
        d.node = _arr(keep, mesh.node, np.int32)
        d.eq = _arr(keep, mesh.eq, np.int32)
        d.param = _arr(keep, mesh.param, np.int32)
        d.nparam = len(materials)
        d.params = C.cast(self._params, C.c_void_p)
        d.nnode = mesh.nnode
        d.coords = _arr(keep, mesh.coords, np.float64)
        d.neq = mesh.neq
        if rayleigh is not None:
            self._damp = (Rayleigh * len(materials))(*[Rayleigh(*r) for r in rayleigh])
            d.damp = C.cast(self._damp, C.c_void_p)
        h = C.c_void_p()
        flags = (FX_BATCH_COMPACT if compact else 0) | (FX_BATCH_FP32 if fp32 else 0) | (FX_BATCH_GMRES if gmres else 0)
        if element == "SSPbrick":
            flags |= FX_BATCH_SSPBRICK
        _check(lib().fx_pimy_batch_create(ctx.h, C.byref(d), flags, C.byref(h)), ctx)
        self.h = h
        ctx.adopt(self)
        return self

    def newmark(self, u, v, a, du, dt, nsteps, gamma=0.5, beta=0.25, test=1, tol=1e-10, max_iter=25, cg_rtol=1e-12,
                cg_max_it=10000, ground_dir=-1, ground_accel=None, max_retries=0, num_sublevels=0, num_substeps=2,
                series=None, ground_factor=1.0, time=0.0, mixed=False, inner_rtol=1e-3, static=False, algorithm=0,
                modified_initial=False, factor_once=False, load_increment=None, load_min_increment=0.0,
                load_max_increment=0.0, load_num_iter=1, recorder=None):
        """nsteps Newmark + Newton steps on device arrays u, v, a, du (in place); returns the
        fx_newmark_result as a dict. test: 0 NormUnbalance, 1 NormDispIncr, 2 EnergyIncr.
        series: the excitation's series (below), evaluated at every attempt's end time (needed to
        substep). mixed: the Newton systems by mixed-precision CG (FP32 batches)."""
        ga = None if ground_accel is None else np.ascontiguousarray(ground_accel, np.float64)
        sv = ts = tv = None
        if series is not None:
            # {"values", "dt"[, "factor", "start", "use_last"]}: PathSeries; {"values", "time"[, "factor",
            # "loc"]}: PathTimeSeries, "loc" an int32 array of one entry holding the interval walk
            sv = np.ascontiguousarray(series["values"], np.float64)
            tv = None if series.get("time") is None else np.ascontiguousarray(series["time"], np.float64)
            loc = series.get("loc")
            if loc is not None:
                assert isinstance(loc, np.ndarray) and loc.dtype == np.int32 and loc.size == 1 and loc.flags.c_contiguous
            ts = TimeSeries(sv.ctypes.data, len(sv), series.get("dt", 0.0), series.get("factor", 1.0),
                            series.get("start", 0.0), int(bool(series.get("use_last", False))),
                            None if tv is None else tv.ctypes.data, None if loc is None else loc.ctypes.data)
        d = NewmarkDesc(gamma, beta, dt, nsteps, test, tol, max_iter, cg_rtol, cg_max_it, ground_dir,
                        None if ga is None else ga.ctypes.data, max_retries, num_sublevels, num_substeps,
                        None if ts is None else C.cast(C.pointer(ts), C.c_void_p), ground_factor, time,
                        int(bool(mixed)), inner_rtol, int(bool(static)), algorithm, int(bool(modified_initial)),
                        int(bool(factor_once)), dt if load_increment is None else load_increment, load_min_increment,
                        load_max_increment, load_num_iter, None if recorder is None else recorder.h)
        r = NewmarkResult()
        _check(lib().fx_newmark(self.h, C.byref(d), u.ptr, v.ptr, a.ptr, du.ptr, C.byref(r)), self.ctx)
        return {f: getattr(r, f) for f, _ in NewmarkResult._fields_}

    def set_stage(self, stage):
        _check(lib().fx_set_material_stage(self.h, stage), self.ctx)

    def store_tangents(self, u, f=None):
        _check(lib().fx_store_tangents(self.h, u.ptr, f.ptr if f is not None else None), self.ctx)

    def commit(self):
        _check(lib().fx_commit_state(self.h), self.ctx)

    def pimy_state(self, committed=True):
        d = DeviceArray(self.ctx, PIMY_NSTATE * self.npts)
        _check(lib().fx_pimy_state(self.h, int(committed), d.ptr), self.ctx)
        return d.get().reshape(PIMY_NSTATE, self.npts).T

    def close(self):
        if self.h and self.ctx.h:
            lib().fx_batch_destroy(self.h)
        self.h = None

    def release(self):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self):
        i = BatchInfo()
        _check(lib().fx_batch_info_get(self.h, C.byref(i)), self.ctx)
        return {f: getattr(i, f) for f, _ in i._fields_}

    def tune(self, variant=-1):
        _check(lib().fx_batch_tune(self.h, variant), self.ctx)
        return self.info()["variant"]

    def apply(self, x, y, k=None):
        _check(lib().fx_operator_apply(self.h, C.byref(k or coeffs()), x.ptr, y.ptr), self.ctx)

    def apply_f32(self, x, y, k=None):
        _check(lib().fx_operator_apply_f32(self.h, C.byref(k or coeffs()), x.ptr, y.ptr), self.ctx)

    def diagonal(self, d, k=None):
        _check(lib().fx_operator_diagonal(self.h, C.byref(k or coeffs()), d.ptr), self.ctx)

    def run_operator(self, out, x=None, y=None, slots=None, slot_offset=0, diagonal=False, unit_dir=-1, k=None):
        """fx_run_operator (InternalForce::Batch::run_operator): out FX_OUT_SLOTS writes this batch's
        element values into slots (a DeviceArray) at slot_offset + a * padded + e; FX_OUT_PLAIN /
        FX_OUT_ATOMIC add into y."""
        _check(lib().fx_run_operator(self.h, int(bool(diagonal)), C.byref(k or coeffs()), out,
                                     x.ptr if x is not None else None, y.ptr if y is not None else None,
                                     slots.ptr if slots is not None else None, slot_offset, unit_dir), self.ctx)

    def set_nodal_mass(self, mass=None, alpha_mass=None):
        m = None if mass is None else np.ascontiguousarray(mass, np.float64)
        am = None if alpha_mass is None else np.ascontiguousarray(alpha_mass, np.float64)
        _check(lib().fx_batch_set_nodal_mass(self.h, None if m is None else m.ctypes.data,
                                             None if am is None else am.ctypes.data), self.ctx)

    def mass_times_ground(self, direction, y):
        _check(lib().fx_mass_times_ground(self.h, direction, y.ptr), self.ctx)

    def time_apply(self, x, y, reps, precision=64, k=None):
        ms = C.c_float()
        _check(lib().fx_operator_time(self.h, C.byref(k or coeffs()), precision, x.ptr, y.ptr, reps, C.byref(ms)),
               self.ctx)
        return ms.value

    def cg(self, b, x, rtol=1e-8, atol=0.0, max_it=10000, k=None):
        o = SolveOptions(rtol, atol, max_it)
        r = SolveResult()
        _check(lib().fx_cg(self.h, C.byref(k or coeffs()), C.byref(o), b.ptr, x.ptr, C.byref(r)), self.ctx)
        return r.as_dict()

    def gmres(self, b, x, rtol=1e-8, atol=0.0, max_it=10000, k=None):
        o = SolveOptions(rtol, atol, max_it)
        r = SolveResult()
        _check(lib().fx_gmres(self.h, C.byref(k or coeffs()), C.byref(o), b.ptr, x.ptr, C.byref(r)), self.ctx)
        return r.as_dict()

    def krylov(self, b, x, rtol=1e-8, atol=0.0, max_it=10000, k=None):
        o = SolveOptions(rtol, atol, max_it)
        r = SolveResult()
        _check(lib().fx_krylov(self.h, C.byref(k or coeffs()), C.byref(o), b.ptr, x.ptr, C.byref(r)), self.ctx)
        return r.as_dict()

    def cg_mixed(self, b, x, rtol=1e-8, atol=0.0, max_it=100000, inner_rtol=1e-3, k=None):
        o = SolveOptions(rtol, atol, max_it)
        r = SolveResult()
        _check(lib().fx_cg_mixed(self.h, C.byref(k or coeffs()), C.byref(o), inner_rtol, b.ptr, x.ptr, C.byref(r)),
               self.ctx)
        return r.as_dict()

    def lumped_mass(self, m, am=None):
        _check(lib().fx_lumped_mass(self.h, m.ptr, am.ptr if am is not None else None), self.ctx)

    def body_load(self, fb):
        _check(lib().fx_body_load(self.h, fb.ptr), self.ctx)

    def force(self, u, f, v=None):
        _check(lib().fx_force(self.h, u.ptr, v.ptr if v is not None else None, f.ptr), self.ctx)

    def critical_dt(self, scheme=0):
        dt = C.c_double()
        _check(lib().fx_critical_dt(self.h, scheme, C.byref(dt)), self.ctx)
        return dt.value

    def explicit_cd(self, dt, nsteps, u, um, v, a, dash=None, load_eq=None, load_hist=None, scheme=0, ground_dir=-1,
                    ground_accel=None, series=None, ground_factor=1.0, time=0.0, recorder=None):
        """nsteps explicit steps on device arrays u, um, v, a (in place): scheme 0 CentralDifference
        (um = U(n-1)), 1 ExplicitDifference (um = the element velocity); a uniform excitation in
        ground_dir from ground_accel (per step) or series ({"values", "dt"} or {"values", "time"});
        recorder: a Recorder fed after every step."""
        le = np.ascontiguousarray(load_eq if load_eq is not None else np.zeros(0), np.int32)
        lh = np.ascontiguousarray(load_hist if load_hist is not None else np.zeros(0), np.float64)
        ga = None if ground_accel is None else np.ascontiguousarray(ground_accel, np.float64)
        keep = []
        ts = None
        if series is not None:
            sv = np.ascontiguousarray(series["values"], np.float64)
            tv = None if series.get("time") is None else np.ascontiguousarray(series["time"], np.float64)
            keep += [sv, tv]
            ts = TimeSeries(sv.ctypes.data, len(sv), series.get("dt", 0.0), series.get("factor", 1.0),
                            series.get("start", 0.0), int(bool(series.get("use_last", False))),
                            None if tv is None else tv.ctypes.data, None)
        d = CdDesc(dt, nsteps, dash.ptr if dash is not None else None, len(le), le.ctypes.data if len(le) else None,
                   lh.ctypes.data if len(lh) else None, scheme, ground_dir, None if ga is None else ga.ctypes.data,
                   None if ts is None else C.cast(C.pointer(ts), C.c_void_p), ground_factor, time,
                   None if recorder is None else recorder.h)
        _check(lib().fx_explicit_cd(self.h, C.byref(d), u.ptr, um.ptr, v.ptr, a.ptr), self.ctx)


class Recorder:
    """Node recorders (fx_recorder_*): OpenSees `recorder Node` text files written from device
    vectors by a writer thread, the values taken off the device by asynchronous copies."""

    RESPONSES = {"disp": 0, "vel": 1, "accel": 2}

    def __init__(self, ctx):
        self.ctx = ctx
        h = vp()
        _check(lib().fx_recorder_create(ctx.h, C.byref(h)), ctx)
        self.h = h

    def add_node(self, file, response, eqs, time=True, dT=0.0, precision=6):
        """A Node recorder of equations `eqs` (Mesh.recorder_columns); returns its index."""
        e = np.ascontiguousarray(eqs, np.int32)
        idx = C.c_int32()
        _check(lib().fx_recorder_add_node(self.h, os.fsencode(file), self.RESPONSES[response], int(bool(time)),
                                          float(dT), int(precision), len(e), e.ctypes.data if len(e) else None,
                                          C.byref(idx)), self.ctx)
        return idx.value

    def activate(self, index):
        _check(lib().fx_recorder_activate(self.h, index), self.ctx)

    def record(self, time, u=None, v=None, a=None):
        """Records every due recorder from device arrays u, v, a (one DeviceArray at least)."""
        n = next(x.n for x in (u, v, a) if x is not None)
        _check(lib().fx_recorder_record(self.h, float(time), n, *(None if x is None else x.ptr for x in (u, v, a))),
               self.ctx)

    def finish(self):
        _check(lib().fx_recorder_finish(self.h), self.ctx)

    @property
    def records(self):
        return lib().fx_recorder_records(self.h)

    def close(self):
        if self.h:
            lib().fx_recorder_destroy(self.h)
            self.h = None


def copy(ctx, src, dst):
    _check(lib().fx_copy(ctx.h, src.n, src.ptr, dst.ptr), ctx)


def dot(ctx, x, y):
    out = C.c_double()
    _check(lib().fx_dot(ctx.h, x.n, x.ptr, y.ptr, C.byref(out)), ctx)
    return out.value


class ColumnPool:
    """fx_column_pool: independent 1 x 1 x N PressureIndependMultiYield columns (laminar ties,
    base fixed), each with its own Newmark + Newton + PCG analysis, advanced together."""

    def __init__(self, ctx, max_elements, max_columns, dt, max_materials=None, max_record=None, gamma=0.5,
                 beta=0.25, test=1, tol=1e-10, max_iter=25, cg_rtol=1e-12, cg_max_it=10000, graph=True, mixed=False,
                 inner_rtol=1e-5, compact=False, waves=0, one_stream=False, max_retries=0, num_sublevels=0,
                 num_substeps=2, tune=False):
        self.ctx = ctx
        flags = ((0 if graph else FX_COLUMN_NO_GRAPH) | (FX_COLUMN_MIXED if mixed else 0) |
                 (FX_COLUMN_COMPACT if compact else 0) | (int(waves) << 8) | (8 if one_stream else 0) |
                 (FX_COLUMN_TUNE if tune else 0))
        d = ColumnPoolDesc(int(max_elements), int(max_columns), int(max_materials or 16 * max_columns),
                           int(max_record or 1 << 20), gamma, beta, dt, test, tol, max_iter, cg_rtol, cg_max_it, flags,
                           inner_rtol, max_retries, num_sublevels, num_substeps)
        h = C.c_void_p()
        _check(lib().fx_column_pool_create(ctx.h, C.byref(d), C.byref(h)), ctx)
        self.h = h
        self.columns = {}  # id -> (nelem, nsteps, record)
        ctx.adopt(self)

    def add(self, mats, mat_of, nsteps, ground_accel, h=(1.0, 1.0, 1.0), stage=1, ground_dir=0, record=True):
        """Admit a column (mats: PimyMaterial list, mat_of: parameter set per element bottom to
        top). Returns its id, or None when the pool has no room now (FX_EFULL)."""
        mo = np.ascontiguousarray(mat_of, np.int32)
        ga = np.ascontiguousarray(ground_accel, np.float64)
        arr = (PimyMaterial * len(mats))(*mats)
        d = ColumnDesc(len(mo), h[0], h[1], h[2], len(mats), C.cast(arr, C.c_void_p), mo.ctypes.data, stage,
                       int(nsteps), ground_dir, ga.ctypes.data, int(bool(record)))
        cid = C.c_int32()
        rc = lib().fx_column_add(self.h, C.byref(d), C.byref(cid))
        if rc == FX_EFULL:
            return None
        _check(rc, self.ctx)
        self.columns[cid.value] = (len(mo), int(nsteps), bool(record))
        return cid.value

    def run(self, iterations=1):
        """Newton iterations of every running column; returns the number still running."""
        n = C.c_int32()
        _check(lib().fx_column_run(self.h, int(iterations), C.byref(n)), self.ctx)
        return n.value

    def finished(self):
        """Ids of the columns finished or failed at the end of the last run (not removed)."""
        cap = max(1, len(self.columns))
        ids = np.zeros(cap, np.int32)
        n = C.c_int32()
        _check(lib().fx_column_finished(self.h, ids.ctypes.data, cap, C.byref(n)), self.ctx)
        return [int(i) for i in ids[:min(n.value, cap)]]

    def status(self, cid):
        r = ColumnResult()
        _check(lib().fx_column_status(self.h, cid, C.byref(r)), self.ctx)
        return {f: getattr(r, f) for f, _ in ColumnResult._fields_}

    def fetch(self, cid, history=True):
        n, nsteps, rec = self.columns[cid]
        u, v, a = np.zeros(3 * n), np.zeros(3 * n), np.zeros(3 * n)
        hist = np.zeros(9 * nsteps) if (history and rec) else None
        _check(lib().fx_column_fetch(self.h, cid, u.ctypes.data, v.ctypes.data, a.ctypes.data,
                                     hist.ctypes.data if hist is not None else None), self.ctx)
        out = dict(u=u.reshape(n, 3), v=v.reshape(n, 3), a=a.reshape(n, 3))
        if hist is not None:
            out["history"] = hist.reshape(nsteps, 9)
        return out

    def state(self, cid):
        n = self.columns[cid][0]
        s = np.zeros(PIMY_NSTATE * n)
        _check(lib().fx_column_state(self.h, cid, s.ctypes.data), self.ctx)
        return s.reshape(PIMY_NSTATE, n).T

    def remove(self, cid):
        _check(lib().fx_column_remove(self.h, cid), self.ctx)
        del self.columns[cid]

    def close(self):
        if self.h and self.ctx.h:
            lib().fx_column_pool_destroy(self.h)
        self.h = None

    def release(self):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
