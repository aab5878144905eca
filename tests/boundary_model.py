"""A model of two element batches over one equation space, as the reference's compile step
makes one batch per (element type, material) pair (compile/batches.cpp:54-227): the lower layer
stdBrick<ElasticIsotropic>, the upper layer stdBrick<PressureIndependMultiYield>, on a distorted
4^3 box (general geometry), with nodal masses on some nodes. The oracle's expectations follow the
reference's InternalForce::apply_operator / operator_diagonal / mass_times_ground
(internal_force.cpp:442-448, 545-626): every batch's element values in (batch, element, DOF)
order (SlotGather, assembly.cpp:30-38), then the nodal masses.
"""
import ctypes as C
import types

import numpy as np

from femorax_amd import core
from oracle import fxo
from tests import meshes

PAD = 64


def split_batches(mesh):
    """The compiled mesh's elements split by group into batches (group 0, group 1), each with its
    own padded maps value[k * padded + e] and parameter sets."""
    P = mesh.padded
    out = []
    for g in (0, 1):
        sets = [i for i in range(mesh.nparam) if mesh.param_group[i] == g]
        local = {s: j for j, s in enumerate(sets)}
        elems = [e for e in range(mesh.nelem) if mesh.param_group[mesh.param[e]] == g]
        n = len(elems)
        Pb = max(PAD, (n + PAD - 1) // PAD * PAD)
        node = np.full(8 * Pb, -1, np.int32)
        eq = np.full(24 * Pb, -1, np.int32)
        param = np.zeros(Pb, np.int32)
        for i, e in enumerate(elems):
            for k in range(8):
                node[k * Pb + i] = mesh.node[k * P + e]
            for a in range(24):
                eq[a * Pb + i] = mesh.eq[a * P + e]
            param[i] = local[mesh.param[e]]
        out.append(types.SimpleNamespace(nelem=n, padded=Pb, node=node, eq=eq, param=param, nnode=mesh.nnode,
                                         coords=mesh.coords, neq=mesh.neq, sets=sets))
    return out


def soil(n):
    out = []
    for i in range(n):
        G = 40e6 + 30e6 * i
        out.append(dict(rho=1900.0, G=G, K=2.17 * G, cohesion=30e3 + 10e3 * i, gamma_peak=0.1, phi=0.0, pref=100e3,
                        pdc=0.0, nsurf=20, b=(0.0, 0.0, -1900.0 * 9.81), mass_type=0))
    return out


class TwoBatchModel:
    """Host data of the model, its oracle element functions and its expectations."""

    def __init__(self, n=4, seed=5):
        self.model = meshes.distorted_box(n, layers=2, seed=seed)
        self.mesh = core.Mesh(self.model)
        self.neq = self.mesh.neq
        self.elastic_mesh, self.pimy_mesh = split_batches(self.mesh)
        allp = self.mesh.params_for(self.model)
        self.elastic_params = [allp[s] for s in self.elastic_mesh.sets]
        # Rayleigh damping on the elastic sets, so cC terms are exercised
        for i, p in enumerate(self.elastic_params):
            p["rayleigh"] = (0.3 + 0.1 * i, 2e-3, 0.0, 0.0)
        self.soil = soil(len(self.pimy_mesh.sets))
        rng = np.random.default_rng(seed)
        self.nodal_m = np.where(rng.uniform(size=self.neq) < 0.2, rng.uniform(50.0, 150.0, self.neq), 0.0)
        self.nodal_am = 0.4 * self.nodal_m
        self.x = meshes.random_vector(self.neq, seed + 1)
        self.u = meshes.random_vector(self.neq, seed + 2) * 1e-3  # the PIMY tangents' displacement
        N = self.mesh.nnode
        self.xyz = self.mesh.coords.reshape(3, N).T.copy()
        self._oracle_setup()

    # ---- the oracle, element by element -------------------------------------------------------
    def _oracle_setup(self):
        L = fxo.lib()
        vp = C.c_void_p
        L.fxo_brick_force_pimy.argtypes = [vp, fxo.f64p, fxo.f64p, vp, vp, vp, fxo.f64p, fxo.f64p]
        L.fxo_brick_operator_apply_gp.argtypes = [vp, C.POINTER(fxo.Coeffs), fxo.f64p, fxo.f64p, vp, fxo.f64p,
                                                  fxo.f64p]
        L.fxo_brick_operator_diagonal_gp.argtypes = [vp, C.POINTER(fxo.Coeffs), fxo.f64p, fxo.f64p, vp, fxo.i32p,
                                                     fxo.f64p]
        self.oe = fxo.elastic_params(self.elastic_params)
        self.op = fxo.pimy_elems(self.soil, stage=1)
        self.D = []
        for p in self.elastic_params:
            D = np.zeros(36)
            L.fxo_elastic_tangent(p["E"], p["nu"], D)
            self.D.append(D)
        # the PIMY points' tangents at u from their initial state (stored tangents), and their
        # committed histories (at rest)
        b = self.pimy_mesh
        self.Dgp = []
        self.C0 = []
        for e in range(b.nelem):
            p = self.op[b.param[e]]
            Cs = []
            for _ in range(8):
                s = np.zeros(fxo.PIMY_NSTATE)
                L.fxo_pimy_initial_state(C.byref(p.mat), s.ctypes.data)
                Cs.append(s)
            Ts = [c.copy() for c in Cs]
            Cp = (C.c_void_p * 8)(*[c.ctypes.data for c in Cs])
            Tp = (C.c_void_p * 8)(*[t.ctypes.data for t in Ts])
            xe = self.element_xe(b, e)
            ue = self.element_values(b, e, self.u)
            fe = np.zeros(24)
            Dgp = np.zeros(8 * 36)
            L.fxo_brick_force_pimy(C.cast(C.pointer(p), vp), xe, ue, None, C.cast(Cp, vp), C.cast(Tp, vp), fe, Dgp)
            self.Dgp.append(Dgp)
            self.C0.append(Cs)

    def element_xe(self, b, e):
        return np.ascontiguousarray([self.xyz[b.node[k * b.padded + e]] for k in range(8)], np.float64)

    def element_values(self, b, e, x, unit_dir=-1):
        v = np.zeros(24)
        for a in range(24):
            q = b.eq[a * b.padded + e]
            if unit_dir >= 0:
                v[a] = 1.0 if a % 3 == unit_dir else 0.0
            else:
                v[a] = x[q] if q >= 0 else 0.0
        return v

    def element_eq(self, b, e):
        return np.array([b.eq[a * b.padded + e] for a in range(24)], np.int32)

    def oracle_element(self, which, e, k, diagonal=False, x=None, unit_dir=-1):
        """The 24 values of element e of batch `which` (0 elastic, 1 PIMY)."""
        L = fxo.lib()
        vp = C.c_void_p
        b = self.elastic_mesh if which == 0 else self.pimy_mesh
        xe = self.element_xe(b, e)
        ye = np.zeros(24)
        kk = fxo.coeffs(k.stiffness, k.damping, k.mass, bool(k.initial))
        if which == 0:
            p = self.oe[b.param[e]]
            D = self.D[b.param[e]]
            if diagonal:
                L.fxo_brick_operator_diagonal(C.byref(p), C.byref(kk), xe, D, self.element_eq(b, e), ye)
            else:
                L.fxo_brick_operator_apply(C.byref(p), C.byref(kk), xe, D, self.element_values(b, e, x, unit_dir), ye)
        else:
            p = self.op[b.param[e]]
            Cp = (C.c_void_p * 8)(*[c.ctypes.data for c in self.C0[e]])
            if diagonal:
                L.fxo_brick_operator_diagonal_gp(C.cast(C.pointer(p), vp), C.byref(kk), xe, self.Dgp[e], C.cast(Cp, vp),
                                                 self.element_eq(b, e), ye)
            else:
                L.fxo_brick_operator_apply_gp(C.cast(C.pointer(p), vp), C.byref(kk), xe, self.Dgp[e], C.cast(Cp, vp),
                                              self.element_values(b, e, x, unit_dir), ye)
        return ye

    def expected(self, k, diagonal=False, unit_dir=-1):
        """InternalForce::apply_operator / operator_diagonal (unit_dir < 0) or mass_times_ground
        (unit_dir >= 0, k = {0, 0, 1}): the batches' element values gathered in (batch, element,
        DOF) order, then the nodal masses."""
        y = np.zeros(self.neq)
        for which, b in enumerate((self.elastic_mesh, self.pimy_mesh)):
            for e in range(b.nelem):
                ye = self.oracle_element(which, e, k, diagonal, self.x, unit_dir)
                for a in range(24):
                    q = b.eq[a * b.padded + e]
                    if q >= 0:
                        y[q] += ye[a]
        if unit_dir >= 0:
            dirs = self.eq_dir()
            y += np.where(dirs == unit_dir, self.nodal_m, 0.0)
        else:
            w = np.ones(self.neq) if diagonal else self.x
            if k.mass != 0.0:
                y += k.mass * self.nodal_m * w
            if k.damping != 0.0:
                y += k.damping * self.nodal_am * w
        return y

    def eq_dir(self):
        d = np.full(self.neq, -1)
        for q3, q in enumerate(self.mesh.node_eq):
            if q >= 0:
                d[q] = q3 % 3
        return d

    # ---- the device batches ------------------------------------------------------------------
    def device_batches(self, ctx, slots=True):
        eb = core.Batch(ctx, self.elastic_mesh, self.elastic_params, general=True, slots=slots, tune=False)
        keep = []
        mats = [core.pimy_material(m["rho"], m["G"], m["K"], m["cohesion"], m["gamma_peak"], m["phi"], m["pref"],
                                   m["pdc"], nsurf=m["nsurf"], b=m["b"], mass_type=m["mass_type"], keep=keep)
                for m in self.soil]
        pb = core.Batch.pimy(ctx, self.pimy_mesh, mats)
        pb.set_stage(1)
        du = core.DeviceArray.from_host(ctx, self.u)
        pb.store_tangents(du)
        du.free()
        return eb, pb
