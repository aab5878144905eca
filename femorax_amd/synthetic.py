"""Synthetic models, built in memory the way the reference's tests build them
(tests/common/synthetic_models.hpp:24-94): structured stdBrick grids with fixed bases and
laminar equalDOF ties, materials per layer from the reference's 64-bit LCG
(tests/reference/drivers/driver_common.hpp:21-32), and the load vectors the configurations
need (consistent surface tractions, body force). There is no model file on this path: the
reference's own tests and benchmarks build their right-hand sides in C++ the same way
(tests/unit/test_cg.cpp:66-80, benchmarks/bench_operator.cpp:113-132).

A model is a dict:
  nnode, xyz [nnode, 3] (model order: i fastest, then j, then k; tag = index + 1)
  nelem, conn [nelem, 8] (model node indices, OpenSees Brick order), group [nelem]
  fix [n, 2] (node, dof 0..2), tie [n, 3] (retained, constrained, dof 0..2)
  params: list of per-group dicts (E, nu, rho, b, mass_type, rayleigh)
  grid: (nx, ny, nz) element counts, h: (dx, dy, dz)
"""
import numpy as np

SEED_BASE = 20261202


class Lcg:
    """The reference's platform-independent LCG: uniform in [-1, 1)."""

    def __init__(self, seed):
        self.s = seed & 0xFFFFFFFFFFFFFFFF

    def symmetric(self):
        self.s = (self.s * 6364136223846793005 + 1442695040888963407) & 0xFFFFFFFFFFFFFFFF
        return (self.s >> 11) * (2.0 / 9007199254740992.0) - 1.0

    def uniform(self, lo, hi):
        return lo + (hi - lo) * 0.5 * (self.symmetric() + 1.0)

    def vector(self, n):
        return np.array([self.symmetric() for _ in range(n)])


def grid(nx, ny, nz, dx, dy, dz, z0=0.0):
    """Nodes and hex8 connectivity of an nx x ny x nz grid (model order, i fastest)."""
    NX, NY, NZ = nx + 1, ny + 1, nz + 1
    k, j, i = np.meshgrid(np.arange(NZ), np.arange(NY), np.arange(NX), indexing="ij")
    xyz = np.stack([i.ravel() * dx, j.ravel() * dy, z0 + k.ravel() * dz], axis=1).astype(np.float64)
    ek, ej, ei = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    ei, ej, ek = ei.ravel(), ej.ravel(), ek.ravel()

    def nid(a, b, c):
        return a + NX * (b + NY * c)

    conn = np.stack([nid(ei, ej, ek), nid(ei + 1, ej, ek), nid(ei + 1, ej + 1, ek), nid(ei, ej + 1, ek),
                     nid(ei, ej, ek + 1), nid(ei + 1, ej, ek + 1), nid(ei + 1, ej + 1, ek + 1),
                     nid(ei, ej + 1, ek + 1)], axis=1).astype(np.int32)
    return xyz, conn, ek.astype(np.int32)


def fixed_base(nx, ny):
    NX, NY = nx + 1, ny + 1
    base = np.arange(NX * NY, dtype=np.int32)
    return np.stack([np.repeat(base, 3), np.tile(np.arange(3, dtype=np.int32), len(base))], axis=1)


def elastic_box(n=16, L=1.0, E=30e6, nu=0.3, rho=2000.0, layers=None, body=False, seed=None):
    """An n^3 box on [0, L]^3, base fixed. layers: None (one material) or a count of equal
    layers whose E (log-uniform 20-200 MPa) and nu (uniform 0.25-0.45) come from the LCG."""
    h = L / n
    xyz, conn, ek = grid(n, n, n, h, h, h)
    params = []
    if layers:
        rng = Lcg(seed if seed is not None else SEED_BASE + 1)
        per = n // layers
        group = np.minimum(ek // per, layers - 1).astype(np.int32)
        for _ in range(layers):
            Ei = float(np.exp(rng.uniform(np.log(20e6), np.log(200e6))))
            nui = rng.uniform(0.25, 0.45)
            b = (0.0, 0.0, -rho * 9.81) if body else (0.0, 0.0, 0.0)
            params.append(dict(E=Ei, nu=nui, rho=rho, b=b, mass_type=0))
    else:
        group = np.zeros(len(conn), np.int32)
        b = (0.0, 0.0, -rho * 9.81) if body else (0.0, 0.0, 0.0)
        params.append(dict(E=E, nu=nu, rho=rho, b=b, mass_type=0))
    return dict(nnode=len(xyz), xyz=xyz, nelem=len(conn), conn=conn, group=group,
                fix=fixed_base(n, n), tie=np.zeros((0, 3), np.int32), params=params,
                grid=(n, n, n), h=(h, h, h))


def distorted_box(n=128, L=None, amp=0.2, layers=8, seed=SEED_BASE + 5):
    """The layered n^3 box of elastic_box (h = 1 m unless L is given) with every interior node
    moved by up to amp * h in each direction (numpy's generator from `seed`): no element is an
    axis-aligned box, so the operator runs on the general-geometry path."""
    L = float(n) if L is None else L
    m = elastic_box(n, L, layers=layers, body=True, seed=seed)
    h = L / n
    N1 = n + 1
    idx = np.arange(m["nnode"])
    i, j, k = idx % N1, (idx // N1) % N1, idx // (N1 * N1)
    inner = (i > 0) & (i < n) & (j > 0) & (j < n) & (k > 0) & (k < n)
    rng = np.random.default_rng(seed)
    xyz = m["xyz"].copy()
    xyz[inner] += amp * h * rng.uniform(-1.0, 1.0, size=(int(inner.sum()), 3))
    m["xyz"] = xyz
    return m


def soil_column(nz, h=1.0, params=None, layer_of=None, nx=1, ny=1):
    """A 1x1xN (or nx x ny x N) column: base fixed, side nodes of each level tied to the first
    side node in DOFs 1 2 3 (the laminar boundary of synthetic_models.hpp:64-79)."""
    xyz, conn, ek = grid(nx, ny, nz, h, h, h, z0=-nz * h)
    NX, NY = nx + 1, ny + 1
    ties = []
    for k in range(1, nz + 1):
        master = -1
        for j in range(NY):
            for i in range(NX):
                if not (i == 0 or i == nx or j == 0 or j == ny):
                    continue
                t = i + NX * (j + NY * k)
                if master < 0:
                    master = t
                    continue
                for d in range(3):
                    ties.append((master, t, d))
    group = (layer_of(ek) if layer_of is not None else np.zeros(len(conn), np.int32)).astype(np.int32)
    return dict(nnode=len(xyz), xyz=xyz, nelem=len(conn), conn=conn, group=group, fix=fixed_base(nx, ny),
                tie=np.array(ties, np.int32).reshape(-1, 3), params=params or [], grid=(nx, ny, nz),
                h=(h, h, h))


def reference_column(side=40, lumped=False, damping=True):
    """The reference's operator-benchmark model (tests/common/synthetic_models.hpp:24-94 with
    benchmarks/bench_operator.cpp:71-84): side^3 stdBrick grid of 1 x 1 x 0.5 m bricks, base
    fixed, the side nodes of each elevation tied to the first one in DOFs 1 2 3, two
    ElasticIsotropic materials (lower and upper half) with their body forces, consistent mass
    unless lumped, and two Rayleigh regions: one over every element (alphaM 0.6, betaK 5.8e-4)
    and one over the top element layer (alphaM 0.9, betaK 3e-4, betaKinit 1e-4; the later region
    wins). Groups: 0 lower half, 1 upper half below the top layer, 2 the top layer."""
    nx = ny = nz = side
    dz = 0.5
    xyz, conn, ek = grid(nx, ny, nz, 1.0, 1.0, dz, z0=-nz * dz)
    NX, NY = nx + 1, ny + 1
    ties = []
    for k in range(1, nz + 1):
        master = -1
        for j in range(NY):
            for i in range(NX):
                if not (i == 0 or i == nx or j == 0 or j == ny):
                    continue
                t = i + NX * (j + NY * k)
                if master < 0:
                    master = t
                    continue
                for d in range(3):
                    ties.append((master, t, d))
    group = np.where(ek < nz // 2, 0, np.where(ek < nz - 1, 1, 2)).astype(np.int32)
    mass_type = 1 if lumped else 0
    ray_all = (0.6, 5.8e-4, 0.0, 0.0) if damping else (0.0, 0.0, 0.0, 0.0)
    ray_top = (0.9, 3e-4, 1e-4, 0.0) if damping else (0.0, 0.0, 0.0, 0.0)
    lower = dict(E=377000.0, nu=0.3, rho=2.0285, b=(0.0, 0.0, -19.9), mass_type=mass_type, rayleigh=ray_all)
    upper = dict(E=109200.0, nu=0.3, rho=2.0183, b=(0.0, 0.0, -19.8), mass_type=mass_type, rayleigh=ray_all)
    top = dict(upper, rayleigh=ray_top)
    return dict(nnode=len(xyz), xyz=xyz, nelem=len(conn), conn=conn, group=group, fix=fixed_base(nx, ny),
                tie=np.array(ties, np.int32).reshape(-1, 3), params=[lower, upper, top], grid=(nx, ny, nz),
                h=(1.0, 1.0, dz))


# the operator coefficients of the reference's operator benchmark (bench_operator.cpp:84)
REFERENCE_COLUMN_COEFFS = (1.0, 2e3, 1.6e7)


def top_traction_nodal(model, traction_z):
    """Consistent nodal loads of a -z traction on the top face (z = max): each top quad gives
    t * A / 4 to its 4 corners. traction_z: scalar or per-top-node array (model order of the
    top layer). Returns [nnode, 3]."""
    nx, ny, nz = model["grid"]
    dx, dy, _ = model["h"]
    NX, NY = nx + 1, ny + 1
    F = np.zeros((model["nnode"], 3))
    top0 = NX * NY * nz
    t = np.broadcast_to(np.asarray(traction_z, dtype=np.float64), (NX * NY,))
    A = dx * dy
    for j in range(ny):
        for i in range(nx):
            corners = [i + NX * j, i + 1 + NX * j, i + 1 + NX * (j + 1), i + NX * (j + 1)]
            tq = np.mean(t[corners])
            for c in corners:
                F[top0 + c, 2] += -tq * A / 4.0
    return F


def nodal_to_equations(F, node_rank, node_eq, neq):
    """Sum nodal loads [nnode, 3] (model order) into equation space; fixed DOFs drop out and
    tied DOFs add onto their shared equation."""
    b = np.zeros(neq)
    eqs = node_eq.reshape(-1, 3)[node_rank]  # model node -> its equations
    for d in range(3):
        q = eqs[:, d]
        m = q >= 0
        np.add.at(b, q[m], F[m, d])
    return b


def config_block(n):
    """n^3 static box: E 30 MPa, nu 0.3, 1 m cube, base fixed, 10 kPa top traction (the C1 loads)."""
    m = elastic_box(n, 1.0, 30e6, 0.3, 2000.0)
    m["nodal_load"] = top_traction_nodal(m, 10e3)
    return m


def config_c1():
    """16^3 static box: E 30 MPa, nu 0.3, 1 m cube, base fixed, 10 kPa top traction."""
    return config_block(16)


def config_c2(n=128, seed=SEED_BASE + 1):
    """n^3 layered box on [0, n]^3 m (8 layers, E log-uniform 20-200 MPa, nu 0.25-0.45 from
    the LCG), body force, random top traction 10 kPa (1 + 0.5 u)."""
    m = elastic_box(n, float(n), layers=8, body=True, seed=seed)
    rng = Lcg(seed + 1000)
    NX = n + 1
    t = np.array([1e4 * (1.0 + 0.5 * rng.symmetric()) for _ in range(NX * NX)])
    m["nodal_load"] = top_traction_nodal(m, t)
    return m


def lysmer_dashpots(model, vs_of_group, nu=0.3):
    """Lysmer-Kuhlemeyer boundary dashpots on the four sides and the bottom of a grid model:
    each boundary face of an element gives its 4 corners rho Vp A/4 in the face's normal
    direction and rho Vs A/4 in the two tangential ones (Vp = Vs sqrt(2(1-nu)/(1-2nu))).
    Returns nodal dashpot coefficients [nnode, 3] (model order)."""
    nx, ny, nz = model["grid"]
    dx, dy, dz = model["h"]
    NX, NY = nx + 1, ny + 1
    C = np.zeros((model["nnode"], 3))
    ratio = np.sqrt(2.0 * (1.0 - nu) / (1.0 - 2.0 * nu))
    conn = model["conn"]
    group = model["group"]
    params = model["params"]
    ek = np.arange(model["nelem"]) // (nx * ny)
    ej = (np.arange(model["nelem"]) // nx) % ny
    ei = np.arange(model["nelem"]) % nx
    # (selector, face corners in Brick order, normal axis, face area)
    faces = [(ei == 0, [0, 3, 7, 4], 0, dy * dz), (ei == nx - 1, [1, 2, 6, 5], 0, dy * dz),
             (ej == 0, [0, 1, 5, 4], 1, dx * dz), (ej == ny - 1, [3, 2, 6, 7], 1, dx * dz),
             (ek == 0, [0, 1, 2, 3], 2, dx * dy)]
    for sel, corners, axis, area in faces:
        for e in np.nonzero(sel)[0]:
            g = group[e]
            rho = params[g]["rho"]
            vs = vs_of_group[g]
            for c in corners:
                n = conn[e, c]
                for d in range(3):
                    C[n, d] += rho * (vs * ratio if d == axis else vs) * area / 4.0
    return C


def ricker(t, f=5.0, amplitude=1.0e6):
    """Ricker wavelet f0 (1 - 2 pi^2 f^2 tau^2) exp(-pi^2 f^2 tau^2), tau = t - 1.2 / f."""
    tau = np.asarray(t) - 1.2 / f
    a = (np.pi * f * tau) ** 2
    return amplitude * (1.0 - 2.0 * a) * np.exp(-a)


def config_c3(nx=200, ny=200, nz=100, h=5.0, layers=10, seed=SEED_BASE + 2):
    """Explicit wave propagation box: nx x ny x nz bricks of h, `layers` equal layers with
    Vs ~ U(150, 800) m/s, rho ~ U(1800, 2200) kg/m^3, nu 0.3 (G = rho Vs^2, E = 2G(1+nu)),
    lumped mass, no fixed DOFs, Lysmer dashpots on the sides and bottom, a 5 Hz Ricker point
    force (z) at the centre node of the surface."""
    xyz, conn, ek = grid(nx, ny, nz, h, h, h, z0=-nz * h)
    rng = Lcg(seed)
    per = max(1, nz // layers)
    # the top layer is group layers-1 counted from the base, as the grid counts k from the base
    group = np.minimum(ek // per, layers - 1).astype(np.int32)
    params, vs = [], []
    for _ in range(layers):
        v = rng.uniform(150.0, 800.0)
        rho = rng.uniform(1800.0, 2200.0)
        G = rho * v * v
        params.append(dict(E=2.0 * G * 1.3, nu=0.3, rho=rho, b=(0.0, 0.0, 0.0), mass_type=1))
        vs.append(v)
    m = dict(nnode=len(xyz), xyz=xyz, nelem=len(conn), conn=conn, group=group, fix=np.zeros((0, 2), np.int32),
             tie=np.zeros((0, 3), np.int32), params=params, grid=(nx, ny, nz), h=(h, h, h))
    m["vs"] = vs
    m["dashpots"] = lysmer_dashpots(m, vs)
    NX, NY = nx + 1, ny + 1
    m["source_node"] = (nx // 2) + NX * ((ny // 2) + NY * nz)
    return m


def ground_motion(nsteps, dt, seed, pga=0.3 * 9.81, band=(0.1, 20.0)):
    """Base acceleration: LCG white noise, 4th-order Butterworth band-pass (zero-phase: forward
    and backward), scaled to the peak ground acceleration `pga`. Values at the end of each
    step (t = (k + 1) dt)."""
    from scipy.signal import butter, sosfiltfilt

    rng = Lcg(seed)
    length = max(nsteps, 1024)  # short records: the first nsteps values of a 1024-step record
    w = np.array([rng.symmetric() for _ in range(length)])
    sos = butter(4, [band[0], band[1]], btype="bandpass", fs=1.0 / dt, output="sos")
    g = sosfiltfilt(sos, w)[:nsteps]
    return g * (pga / np.max(np.abs(g)))


def soil_layers(rng, nlayer=10):
    """PressureIndependMultiYield layers: Gmax ~ U(10, 300) MPa, K = 2.17 Gmax, cohesion
    ~ U(20, 150) kPa, rho 2000, phi 0, gamma_peak 0.1, 20 automatic surfaces, lumped mass, no
    body force."""
    out = []
    for _ in range(nlayer):
        G = rng.uniform(10e6, 300e6)
        out.append(dict(rho=2000.0, G=G, K=2.17 * G, cohesion=rng.uniform(20e3, 150e3), gamma_peak=0.1, phi=0.0,
                        pref=100e3, pdc=0.0, nsurf=20, b=(0.0, 0.0, 0.0), mass_type=1))
    return out


def column_layer_of(nelem, nlayer=10):
    """Layer of each element of a column, bottom to top: `nlayer` nearly equal layers."""
    return np.minimum(np.arange(nelem) * nlayer // nelem, nlayer - 1).astype(np.int32)


def config_c4(ncol=4096, nsteps=4000, dt=0.005, seed=SEED_BASE + 3):
    """Batched site response: `ncol` columns 1 x 1 x N (h = 1 m), N ~ U{20..400}, 10 layers per
    column (soil_layers), stage 1 from t = 0, one base record (ground_motion, 0.3 g) in x.
    Returns (columns, record): columns = list of dict(nelem, mats, mat_of)."""
    rng = Lcg(seed)
    cols = []
    for _ in range(ncol):
        n = 20 + int((rng.symmetric() + 1.0) * 0.5 * 381) % 381
        mats = soil_layers(rng, 10)
        cols.append(dict(nelem=n, mats=mats, mat_of=column_layer_of(n, 10)))
    return cols, ground_motion(nsteps, dt, seed + 1000)
