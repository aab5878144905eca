"""Small seeded meshes shared by the tests (sizes the oracle finishes in well under a second)."""
import numpy as np

from femorax_amd import synthetic as S


def distorted_box(n=6, seed=7, amp=0.18, layers=2):
    """An n^3 grid whose interior nodes are moved by up to amp*h (LCG), base fixed, with
    several materials: exercises the general (non-affine) geometry."""
    m = S.elastic_box(n, 1.0, layers=layers, seed=seed, body=True)
    rng = S.Lcg(seed + 99)
    h = 1.0 / n
    xyz = m["xyz"].copy()
    N1 = n + 1
    for idx in range(len(xyz)):
        i, j, k = idx % N1, (idx // N1) % N1, idx // (N1 * N1)
        if 0 < i < n and 0 < j < n and 0 < k < n:
            xyz[idx] += amp * h * np.array([rng.symmetric(), rng.symmetric(), rng.symmetric()])
    m["xyz"] = xyz
    return m


def column(nz=24, seed=11):
    """A 1x1xN tied soil column with 3 layers (laminar equalDOF ties, base fixed)."""
    rng = S.Lcg(seed)
    params = [dict(E=rng.uniform(50e6, 300e6), nu=0.3, rho=2000.0, b=(0.0, 0.0, -19.6e3), mass_type=1)
              for _ in range(3)]
    return S.soil_column(nz, 1.0, params=params, layer_of=lambda k: np.minimum(k * 3 // nz, 2))


def partly_exploded(n=12, lo=0.3, hi=0.7):
    """The distorted n^3 box whose elements with centroid height in [lo, hi) get their own copies
    of their eight nodes (no node shared with any other element): a block of such elements has
    8 local nodes per element, more than the general path stages in LDS, so one apply runs both
    the staged blocks and the per-element-gather blocks (and blocks with some of each)."""
    m = distorted_box(n)
    xyz = m["xyz"]
    conn = m["conn"].copy()
    zc = xyz[conn, 2].mean(axis=1)
    cut = np.nonzero((zc >= lo) & (zc < hi))[0]
    extra = xyz[conn[cut].ravel()]
    conn[cut] = m["nnode"] + np.arange(8 * len(cut), dtype=conn.dtype).reshape(-1, 8)
    m["xyz"] = np.concatenate([xyz, extra])
    m["nnode"] = len(m["xyz"])
    m["conn"] = conn
    return m


def random_vector(n, seed):
    rng = S.Lcg(seed)
    return np.array([rng.symmetric() for _ in range(n)])
