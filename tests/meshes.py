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


def random_vector(n, seed):
    rng = S.Lcg(seed)
    return np.array([rng.symmetric() for _ in range(n)])
