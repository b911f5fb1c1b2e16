"""Seeded synthetic water-cluster geometries for the benchmark workloads.

The reference only reads geometries from its input file (src/GetInput.cpp:256-310,
Waters group: O, H1, H2 lines in Angstrom).  This module writes geometries of
that form: rigid monomers (O-H 0.9572 A, H-O-H 104.52 deg, the TTM values of
src/ttm_data.h:8-9), O positions uniform in a sphere with a minimum O-O distance,
Haar-uniform orientations, all from one numpy seed.
"""
import numpy as np

ROH = 0.9572
HOH = np.deg2rad(104.52)
BOHR2ANGS = 0.52917720859
ANGS2BOHR = 1.0 / BOHR2ANGS


def _monomer():
    half = 0.5 * HOH
    o = np.zeros(3)
    h1 = ROH * np.array([0.0, np.sin(half), -np.cos(half)])
    h2 = ROH * np.array([0.0, -np.sin(half), -np.cos(half)])
    return np.stack([o, h1, h2])


def _haar_rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])


def _in_sphere(rng, radius):
    while True:
        p = rng.uniform(-radius, radius, size=3)
        if p @ p <= radius * radius:
            return p


def water_cluster(nw, radius, seed, min_oo=2.6, ion_min=None, max_tries=200000):
    """Return (xyz[nw, 9] in Angstrom (O,H1,H2), ion position in Angstrom or None).

    ion_min: if given, also place one ion at a random interior point at least
    ion_min Angstrom from every O.
    """
    rng = np.random.default_rng(seed)
    mono = _monomer()
    opos = []
    tries = 0
    while len(opos) < nw:
        tries += 1
        if tries > max_tries:
            raise RuntimeError("cannot place %d waters in a %.2f A sphere" % (nw, radius))
        p = _in_sphere(rng, radius)
        if all(np.linalg.norm(p - q) >= min_oo for q in opos):
            opos.append(p)
    xyz = np.zeros((nw, 9))
    for w, o in enumerate(opos):
        R = _haar_rotation(rng)
        xyz[w] = (mono @ R.T + o).reshape(-1)
    ion = None
    if ion_min is not None:
        while True:
            p = _in_sphere(rng, radius)
            if all(np.linalg.norm(p - q) >= ion_min for q in opos):
                ion = p
                break
    return xyz, ion


def perturb(xyz, rng, sigma_t=0.05, sigma_deg=2.0):
    """Rigid-body Gaussian move of every monomer (MD-like chain step)."""
    out = np.array(xyz, dtype=np.float64).reshape(-1, 3, 3).copy()
    for w in range(out.shape[0]):
        com = (16.0 * out[w, 0] + out[w, 1] + out[w, 2]) / 18.0
        axis = rng.normal(size=3)
        axis /= np.linalg.norm(axis)
        ang = np.deg2rad(sigma_deg) * rng.normal()
        K = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]], [-axis[1], axis[0], 0]])
        R = np.eye(3) + np.sin(ang) * K + (1 - np.cos(ang)) * K @ K
        out[w] = (out[w] - com) @ R.T + com + sigma_t * rng.normal(size=3)
    return out.reshape(-1, 9)


def cluster_radius_bohr(xyz):
    p = np.asarray(xyz).reshape(-1, 3)
    return float(np.max(np.linalg.norm(p, axis=1))) * ANGS2BOHR


# Batched and served mixes.  Grids (N, L bohr) by water count (<= 12, <= 22,
# more), moved up while the box is shorter than the cluster radius + 20 bohr.
MIX_GRIDS = [(32, 32.0), (48, 38.0), (64, 44.0)]
MIX_BIG = (96, 56.0)
SEED_MIX = 4000


def mixed_snapshot(i, rng, seed0=SEED_MIX):
    """Snapshot i of the uneven batched mix: water count uniform in [4, 32],
    placement as for the other workloads (sphere radius growing with the
    count), 1-4 states.  Returns (xyz, n, L, nroots)."""
    nw = int(rng.integers(4, 33))
    radius = max(3.0, 2.1 * nw ** (1.0 / 3.0))
    xyz, _ = water_cluster(nw, radius, seed0 + i)
    rc = cluster_radius_bohr(xyz)
    g = 0 if nw <= 12 else (1 if nw <= 22 else 2)
    while g < len(MIX_GRIDS) - 1 and MIX_GRIDS[g][1] < rc + 20.0:
        g += 1
    n, L = MIX_GRIDS[g]
    return xyz, n, L, int(rng.integers(1, 5))
