"""Synthetic inputs for the five benchmark workloads (atomic units: Bohr, e, Hartree).

Every generator is seeded (numpy PCG64, seed 20260000 + workload index unless given) and
returns plain numpy arrays in the layouts of the C ABI: coords ``(n, 3)``, lattice ``(3, 3)``
with row i the lattice vector a_i, charges ``(n,)``, species ``(n,)`` int32 and per-species
Hubbard U.  Nothing here reads the reference; the shapes follow the workload statements:

* ``nacl``     NaCl rocksalt 4x4x4 conventional cells, 512 atoms, a = 22.56 A.
* ``water``    4096 triatomics (12,288 atoms) in a cubic 49.7 A cell.
* ``triclinic`` 98,304 atoms, a = b = c = 99.4 A, alpha = beta = 90, gamma = 75 degrees,
               Poisson-disc positions (d_min 1 A), charges U[-1, 1].
* ``batch``    structures with N log-uniform in [64, 8192], density 0.1 A^-3, angles U[70,110].
* ``serving``  resident structures with N log-uniform in [100, 20000] built like ``triclinic``.
"""
from __future__ import annotations

import math

import numpy as np

#: Angstrom -> Bohr, the reference's constant AA__Bohr = 1/0.529177249
#: (src/dftbp/common/constants.F90).
AA = 1.0 / 0.529177249

SEED0 = 20260000


def lattice_from_parameters(a, b, c, alpha_deg, beta_deg, gamma_deg) -> np.ndarray:
    """Rows a1 = (a,0,0), a2 in the xy-plane, a3 completing the given angles."""
    al, be, ga = (math.radians(x) for x in (alpha_deg, beta_deg, gamma_deg))
    a1 = [a, 0.0, 0.0]
    a2 = [b * math.cos(ga), b * math.sin(ga), 0.0]
    cx = c * math.cos(be)
    cy = c * (math.cos(al) - math.cos(be) * math.cos(ga)) / math.sin(ga)
    cz = math.sqrt(max(c * c - cx * cx - cy * cy, 0.0))
    return np.array([a1, a2, [cx, cy, cz]], dtype=np.float64)


def neutralise(q: np.ndarray) -> np.ndarray:
    return q - q.mean()


def _wrap_cartesian(frac: np.ndarray, lat: np.ndarray) -> np.ndarray:
    return (frac - np.floor(frac)) @ lat


def poisson_disc(n: int, lat: np.ndarray, dmin: float, rng: np.random.Generator,
                 max_rounds: int = 200) -> np.ndarray:
    """n uniform random points in the periodic cell, redrawn until no two (or a point and an
    image) are closer than dmin.  Returns Cartesian coordinates inside the cell."""
    from scipy.spatial import cKDTree

    inv = np.linalg.inv(lat)
    frac = rng.random((n, 3))
    # margin in fractional units that covers dmin along every axis
    reclen = np.linalg.norm(inv.T, axis=1)
    margin = dmin * reclen
    shifts = np.array([[i, j, k] for i in (-1, 0, 1) for j in (-1, 0, 1) for k in (-1, 0, 1)],
                      dtype=np.float64)
    for _ in range(max_rounds):
        # replicate the points that sit near a face, with the image shift that brings them in
        pts, owner = [frac], [np.arange(n)]
        for s in shifts:
            if not s.any():
                continue
            f = frac + s
            keep = np.all((f > -margin) & (f < 1.0 + margin), axis=1)
            if keep.any():
                pts.append(f[keep])
                owner.append(np.nonzero(keep)[0])
        allf = np.concatenate(pts)
        own = np.concatenate(owner)
        tree = cKDTree(allf @ lat)
        pairs = tree.query_pairs(dmin, output_type="ndarray")
        if len(pairs) == 0:
            return frac @ lat
        a, b = own[pairs[:, 0]], own[pairs[:, 1]]
        bad = np.unique(np.maximum(a, b))  # redraw one point of each close pair
        bad = bad[bad >= 0]
        frac[bad] = rng.random((len(bad), 3))
    raise RuntimeError("poisson_disc did not converge")


def nacl(seed: int = SEED0 + 1, reps: int = 4, a_conv: float = 5.64):
    """NaCl rocksalt reps^3 conventional cells; charges +-0.8 + N(0, 0.05), neutralised.
    U(Na) = 0.20 Ha, U(Cl) = 0.30 Ha."""
    rng = np.random.default_rng(seed)
    a = a_conv * AA
    fcc = np.array([[0, 0, 0], [0, 0.5, 0.5], [0.5, 0, 0.5], [0.5, 0.5, 0]])
    cells = np.array([[i, j, k] for i in range(reps) for j in range(reps) for k in range(reps)])
    na = (cells[:, None, :] + fcc[None, :, :]).reshape(-1, 3) * a
    cl = na + np.array([0.5 * a, 0.0, 0.0])
    coords = np.empty((2 * len(na), 3))
    coords[0::2] = na
    coords[1::2] = cl
    species = np.tile(np.array([0, 1], dtype=np.int32), len(na))
    q = np.where(species == 0, 0.8, -0.8) + rng.normal(0.0, 0.05, len(species))
    lat = np.eye(3) * (reps * a)
    return {"name": "nacl", "coords": coords, "lattice": lat, "charges": neutralise(q),
            "species": species, "hubbard_u": np.array([0.20, 0.30])}


def nacl_ideal(reps: int = 4, a_conv: float = 5.64):
    """The same rocksalt lattice with exact +-1 charges (Madelung known answer)."""
    w = nacl(reps=reps, a_conv=a_conv)
    w["charges"] = np.where(w["species"] == 0, 1.0, -1.0).astype(np.float64)
    return w


def water(seed: int = SEED0 + 2, n_side: int = 16, box: float = 49.7):
    """n_side^3 triatomics: O on a jittered lattice (+-0.5 A), H at 0.96 A and 104.5 degrees,
    random orientation; q(O) = -0.8, q(H) = +0.4, plus N(0, 0.02), neutralised.
    U(O) = 0.4 Ha, U(H) = 0.3 Ha."""
    rng = np.random.default_rng(seed)
    L = box * AA
    h = L / n_side
    g = np.array([[i, j, k] for i in range(n_side) for j in range(n_side) for k in range(n_side)],
                 dtype=np.float64)
    o = (g + 0.5) * h + rng.uniform(-0.5 * AA, 0.5 * AA, g.shape)
    nmol = len(o)
    # random rotations from normalised Gaussian quaternions
    qt = rng.normal(size=(nmol, 4))
    qt /= np.linalg.norm(qt, axis=1, keepdims=True)
    w, x, y, z = qt.T
    rot = np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)], -1),
        np.stack([2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)], -1),
        np.stack([2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], -1),
    ], 1)
    half = math.radians(104.5) / 2.0
    doh = 0.96 * AA
    h_local = np.array([[doh * math.sin(half), doh * math.cos(half), 0.0],
                        [-doh * math.sin(half), doh * math.cos(half), 0.0]])
    hs = np.einsum("mij,hj->mhi", rot, h_local) + o[:, None, :]
    coords = np.empty((3 * nmol, 3))
    coords[0::3] = o
    coords[1::3] = hs[:, 0]
    coords[2::3] = hs[:, 1]
    species = np.tile(np.array([0, 1, 1], dtype=np.int32), nmol)
    q = np.where(species == 0, -0.8, 0.4) + rng.normal(0.0, 0.02, len(species))
    return {"name": "water", "coords": coords, "lattice": np.eye(3) * L,
            "charges": neutralise(q), "species": species, "hubbard_u": np.array([0.4, 0.3]),
            "grid": (64, 64, 64), "order": 6, "r_cut": 10.0 * AA}


def triclinic(seed: int = SEED0 + 3, n: int = 98304, side: float = 99.4, gamma: float = 75.0,
              dmin: float = 1.0):
    """Poisson-disc charges U[-1, 1] (neutralised) in a triclinic cell (a = b = c = side,
    alpha = beta = 90, gamma)."""
    rng = np.random.default_rng(seed)
    lat = lattice_from_parameters(side * AA, side * AA, side * AA, 90.0, 90.0, gamma)
    coords = poisson_disc(n, lat, dmin * AA, rng)
    q = neutralise(rng.uniform(-1.0, 1.0, n))
    return {"name": "triclinic", "coords": coords, "lattice": lat, "charges": q,
            "species": None, "hubbard_u": None, "grid": (128, 128, 128), "order": 6,
            "r_cut": 10.0 * AA}


def smooth235(n: int) -> int:
    """Smallest 2^a 3^b 5^c >= n."""
    m = max(n, 1)
    while True:
        r = m
        for p in (2, 3, 5):
            while r % p == 0:
                r //= p
        if r == 1:
            return m
        m += 1


def cell_for_density(n: int, density: float, angles, rng=None):
    """Equal edges with the given angles, sized for `density` atoms per A^3 (returns A)."""
    al, be, ga = (math.radians(x) for x in angles)
    f = math.sqrt(1 - math.cos(al) ** 2 - math.cos(be) ** 2 - math.cos(ga) ** 2
                  + 2 * math.cos(al) * math.cos(be) * math.cos(ga))
    return (n / density / f) ** (1.0 / 3.0)


def batch(seed: int = SEED0 + 4, count: int = 256, nmin: int = 64, nmax: int = 8192,
          density: float = 0.1, spacing: float = 0.8, order: int = 6, dmin: float = 1.0):
    """Independent triclinic structures; grids at ~`spacing` A rounded up to 2^a 3^b 5^c."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n = int(round(math.exp(rng.uniform(math.log(nmin), math.log(nmax)))))
        ang = rng.uniform(70.0, 110.0, 3)
        side = cell_for_density(n, density, ang)
        lat = lattice_from_parameters(side * AA, side * AA, side * AA, *ang)
        coords = poisson_disc(n, lat, dmin * AA, rng)
        q = neutralise(rng.uniform(-1.0, 1.0, n))
        K = tuple(smooth235(int(math.ceil(side / spacing))) for _ in range(3))
        out.append({"name": "batch", "coords": coords, "lattice": lat, "charges": q,
                    "species": None, "hubbard_u": None, "grid": K, "order": order,
                    "r_cut": 10.0 * AA})
    return out


def _serving_structure(args):
    n, seed, density, gamma, spacing, order, dmin = args
    rng = np.random.default_rng(seed)
    side = cell_for_density(n, density, (90.0, 90.0, gamma))
    lat = lattice_from_parameters(side * AA, side * AA, side * AA, 90.0, 90.0, gamma)
    coords = poisson_disc(n, lat, dmin * AA, rng)
    q = neutralise(rng.uniform(-1.0, 1.0, n))
    K = tuple(smooth235(int(math.ceil(side / spacing))) for _ in range(3))
    return {"name": "serving", "coords": coords, "lattice": lat, "charges": q,
            "species": None, "hubbard_u": None, "grid": K, "order": order,
            "r_cut": 10.0 * AA}


def serving(seed: int = SEED0 + 5, count: int = 1000, nmin: int = 100, nmax: int = 20000,
            density: float = 0.1036, gamma: float = 75.0, spacing: float = 0.777,
            order: int = 6, dmin: float = 1.0, workers: int = 0):
    """Resident structures built like ``triclinic`` with N log-uniform in [nmin, nmax].

    Sizes come from one generator seeded with ``seed``; structure k is drawn from its own
    generator (the k-th child of ``seed``), so the set does not depend on ``workers``
    (processes used to draw the positions; 0: serial)."""
    rng = np.random.default_rng(seed)
    ns = [int(round(math.exp(rng.uniform(math.log(nmin), math.log(nmax)))))
          for _ in range(count)]
    children = np.random.SeedSequence(seed).spawn(count)
    jobs = [(n, ch, density, gamma, spacing, order, dmin) for n, ch in zip(ns, children)]
    if workers and workers > 1:
        import concurrent.futures as cf
        import multiprocessing as mp
        with cf.ProcessPoolExecutor(max_workers=workers,
                                    mp_context=mp.get_context("spawn")) as ex:
            return list(ex.map(_serving_structure, jobs, chunksize=8))
    return [_serving_structure(j) for j in jobs]


def scc_update(q: np.ndarray, rng: np.random.Generator, sigma: float = 0.01) -> np.ndarray:
    """One SCC-iteration charge update: q + N(0, sigma), neutralised."""
    return neutralise(q + rng.normal(0.0, sigma, q.shape))
