"""Synthetic water systems and ring-polymer initial conditions.

Units follow the reference: positions in Angstrom, momenta in amu*Angstrom/ps,
energies in kcal/mol, time in ps (consts.f90:22-42).

The ring-polymer sampling restates gen_rand_ring / gen_rand_ring_momenta
(NormalModes.f90:224-275) with a seeded numpy generator instead of the
reference's clock-seeded Fortran RNG (InputOutput.f90:354-358).
"""
from __future__ import annotations

import numpy as np

# consts.f90:22-42
CHARGECON = 18.22234397655801030455
DEBYE = 4.8033324
MASSCON = 418.4
KB_AMUA2PS2PERK = 0.831447148
HBAR = 6.35077993041
PI = 3.14159265358979323844

OH_BOND = 0.9572          # Angstrom
HOH_ANGLE_DEG = 104.52
MASS_O = 15.995           # amu (15.994 is the reference default)
MASS_H = 1.008
WATER_MASS_G = (MASS_O + 2 * MASS_H) * 1.66053892e-24


def box_edge_for_density(nwaters: int, density_g_cm3: float = 0.997) -> float:
    """Cubic box edge (Angstrom) holding nwaters molecules at the given density."""
    vol_cm3 = nwaters * WATER_MASS_G / density_g_cm3
    return (vol_cm3 * 1e24) ** (1.0 / 3.0)


def lattice_for(nwaters: int) -> tuple[int, int, int]:
    """Lattice shapes used by the benchmark systems (4x2x4, 4x4x8, 8x8x8, ...)."""
    table = {32: (4, 2, 4), 64: (4, 4, 4), 128: (4, 4, 8), 256: (8, 4, 8), 512: (8, 8, 8)}
    if nwaters in table:
        return table[nwaters]
    a = int(round(nwaters ** (1 / 3)))
    for nx in range(max(a, 1), 0, -1):
        if nwaters % nx == 0:
            rest = nwaters // nx
            for ny in range(int(np.sqrt(rest)), 0, -1):
                if rest % ny == 0:
                    return (nx, ny, rest // ny)
    return (nwaters, 1, 1)


def _random_rotation(rng: np.random.Generator) -> np.ndarray:
    """Haar-uniform rotation matrix from a normalised Gaussian quaternion."""
    q = rng.standard_normal(4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])


def monomer_geometry() -> np.ndarray:
    """O, H1, H2 of a rigid monomer with O at the origin, shape (3, 3)."""
    half = np.deg2rad(HOH_ANGLE_DEG) / 2.0
    return np.array([
        [0.0, 0.0, 0.0],
        [OH_BOND * np.sin(half), 0.0, OH_BOND * np.cos(half)],
        [-OH_BOND * np.sin(half), 0.0, OH_BOND * np.cos(half)],
    ])


def water_box(nwaters: int, box: float | None = None, lattice=None, jitter: float = 0.3,
              seed: int = 0) -> tuple[np.ndarray, float]:
    """Centroid coordinates (Natoms, 3) in OHH order and the cubic box edge.

    O atoms sit at lattice cell centres in [-L/2, L/2] (the reference centres
    its coordinates, InputOutput.f90:449-452) with U(-jitter, jitter) noise,
    each monomer rotated by a Haar-uniform rotation.
    """
    rng = np.random.default_rng(seed)
    if box is None:
        box = box_edge_for_density(nwaters)
    if lattice is None:
        lattice = lattice_for(nwaters)
    nx, ny, nz = lattice
    if nx * ny * nz != nwaters:
        raise ValueError(f"lattice {lattice} does not hold {nwaters} molecules")
    mono = monomer_geometry()
    out = np.empty((nwaters, 3, 3))
    iw = 0
    for ix in range(nx):
        for iy in range(ny):
            for iz in range(nz):
                centre = np.array([(ix + 0.5) * box / nx, (iy + 0.5) * box / ny,
                                   (iz + 0.5) * box / nz]) - box / 2.0
                centre += rng.uniform(-jitter, jitter, 3)
                rot = _random_rotation(rng)
                out[iw] = centre + mono @ rot.T
                iw += 1
    return out.reshape(3 * nwaters, 3), float(box)


def nm_matrix(nbeads: int) -> np.ndarray:
    """Real normal-mode matrix C (P x P), rows are modes (NormalModes.f90:100-112)."""
    P = nbeads
    l = (P - 2) // 2 if P % 2 == 0 else (P - 1) // 2
    C = np.zeros((P, P))
    for i in range(1, P + 1):
        C[0, i - 1] = np.sqrt(1.0 / P)
        for k in range(1, l + 1):
            C[k, i - 1] = np.sqrt(2.0 / P) * np.cos(k * i * 2.0 * PI / P)
            C[k + l, i - 1] = np.sqrt(2.0 / P) * np.sin(k * i * 2.0 * PI / P)
        if P % 2 == 0:
            C[2 * l + 1, i - 1] = np.sqrt(1.0 / P) * (-1) ** i
    return C


def nm_frequencies(nbeads: int, temp: float) -> np.ndarray:
    """Free ring-polymer mode frequencies in the C row order (NormalModes.f90:118-153)."""
    P = nbeads
    omegan = KB_AMUA2PS2PERK * temp * P / HBAR
    l = (P - 2) // 2 if P % 2 == 0 else (P - 1) // 2
    om = np.zeros(P)
    for k in range(1, l + 1):
        w = 2 * omegan * np.sin(PI * float(np.float32(k) / np.float32(P)))
        om[k] = w
        om[k + l] = w
    if P % 2 == 0:
        om[2 * l + 1] = 2 * omegan * np.sin(PI * float(np.float32(l + 1) / np.float32(P)))
    return om


def atom_masses(nwaters: int, mass_o: float = MASS_O, mass_h: float = MASS_H) -> np.ndarray:
    m = np.empty(3 * nwaters)
    m[0::3] = mass_o
    m[1::3] = mass_h
    m[2::3] = mass_h
    return m


def ring_polymer(centroids: np.ndarray, nbeads: int, temp: float = 300.0, seed: int = 0,
                 mass_o: float = MASS_O, mass_h: float = MASS_H):
    """Bead positions and momenta, each (P, Natoms, 3).

    Offsets from the centroid follow the free ring-polymer normal-mode
    Gaussian at temperature temp (x~_k ~ N(0, kB T P / (m w_k^2)), centroid
    mode zero); momenta p~_k ~ N(0, m kB T P).  Both are mapped back with C^T.
    """
    rng = np.random.default_rng(seed)
    natoms = centroids.shape[0]
    masses = atom_masses(natoms // 3, mass_o, mass_h)
    P = nbeads
    C = nm_matrix(P)
    om = nm_frequencies(P, temp)
    ktp = KB_AMUA2PS2PERK * temp * P
    xt = np.zeros((P, natoms, 3))
    for k in range(1, P):
        sd = np.sqrt(ktp / (masses * om[k] ** 2))
        xt[k] = rng.standard_normal((natoms, 3)) * sd[:, None]
    pt = rng.standard_normal((P, natoms, 3)) * np.sqrt(masses * ktp)[None, :, None]
    RRt = np.einsum("kj,kan->jan", C, xt) + centroids[None]
    PPt = np.einsum("kj,kan->jan", C, pt)
    return np.ascontiguousarray(RRt), np.ascontiguousarray(PPt)


def spaced_water_box(nwaters: int, box: float | None = None, min_contact: float = 1.9,
                     tries: int = 200, seed: int = 0) -> tuple[np.ndarray, float]:
    """Centroid coordinates (Natoms, 3) and box edge with no short contacts:
    the oxygens on nwaters randomly chosen sites of the smallest n^3 grid that
    holds them (spacing L/n, at least 2.4 A at liquid density for the sizes
    used here), each monomer rotated by the first of up to `tries`
    Haar-uniform rotations whose atoms keep min_contact (A) from every atom of
    the molecules already placed (minimum image), else the best one tried.
    A lattice with a short axis (water_box with 4 x 4 x 8 at 128 waters puts
    oxygens 1.96 A apart) throws TTM2.1-F, whose H sites are polarisable, into
    the polarisation catastrophe; these boxes do not."""
    rng = np.random.default_rng(seed)
    if box is None:
        box = box_edge_for_density(nwaters)
    n = 1
    while n ** 3 < nwaters:
        n += 1
    sites = rng.choice(n ** 3, size=nwaters, replace=False)
    mono = monomer_geometry()
    placed = np.empty((0, 3))
    out = np.empty((nwaters, 3, 3))
    for iw, sidx in enumerate(sites):
        ix, iy, iz = sidx // (n * n), (sidx // n) % n, sidx % n
        centre = (np.array([ix, iy, iz]) + 0.5) * box / n - box / 2.0
        best, best_d = None, -1.0
        for _ in range(tries):
            cand = centre + mono @ _random_rotation(rng).T
            if placed.shape[0] == 0:
                best = cand
                break
            d = cand[:, None, :] - placed[None, :, :]
            d -= box * np.rint(d / box)
            dmin = float(np.sqrt(np.min(np.sum(d * d, axis=-1))))
            if dmin > best_d:
                best, best_d = cand, dmin
            if dmin >= min_contact:
                break
        out[iw] = best
        placed = np.vstack([placed, best])
    return out.reshape(3 * nwaters, 3), float(box)


def make_system(nwaters: int, nbeads: int, temp: float = 300.0, seed: int = 0, box=None,
                lattice=None, spaced: bool = False):
    """Centroid box + ring polymer: returns (RRt, PPt, RRc, box_edge).  spaced:
    the centroids from spaced_water_box instead of water_box."""
    if spaced:
        cen, L = spaced_water_box(nwaters, box=box, seed=seed)
    else:
        cen, L = water_box(nwaters, box=box, lattice=lattice, seed=seed)
    RRt, PPt = ring_polymer(cen, nbeads, temp=temp, seed=seed + 7919)
    RRc = RRt.mean(axis=0)
    return RRt, PPt, RRc, L


def _monomer(rng: np.random.Generator, roh_sd: float, ang_sd_deg: float) -> np.ndarray:
    """A monomer about its oxygen with Gaussian-distorted bonds and angle,
    randomly oriented: rows O, H1, H2."""
    r1, r2 = OH_BOND + roh_sd * rng.standard_normal(2)
    th = np.deg2rad(HOH_ANGLE_DEG + ang_sd_deg * rng.standard_normal())
    m = np.array([[0.0, 0.0, 0.0], [r1, 0.0, 0.0], [r2 * np.cos(th), r2 * np.sin(th), 0.0]])
    return m @ _random_rotation(rng).T


def water_cluster(nwaters: int = 5, seed: int = 0, oo_min: float = 2.6, oo_max: float = 3.2,
                  roh_sd: float = 0.02, ang_sd_deg: float = 3.0) -> np.ndarray:
    """A gas-phase cluster: each new oxygen sits oo_min..oo_max from a random
    earlier one and no closer than oo_min to any; monomers randomly oriented
    with distorted geometry.  Returns (3*nwaters, 3) coordinates, O H H per
    molecule, centred on the origin (for a large periodic box)."""
    rng = np.random.default_rng(seed)
    oxy = [np.zeros(3)]
    while len(oxy) < nwaters:
        base = oxy[rng.integers(len(oxy))]
        d = rng.standard_normal(3)
        cand = base + rng.uniform(oo_min, oo_max) * d / np.linalg.norm(d)
        if all(np.linalg.norm(cand - o) >= oo_min for o in oxy):
            oxy.append(cand)
    oxy = np.array(oxy) - np.mean(oxy, axis=0)
    out = np.zeros((3 * nwaters, 3))
    for i, o in enumerate(oxy):
        out[3 * i:3 * i + 3] = o + _monomer(rng, roh_sd, ang_sd_deg)
    return out


def distorted_box(nwaters: int, seed: int = 0, roh_sd: float = 0.02,
                  ang_sd_deg: float = 3.0) -> tuple[np.ndarray, float]:
    """The periodic water box of the benchmark construction with each monomer's
    bonds and angle distorted as in water_cluster: (coordinates, box edge)."""
    rng = np.random.default_rng(seed + 7919)
    cen, L = water_box(nwaters, seed=seed)
    out = cen.copy()
    for i in range(nwaters):
        o = cen[3 * i]
        out[3 * i:3 * i + 3] = o + _monomer(rng, roh_sd, ang_sd_deg)
    return out, L
