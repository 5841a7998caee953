"""Synthetic systems of the benchmark configurations.

Every input here is generated from a seed, so both the HIP path and any
checker can rebuild identical data:

* cells and atoms: diamond C8 (a = 6.74), Si64 = 2x2x2 conventional diamond
  (L = 20.52), fcc Al4 (a = 7.65), Si216 = 3x3x3 (L = 30.78); standard
  diamond fractional sites (0,0,0), (0,1/2,1/2), (1/2,0,1/2), (1/2,1/2,0),
  (1/4,1/4,1/4), (1/4,3/4,3/4), (3/4,1/4,3/4), (3/4,3/4,1/4);
* local potential V(r) = sum_a depth * exp(-|r - R_a|^2 / (2 sigma^2)) over
  the nearest periodic image, on the grid points r in the symmetric range
  (src/basis/real_space.hpp:126-141);
* KB projectors: s, p_y, p_z, p_x per atom, radial exp(-r^2 / (2 w^2)), cut at
  r_c; coupling matrix D diagonal with entries U[-2, 2] Ha;
* states: N(0,1) complex coefficients on the G sphere, orthonormalised with
  the reference's overlap -> Cholesky -> rotate (src/operations/orthogonalize.hpp:25-32).

Convention declared for the unpinned "Gaussian width" of the configs: a width
sigma means exp(-r^2 / (2 sigma^2)) for both V and the projectors.
"""
from dataclasses import dataclass, field

import numpy as np

SEED_BASE = 20261202

DIAMOND_SITES = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0],
                          [.25, .25, .25], [.25, .75, .75], [.75, .25, .75], [.75, .75, .25]])
FCC_SITES = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0]])


@dataclass
class System:
    name: str
    config_id: int
    lattice: np.ndarray          # rows = cell vectors (bohr)
    n: tuple                     # real-space grid
    positions: np.ndarray        # [natoms, 3] cartesian (bohr)
    ecut: float                  # Hartree
    nst: int
    v_depth: float = -4.0
    v_sigma: float = 0.6
    kb_width: float = 0.8
    kb_radius: float = 2.0
    nlm: int = 4
    kpoints: np.ndarray = field(default_factory=lambda: np.zeros((1, 3)))  # covariant
    nspin: int = 1
    ecut_jitter: float = 0.0     # per-block cutoff E_k = ecut (1 + jitter u_k), u in U[-1/2, 1/2]

    @property
    def seed(self):
        return SEED_BASE + self.config_id

    @property
    def N(self):
        return int(np.prod(self.n))

    @property
    def volume(self):
        return float(abs(np.linalg.det(self.lattice)))

    @property
    def natoms(self):
        return len(self.positions)

    def kb_coeffs(self):
        rng = np.random.default_rng(self.seed + 1000)
        return [np.diag(rng.uniform(-2.0, 2.0, self.nlm)) for _ in range(self.natoms)]

    def block_ecuts(self):
        nb = len(self.kpoints) * self.nspin
        if self.ecut_jitter == 0.0:
            return np.full(nb, self.ecut)
        rng = np.random.default_rng(self.seed + 2000)
        u = rng.uniform(-0.5, 0.5, nb)
        return self.ecut * (1.0 + self.ecut_jitter * u)

    def blocks(self):
        """(kpoint, spin, ecut) of every block, k-points outer, spin inner."""
        ec = self.block_ecuts()
        out = []
        i = 0
        for k in self.kpoints:
            for s in range(self.nspin):
                out.append((np.asarray(k, dtype=np.float64), s, float(ec[i])))
                i += 1
        return out


def _supercell(sites, a, rep):
    pos = []
    for i in range(rep):
        for j in range(rep):
            for k in range(rep):
                for s in sites:
                    pos.append((np.asarray(s) + [i, j, k]) * a)
    return np.array(pos, dtype=np.float64)


def monkhorst_pack(dims, shifted=False):
    """ionic::brillouin k-points of an unreduced grid (src/ionic/brillouin.hpp:72-86), covariant."""
    ks = []
    sh = 0.5 if shifted else 0.0
    for i in range(dims[0]):
        for j in range(dims[1]):
            for k in range(dims[2]):
                f = np.array([(i + sh) / dims[0], (j + sh) / dims[1], (k + sh) / dims[2]])
                f = np.where(f >= 0.5, f - 1.0, f)
                ks.append(2 * np.pi * f)
    return np.array(ks)


def config(cid):
    if cid == 1:
        a = 6.74
        return System("C8 diamond 32^3", 1, np.eye(3) * a, (32, 32, 32), DIAMOND_SITES * a, 25.0, 16,
                      v_sigma=0.6, kb_width=0.8)
    if cid in (2, 4):
        a = 10.26
        return System("Si64 72^3" if cid == 2 else "Si64 72^3 TDDFT", cid, np.eye(3) * 2 * a, (72, 72, 72),
                      _supercell(DIAMOND_SITES, a, 2), 30.0, 128, v_sigma=0.7, kb_width=0.9)
    if cid == 3:
        a = 7.65
        return System("Al4 fcc 24^3, 64 k x 2 spins", 3, np.eye(3) * a, (24, 24, 24), FCC_SITES * a, 20.0, 12,
                      v_sigma=0.7, kb_width=0.9, kpoints=monkhorst_pack((4, 4, 4)), nspin=2, ecut_jitter=0.1)
    if cid == 5:
        a = 10.26
        return System("Si216 96^3 ACE", 5, np.eye(3) * 3 * a, (96, 96, 96), _supercell(DIAMOND_SITES, a, 3), 30.0, 432,
                      v_sigma=0.7, kb_width=0.9)
    raise ValueError(cid)


def small_config(seed=7, n=(16, 16, 16), a=6.0, natoms=3, nst=5, ecut=6.0):
    """A small random system for parity tests."""
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-0.5, 0.5, (natoms, 3)) * a
    return System(f"small{n[0]}", 100 + seed, np.eye(3) * a, n, pos, ecut, nst, v_sigma=0.6, kb_width=0.7,
                  kb_radius=1.6)


def rgrid(lattice, n):
    """cartesian position of every grid point (symmetric range), [N, 3]."""
    idx = [np.arange(k) for k in n]
    sym = [np.where(i >= (k + 1) // 2, i - k, i) / k for i, k in zip(idx, n)]
    f0, f1, f2 = np.meshgrid(*sym, indexing="ij")
    f = np.stack([f0, f1, f2], -1).reshape(-1, 3)
    return f @ np.asarray(lattice)


def local_potential(sysm):
    r = rgrid(sysm.lattice, sysm.n)
    inv = np.linalg.inv(sysm.lattice)
    v = np.zeros(len(r))
    for p in sysm.positions:
        d = r - p
        f = d @ inv
        f -= np.round(f)
        d = f @ sysm.lattice
        v += sysm.v_depth * np.exp(-np.sum(d * d, 1) / (2.0 * sysm.v_sigma ** 2))
    return v


def sphere_volume_element(sysm):
    """fourier_space volume element V / N^2 (src/basis/fourier_space.hpp:48-50)."""
    return sysm.volume / float(sysm.N) ** 2


def orthonormalize(c, vol_el):
    """overlap S = vol_el conj(c) c^T -> Cholesky S = L L^H -> Phi <- Phi L^{-H},
    i.e. c <- conj(L)^{-1} c on band-major states."""
    s = vol_el * (c.conj() @ c.T)
    L = np.linalg.cholesky(s)
    return np.linalg.solve(L.conj(), c)


def random_states(npw, nst, seed, vol_el):
    rng = np.random.default_rng(seed)
    c = rng.standard_normal((nst, npw)) + 1j * rng.standard_normal((nst, npw))
    return orthonormalize(c, vol_el)
