"""Synthetic molten-NaCl workloads and unit conversions.

Geometry: rock-salt ordering on an m^3 simple-cubic lattice at a given mass density,
species by the parity of (i+j+k), uniform jitter of +-0.1 A per coordinate, wrapped
into [0, L); Maxwell-Boltzmann velocities with the centre-of-mass velocity removed.
All arrays are in the atomic units the C-ABI expects: positions in Bohr, masses in
electron masses, velocities in Bohr per a.u. time.

Unit conversions follow the reference driver: positions A -> Bohr
(``maze_poisson/solver.py:310``), masses amu * conv_mass (``maze_poisson/solver.py:312``),
Tosi-Fumi parameters (``maze_poisson/solver.py:198-202``), dt fs -> a.u.
(``maze_poisson/myio/input/base_file_input.py:140-142``), constants
(``maze_poisson/constants.py:1-23``).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

A0 = 0.529177210903            # Bohr radius in A
T_AU = 2.4188843265857e-2      # fs per a.u. of time
CONV_MASS = 1.66054e-27 / 9.1093837e-31
KB = 3.1668e-6                 # Hartree / K
KJMOL_TO_HARTREE = 0.0003808798
AMU_G = 1.66054e-24            # g per amu

M_NA, M_CL = 22.99, 35.45

# Fumi-Tosi Born-Mayer-Huggins parameters for NaCl (Tosi & Fumi, J. Phys. Chem.
# Solids 25, 45 (1964); Fumi & Tosi, ibid. 25, 31 (1964)), in the reference's input
# units (maze_poisson/solver.py:194-202): A kJ/mol, B 1/A, C kJ/mol A^6,
# D kJ/mol A^8, sigma A.  A_ij = beta_ij * b with b = 0.338e-19 J and Pauling
# factors beta = 1.25 / 1.00 / 0.75; B = 1/rho, rho = 0.317 A.  The reference ships
# no values; parity checks are insensitive to them (oracle and GPU read the same).
TOSI_FUMI_NACL = {
    # (type1, type2): (A, B, C, D, sigma)
    ("Na", "Na"): (25.4435, 1.0 / 0.317, 101.1719, 48.1771, 2.340),
    ("Na", "Cl"): (20.3548, 1.0 / 0.317, 674.4798, 837.0770, 2.755),
    ("Cl", "Cl"): (15.2661, 1.0 / 0.317, 6985.6979, 14031.5760, 3.170),
}


def tf_params_au(species=("Na", "Cl")) -> np.ndarray:
    """(n_typ, n_typ, 5) flattened, converted as in maze_poisson/solver.py:186-203."""
    nt = len(species)
    out = np.full((nt, nt, 5), np.nan)
    for (a, b), vals in TOSI_FUMI_NACL.items():
        i, j = species.index(a), species.index(b)
        out[i, j] = vals
        out[j, i] = vals
    out[:, :, 0] *= KJMOL_TO_HARTREE
    out[:, :, 1] *= A0
    out[:, :, 2] *= KJMOL_TO_HARTREE / A0**6
    out[:, :, 3] *= KJMOL_TO_HARTREE / A0**8
    out[:, :, 4] /= A0
    return np.ascontiguousarray(out.reshape(-1), dtype=np.float64)


def box_length_ang(n_ions: int, density: float = 1.379) -> float:
    """Cubic box side in A for n_ions (half Na, half Cl) at `density` g/cm^3."""
    vol_cm3 = (n_ions / 2) * (M_NA + M_CL) * AMU_G / density
    return vol_cm3 ** (1.0 / 3.0) * 1e8


def dt_au(dt_fs: float) -> float:
    return dt_fs / T_AU


@dataclass
class System:
    """One replica: particles plus grid geometry, in atomic units."""
    n_grid: int
    L: float                      # Bohr
    types: np.ndarray             # int32 (N,)
    pos: np.ndarray               # (N, 3) Bohr
    vel: np.ndarray               # (N, 3)
    mass: np.ndarray              # (N,) electron masses
    charges: np.ndarray           # (N,)
    tf: np.ndarray = field(default_factory=tf_params_au)
    n_typ: int = 2

    @property
    def h(self) -> float:
        return self.L / self.n_grid

    @property
    def n_p(self) -> int:
        return len(self.types)


def nacl_system(n_ions: int, n_grid: int, seed: int, T: float = 1500.0,
                density: float = 1.379, jitter_ang: float = 0.1) -> System:
    """Seeded jittered rock-salt melt.  When n_ions is not 2 m^3 on a full lattice
    (e.g. 2048), the smallest m with m^3 >= n_ions is used and the first n_ions/2
    sites of each species (in lattice order) are kept."""
    rng = np.random.default_rng(seed)
    L_ang = box_length_ang(n_ions, density)
    m = round(n_ions ** (1.0 / 3.0))
    if m ** 3 < n_ions:
        m += 1
    a = L_ang / m
    idx = np.array([(i, j, k) for i in range(m) for j in range(m) for k in range(m)])
    parity = idx.sum(axis=1) % 2
    half = n_ions // 2
    na_sites = idx[parity == 0][:half]
    cl_sites = idx[parity == 1][:n_ions - half]
    sites = np.concatenate([na_sites, cl_sites])
    types = np.concatenate([np.zeros(len(na_sites)), np.ones(len(cl_sites))]).astype(np.int32)
    order = np.lexsort((sites[:, 2], sites[:, 1], sites[:, 0]))
    sites, types = sites[order], types[order]
    pos_ang = sites * a + rng.uniform(-jitter_ang, jitter_ang, size=sites.shape)
    pos_ang = np.mod(pos_ang, L_ang)
    mass = np.where(types == 0, M_NA, M_CL) * CONV_MASS
    charges = np.where(types == 0, 1.0, -1.0)
    vel = rng.normal(0.0, 1.0, size=(n_ions, 3)) * np.sqrt(KB * T / mass)[:, None]
    vel -= (mass[:, None] * vel).sum(axis=0) / mass.sum()
    return System(
        n_grid=n_grid,
        L=L_ang / A0,
        types=np.ascontiguousarray(types),
        pos=np.ascontiguousarray(pos_ang / A0),
        vel=np.ascontiguousarray(vel),
        mass=np.ascontiguousarray(mass),
        charges=np.ascontiguousarray(charges.astype(np.float64)),
    )


def grid_for(n_ions: int, spacing_ang: float = 0.41) -> int:
    """Even grid size giving about `spacing_ang` A spacing at density 1.379."""
    n = round(box_length_ang(n_ions) / spacing_ang)
    return n + (n % 2)


def poisson_tol(sys_: System, rtol: float = 1e-8, eps_s: float = 1.0) -> float:
    """Absolute infinity-norm tolerance equivalent to `rtol` relative to the scale of
    the Poisson right-hand side, rtol * max|4 pi q_grid / (h eps)|, estimated from the
    largest single-ion grid charge (|q| <= 1 on one node)."""
    qmax = float(np.max(np.abs(sys_.charges))) if sys_.n_p else 1.0
    return rtol * 4.0 * math.pi * qmax / (sys_.h * eps_s)


# ---- linearized Poisson-Boltzmann inputs -----------------------------------------
# SI constants of the screening term, maze_poisson/constants.py:14-18.
EPS0_SI = 8.854187817e-12      # F/m
KB_SI = 1.380649e-23           # J/K
NA_SI = 6.02214076e23          # 1/mol
EC_SI = 1.602176634e-19        # C
BR_SI = 5.29177210903e-11      # m


def kbar2(ionic_molar: float, eps_s: float, T: float, h_bohr: float) -> float:
    """Grid-scaled screening constant kbar^2 = kappa^2 h^2 for an ionic strength in
    mol/L, ``maze_poisson/solver.py:149-153``."""
    return ((8 * np.pi * NA_SI * EC_SI ** 2 * ionic_molar * 1e3)
            / (eps_s * EPS0_SI * KB_SI * T) * BR_SI ** 2 * h_bohr ** 2)


def pb_eps_field(n: int, L: float, seed: int, lo: float = 2.0, hi: float = 80.0,
                 n_gauss: int = 8) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Smooth synthetic dielectric on the faces of an n^3 periodic grid of side L:
    eps = lo + (hi - lo) S / max S, with S a sum of `n_gauss` Gaussians (centres
    uniform in the box, widths uniform in [0.05, 0.2] L, minimum-image distance),
    evaluated at the +x, +y and +z face centres (eps_x[i,j,k] sits at
    ((i + 1/2) h, j h, k h), and so on)."""
    g = np.random.default_rng(seed)
    centres = g.uniform(0.0, L, (n_gauss, 3))
    widths = g.uniform(0.05, 0.2, n_gauss) * L
    h = L / n
    idx = np.arange(n) * h
    S = []
    for axis in range(3):
        coords = [idx, idx, idx]
        coords[axis] = idx + 0.5 * h
        acc = np.zeros((n, n, n))
        for c, w in zip(centres, widths):
            d2 = 0.0
            for a in range(3):
                d = coords[a] - c[a]
                d -= L * np.round(d / L)
                shape = [1, 1, 1]
                shape[a] = n
                d2 = d2 + (d * d).reshape(shape)
            acc += np.exp(-0.5 * d2 / (w * w))
        S.append(acc)
    smax = max(s.max() for s in S)
    return tuple(np.ascontiguousarray(lo + (hi - lo) * s / smax) for s in S)


# ---- many-replica workloads ------------------------------------------------------
ENSEMBLE_SIZES = (216, 512, 1000, 2048, 4096)


def even_round(x: float) -> int:
    """The even integer nearest to x."""
    return 2 * int(round(x / 2.0))


def ensemble_systems(count: int = 64, seed0: int = 4000, spacing_ang: float = 0.41,
                     sizes=ENSEMBLE_SIZES, density_range=(1.30, 1.40)) -> list:
    """Replicas of uneven size: replica r draws its ion count from `sizes` and its
    density uniformly from `density_range` with generator seed0 + r, and gets the
    even grid nearest to L / spacing_ang."""
    out = []
    for r in range(count):
        g = np.random.default_rng(seed0 + r)
        n_ions = int(g.choice(sizes))
        rho = float(g.uniform(*density_range))
        n_grid = even_round(box_length_ang(n_ions, rho) / spacing_ang)
        out.append(nacl_system(n_ions, n_grid, seed=seed0 + r, density=rho))
    return out


@dataclass
class Request:
    arrival: float      # s after the start of the run
    system: System
    steps: int


def serving_trace(count: int, rate_hz: float, seed: int = 5, n_min: int = 216,
                  n_max: int = 32768, steps_range=(100, 2000), spacing_ang: float = 0.41,
                  density: float = 1.379) -> list:
    """Poisson arrivals at `rate_hz`; ion counts log-uniform in [n_min, n_max] (rounded
    to even); steps uniform in steps_range (inclusive); even grid nearest to
    L / spacing_ang."""
    g = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(count):
        t += float(g.exponential(1.0 / rate_hz))
        n_ions = even_round(math.exp(g.uniform(math.log(n_min), math.log(n_max))))
        steps = int(g.integers(steps_range[0], steps_range[1] + 1))
        n_grid = even_round(box_length_ang(n_ions, density) / spacing_ang)
        sys_ = nacl_system(n_ions, n_grid, seed=seed * 100003 + i, density=density)
        out.append(Request(t, sys_, steps))
    return out
