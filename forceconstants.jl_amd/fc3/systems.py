"""Systems and potentials: the argument types of the drop-in entry points.

Mirrors the reference's data model:
  SuperCellSystem   src/types.jl:20-50   (positions, masses, cubic box_sizes_SC)
  LJ                src/interactions/LennardJones.jl:6-23
  StillingerWeberSilicon  src/interactions/SW.jl:3-44 (LAMMPS Si.sw parameters)
  AnalyticalCalculator / AutoDiffCalculator  src/types.jl:193-201

Units follow the reference's choices for these potentials: LJ in Angstrom and
kcal/mol, SW in Angstrom and eV, masses in amu.  Values are plain floats (the
reference strips units inside its loops and keeps them in `.units`).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

# FCC conventional basis in the order SimpleCrystals uses; the order of the
# first cell is pinned by the reference's known-answer test (runtests.jl:109-136).
FCC_BASIS = np.array([[0.0, 0.0, 0.0], [0.5, 0.5, 0.0], [0.5, 0.0, 0.5], [0.0, 0.5, 0.5]])
# Diamond: the FCC basis followed by the same four sites shifted by (1/4,1/4,1/4).
# SimpleCrystals' diamond order is not pinned by any reference test; this order is
# this library's own and is fed identically to the oracle and to the GPU.
DIAMOND_BASIS = np.vstack([FCC_BASIS, FCC_BASIS + 0.25])


def lattice_positions(a: float, cells: Sequence[int], basis: np.ndarray) -> np.ndarray:
    """Positions (N x 3, Angstrom): cells outermost (x slowest, z fastest), basis innermost."""
    nx, ny, nz = (int(c) for c in cells)
    out = []
    for ix in range(nx):
        for iy in range(ny):
            for iz in range(nz):
                cell = np.array([ix, iy, iz], dtype=np.float64)
                for b in basis:
                    out.append(a * (cell + b))
    return np.asarray(out, dtype=np.float64)


@dataclass
class SuperCellSystem:
    """Atoms in a periodic orthorhombic box (types.jl:20-50)."""

    positions: np.ndarray            # N x 3, Angstrom
    masses: np.ndarray               # N, amu
    box_sizes: np.ndarray            # 3, Angstrom

    def __post_init__(self):
        self.positions = np.ascontiguousarray(self.positions, dtype=np.float64).reshape(-1, 3)
        self.masses = np.ascontiguousarray(self.masses, dtype=np.float64).reshape(-1)
        self.box_sizes = np.ascontiguousarray(self.box_sizes, dtype=np.float64).reshape(3)
        if self.masses.shape[0] != self.positions.shape[0]:
            raise ValueError("masses and positions disagree on the atom count")

    @property
    def n_atoms(self) -> int:
        return int(self.positions.shape[0])

    @property
    def n_dof(self) -> int:
        return 3 * self.n_atoms


def fcc(a: float, cells, mass: float, displacement_sigma: float = 0.0,
        seed: Optional[int] = None) -> SuperCellSystem:
    """FCC supercell; optional i.i.d. Gaussian displacements N(0, sigma^2) per
    Cartesian component from numpy's PCG64 stream seeded with `seed`.  Displaced
    positions are not re-wrapped into the box."""
    pos = lattice_positions(a, cells, FCC_BASIS)
    if displacement_sigma > 0.0:
        rng = np.random.default_rng(seed)
        pos = pos + rng.normal(0.0, displacement_sigma, size=pos.shape)
    L = np.array([a * c for c in cells], dtype=np.float64)
    return SuperCellSystem(pos, np.full(pos.shape[0], mass), L)


def diamond(a: float, cells, mass: float, displacement_sigma: float = 0.0,
            seed: Optional[int] = None) -> SuperCellSystem:
    pos = lattice_positions(a, cells, DIAMOND_BASIS)
    if displacement_sigma > 0.0:
        rng = np.random.default_rng(seed)
        pos = pos + rng.normal(0.0, displacement_sigma, size=pos.shape)
    L = np.array([a * c for c in cells], dtype=np.float64)
    return SuperCellSystem(pos, np.full(pos.shape[0], mass), L)


@dataclass(frozen=True)
class LJ:
    """12-6 Lennard-Jones, unshifted (LennardJones.jl:6-56)."""

    sigma: float
    epsilon: float
    r_cut: float


# cos(theta0) as baked into the reference's generated SW derivative blocks
GENERATED_COS_THETA0 = -0.333333333333


@dataclass(frozen=True)
class SWParams:
    """SW_Params (SW.jl:3-14)."""

    epsilon: float = 2.1683
    sigma: float = 2.0951
    a: float = 1.80
    lam: float = 21.0
    gamma: float = 1.20
    cos_theta0: float = -1.0 / 3.0
    A: float = 7.049556277
    B: float = 0.6022245584
    p: float = 4.0
    q: float = 0.0

    def as_array(self) -> np.ndarray:
        return np.array([self.epsilon, self.sigma, self.a, self.lam, self.gamma,
                         self.cos_theta0, self.A, self.B, self.p, self.q], dtype=np.float64)

    def as_abi_array(self, generated_cos_theta0: Optional[float] = None) -> np.ndarray:
        """The 11 doubles of the C ABI: the 10 parameters followed by the cos(theta0)
        of the generated derivative blocks.  By default that is the constant the
        reference's AutoDiffCalculator bakes in (SW/H3_exec_ijk.jl:39) when the
        potential uses cos(theta0) = -1/3, and cos_theta0 otherwise."""
        if generated_cos_theta0 is None:
            generated_cos_theta0 = (GENERATED_COS_THETA0 if self.cos_theta0 == -1.0 / 3.0
                                    else self.cos_theta0)
        return np.append(self.as_array(), generated_cos_theta0)


@dataclass(frozen=True)
class StillingerWeberSilicon:
    """StillingerWeberSilicon (SW.jl:21-44): r_cut = a*sigma."""

    params: SWParams = field(default_factory=SWParams)

    @property
    def r_cut(self) -> float:
        return self.params.a * self.params.sigma


@dataclass(frozen=True)
class AnalyticalCalculator:
    tol: float
    r_cut: float


@dataclass(frozen=True)
class AutoDiffCalculator:
    tol: float
    r_cut: float


# Workloads used across tests and the benchmark (Angstrom, kcal/mol, eV, amu).
ARGON_LJ = LJ(sigma=3.4, epsilon=0.24037, r_cut=8.5)
ARGON_A = 5.2468
ARGON_MASS = 39.95
SI_A = 5.43
SI_MASS = 28.0855
