"""Synthetic molecular systems in SwinMD's units (nm, ps, amu, kJ/mol, e).

A :class:`System` carries what SwinMD's ``Parameters``/``Ensemble`` hold for
the nonbonded path: per-atom type, molecule, charge and mass, the molecule
layout (``Molecule::myAtoms`` order), bonds and angles, the exclusion table
built like ``Ensemble::build_exclusion_table`` (src/Ensemble.cpp:323-367) and
the LJ pair table built like ``Parameters::set_lj_param``
(src/Parameters.cpp:1360-1372).  The box is derived from the number density
exactly as ``Ensemble::compute_bound_box`` does (src/Ensemble.cpp:405-414).

Builders follow the synthetic workloads the benchmarks are quoted on: an fcc
Lennard-Jones crystal, rigid 3-site point-charge water on a jittered lattice
with random orientations, polarizable-model water and linear CO2.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

SQRT_COULOMB = 11.78708976   # src/NEMD_defs.h:103
COULOMB = 138.93548501       # src/NEMD_defs.h:104
BOLTZMANN = 0.0083145112     # src/NEMD_defs.h:108
PI = 3.14159265358979323846  # src/NEMD_defs.h:77


@dataclass
class AtomType:
    mass: float
    charge: float
    eps: float
    sigma: float
    dip: tuple = (0.0, 0.0, 0.0)
    quad: tuple = (0.0, 0.0, 0.0, 0.0, 0.0, 0.0)   # xx xy xz yy yz zz
    polar: float = 0.0
    type_name: str = "X"
    atom_name: str = "X"


@dataclass
class System:
    name: str
    types: list                 # list[AtomType]
    mol_type_name: str
    atoms_per_mol: int
    pos: np.ndarray             # (N, 3)
    type: np.ndarray            # (N,) int32
    mol: np.ndarray             # (N,) int32
    density: float              # atoms per nm^3 (reference config value)
    bonds: list = field(default_factory=list)          # (a1, a2, bond_type)
    angles: list = field(default_factory=list)         # (a1, a2, a3, angle_type)
    bond_params: list = field(default_factory=list)    # (t1, t2, r0, k)
    angle_params: list = field(default_factory=list)   # (t1, t2, t3, theta_deg, k)

    # ---- derived ----------------------------------------------------------
    @property
    def n_atoms(self) -> int:
        return int(self.pos.shape[0])

    @property
    def n_mols(self) -> int:
        return int(self.mol.max()) + 1 if self.n_atoms else 0

    @property
    def box(self) -> np.ndarray:
        # vol = nAtoms/density; boxLen = pow(vol, 1.0/3.0)
        L = (self.n_atoms / self.density) ** (1.0 / 3.0)
        return np.array([L, L, L], dtype=np.float64)

    @property
    def volume(self) -> float:
        b = self.box
        return float(b[0] * b[1] * b[2])

    @property
    def mol_start(self) -> np.ndarray:
        counts = np.bincount(self.mol, minlength=self.n_mols)
        out = np.zeros(self.n_mols + 1, dtype=np.int32)
        out[1:] = np.cumsum(counts)
        return out

    @property
    def mass(self) -> np.ndarray:
        return np.array([self.types[t].mass for t in self.type], dtype=np.float64)

    @property
    def charge(self) -> np.ndarray:
        return np.array([self.types[t].charge for t in self.type], dtype=np.float64)

    @property
    def scaled_charge(self) -> np.ndarray:
        # atoms[i].scaledCharge = charge*SQRTCOULOMBCONSTANT (src/Ensemble.cpp:191)
        return self.charge * SQRT_COULOMB

    def exclusions(self):
        """CSR exclusion table: bonded partners plus angle 1-3 partners."""
        sets = [set() for _ in range(self.n_atoms)]
        for a1, a2, _ in self.bonds:
            sets[a1].add(a2)
            sets[a2].add(a1)
        for a1, _, a3, _ in self.angles:
            sets[a1].add(a3)
            sets[a3].add(a1)
        off = np.zeros(self.n_atoms + 1, dtype=np.int32)
        flat = []
        for i, s in enumerate(sets):
            row = sorted(s)
            flat.extend(row)
            off[i + 1] = len(flat)
        return off, np.array(flat, dtype=np.int32)

    def lj_table(self):
        """sigma^2 and 4*eps per type pair (Parameters::set_lj_param)."""
        n = len(self.types)
        sig2 = np.zeros(n * n)
        eps4 = np.zeros(n * n)
        for i, a in enumerate(self.types):
            for j, b in enumerate(self.types):
                sigma = (a.sigma + b.sigma) / 2
                eps = math.sqrt(a.eps * b.eps)
                sig2[i * n + j] = sigma * sigma
                eps4[i * n + j] = 4.0 * eps
        return sig2, eps4

    def com(self, pos=None) -> np.ndarray:
        """Molecule::mass_center for every molecule (src/Molecule.h:61-67)."""
        pos = self.pos if pos is None else pos
        ms = self.mol_start
        m = self.mass
        out = np.zeros((self.n_mols, 3))
        for k in range(self.n_mols):
            c = np.zeros(3)
            M = 0.0
            for a in range(ms[k], ms[k + 1]):
                M += m[a]
            for a in range(ms[k], ms[k + 1]):
                c = c + m[a] * pos[a]
            out[k] = c / M
        return out

    def constraint_lengths2(self) -> np.ndarray:
        """Squared SHAKE lengths per molecule: H1-O, O-H2 and the H-H distance
        from the law of cosines (LFIntegrator::initialise,
        src/integrators/LFIntegrator.cpp:156-167; Parameters::get_bond_length
        src/Parameters.h:204-227)."""
        if self.atoms_per_mol != 3 or not self.bond_params or not self.angle_params:
            raise ValueError("constraint lengths need 3-site molecules with bond and angle types")
        r0 = self.bond_params[0][2]
        theta = self.angle_params[0][3] * 2 * PI / 360.0
        l1 = r0
        l2 = r0
        l3 = math.sqrt(l1 * l1 + l2 * l2 - 2 * l1 * l2 * math.cos(theta))
        one = np.array([l1 * l1, l2 * l2, l3 * l3])
        return np.tile(one, self.n_mols)


# ---------------------------------------------------------------------------
# builders
# ---------------------------------------------------------------------------

def _random_rotations(rng: np.random.Generator, n: int) -> np.ndarray:
    """Uniform random rotation matrices from unit quaternions (Shoemake)."""
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    q = np.stack([
        np.sqrt(1 - u1) * np.sin(2 * np.pi * u2),
        np.sqrt(1 - u1) * np.cos(2 * np.pi * u2),
        np.sqrt(u1) * np.sin(2 * np.pi * u3),
        np.sqrt(u1) * np.cos(2 * np.pi * u3),
    ], axis=1)
    x, y, z, w = q.T
    R = np.empty((n, 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z)
    R[:, 0, 1] = 2 * (x * y - z * w)
    R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w)
    R[:, 1, 1] = 1 - 2 * (x * x + z * z)
    R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w)
    R[:, 2, 1] = 2 * (y * z + x * w)
    R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def _lattice_sites(dims, L, n_sites, rng, jitter, subset=False):
    nx, ny, nz = dims
    g = np.stack(np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij"), -1).reshape(-1, 3)
    if subset:
        idx = np.sort(rng.choice(len(g), size=n_sites, replace=False))
        g = g[idx]
    else:
        g = g[:n_sites]
    spacing = np.array([L / nx, L / ny, L / nz])
    sites = -L / 2 + (g + 0.5) * spacing
    sites = sites + rng.uniform(-jitter, jitter, size=sites.shape)
    return sites


SPCE_TYPES = [
    AtomType(mass=1.008, charge=0.4238, eps=0.0, sigma=0.0, type_name="H", atom_name="Hydrogen"),
    AtomType(mass=15.999, charge=-0.8476, eps=0.6502, sigma=0.3166, type_name="O", atom_name="Oxygen"),
]

# Polarizable water site parameters of the reference's example input
# (test/sysData.in:12-13): charge, buffered 14-7 eps and R, dipole,
# traceless quadrupole and polarizability in the reference's units.
AMOEBA_TYPES = [
    AtomType(mass=1.008, charge=0.25983, eps=0.0565245, sigma=0.2655,
             dip=(-0.002042095, 0.0, -0.003078753),
             quad=(-0.0001028545, 0.0, -5.684578967e-06, -0.000300723, 0.0, 0.0004035771),
             polar=0.000496, type_name="H", atom_name="Hydrogen"),
    AtomType(mass=15.999, charge=-0.51966, eps=0.46057, sigma=0.3405,
             dip=(0.0, 0.0, 0.0075561214),
             quad=(0.0010620922, 0.0, 0.0, -0.00117077124, 0.0, 0.000108679),
             polar=0.000837, type_name="O", atom_name="Oxygen"),
]


def water(n_mol: int, density_mol: float | None = None, box: float | None = None, *, seed: int = 1,
          lattice=None, jitter: float = 0.02, r_oh: float = 0.1, hoh_deg: float = 109.47,
          types=None, subset: bool = False, name: str = "water") -> System:
    """Rigid 3-site water (site order H1, O, H2; types H=0, O=1) on a jittered
    lattice with uniformly random orientations.  Give either the molecular
    number density (molecules/nm^3) or the box edge (nm)."""
    types = list(SPCE_TYPES if types is None else types)
    n_atoms = 3 * n_mol
    if box is not None:
        density = n_atoms / (box ** 3)
    elif density_mol is not None:
        density = 3.0 * density_mol
    else:
        raise ValueError("need density_mol or box")
    L = (n_atoms / density) ** (1.0 / 3.0)
    rng = np.random.default_rng(seed)
    if lattice is None:
        k = int(math.ceil(n_mol ** (1.0 / 3.0) - 1e-9))
        lattice = (k, k, k)
    sites = _lattice_sites(lattice, L, n_mol, rng, jitter, subset=subset)
    theta = hoh_deg * 2 * PI / 360.0
    half = 0.5 * theta
    # body frame: O at origin, H in the xz-plane symmetric about z
    body = np.array([[r_oh * math.sin(half), 0.0, r_oh * math.cos(half)],
                     [0.0, 0.0, 0.0],
                     [-r_oh * math.sin(half), 0.0, r_oh * math.cos(half)]])
    m = np.array([types[0].mass, types[1].mass, types[0].mass])
    body = body - (m[:, None] * body).sum(0) / m.sum()
    R = _random_rotations(rng, n_mol)
    pos = sites[:, None, :] + np.einsum("mij,aj->mai", R, body)
    pos = pos.reshape(-1, 3).astype(np.float64)
    t = np.tile(np.array([0, 1, 0], dtype=np.int32), n_mol)
    mol = np.repeat(np.arange(n_mol, dtype=np.int32), 3)
    bonds, angles = [], []
    for k in range(n_mol):
        h1, o, h2 = 3 * k, 3 * k + 1, 3 * k + 2
        bonds.append((h1, o, 0))
        bonds.append((o, h2, 0))
        angles.append((h1, o, h2, 0))
    return System(name=name, types=types, mol_type_name="WAT", atoms_per_mol=3, pos=pos, type=t,
                  mol=mol, density=density, bonds=bonds, angles=angles,
                  bond_params=[(0, 1, r_oh, 221743.52)],
                  angle_params=[(0, 1, 0, hoh_deg, 142.56735)])


def lj_fcc(cells: int = 6, rho: float = 0.8442, sigma: float = 1.0, eps: float = 1.0,
           mass: float = 1.0, jitter: float = 0.0, seed: int = 1) -> System:
    """Lennard-Jones atoms on an fcc lattice (4 atoms per cubic cell), one atom
    per molecule."""
    n = 4 * cells ** 3
    L = (n / rho) ** (1.0 / 3.0)
    a = L / cells
    basis = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    g = np.stack(np.meshgrid(np.arange(cells), np.arange(cells), np.arange(cells), indexing="ij"), -1).reshape(-1, 3)
    pos = ((g[:, None, :] + basis[None]) * a).reshape(-1, 3) - L / 2 + 0.25 * a
    if jitter:
        pos = pos + np.random.default_rng(seed).uniform(-jitter, jitter, size=pos.shape)
    types = [AtomType(mass=mass, charge=0.0, eps=eps, sigma=sigma, type_name="A", atom_name="Argon")]
    return System(name="lj_fcc", types=types, mol_type_name="LJ", atoms_per_mol=1,
                  pos=pos.astype(np.float64), type=np.zeros(n, dtype=np.int32),
                  mol=np.arange(n, dtype=np.int32), density=rho)


CO2_TYPES = [
    AtomType(mass=15.999, charge=-0.3256, eps=0.0, sigma=0.0, type_name="O", atom_name="Oxygen"),
    AtomType(mass=15.999, charge=-0.3256, eps=0.0, sigma=0.0, type_name="O", atom_name="Oxygen"),
    AtomType(mass=12.011, charge=0.6512, eps=0.0, sigma=0.0, type_name="C", atom_name="Carbon"),
]


def co2(n_mol: int, density_mol: float, seed: int = 1, r_co: float = 0.116, jitter: float = 0.02) -> System:
    """Linear CO2 with site order O0, O1, C (types 0, 1, 2), the layout the
    reference's Buckingham term assumes (src/forces/BuckinghamForce.cpp:379-388)."""
    n_atoms = 3 * n_mol
    density = 3.0 * density_mol
    L = (n_atoms / density) ** (1.0 / 3.0)
    rng = np.random.default_rng(seed)
    k = int(math.ceil(n_mol ** (1.0 / 3.0) - 1e-9))
    sites = _lattice_sites((k, k, k), L, n_mol, rng, jitter)
    body = np.array([[0.0, 0.0, r_co], [0.0, 0.0, -r_co], [0.0, 0.0, 0.0]])
    R = _random_rotations(rng, n_mol)
    pos = (sites[:, None, :] + np.einsum("mij,aj->mai", R, body)).reshape(-1, 3)
    t = np.tile(np.array([0, 1, 2], dtype=np.int32), n_mol)
    mol = np.repeat(np.arange(n_mol, dtype=np.int32), 3)
    bonds, angles = [], []
    for m in range(n_mol):
        o0, o1, c = 3 * m, 3 * m + 1, 3 * m + 2
        bonds.append((o0, c, 0))
        bonds.append((o1, c, 0))
        angles.append((o0, c, o1, 0))
    return System(name="co2", types=list(CO2_TYPES), mol_type_name="CO2", atoms_per_mol=3,
                  pos=pos.astype(np.float64), type=t, mol=mol, density=density, bonds=bonds,
                  angles=angles, bond_params=[(0, 2, 2 * r_co, 1.0e5)],
                  angle_params=[(0, 2, 1, 180.0, 1.0e3)])


# Buckingham table of the reference's CO2 model (src/forces/BuckinghamForce.h:20-28,
# BuckinghamForce.cpp:410-470), [type_i*3 + type_j].
def co2_buckingham_table():
    AAcc, AAco, AAoo = 1.489504e-03, 1.556448e-03, 1.627576e-03
    BBcc, BBco, BBoo = 907.928, 36651.84, 1481136.0
    CCcc, CCco, CCoo = 22.7, 31.6, 44.0
    A = np.array([AAoo, AAoo, AAco, AAoo, AAoo, AAco, AAco, AAco, AAcc])
    B = np.array([BBoo, BBoo, BBco, BBoo, BBoo, BBco, BBco, BBco, BBcc])
    C = np.array([CCoo, CCoo, CCco, CCoo, CCoo, CCco, CCco, CCco, CCcc])
    return A, B, C


# The polarizable sites with the buffered 14-7 on O only (H eps = 0), as the
# polarizable config names them.  Without a repulsive H site this model
# collapses: hydrogens fall into neighbouring oxygens within ~20 steps, even
# from an equilibrated liquid (the MD workloads therefore use AMOEBA_TYPES,
# the reference example's full set; this set stays for force-level fixtures).
POLAR_TYPES = [AtomType(mass=AMOEBA_TYPES[0].mass, charge=AMOEBA_TYPES[0].charge, eps=0.0,
                        sigma=AMOEBA_TYPES[0].sigma, dip=AMOEBA_TYPES[0].dip, quad=AMOEBA_TYPES[0].quad,
                        polar=AMOEBA_TYPES[0].polar, type_name="H", atom_name="Hydrogen"),
               AMOEBA_TYPES[1]]


def c2_water(seed: int = 2024) -> System:
    """4,096 rigid SPC/E-charge waters, 16^3 lattice, L = 4.9669 nm."""
    return water(4096, box=4.9669, seed=seed, lattice=(16, 16, 16), jitter=0.02, name="c2_water")


def polar_water(n_mol: int = 1024, seed: int = 2025, o_only: bool = False) -> System:
    """Polarizable water at 1.0 g/cm^3 with the 3-site water's geometry
    (O-H 0.1 nm, H-O-H 109.47 deg): site charges, dipoles, quadrupoles,
    polarizabilities and buffered 14-7 of the reference's example
    (AMOEBA_TYPES, test/sysData.in:12-13); o_only drops the H 14-7 term
    (POLAR_TYPES), a model that does not survive MD.
    1,024 molecules: a seeded subset of an 11^3 lattice in L = 3.1289 nm;
    8,192 molecules: a seeded subset of a 21^3 lattice in L = 6.2579 nm
    (0.298 nm spacing; a 16 x 16 x 32 lattice in the cubic box would put
    neighbours 0.196 nm apart along z, a start no MD step survives)."""
    types = list(POLAR_TYPES if o_only else AMOEBA_TYPES)
    if n_mol == 1024:
        return water(1024, box=3.1289, seed=seed, lattice=(11, 11, 11), subset=True, types=types,
                     name="polar_water1024")
    if n_mol == 8192:
        return water(8192, box=6.2579, seed=seed, lattice=(21, 21, 21), subset=True, types=types,
                     name="polar_water8192")
    raise ValueError("polar_water: 1024 or 8192 molecules")


def point_charge_twin(system: System) -> System:
    """The same water (coordinates, topology, masses) with the rigid
    point-charge sites (SPCE_TYPES): the model that relaxes a lattice start
    before a polarizable run."""
    import dataclasses
    return dataclasses.replace(system, types=list(SPCE_TYPES), name=system.name + "_point_charge")


def kcut_for(n2max: float, L: float) -> float:
    """k cutoff selecting |n|^2 <= n2max in a cubic box: kCut = 2*pi*sqrt(n2max+0.5)/L."""
    return 2 * PI * math.sqrt(n2max + 0.5) / L
