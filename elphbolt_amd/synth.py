"""Seeded synthetic Wannier-space inputs in elphbolt's EPW data format.

The generator builds everything the vertex path reads
(src/wannier.f90:289-428 `read_EPW_Wannier`):

* Wigner-Seitz cell sets with integer degeneracies, by the EPW rule that
  writes `rcells_*`/`wsdeg_*` (thirdparty/EPW/src/ephwann_shuffle.f90:383-421):
  a lattice vector n of the coarse-grid supercell belongs to the set when no
  supercell image n - T*N is shorter, and its degeneracy is the number of
  images at the same minimal length.  Sum_n 1/deg(n) = N1*N2*N3.
* H(Re): complex Hermitian-pair blocks, H(-R) = H(R)^dagger, entries
  N(0,1) * 2 eV * exp(-|R|/2 A), stored in Ry.
* Force constants from a central-spring model with k(d) ~ exp(-|d|/1.5 A),
  cut below half the shortest supercell vector so every spring has a unique
  minimal image.  With that cut the WS-wrapped interpolation of `ph_wann`
  (src/wannier.f90:631-665) reproduces the spring model exactly, so D(q) is
  positive semidefinite and obeys the acoustic sum rule (no phonon energy
  below -5 meV, src/wannier.f90:774-780).  Scaled so max omega = 62 meV.
* g(np, mp, Re, sp, Rp): |g| = 0.1 eV * exp(-|Re|/2 A - |Rp|/2 A), uniform
  random phase, stored in Ry, Fortran (column-major) layout
  (nb, nb, nwsk, nbr, nwsg) as EPW's `epmatwp1` (ephwann_shuffle.f90:472).
* rcells_g == rcells_q and wsdeg_g == wsdeg_q, as `g2(..., 'ph')` requires
  (it sums nwsg cells with rcells_q/phwsdeg, src/wannier.f90:986-996).

Every random stream is a numpy PCG64 generator keyed by (seed, tensor id), so
the same seed gives the same material on every machine with this numpy.
"""
from __future__ import annotations

import dataclasses
import math
import os
from typing import Optional

import numpy as np

from . import constants as C

ANGSTROM_PER_NM = 10.0

# tensor ids for the keyed random streams
_TID_H, _TID_G, _TID_KWIN = 1, 3, 4


@dataclasses.dataclass
class Material:
    """Wannier-space tensors of one material (fields named as `type Wannier`,
    src/wannier.f90:46-93, plus the crystal data the path reads)."""

    name: str
    lattvecs: np.ndarray          # (3,3) nm, columns are lattice vectors
    basis: np.ndarray             # (3, natoms) fractional
    masses: np.ndarray            # (natoms,) amu, per atom
    coarse_qmesh: np.ndarray      # (3,) int64
    rcells_k: np.ndarray          # (nwsk, 3) int64
    elwsdeg: np.ndarray           # (nwsk,) int64
    rcells_q: np.ndarray          # (nwsq, 3) int64
    phwsdeg: np.ndarray           # (nwsq,) int64
    rcells_g: np.ndarray          # (nwsg, 3) int64
    gwsdeg: np.ndarray            # (nwsg,) int64
    Hwann: np.ndarray             # (nwsk, nb, nb) complex128, Ry
    Dphwann: np.ndarray           # (nwsq, nbr, nbr) complex128
    gwann: Optional[np.ndarray]   # (nb, nb, nwsk, nbr, nwsg) complex128, Fortran order; None: made in HBM
    # window / mesh settings that come with the config
    wvmesh: np.ndarray = dataclasses.field(default_factory=lambda: np.array([4, 4, 4], np.int64))
    mesh_ref: int = 1
    enref: float = 0.0
    fsthick: float = 0.5
    polar: bool = False
    epsilon: Optional[np.ndarray] = None      # (3,3)
    born: Optional[np.ndarray] = None         # (3,3,natoms)
    vertex_seed: Optional[int] = None         # seed of the device-generated vertex (gwann None)

    @property
    def gwann_bytes(self) -> int:
        return 16 * self.nb * self.nb * self.nbr * self.nwsk * self.nwsg

    @property
    def nb(self) -> int:
        return int(self.Hwann.shape[1])

    @property
    def nbr(self) -> int:
        return int(self.Dphwann.shape[1])

    @property
    def natoms(self) -> int:
        return int(self.basis.shape[1])

    @property
    def nwsk(self) -> int:
        return int(self.rcells_k.shape[0])

    @property
    def nwsq(self) -> int:
        return int(self.rcells_q.shape[0])

    @property
    def nwsg(self) -> int:
        return int(self.rcells_g.shape[0])

    @property
    def phmesh(self) -> np.ndarray:
        return self.wvmesh // self.mesh_ref


def fcc_lattice(a_angstrom: float) -> np.ndarray:
    """fcc primitive vectors in nm as columns, in the orientation of
    examples/Si/input/input.nml: (-1,0,1), (0,1,1), (-1,1,0) times a/2."""
    h = 0.5 * a_angstrom / ANGSTROM_PER_NM
    return h * np.array([[-1.0, 0.0, -1.0],
                         [0.0, 1.0, 1.0],
                         [1.0, 1.0, 0.0]])


def ws_cells(lattvecs: np.ndarray, mesh) -> tuple[np.ndarray, np.ndarray]:
    """Wigner-Seitz lattice vectors of the mesh supercell and their
    degeneracies (EPW rule, loop order n1 slowest, n3 fastest)."""
    mesh = np.asarray(mesh, dtype=np.int64)
    metric = lattvecs.T @ lattvecs
    rng = [np.arange(-2 * m, 2 * m + 1) for m in mesh]
    n = np.stack(np.meshgrid(*rng, indexing="ij"), axis=-1).reshape(-1, 3)
    t = np.stack(np.meshgrid(*[np.arange(-2, 3)] * 3, indexing="ij"), axis=-1).reshape(-1, 3)
    shifts = t * mesh  # (125, 3)
    cells, degs = [], []
    scale = float(np.trace(metric)) / 3.0
    eps = 1e-8 * scale
    for chunk in np.array_split(n, max(1, len(n) // 4096)):
        d = chunk[:, None, :] - shifts[None, :, :]
        dist = np.einsum("nti,ij,ntj->nt", d, metric, d)
        mind = dist.min(axis=1)
        i0 = np.where((t == 0).all(axis=1))[0][0]
        keep = np.abs(dist[:, i0] - mind) < eps
        deg = (np.abs(dist - mind[:, None]) < eps).sum(axis=1)
        cells.append(chunk[keep])
        degs.append(deg[keep])
    cells = np.concatenate(cells).astype(np.int64)
    degs = np.concatenate(degs).astype(np.int64)
    return cells, degs


def _cart_len_angstrom(lattvecs, cells_frac) -> np.ndarray:
    return np.linalg.norm(cells_frac @ lattvecs.T, axis=1) * ANGSTROM_PER_NM


def _partner_index(cells: np.ndarray) -> np.ndarray:
    lut = {tuple(c): i for i, c in enumerate(cells)}
    return np.array([lut[tuple(-c)] for c in cells], dtype=np.int64)


def make_hamiltonian(cells, lattvecs, nb, seed) -> np.ndarray:
    rng = np.random.default_rng([seed, _TID_H])
    nws = cells.shape[0]
    env = 2.0 * np.exp(-_cart_len_angstrom(lattvecs, cells) / 2.0)  # eV
    partner = _partner_index(cells)
    raw = rng.standard_normal((nws, nb, nb)) + 1j * rng.standard_normal((nws, nb, nb))
    H = np.empty((nws, nb, nb), np.complex128)
    for i in range(nws):
        p = partner[i]
        if p == i:  # R = 0
            A = raw[i] * env[i]
            H[i] = 0.5 * (A + A.conj().T)
        elif i < p:
            H[i] = raw[i] * env[i]
            H[p] = H[i].conj().T
    return H / C.RYD2EV


_T_GRID = np.stack(np.meshgrid(*[np.arange(-2, 3)] * 3, indexing="ij"), axis=-1).reshape(-1, 3)


def _min_image(lattvecs, cell, dvec_frac, mesh):
    t = _T_GRID
    cand = cell[None, :] + t * mesh[None, :]
    d = (cand + dvec_frac[None, :]) @ lattvecs.T
    lens = np.linalg.norm(d, axis=1)
    j = int(np.argmin(lens))
    return cand[j], d[j], lens[j]


def make_force_constants(cells, lattvecs, basis, masses, mesh, omega_max_ev=0.062):
    """Central-spring force constants on the WS cell set (see module doc)."""
    mesh = np.asarray(mesh, np.int64)
    nat = basis.shape[1]
    nbr = 3 * nat
    superv = lattvecs * mesh[None, :]
    rcut = 0.45 * min(np.linalg.norm(superv, axis=0))
    classes = {}
    springs = []  # (i, j, T, e, k)
    for cell in cells:
        key = tuple(np.mod(cell, mesh))
        if key in classes:
            continue
        blk = np.zeros((nbr, nbr))
        for i in range(nat):
            for j in range(nat):
                T, d, L = _min_image(lattvecs, cell, basis[:, j] - basis[:, i], mesh)
                if L < 1e-12 or L >= rcut:
                    continue
                k = math.exp(-(L * ANGSTROM_PER_NM) / 1.5)
                e = d / L
                blk[3 * i:3 * i + 3, 3 * j:3 * j + 3] = -k * np.outer(e, e)
                springs.append((i, j, T.copy(), e, k))
        classes[key] = blk
    # self terms balance every spring of atom i (acoustic sum rule)
    key0 = (0, 0, 0)
    for (i, j, T, e, k) in springs:
        classes[key0][3 * i:3 * i + 3, 3 * i:3 * i + 3] += k * np.outer(e, e)

    if not springs:
        raise ValueError("no spring shorter than 0.45 x the shortest supercell vector: use a larger WS grid")
    # scale so that max omega (on a 4^3 q sample) equals omega_max_ev
    qs = np.stack(np.meshgrid(*[np.arange(4) / 4.0] * 3, indexing="ij"), -1).reshape(-1, 3)
    si = np.array([sp[0] for sp in springs])
    sj = np.array([sp[1] for sp in springs])
    sT = np.array([sp[2] for sp in springs], np.float64)
    sblk = np.array([-sp[4] * np.outer(sp[3], sp[3]) for sp in springs])
    ph = np.exp(2j * np.pi * (qs @ sT.T))                        # (nq, nspring)
    Dq = np.zeros((qs.shape[0], nat, nat, 3, 3), np.complex128)
    np.add.at(Dq, (slice(None), si, sj), ph[:, :, None, None] * sblk[None])
    Dq = Dq.transpose(0, 1, 3, 2, 4).reshape(qs.shape[0], nbr, nbr) + classes[key0][None]
    mvec = np.repeat(masses, 3)
    mnorm = C.RYD2AMU / np.sqrt(np.outer(mvec, mvec))
    wmax2 = float(np.linalg.eigvalsh(Dq * mnorm[None]).max())
    scale = (omega_max_ev / C.RYD2EV) ** 2 / wmax2
    D = np.empty((cells.shape[0], nbr, nbr), np.complex128)
    for iuc, cell in enumerate(cells):
        D[iuc] = classes[tuple(np.mod(cell, mesh))] * scale
    return D


def make_vertex(cells_k, cells_g, lattvecs, nb, nbr, seed, chunk_rp: int = 64) -> np.ndarray:
    """g(np, mp, Re, sp, Rp) in Fortran order, Ry.  Returned as the numpy
    view with shape (nb, nb, nwsk, nbr, nwsg) whose memory is column-major."""
    nwsk, nwsg = cells_k.shape[0], cells_g.shape[0]
    ek = np.exp(-_cart_len_angstrom(lattvecs, cells_k) / 2.0)
    eg = np.exp(-_cart_len_angstrom(lattvecs, cells_g) / 2.0)
    amp = 0.1 / C.RYD2EV
    # C-order array with reversed axes == Fortran order of (nb,nb,nwsk,nbr,nwsg)
    out = np.empty((nwsg, nbr, nwsk, nb, nb), np.complex128)
    shape = (nbr, nwsk, nb, nb)
    th = np.empty(shape)
    tmp = np.empty(shape)
    env = (amp * ek)[None, :, None, None] * np.ones(shape)
    for rp in range(nwsg):
        # one stream per Rp cell (chunk_rp cells share a key, as a block)
        r0 = (rp // chunk_rp) * chunk_rp
        if rp == r0:
            rng = np.random.default_rng([seed, _TID_G, r0])
        rng.random(out=th)
        th *= 2.0 * np.pi
        o = out[rp]
        np.cos(th, out=tmp)
        np.multiply(tmp, env, out=tmp)
        tmp *= eg[rp]
        o.real = tmp
        np.sin(th, out=tmp)
        np.multiply(tmp, env, out=tmp)
        tmp *= eg[rp]
        o.imag = tmp
    return out.T  # shape (nb, nb, nwsk, nbr, nwsg), F-contiguous


def band_energies(mat: Material, kfrac: np.ndarray) -> np.ndarray:
    """Plain numpy H(k) eigenvalues (eV), used only to place enref/fsthick."""
    ph = np.exp(2j * np.pi * kfrac @ mat.rcells_k.T) / mat.elwsdeg[None, :]
    Hk = np.einsum("kr,rij->kij", ph, mat.Hwann)
    Hk = 0.5 * (Hk + np.conj(np.swapaxes(Hk, 1, 2)))
    return np.linalg.eigvalsh(Hk) * C.RYD2EV


def choose_window(mat: Material, bands_in_window: float, rng_seed: int, fsthick=None):
    """enref = median of (e4+e5)/2 over a 6^3 k sample; fsthick either given or
    the quantile of |e - enref| keeping `bands_in_window` bands per k on average."""
    g = np.stack(np.meshgrid(*[np.arange(6) / 6.0] * 3, indexing="ij"), -1).reshape(-1, 3)
    e = band_energies(mat, g)
    nb = e.shape[1]
    lo, hi = max(0, nb // 2 - 1), min(nb - 1, nb // 2)
    enref = float(np.median(0.5 * (e[:, lo] + e[:, hi])))
    if fsthick is None:
        dev = np.sort(np.abs(e - enref).ravel())
        frac = min(1.0, bands_in_window / nb)
        fsthick = float(dev[min(len(dev) - 1, int(frac * len(dev)))])
    return enref, float(fsthick)


def make_material(name: str, a_angstrom: float, masses, nb: int, ws_mesh, wvmesh,
                  seed: int, mesh_ref: int = 1, fsthick=None, bands_in_window=4.0,
                  basis=None, vertex: str = "host", window: bool = True) -> Material:
    """vertex "host": gwann generated here (make_vertex); "device": left None
    and generated in HBM by the library (elph_synth_vertex, same amplitude law,
    different random phases).  window False skips choose_window."""
    lat = fcc_lattice(a_angstrom)
    masses = np.asarray(masses, np.float64)
    nat = masses.shape[0]
    if basis is None:
        basis = np.zeros((3, nat))
        for i in range(nat):
            basis[:, i] = 0.25 * i
    ws_mesh = np.asarray(ws_mesh, np.int64)
    cells, degs = ws_cells(lat, ws_mesh)
    H = make_hamiltonian(cells, lat, nb, seed)
    D = make_force_constants(cells, lat, basis, masses, ws_mesh)
    g = make_vertex(cells, cells, lat, nb, 3 * nat, seed) if vertex == "host" else None
    mat = Material(name=name, lattvecs=lat, basis=basis, masses=masses,
                   coarse_qmesh=ws_mesh.copy(),
                   rcells_k=cells, elwsdeg=degs, rcells_q=cells.copy(), phwsdeg=degs.copy(),
                   rcells_g=cells.copy(), gwsdeg=degs.copy(), Hwann=H, Dphwann=D, gwann=g,
                   wvmesh=np.asarray(wvmesh, np.int64), mesh_ref=mesh_ref,
                   vertex_seed=None if g is not None else seed)
    if window:
        mat.enref, mat.fsthick = choose_window(mat, bands_in_window, seed, fsthick)
    return mat


SEED_BASE = 0xE1B01700


def config(idx: int) -> Material:
    """The named workloads.  0: Si-like, 6^3 WS grid, 12^3 k = q, +-0.5 eV.
    1/2: 3C-SiC-like, 8^3 WS grid, 24^3 k = q, window keeping ~4 of 8 bands."""
    if idx == 0:
        return make_material("si_like_ws6_k12", 5.43, [27.976928, 27.976928], 8,
                             [6, 6, 6], [12, 12, 12], SEED_BASE + 0, fsthick=0.5)
    if idx in (1, 2):
        return make_material("sic_like_ws8_k24", 4.36, [28.0855, 12.0107], 8,
                             [8, 8, 8], [24, 24, 24], SEED_BASE + 1, bands_in_window=4.0)
    raise ValueError(f"unknown config {idx}")


MULTI_SEED_BASE = 0xE1B03000
_WS_COUNT: dict = {}
_MASSES = [28.0855, 12.0107, 15.999, 24.305, 69.723, 74.92, 10.81, 30.97, 26.98, 14.007, 72.63, 118.71]


def ws_count(n: int) -> int:
    """Number of cells of the WS set of an fcc n^3 grid (independent of a)."""
    if n not in _WS_COUNT:
        _WS_COUNT[n] = ws_cells(fcc_lattice(4.36), np.array([n, n, n], np.int64))[0].shape[0]
    return _WS_COUNT[n]


def multi_material_draw(i: int, budget_bytes: float = 4e9, seed_base: int = MULTI_SEED_BASE,
                        pair_floor: bool = False):
    """The size draw of material i of the multi-material generator, without
    building it: dict(nat, nb, n, wv, a, nws, rng) where rng is the stream
    positioned after the draw (the basis is drawn from it next).  The WS grid
    n^3 (n uniform in 4..9), the fine mesh and the lattice constant are drawn
    once; atoms and bands are drawn again until the vertex fits the budget, so
    the grid sizes stay uniform.  pair_floor: also drawn again for one atom on
    a 12^3 mesh -- a single atom keeps the full cubic group, whose 72 IBZ k
    give at most 72 * 1728 = 124416 k-q pairs, below the batched workload's
    floor of 0.5e6 (every other draw reaches it)."""
    rng = np.random.default_rng(seed_base + i)
    n = int(rng.integers(4, 10))
    wv = int(rng.choice([12, 16, 20, 24]))
    a = float(rng.uniform(4.0, 6.0))
    nws = ws_count(n)
    for _ in range(10000):
        nat = int(rng.integers(1, 9))
        nb = int(rng.integers(4, 27))
        if pair_floor and nat == 1 and wv == 12:
            continue
        if 16 * nb * nb * 3 * nat * nws * nws <= budget_bytes:
            return dict(nat=nat, nb=nb, n=n, wv=wv, a=a, nws=nws, rng=rng)
    raise ValueError("no draw within the budget")


def multi_material(i: int, budget_bytes: float = 4e9, vertex: str = "device", window: bool = True,
                   seed_base: int = MULTI_SEED_BASE, bands_in_window=None, pair_floor: bool = False) -> Material:
    """Material i of the multi-material generator: atoms per cell uniform in
    1..8 (3..24 branches), Wannier bands uniform in 4..26, WS grid n^3 with n
    uniform in 4..9 (93..857 cells for fcc), fine mesh from {12, 16, 20, 24}^3.
    Atoms and bands are drawn again (same stream) until 16 nb^2 nbr nws^2 <=
    budget_bytes.
    The vertex follows the law of config 0 (device-generated by default)."""
    d = multi_material_draw(i, budget_bytes, seed_base, pair_floor)
    nat, nb, n, wv, a, rng = d["nat"], d["nb"], d["n"], d["wv"], d["a"], d["rng"]
    masses = [_MASSES[j % len(_MASSES)] for j in range(nat)]
    basis = np.zeros((3, nat))
    for j in range(nat):
        basis[:, j] = (j / nat) * np.ones(3) + 0.05 * rng.standard_normal(3)
    return make_material(f"multi{i}_nb{nb}_na{nat}_ws{n}_k{wv}", a, masses, nb, [n] * 3, [wv] * 3,
                         seed_base + i, basis=basis, vertex=vertex, window=window,
                         bands_in_window=nb / 2.0 if bands_in_window is None else bands_in_window)


SERVING_BUDGET = 8e9
SERVING_MIN_BYTES = 2e9


def serving_indices(count: int = 8, budget_bytes: float = SERVING_BUDGET,
                    min_bytes: float = SERVING_MIN_BYTES) -> list:
    """The resident set of the serving workload: the first `count` draws of the
    multi-material generator (vertex cap `budget_bytes`) whose Wannier vertex
    holds at least `min_bytes` (eight draws: 40.3 GB in total)."""
    out, i = [], 0
    while len(out) < count:
        d = multi_material_draw(i, budget_bytes)
        if 16 * d["nb"] ** 2 * 3 * d["nat"] * d["nws"] ** 2 >= min_bytes:
            out.append(i)
        i += 1
    return out


def tiny(nb=2, natoms=1, ws=3, wv=4, seed=7, bands_in_window=None, mesh_ref=1) -> Material:
    """Small materials for parity tests (nb <= 8, few WS cells)."""
    masses = [28.0855, 12.0107, 15.999, 24.305, 69.723, 74.92, 10.81, 30.97][:natoms]
    biw = nb / 2.0 if bands_in_window is None else bands_in_window
    return make_material(f"tiny_nb{nb}_na{natoms}_ws{ws}_k{wv}", 4.36, masses, nb,
                         [ws] * 3, [wv] * 3, seed, mesh_ref=mesh_ref, bands_in_window=biw)


def with_polar(mat: Material, kind: str = "sic", seed: int = 0) -> Material:
    """A polar copy of `mat` (Born effective charges and dielectric tensor).
    kind "sic": the 3C-SiC values of the reference's test/3C-SiC/input.nml
    (epsilon = 7.4190074 I, Z*_1 = -Z*_2 = 2.7514963 I; two atoms).
    kind "random": a symmetric positive-definite anisotropic epsilon and
    non-symmetric Born tensors with sum_a Z*_a = 0 (charge neutrality)."""
    nat = mat.natoms
    if kind == "sic":
        if nat != 2:
            raise ValueError("the 3C-SiC charges need two atoms")
        eps = 7.4190074 * np.eye(3)
        born = np.zeros((3, 3, 2))
        born[:, :, 0] = 2.7514963 * np.eye(3)
        born[:, :, 1] = -2.7514963 * np.eye(3)
    elif kind == "random":
        rng = np.random.default_rng(seed)
        a = rng.standard_normal((3, 3)) * 0.4
        eps = 6.0 * np.eye(3) + a @ a.T
        born = rng.standard_normal((3, 3, nat)) * 0.5
        for i in range(nat):
            born[:, :, i] += (2.0 if i % 2 == 0 else -2.0) * np.eye(3)
        born -= born.mean(axis=2, keepdims=True)
    else:
        raise ValueError(f"unknown polar kind {kind!r}")
    return dataclasses.replace(mat, name=mat.name + f"_polar_{kind}", polar=True,
                               epsilon=np.asfortranarray(eps), born=np.asfortranarray(born))


# --------------------------------------------------------------------------
# EPW file format writer (the reader is the product's, in the C++ library)
# --------------------------------------------------------------------------

def _fmt_c(z: complex) -> str:
    return f"({z.real:.17E},{z.imag:.17E})"


def _fmt_c_fortran(z: complex, i: int) -> str:
    """Other forms a Fortran list-directed read accepts for the same value:
    blanks around both parts, D exponents, the pair split over two records,
    values after the first on the record (ignored by the read)."""
    re = f"{z.real:.17E}".replace("E", "D")
    im = f"{z.imag:.17E}".replace("E", "d") if i % 2 else f"{z.imag:.17E}"
    if i % 3 == 0:
        return f"  ( {re} ,\n   {im} )  9.9 -1"
    return f" (  {re}  , {im}  )   ,7.0"


def write_epw(mat: Material, outdir: str, ef: float = 0.0, style: str = "plain") -> None:
    """Write epwdata.fmt, epmatwp1, rcells_{k,q,g}, wsdeg_{k,q,g}
    (layout of src/wannier.f90:313-427).  style "fortran" writes the text
    files in other forms that the reference's list-directed reads accept
    (blanks, D exponents, comma separators, records split and continued,
    trailing values), with the same values."""
    os.makedirs(outdir, exist_ok=True)
    nb, nbr = mat.nb, mat.nbr
    fort = style == "fortran"
    ndummy = (nbr // 3 + 1) * 9
    with open(os.path.join(outdir, "epwdata.fmt"), "w") as f:
        if fort:
            f.write(f"  {ef:.17E}   123.0  unused\n".replace("E", "D"))
            f.write(f" {nb},{mat.nwsk} , {nbr}\n   {mat.nwsq}   {mat.nwsg}  99\n")
            vals = ["0.0D0"] * ndummy
            for j in range(0, ndummy, 4):
                f.write(" , ".join(vals[j:j + 4]) + "\n")
        else:
            f.write(f"{ef:.17E}\n")
            f.write(f"{nb} {mat.nwsk} {nbr} {mat.nwsq} {mat.nwsg}\n")
            f.write(" ".join(["0.0"] * ndummy) + "\n")
        i = 0
        for ib in range(nb):
            for jb in range(nb):
                for iuc in range(mat.nwsk):
                    z = mat.Hwann[iuc, ib, jb]
                    f.write((_fmt_c_fortran(z, i) if fort else _fmt_c(z)) + "\n")
                    i += 1
        for ib in range(nbr):
            for jb in range(nbr):
                for iuc in range(mat.nwsq):
                    z = mat.Dphwann[iuc, ib, jb]
                    f.write((_fmt_c_fortran(z, i) if fort else _fmt_c(z)) + "\n")
                    i += 1
    np.asfortranarray(mat.gwann).ravel(order="F").tofile(os.path.join(outdir, "epmatwp1"))
    for tag, cells, deg in (("k", mat.rcells_k, mat.elwsdeg), ("q", mat.rcells_q, mat.phwsdeg),
                            ("g", mat.rcells_g, mat.gwsdeg)):
        with open(os.path.join(outdir, f"rcells_{tag}"), "w") as f:
            for c in cells:
                f.write(f" {c[0]}, {c[1]},{c[2]}  0\n" if fort else f"{c[0]:10d} {c[1]:10d} {c[2]:10d}\n")
        with open(os.path.join(outdir, f"wsdeg_{tag}"), "w") as f:
            for d in deg:
                f.write(f"  {d} 5\n" if fort else f"{d:10d}\n")
