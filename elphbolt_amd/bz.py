"""Symmetry, irreducible wedge and the electron/phonon state the vertex drivers
consume (src/symmetry.f90, src/electron.f90:321-595, src/phonon.f90:159-310),
computed by libelph_vertex.so (host bookkeeping + device band structures)."""
from __future__ import annotations

import ctypes as ct

import numpy as np

from . import _lib
from ._lib import ElphElectrons, ElphError, ElphPhonons, c_dblp, c_i64p
from .wannier import Wannier, _p, f64, i64


def atom_types(masses):
    """Species index per atom (1-based), equal masses = same species."""
    types, seen = [], {}
    for m in masses:
        key = round(float(m), 6)
        if key not in seen:
            seen[key] = len(seen) + 1
        types.append(seen[key])
    return np.array(types, np.int64)


class Symmetry:
    """`sym%calculate_symmetries` (src/symmetry.f90:66-212)."""

    def __init__(self, lattvecs, basis, masses, mesh, symprec=1e-5):
        lib = _lib.load()
        L, b = f64(lattvecs), f64(basis)
        t = atom_types(masses)
        m = i64(mesh)
        ns, nsr = ct.c_int64(), ct.c_int64()
        qr = np.zeros(9 * 96)
        st = lib.elph_symmetry(_p(L), len(masses), _p(b), _p(t, c_i64p), symprec, _p(m, c_i64p), ct.byref(ns),
                               ct.byref(nsr), _p(qr))
        if st != 0:
            raise ElphError(st, "symmetry search failed")
        self.nsymm = ns.value
        self.nsymm_rot = nsr.value
        self.qrotations = qr[:9 * nsr.value].copy()
        # Cartesian operations (src/symmetry.f90:131-137): crotations_orig
        # (3, 3, nsymm) and crotations (3, 3, nsymm_rot) in the order of qrotations
        co, cr = np.zeros(9 * 48), np.zeros(9 * 96)
        ns2, nsr2 = ct.c_int64(), ct.c_int64()
        st = lib.elph_symmetry_cartesian(_p(L), len(masses), _p(b), _p(t, c_i64p), symprec, _p(m, c_i64p),
                                         ct.byref(ns2), ct.byref(nsr2), _p(co), _p(cr))
        if st != 0 or ns2.value != self.nsymm or nsr2.value != self.nsymm_rot:
            raise ElphError(st or _lib.ELPH_EINVAL, "Cartesian symmetry operations failed")
        self.crotations_orig = co[:9 * self.nsymm].reshape((3, 3, self.nsymm), order="F").copy(order="F")
        self.crotations = cr[:9 * self.nsymm_rot].reshape((3, 3, self.nsymm_rot), order="F").copy(order="F")


def symmetrizers(mesh, sym: Symmetry, indexlist=None):
    """`el%symmetrizers` (src/electron.f90:462-476): (3, 3, nk) for the points
    of indexlist (the whole mesh when None)."""
    lib = _lib.load()
    m = i64(mesh)
    il = i64(indexlist) if indexlist is not None else None
    nk = int(np.prod(m)) if indexlist is None else len(indexlist)
    out = np.zeros((3, 3, nk), order="F")
    co = np.asfortranarray(sym.crotations_orig)
    st = lib.elph_symmetrizers(_p(m, c_i64p), sym.nsymm, _p(sym.qrotations), _p(co), nk, _p(il, c_i64p), _p(out))
    if st != 0:
        raise ElphError(st, "symmetrizers failed")
    return out


def irred_wedge(mesh, sym: Symmetry, indexlist=None, maps=False):
    """`find_irred_wedge` (src/symmetry.f90:303-451): (indexlist_irred, nequiv)
    and, with maps, ibz2fbz_map (nsymm_rot, nk_irred, 2) column-major."""
    lib = _lib.load()
    m = i64(mesh)
    nlist = int(np.prod(m)) if indexlist is None else len(indexlist)
    irr = np.zeros(nlist, np.int64)
    neq = np.zeros(nlist, np.int64)
    il = i64(indexlist) if indexlist is not None else None
    mp = np.zeros((sym.nsymm_rot, nlist, 2), np.int64, order="F") if maps else None
    n = ct.c_int64()
    st = lib.elph_irred_wedge(_p(m, c_i64p), sym.nsymm_rot, _p(sym.qrotations), nlist, _p(il, c_i64p),
                              ct.byref(n), _p(irr, c_i64p), _p(neq, c_i64p), _p(mp, c_i64p) if maps else None)
    if st != 0:
        raise ElphError(st, "Severe error: Could not find irreducible wedge.")
    if maps:
        return irr[:n.value].copy(), neq[:n.value].copy(), np.asfortranarray(mp[:, :n.value, :])
    return irr[:n.value].copy(), neq[:n.value].copy()


def _view_array(ptr, shape, complex_=False, dtype=np.float64):
    n = int(np.prod(shape))
    if dtype == np.int64:
        buf = np.ctypeslib.as_array(ptr, shape=(n,))
        return np.array(buf, copy=True)
    buf = np.ctypeslib.as_array(ptr, shape=(2 * n if complex_ else n,))
    if complex_:
        buf = buf.view(np.complex128)
    return np.array(buf.reshape(shape, order="F"), order="F")


class Electrons:
    """Fields of `type electron` the vertex path reads, from
    `calculate_electrons` (src/electron.f90:321-595)."""

    def __init__(self, wann: Wannier, sym: Symmetry, wvmesh, enref, fsthick):
        lib = _lib.load()
        self.wvmesh = np.array(wvmesh, np.int64)
        self.enref, self.fsthick = float(enref), float(fsthick)
        m = i64(wvmesh)
        state = ct.c_void_p()
        v = ElphElectrons()
        st = lib.elph_calculate_electrons(wann.ctx, sym.nsymm_rot, _p(sym.qrotations), _p(m, c_i64p),
                                          self.enref, self.fsthick, ct.byref(state), ct.byref(v))
        if st == _lib.ELPH_EWINDOW:
            raise ElphError(st, "No states found within Fermi window.")
        if st != 0:
            raise ElphError(st, "calculate_electrons failed")
        nb = wann.numwannbands
        self.nb = nb
        self.nwv, self.nwv_irred, self.nstates_inwindow = v.nwv, v.nwv_irred, v.nstates_inwindow
        self.indexlist = _view_array(v.indexlist, (v.nwv,), dtype=np.int64)
        self.indexlist_irred = _view_array(v.indexlist_irred, (v.nwv_irred,), dtype=np.int64)
        self.ens = _view_array(v.ens, (v.nwv, nb))
        self.ens_irred = _view_array(v.ens_irred, (v.nwv_irred, nb))
        self.evecs = _view_array(v.evecs, (v.nwv, nb, nb), complex_=True)
        self.evecs_irred = _view_array(v.evecs_irred, (v.nwv_irred, nb, nb), complex_=True)
        lib.elph_state_free(state)

    @classmethod
    def from_arrays(cls, **kw):
        obj = cls.__new__(cls)
        for k, val in kw.items():
            setattr(obj, k, val)
        return obj

    def transport_inputs(self, wann: Wannier, sym: Symmetry):
        """What the electron RTA step reads besides the rates, built as
        calculate_electrons leaves them: the IBZ wedge of the window's list
        (nequiv, ibz2fbz_map), the symmetrizers (src/electron.f90:462-476) and
        the velocities (nwv, nb, 3) km/s -- each list point takes its IBZ
        point's el_wann velocities rotated by the mapping operation and
        multiplied by its symmetrizer (500-522)."""
        irr, neq, m = irred_wedge(self.wvmesh, sym, indexlist=self.indexlist, maps=True)
        if not np.array_equal(irr, self.indexlist_irred):
            raise ElphError(_lib.ELPH_EWEDGE, "IBZ of the list differs from the electron state's")
        S = symmetrizers(self.wvmesh, sym, self.indexlist)
        mesh = self.wvmesh
        i0 = irr - 1
        kirr = np.stack([i0 % mesh[0], (i0 // mesh[0]) % mesh[1], i0 // (mesh[0] * mesh[1])], axis=1) / mesh
        _, _, v_irr = wann.el_wann(kirr, velocities=True, evecs=False)
        vels = np.zeros((self.nwv, self.nb, 3), order="F")
        for i in range(len(irr)):
            il = m[:neq[i], i, 1]
            rot = m[:neq[i], i, 0] - 1
            aux = np.searchsorted(self.indexlist, il)
            v = np.einsum("cdl,bd->lbc", sym.crotations[:, :, rot], v_irr[i])        # (l, nb, 3)
            vels[aux] = np.einsum("cdl,lbd->lbc", S[:, :, aux], v)
        return dict(nequiv=neq, ibz2fbz_map=m, symmetrizers=S, vels=vels, crotations=sym.crotations)


class Phonons:
    """Fields of `type phonon` the vertex path reads (Wannier branch of
    `calculate_phonons`, src/phonon.f90:159-310)."""

    def __init__(self, wann: Wannier, sym: Symmetry, wvmesh):
        lib = _lib.load()
        self.wvmesh = np.array(wvmesh, np.int64)
        m = i64(wvmesh)
        state = ct.c_void_p()
        v = ElphPhonons()
        st = lib.elph_calculate_phonons(wann.ctx, sym.nsymm_rot, _p(sym.qrotations), _p(m, c_i64p),
                                        ct.byref(state), ct.byref(v))
        if st == _lib.ELPH_ENEGPHONON:
            raise ElphError(st, "Large negative phonon energy found! Stopping!")
        if st != 0:
            raise ElphError(st, "calculate_phonons failed")
        nbr = wann.numbranches
        self.nwv, self.nwv_irred = v.nwv, v.nwv_irred
        self.ens = _view_array(v.ens, (v.nwv, nbr))
        self.evecs = _view_array(v.evecs, (v.nwv, nbr, nbr), complex_=True)
        self.indexlist_irred = _view_array(v.indexlist_irred, (v.nwv_irred,), dtype=np.int64)
        lib.elph_state_free(state)

    @classmethod
    def from_arrays(cls, **kw):
        obj = cls.__new__(cls)
        for k, val in kw.items():
            setattr(obj, k, val)
        return obj
