"""ctypes front end of the CPU oracle (test infrastructure only).

Loaded by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg as
the checker; the product never imports it.  Array conventions follow the
reference (column-major, 1-based indices where the reference returns them).
"""
from __future__ import annotations

import ctypes as ct
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

c_i64p = ct.POINTER(ct.c_int64)
c_dblp = ct.POINTER(ct.c_double)
c_i32p = ct.POINTER(ct.c_int32)


class OrcWannier(ct.Structure):
    _fields_ = [
        ("nb", ct.c_int64), ("nbr", ct.c_int64), ("nwsk", ct.c_int64), ("nwsq", ct.c_int64),
        ("nwsg", ct.c_int64), ("natoms", ct.c_int64), ("coarse_qmesh", ct.c_int64 * 3),
        ("rcells_k", c_i64p), ("rcells_q", c_i64p), ("rcells_g", c_i64p),
        ("elwsdeg", c_i64p), ("phwsdeg", c_i64p), ("gwsdeg", c_i64p),
        ("Hwann", c_dblp), ("Dphwann", c_dblp), ("gwann", c_dblp),
        ("lattvecs", ct.c_double * 9), ("basis", c_dblp), ("masses", c_dblp),
        ("polar", ct.c_int32), ("epsilon", ct.c_double * 9), ("born", c_dblp),
    ]


class OrcVec(ct.Structure):
    _fields_ = [("frac", ct.c_double * 3), ("cart", ct.c_double * 3),
                ("ints", ct.c_int64 * 3), ("muxed", ct.c_int64)]


class OrcElectrons(ct.Structure):
    _fields_ = [
        ("wvmesh", ct.c_int64 * 3), ("enref", ct.c_double), ("fsthick", ct.c_double),
        ("nwv_irred", ct.c_int64), ("nwv", ct.c_int64), ("nstates_inwindow", ct.c_int64),
        ("indexlist_irred", c_i64p), ("ens_irred", c_dblp), ("evecs_irred", c_dblp),
        ("indexlist", c_i64p), ("ens", c_dblp), ("evecs", c_dblp),
    ]


class OrcPhonons(ct.Structure):
    _fields_ = [("wvmesh", ct.c_int64 * 3), ("nwv", ct.c_int64), ("ens", c_dblp), ("evecs", c_dblp),
                ("nwv_irred", ct.c_int64), ("indexlist_irred", c_i64p)]


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ct.CDLL(_LIB_PATH)
        L.orc_mux_vector.restype = ct.c_int64
        L.orc_mux_state.restype = ct.c_int64
        L.orc_binsearch.restype = ct.c_int64
        L.orc_g2.restype = ct.c_double
        L.orc_qrotations.restype = ct.c_int64
        L.orc_find_irred_wedge.restype = ct.c_int64
        L.orc_delta_fn_tetra.restype = ct.c_double
        L.orc_delta_fn_triang.restype = ct.c_double
        L.orc_bose.restype = ct.c_double
        L.orc_fermi.restype = ct.c_double
        _lib = L
    return _lib


def _p(a, t=c_dblp):
    return a.ctypes.data_as(t)


def _f64(a):
    return np.asfortranarray(np.asarray(a, dtype=np.float64))


def _i64(a):
    return np.asfortranarray(np.asarray(a, dtype=np.int64))


def _c128(a):
    return np.asfortranarray(np.asarray(a, dtype=np.complex128))


class Wannier:
    """Holds the oracle's view of a synth.Material (keeps arrays alive)."""

    def __init__(self, mat):
        self.mat = mat
        self._keep = []
        s = OrcWannier()
        s.nb, s.nbr, s.nwsk, s.nwsq, s.nwsg, s.natoms = (mat.nb, mat.nbr, mat.nwsk, mat.nwsq,
                                                         mat.nwsg, mat.natoms)
        for i in range(3):
            s.coarse_qmesh[i] = int(mat.coarse_qmesh[i])
        for name in ("rcells_k", "rcells_q", "rcells_g"):
            a = _i64(getattr(mat, name))
            self._keep.append(a)
            setattr(s, name, _p(a, c_i64p))
        for name in ("elwsdeg", "phwsdeg", "gwsdeg"):
            a = _i64(getattr(mat, name))
            self._keep.append(a)
            setattr(s, name, _p(a, c_i64p))
        for name in ("Hwann", "Dphwann", "gwann"):
            a = _c128(getattr(mat, name))
            self._keep.append(a)
            setattr(s, name, _p(a))
        L = _f64(mat.lattvecs)
        for i in range(9):
            s.lattvecs[i] = L.ravel(order="F")[i]
        b = _f64(mat.basis)
        m = _f64(mat.masses)
        self._keep += [b, m]
        s.basis, s.masses = _p(b), _p(m)
        s.polar = 1 if mat.polar else 0
        if mat.polar:
            e = _f64(mat.epsilon)
            for i in range(9):
                s.epsilon[i] = e.ravel(order="F")[i]
            bo = _f64(mat.born)
            self._keep.append(bo)
            s.born = _p(bo)
        self.s = s

    # ---- wannier.f90 ----
    def el_wann(self, kvecs, velocities=True, scissor=None):
        k = _f64(np.atleast_2d(kvecs))
        nk, nb = k.shape[0], self.mat.nb
        en = np.zeros((nk, nb), order="F")
        ev = np.zeros((nk, nb, nb), np.complex128, order="F")
        vel = np.zeros((nk, nb, 3), order="F") if velocities else None
        sc = _f64(scissor) if scissor is not None else None
        st = lib().orc_el_wann(ct.byref(self.s), ct.c_int64(nk), _p(k), _p(en),
                               _p(vel) if vel is not None else None, _p(ev),
                               _p(sc) if sc is not None else None)
        assert st == 0
        return en, ev, vel

    def ph_wann(self, qvecs, velocities=False):
        q = _f64(np.atleast_2d(qvecs))
        nq, nbr = q.shape[0], self.mat.nbr
        en = np.zeros((nq, nbr), order="F")
        ev = np.zeros((nq, nbr, nbr), np.complex128, order="F")
        vel = np.zeros((nq, nbr, 3), order="F") if velocities else None
        st = lib().orc_ph_wann(ct.byref(self.s), ct.c_int64(nq), _p(q), _p(en), _p(ev),
                               _p(vel) if vel is not None else None)
        if st == 2:
            raise RuntimeError("Large negative phonon energy found! Stopping!")
        assert st == 0
        return en, ev, vel

    def dyn_nonanalytic(self, qcart, derivative=False):
        """D_L(q) (nbr,nbr) and optionally dD_L/dq (nbr,nbr,3); q Cartesian, 1/bohr."""
        nbr = self.mat.nbr
        q = _f64(qcart)
        d = np.zeros((nbr, nbr), np.complex128, order="F")
        dd = np.zeros((nbr, nbr, 3), np.complex128, order="F") if derivative else None
        lib().orc_dyn_nonanalytic(ct.byref(self.s), _p(q), _p(d), _p(dd) if dd is not None else None)
        return (d, dd) if derivative else d

    def long_range_prefac(self, qcart, u):
        """glprefac for q Cartesian (1/bohr) and u = e_q / sqrt(M) (nbr,)."""
        q, uu = _f64(qcart), _c128(u)
        out = np.zeros(1, np.complex128)
        lib().orc_long_range_prefac(ct.byref(self.s), _p(q), _p(uu), _p(out))
        return complex(out[0])

    def gkRp(self, kvec):
        m = self.mat
        out = np.zeros((m.nb, m.nb, m.nbr, m.nwsq), np.complex128, order="F")
        k = _f64(kvec)
        lib().orc_gkRp(ct.byref(self.s), _p(k), _p(out))
        return out

    def gReq(self, qvec):
        m = self.mat
        out = np.zeros((m.nb, m.nb, m.nbr, m.nwsk), np.complex128, order="F")
        q = _f64(qvec)
        lib().orc_gReq(ct.byref(self.s), _p(q), _p(out))
        return out

    def g2(self, kvec, qvec, el_evec_k, el_evec_kp, ph_evec_q, ph_en, gmixed, wannspace):
        if wannspace not in ("el", "ph"):
            raise ValueError("Invalid value of wannspace in call to g2_epw. Exiting.")
        k, q = _f64(kvec), _f64(qvec)
        a, b, c = _c128(el_evec_k), _c128(el_evec_kp), _c128(ph_evec_q)
        g = _c128(gmixed)
        return lib().orc_g2(ct.byref(self.s), _p(k), _p(q), _p(a), ct.c_int64(1), _p(b), ct.c_int64(1),
                            _p(c), ct.c_int64(1), ct.c_double(ph_en), _p(g),
                            ct.c_int(1 if wannspace == "ph" else 0))


# ---- symmetry ----
def find_rotations(mat, symprec=1e-5):
    types = np.array(atom_types(mat.masses), np.int64)
    rot = np.zeros(9 * 48, np.int32)
    b = _f64(mat.basis)
    L = _f64(mat.lattvecs)
    n = lib().orc_find_rotations(_p(L), _p(b), _p(types, c_i64p), ct.c_int64(mat.natoms),
                                 ct.c_double(symprec), _p(rot, c_i32p))
    return rot[:9 * n].reshape(n, 9)


def atom_types(masses):
    types, seen = [], {}
    for m in masses:
        key = round(float(m), 6)
        if key not in seen:
            seen[key] = len(seen) + 1
        types.append(seen[key])
    return types


def qrotations(mat, mesh):
    rot = find_rotations(mat)
    L = _f64(mat.lattvecs)
    q = np.zeros(9 * 2 * rot.shape[0])
    r = np.ascontiguousarray(rot.ravel(), np.int32)
    mesh = _i64(mesh)
    n = lib().orc_qrotations(_p(L), _p(r, c_i32p), ct.c_int64(rot.shape[0]), _p(mesh, c_i64p), _p(q), None)
    return q[:9 * n].copy(), int(n), int(rot.shape[0])


def cartesian_rotations(mat, mesh):
    """(crotations_orig (3,3,nsymm), crotations (3,3,nsymm_rot) in the order of
    qrotations(mat, mesh)) -- src/symmetry.f90:131-137, 154-159, 191-205."""
    rot = find_rotations(mat)
    L = _f64(mat.lattvecs)
    r = np.ascontiguousarray(rot.ravel(), np.int32)
    ns = rot.shape[0]
    co = np.zeros(9 * ns)
    lib().orc_cartesian_rotations(_p(L), _p(r, c_i32p), ct.c_int64(ns), _p(co))
    q, c = np.zeros(9 * 2 * ns), np.zeros(9 * 2 * ns)
    n = lib().orc_qrotations(_p(L), _p(r, c_i32p), ct.c_int64(ns), _p(_i64(mesh), c_i64p), _p(q), _p(c))
    return (co.reshape((3, 3, ns), order="F").copy(order="F"),
            c[:9 * n].reshape((3, 3, n), order="F").copy(order="F"))


def symmetrizers(mesh, qrot, nsr, nsymm, crot_orig, indexlist=None):
    """el%symmetrizers (src/electron.f90:462-476), (3, 3, nk)."""
    mesh = _i64(mesh)
    il = _i64(indexlist) if indexlist is not None else None
    nk = int(np.prod(mesh)) if indexlist is None else len(indexlist)
    out = np.zeros((3, 3, nk), order="F")
    co = _f64(crot_orig)
    lib().orc_symmetrizers(_p(mesh, c_i64p), ct.c_int64(nsr), _p(qrot), ct.c_int64(nsymm), _p(co), ct.c_int64(nk),
                           _p(il, c_i64p) if il is not None else None, _p(out))
    return out


def irred_wedge(mesh, qrot, nsr, indexlist=None, maps=False):
    """find_irred_wedge: (indexlist_irred, nequiv) and, with maps, ibz2fbz_map
    (nsr, nk_irred, 2) column-major."""
    mesh = _i64(mesh)
    nlist = int(np.prod(mesh)) if indexlist is None else len(indexlist)
    irr = np.zeros(nlist, np.int64)
    neq = np.zeros(nlist, np.int64)
    il = _i64(indexlist) if indexlist is not None else None
    m = np.zeros((nsr, nlist, 2), np.int64, order="F") if maps else None
    n = lib().orc_find_irred_wedge(_p(mesh, c_i64p), ct.c_int64(nsr), _p(qrot),
                                   _p(il, c_i64p) if il is not None else None, ct.c_int64(nlist),
                                   _p(irr, c_i64p), _p(neq, c_i64p), _p(m, c_i64p) if maps else None)
    if maps:
        return irr[:n].copy(), neq[:n].copy(), np.asfortranarray(m[:, :n, :])
    return irr[:n].copy(), neq[:n].copy()


def field_term(field, nequiv, ibz2fbz_map, T, chempot, ens, vels, rates_ibz, indexlist):
    """calculate_field_term for electrons (bte.f90:912-1011): (nk, nb, 3)."""
    ens, vels, rates = _f64(ens), _f64(vels), _f64(rates_ibz)
    m, neq, il = _i64(ibz2fbz_map), _i64(nequiv), _i64(indexlist)
    nk, nb = ens.shape
    out = np.zeros((nk, nb, 3), order="F")
    lib().orc_field_term(ct.c_char(field.encode()), _p(neq, c_i64p), ct.c_int64(len(neq)), ct.c_int64(m.shape[0]),
                         _p(m, c_i64p), ct.c_double(T), ct.c_double(chempot), ct.c_int64(nk), ct.c_int64(nb),
                         _p(ens), _p(vels), _p(rates), _p(il, c_i64p), _p(out))
    return out


def symmetrize_field_term(symmetrizers, field):
    """bte.f90:241-244: F(ik, ib, :) <- S_ik F(ik, ib, :)."""
    F = np.array(field, np.float64, order="F")
    S = _f64(symmetrizers)
    lib().orc_symmetrize_field_term(ct.c_int64(F.shape[0]), ct.c_int64(F.shape[1]), _p(S), _p(F))
    return F


def transport_coeff(field, T, deg, chempot, ens, vels, volume, mesh, response, crot=None):
    """calculate_transport_coeff for electrons (bz_sums.f90:484-611):
    (trans_coeff_hc, trans_coeff_cc), each (nb, 3, 3); crot (3, 3, nrot)."""
    ens, vels, resp = _f64(ens), _f64(vels), _f64(response)
    nk, nb = ens.shape
    mesh = _i64(mesh)
    cr = _f64(crot) if crot is not None else np.zeros((3, 3, 1), order="F")
    nrot = 0 if crot is None else cr.shape[2]
    hc, cc = np.zeros((nb, 3, 3), order="F"), np.zeros((nb, 3, 3), order="F")
    lib().orc_transport_coeff(ct.c_char(field.encode()), ct.c_double(T), ct.c_int64(deg), ct.c_double(chempot),
                              ct.c_int64(nk), ct.c_int64(nb), _p(ens), _p(vels), ct.c_double(volume),
                              _p(mesh, c_i64p), _p(resp), ct.c_int64(nrot), _p(cr), _p(hc), _p(cc))
    return hc, cc


class Electrons:
    def __init__(self, w: Wannier, qrot, nsr, wvmesh, enref, fsthick):
        self.s = OrcElectrons()
        mesh = _i64(wvmesh)
        st = lib().orc_calculate_electrons(ct.byref(w.s), _p(qrot), ct.c_int64(nsr), _p(mesh, c_i64p),
                                           ct.c_double(enref), ct.c_double(fsthick), ct.byref(self.s))
        if st != 0:
            raise RuntimeError(f"orc_calculate_electrons failed ({st})")
        s, nb = self.s, w.mat.nb
        self.nb = nb
        self.wvmesh = np.array(wvmesh, np.int64)
        self.enref, self.fsthick = enref, fsthick
        self.nwv, self.nwv_irred, self.nstates_inwindow = s.nwv, s.nwv_irred, s.nstates_inwindow

        def arr(ptr, shape, dt):
            n = int(np.prod(shape))
            buf = np.ctypeslib.as_array(ptr, shape=(n * (2 if dt == np.complex128 else 1),))
            if dt == np.complex128:
                buf = buf.view(np.complex128)
            return np.array(buf.reshape(shape, order="F"), order="F")
        self.indexlist = arr(s.indexlist, (s.nwv,), np.int64)
        self.indexlist_irred = arr(s.indexlist_irred, (s.nwv_irred,), np.int64)
        self.ens = arr(s.ens, (s.nwv, nb), np.float64)
        self.ens_irred = arr(s.ens_irred, (s.nwv_irred, nb), np.float64)
        self.evecs = arr(s.evecs, (s.nwv, nb, nb), np.complex128)
        self.evecs_irred = arr(s.evecs_irred, (s.nwv_irred, nb, nb), np.complex128)

    def __del__(self):
        try:
            lib().orc_free_electrons(ct.byref(self.s))
        except Exception:
            pass


class Phonons:
    def __init__(self, w: Wannier, qrot, nsr, wvmesh):
        self.s = OrcPhonons()
        mesh = _i64(wvmesh)
        st = lib().orc_calculate_phonons(ct.byref(w.s), _p(qrot), ct.c_int64(nsr), _p(mesh, c_i64p),
                                         ct.byref(self.s))
        if st != 0:
            raise RuntimeError(f"orc_calculate_phonons failed ({st})")
        s, nbr = self.s, w.mat.nbr
        self.wvmesh = np.array(wvmesh, np.int64)
        self.nwv = s.nwv
        self.nwv_irred = s.nwv_irred
        self.ens = np.array(np.ctypeslib.as_array(s.ens, shape=(s.nwv * nbr,)).reshape((s.nwv, nbr), order="F"),
                            order="F")
        ev = np.ctypeslib.as_array(s.evecs, shape=(2 * s.nwv * nbr * nbr,)).view(np.complex128)
        self.evecs = np.array(ev.reshape((s.nwv, nbr, nbr), order="F"), order="F")
        self.indexlist_irred = np.array(np.ctypeslib.as_array(s.indexlist_irred, shape=(s.nwv_irred,)))

    def __del__(self):
        try:
            lib().orc_free_phonons(ct.byref(self.s))
        except Exception:
            pass


def _electrons_struct(w, el):
    """Build an OrcElectrons pointing at numpy arrays of an electron-state object
    (oracle-made or product-made), so both sides can be driven by the same data."""
    s = OrcElectrons()
    keep = []
    for i in range(3):
        s.wvmesh[i] = int(el.wvmesh[i])
    s.enref, s.fsthick = el.enref, el.fsthick
    s.nwv, s.nwv_irred, s.nstates_inwindow = el.nwv, el.nwv_irred, el.nstates_inwindow
    for name, conv, t in (("indexlist", _i64, c_i64p), ("indexlist_irred", _i64, c_i64p),
                          ("ens", _f64, c_dblp), ("ens_irred", _f64, c_dblp),
                          ("evecs", _c128, c_dblp), ("evecs_irred", _c128, c_dblp)):
        a = conv(getattr(el, name))
        keep.append(a)
        setattr(s, name, _p(a, t))
    return s, keep


def _phonons_struct(ph):
    s = OrcPhonons()
    for i in range(3):
        s.wvmesh[i] = int(ph.wvmesh[i])
    s.nwv = ph.nwv
    e, v, irr = _f64(ph.ens), _c128(ph.evecs), _i64(ph.indexlist_irred)
    s.ens, s.evecs = _p(e), _p(v)
    s.nwv_irred = len(irr)
    s.indexlist_irred = _p(irr, c_i64p)
    return s, [e, v, irr]


def ibzk_g2(w: Wannier, el, ph, mesh_ref=1, istate_first=1, istate_last=None, nthreads=0, ikp_limit=-1):
    """calculate_eph_interaction_ibzk('g') for IBZ states istate_first..istate_last.
    Returns (counts (-1 = skipped), g2 rows of length nstates_inwindow*nbr)."""
    nb, nbr = w.mat.nb, w.mat.nbr
    if istate_last is None:
        istate_last = el.nwv_irred * nb
    nst = istate_last - istate_first + 1
    nprocs = el.nstates_inwindow * nbr
    counts = np.zeros(nst, np.int64)
    out = np.zeros((nst, nprocs))
    es, k1 = _electrons_struct(w, el)
    ps, k2 = _phonons_struct(ph)
    st = lib().orc_ibzk_g2_sample(ct.byref(w.s), ct.byref(es), ct.byref(ps), ct.c_int64(mesh_ref),
                                  ct.c_int64(istate_first), ct.c_int64(istate_last), _p(counts, c_i64p), _p(out),
                                  ct.c_int(nthreads), ct.c_int64(ikp_limit))
    if st == 2:
        raise RuntimeError("Large negative phonon energy found! Stopping!")
    assert st == 0, st
    return counts, out


def ibzk_g2_window(w: Wannier, el, ph, istate_first, istate_last, ikp_first, ikp_count, mesh_ref=1, nthreads=0):
    """The ibzk driver over the final wave vectors ikp_first..ikp_first+ikp_count-1
    (1-based) of the window only.  Returns (counts, rows, offsets): rows[t, :counts[t]]
    are the entries that the full row t holds from offsets[t] on."""
    nb, nbr = w.mat.nb, w.mat.nbr
    nst = istate_last - istate_first + 1
    counts = np.zeros(nst, np.int64)
    inwin = np.abs(np.asarray(el.ens) - el.enref) <= el.fsthick          # (nwv, nb)
    before = int(inwin[:ikp_first - 1].sum()) * nbr
    out = np.zeros((nst, el.nstates_inwindow * nbr))     # the driver's row stride
    es, k1 = _electrons_struct(w, el)
    ps, k2 = _phonons_struct(ph)
    st = lib().orc_ibzk_g2_window(ct.byref(w.s), ct.byref(es), ct.byref(ps), ct.c_int64(mesh_ref),
                                  ct.c_int64(istate_first), ct.c_int64(istate_last), _p(counts, c_i64p), _p(out),
                                  ct.c_int(nthreads), ct.c_int64(ikp_first), ct.c_int64(ikp_count))
    if st == 2:
        raise RuntimeError("Large negative phonon energy found! Stopping!")
    assert st == 0, st
    return counts, out, np.full(nst, before, np.int64)


def g2_list(w: Wannier, kvecs, qvecs, el_evec_k, el_evec_kp, ph_evec_q, ph_en, gmixed, gidx, wannspace,
            nthreads=0):
    """Wannier%g2 (orc_g2) for a list of vertices, OpenMP over the list.
    gmixed: (ng, nb, nb, nbr, nws) blocks; vertex i uses block gidx[i]."""
    n = len(ph_en)
    k, q = _f64(np.reshape(kvecs, (n, 3))), _f64(np.reshape(qvecs, (n, 3)))
    a, b, c = _c128(el_evec_k), _c128(el_evec_kp), _c128(ph_evec_q)
    en = _f64(ph_en)
    gm = np.ascontiguousarray(np.stack([np.asfortranarray(g).ravel(order="F") for g in gmixed]))
    gsz = gm.shape[1]
    gi = _i64(gidx)
    out = np.zeros(n)
    lib().orc_g2_list(ct.byref(w.s), ct.c_int64(n), _p(k), _p(q), _p(a), _p(b), _p(c), _p(en), _p(gm),
                      ct.c_int64(gsz), _p(gi, c_i64p), ct.c_int(1 if wannspace == "ph" else 0), _p(out),
                      ct.c_int(nthreads))
    return out


def ibzq_g2(w: Wannier, el, ph, istate_first=1, istate_last=None, nthreads=0):
    nb, nbr = w.mat.nb, w.mat.nbr
    if istate_last is None:
        istate_last = ph.nwv_irred * nbr
    nst = istate_last - istate_first + 1
    stride = el.nstates_inwindow * nb
    counts = np.zeros(nst, np.int64)
    out = np.zeros((nst, stride))
    es, k1 = _electrons_struct(w, el)
    ps, k2 = _phonons_struct(ph)
    st = lib().orc_ibzq_g2(ct.byref(w.s), ct.byref(es), ct.byref(ps), ct.c_int64(istate_first),
                           ct.c_int64(istate_last), _p(counts, c_i64p), _p(out), ct.c_int(nthreads))
    assert st == 0, st
    return counts, out


# ---- consumers of g2 (delta.f90, interactions.f90 key 'X', RTA rates) ----
class Tetrahedra:
    """form_tetrahedra_3d + fill_tetrahedra_3d (delta.f90:494-646) over a wave
    vector list (blocks: muxed sorted indexlist, else the full mesh) with band
    energies evals (nk, nb).  Arrays in the reference's column-major shapes:
    tetra (6 nk, 4), tetracount (nk), tetramap (2, nk, 24), evals (6 nk, nb, 4)."""

    def __init__(self, mesh, evals, indexlist=None):
        mesh = _i64(mesh)
        evals = np.asfortranarray(np.asarray(evals, np.float64))
        nk, nb = evals.shape
        blocks = indexlist is not None
        il = _i64(indexlist if blocks else np.zeros(1))
        self.mesh, self.nk, self.nb, self.ntet = mesh, nk, nb, 6 * nk
        self.tetra = np.zeros((6 * nk, 4), np.int64, order="F")
        self.tetracount = np.zeros(nk, np.int64)
        self.tetramap = np.zeros((2, nk, 24), np.int64, order="F")
        lib().orc_form_tetrahedra_3d(ct.c_int64(nk), _p(mesh, c_i64p), ct.c_int(int(blocks)), _p(il, c_i64p),
                                     _p(self.tetra, c_i64p), _p(self.tetracount, c_i64p),
                                     _p(self.tetramap, c_i64p))
        self.evals = np.zeros((6 * nk, nb, 4), order="F")
        lib().orc_fill_tetrahedra_3d(ct.c_int64(6 * nk), _p(self.tetra, c_i64p), ct.c_int64(nk), ct.c_int64(nb),
                                     _p(evals), _p(self.evals))

    nvert = 4

    @property
    def simplex(self):
        """(map, count, evals) as el%simplex_map/_count/_evals hold them."""
        return self.tetramap, self.tetracount, self.evals

    def delta(self, e, ik, ib):
        """delta_fn_tetra(e, ik, ib) (delta.f90:61-225), ik/ib 1-based."""
        return lib().orc_delta_fn_tetra(ct.c_double(e), ct.c_int64(ik), ct.c_int64(ib), _p(self.mesh, c_i64p),
                                        ct.c_int64(self.nk), ct.c_int64(self.ntet), ct.c_int64(self.nb),
                                        _p(self.tetramap, c_i64p), _p(self.tetracount, c_i64p), _p(self.evals))


class Triangles:
    """form_triangles + fill_triangles (delta.f90:648-788): triang (2 nk, 3),
    triangcount (nk), triangmap (2, nk, 6), evals (2 nk, nb, 3)."""

    nvert = 3

    def __init__(self, mesh, evals, indexlist=None):
        mesh = _i64(mesh)
        evals = np.asfortranarray(np.asarray(evals, np.float64))
        nk, nb = evals.shape
        blocks = indexlist is not None
        il = _i64(indexlist if blocks else np.zeros(1))
        self.mesh, self.nk, self.nb, self.ntri = mesh, nk, nb, 2 * nk
        self.triang = np.zeros((2 * nk, 3), np.int64, order="F")
        self.triangcount = np.zeros(nk, np.int64)
        self.triangmap = np.zeros((2, nk, 6), np.int64, order="F")
        lib().orc_form_triangles(ct.c_int64(nk), _p(mesh, c_i64p), ct.c_int(int(blocks)), _p(il, c_i64p),
                                 _p(self.triang, c_i64p), _p(self.triangcount, c_i64p), _p(self.triangmap, c_i64p))
        self.evals = np.zeros((2 * nk, nb, 3), order="F")
        lib().orc_fill_triangles(ct.c_int64(2 * nk), _p(self.triang, c_i64p), ct.c_int64(nk), ct.c_int64(nb),
                                 _p(evals), _p(self.evals))

    @property
    def simplex(self):
        return self.triangmap, self.triangcount, self.evals

    def delta(self, e, ik, ib):
        """delta_fn_triang(e, ik, ib) (delta.f90:790-882), ik/ib 1-based."""
        return lib().orc_delta_fn_triang(ct.c_double(e), ct.c_int64(ik), ct.c_int64(ib), _p(self.mesh, c_i64p),
                                         ct.c_int64(self.nk), ct.c_int64(self.ntri), ct.c_int64(self.nb),
                                         _p(self.triangmap, c_i64p), _p(self.triangcount, c_i64p), _p(self.evals))


def delta_fn_tetra_raw(e, ik, ib, mesh, tetramap, tetracount, tetra_evals):
    """delta_fn_tetra on caller-made tables (shapes as in Tetrahedra)."""
    mesh = _i64(mesh)
    tm, tc, te = _i64(tetramap), _i64(tetracount), _f64(tetra_evals)
    nk, ntet, nb = tm.shape[1], te.shape[0], te.shape[1]
    return lib().orc_delta_fn_tetra(ct.c_double(e), ct.c_int64(ik), ct.c_int64(ib), _p(mesh, c_i64p),
                                    ct.c_int64(nk), ct.c_int64(ntet), ct.c_int64(nb), _p(tm, c_i64p),
                                    _p(tc, c_i64p), _p(te))


def bose(e, T):
    return lib().orc_bose(ct.c_double(e), ct.c_double(T))


def fermi(e, chempot, T):
    return lib().orc_fermi(ct.c_double(e), ct.c_double(chempot), ct.c_double(T))


def ibzk_X(w: Wannier, el, ph, tet, counts, g2, T, chempot, mesh_ref=1, istate_first=1, nthreads=0):
    """calculate_eph_interaction_ibzk('X') on g2 rows (counts, g2 as ibzk_g2
    returns them); tet: Tetrahedra or Triangles (delta.f90:47-59).  Returns (Xplus, Xminus, istate_el, istate_ph), rows of
    length nstates_inwindow*nbr (entries past each count are 0)."""
    counts = _i64(counts)
    g2 = np.ascontiguousarray(g2, np.float64)
    nst, nprocs = g2.shape
    xp, xm = np.zeros_like(g2), np.zeros_like(g2)
    se, sp = np.zeros(g2.shape, np.int64), np.zeros(g2.shape, np.int64)
    es, k1 = _electrons_struct(w, el)
    ps, k2 = _phonons_struct(ph)
    smap, scount, sev = tet.simplex
    st = lib().orc_ibzk_X(ct.byref(w.s), ct.byref(es), ct.byref(ps), ct.c_int64(mesh_ref), ct.c_int64(istate_first),
                          ct.c_int64(istate_first + nst - 1), _p(counts, c_i64p), _p(g2), ct.c_double(T),
                          ct.c_double(chempot), ct.c_int(tet.nvert), _p(smap, c_i64p), _p(scount, c_i64p), _p(sev),
                          _p(xp), _p(xm), _p(se, c_i64p), _p(sp, c_i64p), ct.c_int(nthreads))
    if st == 2:
        raise RuntimeError("Large negative phonon energy found! Stopping!")
    assert st == 0, st
    return xp, xm, se, sp


def el_rta_rates(counts, Xplus, Xminus):
    """calculate_el_rta_rates, e-ph channel (interactions.f90:2957-3063)."""
    counts = _i64(counts)
    xp, xm = np.ascontiguousarray(Xplus, np.float64), np.ascontiguousarray(Xminus, np.float64)
    out = np.zeros(len(counts))
    lib().orc_el_rta_rates(ct.c_int64(len(counts)), _p(counts, c_i64p), ct.c_int64(xp.shape[1]), _p(xp), _p(xm),
                           _p(out))
    return out


# ---- index algebra (misc.f90, vector.f90) ----
def mux_vector(v, mesh, base):
    v, mesh = _i64(v), _i64(mesh)
    return int(lib().orc_mux_vector(_p(v, c_i64p), _p(mesh, c_i64p), ct.c_int64(base)))


def demux_vector(i, mesh, base):
    v = np.zeros(3, np.int64)
    mesh = _i64(mesh)
    lib().orc_demux_vector(ct.c_int64(i), _p(v, c_i64p), _p(mesh, c_i64p), ct.c_int64(base))
    return v


def mux_state(nbands, iband, ik):
    return int(lib().orc_mux_state(ct.c_int64(nbands), ct.c_int64(iband), ct.c_int64(ik)))


def demux_state(m, nbands):
    a, b = ct.c_int64(), ct.c_int64()
    lib().orc_demux_state(ct.c_int64(m), ct.c_int64(nbands), ct.byref(a), ct.byref(b))
    return a.value, b.value


def binsearch(array, e):
    a = _i64(array)
    return int(lib().orc_binsearch(_p(a, c_i64p), ct.c_int64(len(a)), ct.c_int64(e)))


def add_and_fold(q1, q2):
    a, b, c = _f64(q1), _f64(q2), np.zeros(3)
    lib().orc_add_and_fold(_p(a), _p(b), _p(c))
    return c


def expi(x):
    out = np.zeros(2)
    lib().orc_expi(ct.c_double(x), _p(out))
    return complex(out[0], out[1])


def distribute_points(npts, num_images, image):
    c, s, e, na = ct.c_int64(), ct.c_int64(), ct.c_int64(), ct.c_int64()
    lib().orc_distribute_points(ct.c_int64(npts), ct.c_int64(num_images), ct.c_int64(image),
                                ct.byref(c), ct.byref(s), ct.byref(e), ct.byref(na))
    return c.value, s.value, e.value, na.value


def _vec(v: OrcVec):
    return dict(frac=np.array(v.frac[:]), cart=np.array(v.cart[:]), int=np.array(v.ints[:]), muxed=v.muxed)


def vec_create(ind, grid, prim):
    v = OrcVec()
    g, p = _i64(grid), _f64(prim)
    lib().orc_vec_create(ct.c_int64(ind), _p(g, c_i64p), _p(p), ct.byref(v))
    return v


def vec_add(v1, v2, grid, prim):
    v = OrcVec()
    g, p = _i64(grid), _f64(prim)
    lib().orc_vec_add(ct.byref(v1), ct.byref(v2), _p(g, c_i64p), _p(p), ct.byref(v))
    return v


def vec_sub(v1, v2, grid, prim):
    v = OrcVec()
    g, p = _i64(grid), _f64(prim)
    lib().orc_vec_sub(ct.byref(v1), ct.byref(v2), _p(g, c_i64p), _p(p), ct.byref(v))
    return v


def vec_change_grid(v1, grid):
    v = OrcVec()
    g = _i64(grid)
    lib().orc_vec_change_grid(ct.byref(v1), _p(g, c_i64p), ct.byref(v))
    return v


vec_dict = _vec
