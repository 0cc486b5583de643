"""Host mirror of elphbolt's `type Wannier` (src/wannier.f90:46-93) over the
HIP library.  Method names, argument meaning and error behaviour follow the
reference: el_wann (430-524), ph_wann (585-782), gkRp (1073-1132), gReq
(1134-1188), g2 (903-1011).  Arrays use the reference's shapes and
column-major order; every call runs on the GPU through libelph_vertex.so.
"""
from __future__ import annotations

import ctypes as ct

import numpy as np

from . import _lib
from ._lib import ElphDims, ElphError, c_dblp, c_i64p


def _p(a, t=c_dblp):
    return a.ctypes.data_as(t) if a is not None else None


def f64(a):
    return np.asfortranarray(np.asarray(a, dtype=np.float64))


def i64(a):
    return np.asfortranarray(np.asarray(a, dtype=np.int64))


def c128(a):
    return np.asfortranarray(np.asarray(a, dtype=np.complex128))


class Wannier:
    """Wannier-space tensors resident in one GPU's HBM.

    Build from arrays (`Wannier(material)`, any object with the fields of
    `elphbolt_amd.synth.Material`) or from EPW files (`Wannier.from_epw`)."""

    def __init__(self, mat=None, device: int = 0, _ctx=None):
        self.lib = _lib.load()
        self.device = device
        self.mat = mat
        if _ctx is not None:
            self.ctx = _ctx
        else:
            d = ElphDims()
            d.numwannbands, d.numbranches = mat.nb, mat.nbr
            d.nwsk, d.nwsq, d.nwsg, d.natoms = mat.nwsk, mat.nwsq, mat.nwsg, mat.natoms
            for i in range(3):
                d.coarse_qmesh[i] = int(mat.coarse_qmesh[i])
            L = f64(mat.lattvecs).ravel(order="F")
            for i in range(9):
                d.lattvecs[i] = L[i]
            d.polar = 1 if getattr(mat, "polar", False) else 0
            if d.polar:
                e = f64(mat.epsilon).ravel(order="F")
                for i in range(9):
                    d.epsilon[i] = e[i]
            gw = None if mat.gwann is None else c128(mat.gwann)
            keep = [i64(mat.rcells_k), i64(mat.rcells_q), i64(mat.rcells_g), i64(mat.elwsdeg), i64(mat.phwsdeg),
                    i64(mat.gwsdeg), c128(mat.Hwann), c128(mat.Dphwann), gw, f64(mat.basis),
                    f64(mat.masses), f64(mat.born) if d.polar else None]
            ctx = ct.c_void_p()
            st = self.lib.elph_ctx_create(ct.byref(ctx), device, ct.byref(d), *[_p(a, c_i64p) for a in keep[:6]],
                                          *[_p(a) for a in keep[6:]])
            if st != 0:
                raise ElphError(st, "elph_ctx_create failed")
            self.ctx = ctx
            if mat.gwann is None:
                # vertex generated in HBM (synth.make_material(..., vertex="device"))
                self._check(self.lib.elph_synth_vertex(ctx, int(mat.vertex_seed) & (2 ** 64 - 1)), "synth_vertex")
        dims = np.zeros(6, np.int64)
        self.lib.elph_ctx_dims(self.ctx, _p(dims, c_i64p))
        (self.numwannbands, self.numbranches, self.nwsk, self.nwsq, self.nwsg, self.natoms) = map(int, dims)

    @classmethod
    def from_epw(cls, dirname, natoms, lattvecs, basis, masses, coarse_qmesh, device=0, polar=False,
                 epsilon=None, born=None):
        """`read_EPW_Wannier` (src/wannier.f90:289-428): epwdata.fmt, epmatwp1,
        rcells_{k,q,g}, wsdeg_{k,q,g} from `dirname`."""
        lib = _lib.load()
        ctx = ct.c_void_p()
        L, b, m, cq = f64(lattvecs), f64(basis), f64(masses), i64(coarse_qmesh)
        e = f64(epsilon) if epsilon is not None else None
        bo = f64(born) if born is not None else None
        st = lib.elph_ctx_create_epw(ct.byref(ctx), device, str(dirname).encode(), natoms, _p(L), _p(b), _p(m),
                                     _p(cq, c_i64p), 1 if polar else 0, _p(e), _p(bo))
        if st != 0:
            raise ElphError(st, f"elph_ctx_create_epw({dirname}) failed")
        return cls(None, device=device, _ctx=ctx)

    def close(self):
        if getattr(self, "ctx", None):
            self.lib.elph_ctx_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def device_buffer(self, shape, dtype=np.float64):
        """HBM buffer on this context's GPU (freed when the object is collected)."""
        return DeviceBuffer(self, shape, dtype)

    def synchronize(self):
        self._check(self.lib.elph_ctx_synchronize(self.ctx), "synchronize")

    def _check(self, st, what):
        if st != 0:
            msg = self.lib.elph_last_error(self.ctx)
            raise ElphError(st, f"{what}: {msg.decode() if msg else ''}")

    # ---- src/wannier.f90:430-524 ----
    def el_wann(self, kvecs, velocities=False, evecs=True, scissor=None):
        k = f64(np.atleast_2d(kvecs))
        nk, nb = k.shape[0], self.numwannbands
        en = np.zeros((nk, nb), order="F")
        ev = np.zeros((nk, nb, nb), np.complex128, order="F") if (evecs or velocities) else None
        vel = np.zeros((nk, nb, 3), order="F") if velocities else None
        sc = f64(scissor) if scissor is not None else None
        self._check(self.lib.elph_el_wann(self.ctx, nk, _p(k), _p(en), _p(vel), _p(ev), _p(sc)), "el_wann")
        return en, ev, vel

    # ---- src/wannier.f90:585-782 ----
    def ph_wann(self, qvecs, evecs=True, velocities=False):
        q = f64(np.atleast_2d(qvecs))
        nq, nbr = q.shape[0], self.numbranches
        en = np.zeros((nq, nbr), order="F")
        ev = np.zeros((nq, nbr, nbr), np.complex128, order="F") if (evecs or velocities) else None
        vel = np.zeros((nq, nbr, 3), order="F") if velocities else None
        st = self.lib.elph_ph_wann(self.ctx, nq, _p(q), _p(en), _p(ev), _p(vel))
        if st == _lib.ELPH_ENEGPHONON:
            raise ElphError(st, "Large negative phonon energy found! Stopping!")
        self._check(st, "ph_wann")
        return en, ev, vel

    # ---- src/wannier.f90:1073-1132 (batched over k) ----
    def gkRp(self, kvecs):
        """g(k,Rp) for each row of kvecs: array (nk, nb, nb, nbr, nwsq) whose
        trailing block per k is the Fortran layout of gkRp.ik<ik>."""
        k = f64(np.atleast_2d(kvecs))
        nk = k.shape[0]
        nb, nbr = self.numwannbands, self.numbranches
        out = np.zeros(nk * nb * nb * nbr * self.nwsq, np.complex128)
        self._check(self.lib.elph_gkRp(self.ctx, nk, _p(k), _p(out)), "gkRp")
        return out.reshape((nk, self.nwsq, nbr, nb, nb)).transpose(0, 4, 3, 2, 1)

    # ---- src/wannier.f90:1134-1188 (batched over q) ----
    def gReq(self, qvecs):
        q = f64(np.atleast_2d(qvecs))
        nq = q.shape[0]
        nb, nbr = self.numwannbands, self.numbranches
        out = np.zeros(nq * nb * nb * nbr * self.nwsk, np.complex128)
        self._check(self.lib.elph_gReq(self.ctx, nq, _p(q), _p(out)), "gReq")
        return out.reshape((nq, self.nwsk, nbr, nb, nb)).transpose(0, 4, 3, 2, 1)

    # ---- src/wannier.f90:903-1011 (list form) ----
    def g2(self, kvec, qvec, el_evec_k, el_evec_kp, ph_evec_q, ph_en, gmixed, wannspace):
        """Single vertex with the reference's argument list; see g2_list."""
        return float(self.g2_list(np.atleast_2d(kvec), np.atleast_2d(qvec), np.atleast_2d(el_evec_k),
                                   np.atleast_2d(el_evec_kp), np.atleast_2d(ph_evec_q), np.atleast_1d(ph_en),
                                   np.asarray(gmixed)[None], np.zeros(1, np.int64), wannspace)[0])

    def g2_list(self, kvecs, qvecs, el_evec_k, el_evec_kp, ph_evec_q, ph_en, gmixed, gidx, wannspace):
        if wannspace not in ("el", "ph"):
            raise ElphError(_lib.ELPH_EINVAL, "Invalid value of wannspace in call to g2_epw. Exiting.")
        n = len(ph_en)
        k, q = f64(kvecs), f64(qvecs)
        a, b, c = c128(el_evec_k), c128(el_evec_kp), c128(ph_evec_q)
        en = f64(ph_en)
        gm = np.ascontiguousarray(np.stack([np.asfortranarray(g).ravel(order="F") for g in gmixed]))
        gi = i64(gidx)
        out = np.zeros(n)
        self._check(self.lib.elph_g2(self.ctx, n, _p(k), _p(q), _p(a), _p(b), _p(c), _p(en), _p(gm), len(gm),
                                     _p(gi, c_i64p), b"p" if wannspace == "ph" else b"e", _p(out)), "g2")
        return out


def read_epw(dirname):
    """`read_EPW_Wannier` (src/wannier.f90:289-428) on the host only: a dict of
    the arrays with the reference's shapes (rcells_* (nws,3), *wsdeg (nws),
    Hwann (nwsk,nb,nb), Dphwann (nwsq,nbr,nbr), gwann (nb,nb,nwsk,nbr,nwsg),
    Fortran order), read by the library's list-directed reader."""
    lib = _lib.load()
    d = np.zeros(5, np.int64)
    st = lib.elph_epw_read_dims(str(dirname).encode(), _p(d, c_i64p))
    if st != 0:
        raise ElphError(st, f"reading EPW files in {dirname} failed")
    nb, nwsk, nbr, nwsq, nwsg = map(int, d)
    out = {"rcells_k": np.zeros((nwsk, 3), np.int64, order="F"), "rcells_q": np.zeros((nwsq, 3), np.int64, order="F"),
           "rcells_g": np.zeros((nwsg, 3), np.int64, order="F"), "elwsdeg": np.zeros(nwsk, np.int64),
           "phwsdeg": np.zeros(nwsq, np.int64), "gwsdeg": np.zeros(nwsg, np.int64),
           "Hwann": np.zeros((nwsk, nb, nb), np.complex128, order="F"),
           "Dphwann": np.zeros((nwsq, nbr, nbr), np.complex128, order="F"),
           "gwann": np.zeros((nb, nb, nwsk, nbr, nwsg), np.complex128, order="F")}
    ints = [_p(out[k], c_i64p) for k in ("rcells_k", "rcells_q", "rcells_g", "elwsdeg", "phwsdeg", "gwsdeg")]
    cplx = [_p(out[k]) for k in ("Hwann", "Dphwann", "gwann")]
    st = lib.elph_epw_read(str(dirname).encode(), _p(d, c_i64p), *ints, *cplx)
    if st != 0:
        raise ElphError(st, f"reading EPW files in {dirname} failed")
    return out


class DeviceBuffer:
    """Device memory allocated through the library (no torch in the process:
    torch wheels bundle their own HIP runtime)."""

    def __init__(self, wann: Wannier, shape, dtype=np.float64):
        self.wann, self.shape, self.dtype = wann, tuple(shape), np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = ct.c_void_p()
        wann._check(wann.lib.elph_device_alloc(wann.ctx, self.nbytes, ct.byref(p)), "device_alloc")
        self.ptr = p

    def copy_from(self, other: "DeviceBuffer"):
        """Device-to-device copy of other's bytes into this buffer."""
        n = min(self.nbytes, other.nbytes)
        self.wann._check(self.wann.lib.elph_device_copy(self.wann.ctx, self.ptr, other.ptr, n), "device_copy")

    def upload(self, host):
        """Host-to-device copy of a numpy array of this buffer's size."""
        host = np.ascontiguousarray(host, self.dtype)
        if host.nbytes != self.nbytes:
            raise ValueError("size mismatch")
        self.wann._check(self.wann.lib.elph_host_to_device(self.wann.ctx, self.ptr, host.ctypes.data_as(ct.c_void_p),
                                                           self.nbytes), "host_to_device")
        return self

    def row(self, i: int):
        """Row i of a 2-D buffer (one device-to-host copy of that row)."""
        n = int(np.prod(self.shape[1:]))
        host = np.empty(n, self.dtype)
        off = ct.c_void_p(int(self.ptr.value) + int(i) * n * self.dtype.itemsize)
        self.wann._check(self.wann.lib.elph_device_to_host(self.wann.ctx, host.ctypes.data_as(ct.c_void_p), off,
                                                           n * self.dtype.itemsize), "device_to_host")
        return host

    def numpy(self):
        host = np.empty(self.shape, self.dtype)
        self.wann._check(self.wann.lib.elph_device_to_host(self.wann.ctx, host.ctypes.data_as(ct.c_void_p), self.ptr,
                                                           self.nbytes), "device_to_host")
        return host

    def free(self):
        if getattr(self, "ptr", None) and self.wann.ctx:
            self.wann.lib.elph_device_free(self.wann.ctx, self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
