"""Host mirror of the vertex drivers of src/interactions.f90:
calculate_eph_interaction_ibzk (2095-2400) and calculate_eph_interaction_ibzq
(1673-1930), key 'g', plus calculate_gkRp / calculate_gReq (1603-1671).

The reference streams gkRp.ik* / gReq.iq* files between the two phases; here
the mixed-representation tensors stay in HBM inside one plan run.  Outputs
are the per-state vertex lists in the reference's order and, on request, the
reference's gk2.istate* / gq2.istate* files.
"""
from __future__ import annotations

import ctypes as ct
import os

import numpy as np

from . import _lib
from ._lib import ElphElectrons, ElphError, ElphPhonons, c_dblp, c_i64p
from .wannier import Wannier, _p, c128, f64, i64

PRECISION = {"fp64": _lib.ELPH_FP64, "mixed": _lib.ELPH_MIXED}


def _el_struct(el):
    s = ElphElectrons()
    keep = []
    for i in range(3):
        s.wvmesh[i] = int(el.wvmesh[i])
    s.enref, s.fsthick = float(el.enref), float(el.fsthick)
    s.nwv_irred, s.nwv, s.nstates_inwindow = int(el.nwv_irred), int(el.nwv), int(el.nstates_inwindow)
    for name, conv, t in (("indexlist_irred", i64, c_i64p), ("ens_irred", f64, c_dblp), ("evecs_irred", c128, c_dblp),
                          ("indexlist", i64, c_i64p), ("ens", f64, c_dblp), ("evecs", c128, c_dblp)):
        a = conv(getattr(el, name))
        keep.append(a)
        setattr(s, name, _p(a, t))
    return s, keep


def _ph_struct(ph):
    s = ElphPhonons()
    for i in range(3):
        s.wvmesh[i] = int(ph.wvmesh[i])
    s.nwv = int(ph.nwv)
    e, v, irr = f64(ph.ens), c128(ph.evecs), i64(ph.indexlist_irred)
    s.ens, s.evecs = _p(e), _p(v)
    s.nwv_irred = len(irr)
    s.indexlist_irred = _p(irr, c_i64p)
    return s, [e, v, irr]


def _dev_ptr(out):
    """(pointer, on_device) of an output buffer: a wannier.DeviceBuffer or any
    object exposing a device pointer as `.ptr`/`data_ptr()` (device), or a numpy
    array (host)."""
    if hasattr(out, "ptr") and not isinstance(out, np.ndarray):
        return out.ptr, 1
    if hasattr(out, "data_ptr") and getattr(out, "is_cuda", False):
        return ct.c_void_p(out.data_ptr()), 1
    return out.ctypes.data_as(ct.c_void_p), 0


class IbzkPlan:
    """calculate_eph_interaction_ibzk('g') with HBM-resident inputs."""

    def __init__(self, wann: Wannier, el, ph, mesh_ref: int = 1, precision: str = "fp64"):
        self.wann, self.lib = wann, wann.lib
        self.nb, self.nbr = wann.numwannbands, wann.numbranches
        self.nwv_irred = int(el.nwv_irred)
        self.active = (np.abs(np.asarray(el.ens_irred) - el.enref) <= el.fsthick)  # (nirr, nb) in-window states
        es, k1 = _el_struct(el)
        ps, k2 = _ph_struct(ph)
        plan = ct.c_void_p()
        st = self.lib.elph_ibzk_plan_create(wann.ctx, ct.byref(es), ct.byref(ps), int(mesh_ref),
                                            PRECISION[precision], ct.byref(plan))
        wann._check(st, "ibzk plan")
        self.plan = plan
        self.nprocs = int(self.lib.elph_ibzk_plan_nprocs(plan))

    def run(self, istate_first=1, istate_last=None, out=None):
        if istate_last is None:
            istate_last = self.nwv_irred * self.nb
        nst = istate_last - istate_first + 1
        counts = np.zeros(max(nst, 1), np.int64)
        if out is None:
            out = np.zeros((nst, self.nprocs))
        ptr, dev = _dev_ptr(out)
        st = self.lib.elph_ibzk_plan_run(self.plan, istate_first, istate_last, _p(counts, c_i64p), ptr, dev)
        self.wann._check(st, "ibzk run")
        return counts[:nst], out

    def run_X(self, tet: "Tetrahedra", T: float, chempot: float, istate_first: int, istate_last: int, g2,
              maps: bool = True, probabilities: bool = True, buffers=None):
        """calculate_eph_interaction_ibzk('X') (src/interactions.f90:2278-2354)
        on the device g2 rows that run() wrote for the same states (a
        DeviceBuffer of shape (nstates, nprocs)), plus the e-ph RTA rate of each
        state (calculate_el_rta_rates, 2957-3063).  Returns a dict of
        DeviceBuffers "Xplus", "Xminus" (THz), "istate_el", "istate_ph" (when
        asked for), the host array "rates" (THz, 0 outside the window) and the
        same rates as the DeviceBuffer "rates_device".  `buffers`: a dict of
        that form from an earlier call on as many states, reused in place."""
        from .wannier import DeviceBuffer
        nst = istate_last - istate_first + 1
        shape = (nst, self.nprocs)
        out = {}
        if buffers is not None:
            out = {k: v for k, v in buffers.items() if k in ("Xplus", "Xminus", "istate_el", "istate_ph")}
        if probabilities and "Xplus" not in out:
            out["Xplus"], out["Xminus"] = DeviceBuffer(self.wann, shape), DeviceBuffer(self.wann, shape)
        if maps and "istate_el" not in out:
            out["istate_el"] = DeviceBuffer(self.wann, shape, np.int64)
            out["istate_ph"] = DeviceBuffer(self.wann, shape, np.int64)
        if not probabilities:
            out.pop("Xplus", None), out.pop("Xminus", None)
        if not maps:
            out.pop("istate_el", None), out.pop("istate_ph", None)
        for v in out.values():
            if v.shape != shape:
                raise ValueError("reused buffers have another shape")
        rates = buffers["rates_device"] if buffers is not None and "rates_device" in buffers else \
            DeviceBuffer(self.wann, (nst,))
        ptr = lambda k: out[k].ptr if k in out else None
        g2p, _ = _dev_ptr(g2)
        st = self.lib.elph_ibzk_plan_X(self.plan, tet.h, float(T), float(chempot), istate_first, istate_last, g2p,
                                       ptr("Xplus"), ptr("Xminus"), ptr("istate_el"), ptr("istate_ph"), rates.ptr)
        self.wann._check(st, "ibzk X")
        out["rates"] = rates.numpy()
        out["rates_device"] = rates
        return out

    def timing(self):
        t = np.zeros(3)
        self.lib.elph_last_timing(self.wann.ctx, _p(t))
        return {"dominant_ms": t[0], "launches": int(t[1]), "all_ms": t[2]}

    def close(self):
        if getattr(self, "plan", None):
            self.lib.elph_ibzk_plan_destroy(self.plan)
            self.plan = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Tetrahedra:
    """form_tetrahedra_3d + fill_tetrahedra_3d (src/delta.f90:494-646) on the
    device: over el%indexlist (electron.f90:582-587) when indexlist is given,
    else over the full mesh; ens (nk, nb) eV."""

    _create = "elph_tetra_create"
    nper, nvert, maxper = 6, 4, 24

    def __init__(self, wann: Wannier, mesh, ens, indexlist=None):
        self.wann, self.lib = wann, wann.lib
        ens = np.asfortranarray(np.asarray(ens, np.float64))
        self.nk, self.nb = ens.shape
        self.mesh = np.asarray(mesh, np.int64)
        il = None if indexlist is None else np.ascontiguousarray(indexlist, np.int64)
        h = ct.c_void_p()
        st = getattr(self.lib, self._create)(wann.ctx, self.nk, _p(self.mesh, c_i64p),
                                             None if il is None else _p(il, c_i64p), self.nb, _p(ens), ct.byref(h))
        wann._check(st, "simplices")
        self.h = h

    def tables(self):
        """(tetra, tetracount, tetramap, tetra_evals) in the reference's shapes."""
        nk, nb, npr, nv = self.nk, self.nb, self.nper, self.nvert
        tetra = np.zeros((npr * nk, nv), np.int64, order="F")
        count = np.zeros(nk, np.int64)
        tmap = np.zeros((2, nk, self.maxper), np.int64, order="F")
        ev = np.zeros((npr * nk, nb, nv), order="F")
        st = self.lib.elph_tetra_tables(self.h, _p(tetra, c_i64p), _p(count, c_i64p), _p(tmap, c_i64p), _p(ev))
        self.wann._check(st, "tetrahedra tables")
        return tetra, count, tmap, ev

    def delta(self, e, ik, ib):
        """delta_fn_tetra (src/delta.f90:61-225) on the device; ik, ib 1-based."""
        e = np.ascontiguousarray(e, np.float64)
        ik = np.ascontiguousarray(np.broadcast_to(ik, e.shape), np.int64)
        ib = np.ascontiguousarray(np.broadcast_to(ib, e.shape), np.int64)
        out = np.zeros(e.shape)
        st = self.lib.elph_tetra_delta(self.h, e.size, _p(e), _p(ik, c_i64p), _p(ib, c_i64p), _p(out))
        self.wann._check(st, "delta_fn_tetra")
        return out

    def close(self):
        if getattr(self, "h", None):
            self.lib.elph_tetra_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Triangles(Tetrahedra):
    """form_triangles + fill_triangles (src/delta.f90:648-788) on the device;
    its delta function is delta_fn_triang (790-882)."""

    _create = "elph_triang_create"
    nper, nvert, maxper = 2, 3, 6


class Batch:
    """Several materials' calculate_eph_interaction_ibzk('g') work in one step:
    captured once into a hipGraph with branches on `nstreams` streams, the
    pair-stage tile shape of each size class chosen by timing (autotune).
    Outputs stay in HBM (one DeviceBuffer per added plan)."""

    def __init__(self, nstreams: int = 4, device: int = 0):
        from . import _lib
        self.lib = _lib.load()
        h = ct.c_void_p()
        st = self.lib.elph_batch_create(ct.byref(h), device, int(nstreams))
        if st != 0:
            raise _lib.ElphError(st, "elph_batch_create failed")
        self.h = h
        self.entries = []

    def _check(self, st, what):
        if st != 0:
            from ._lib import ElphError
            msg = self.lib.elph_batch_error(self.h)
            raise ElphError(st, f"{what}: {msg.decode() if msg else ''}")

    def n_active(self, plan: IbzkPlan, istate_first: int, istate_last: int) -> int:
        s = np.arange(istate_first - 1, istate_last)
        return int(plan.active[s // plan.nb, s % plan.nb].sum())

    def add(self, plan: IbzkPlan, istate_first: int = 1, istate_last=None, out=None):
        """Adds a plan's states; returns (counts, device output (n_active, nprocs)):
        one row per in-window state in istate order (counts -1 elsewhere)."""
        if istate_last is None:
            istate_last = plan.nwv_irred * plan.nb
        nst = istate_last - istate_first + 1
        counts = np.zeros(nst, np.int64)
        if out is None:
            out = plan.wann.device_buffer((max(1, self.n_active(plan, istate_first, istate_last)), plan.nprocs))
        self._check(self.lib.elph_batch_add(self.h, plan.plan, istate_first, istate_last, out.ptr,
                                            _p(counts, c_i64p)), "batch add")
        self.entries.append((plan, counts, out))
        return counts, out

    def prepare(self, autotune: bool = True):
        self._check(self.lib.elph_batch_prepare(self.h, 1 if autotune else 0), "batch prepare")

    def run(self) -> float:
        ms = ct.c_double(0)
        self._check(self.lib.elph_batch_run(self.h, ct.byref(ms)), "batch run")
        return ms.value

    def classes(self):
        from ._lib import ElphBatchClass
        buf = (ElphBatchClass * 64)()
        n = int(self.lib.elph_batch_classes(self.h, buf, 64))
        return [{f: getattr(buf[i], f) for f, _ in ElphBatchClass._fields_} for i in range(min(n, 64))]

    def close(self):
        if getattr(self, "h", None):
            self.lib.elph_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class IbzqPlan:
    """calculate_eph_interaction_ibzq('g') with HBM-resident inputs."""

    def __init__(self, wann: Wannier, el, ph, precision: str = "fp64"):
        self.wann, self.lib = wann, wann.lib
        self.nb, self.nbr = wann.numwannbands, wann.numbranches
        self.nwv_irred_q = int(ph.nwv_irred)
        es, k1 = _el_struct(el)
        ps, k2 = _ph_struct(ph)
        plan = ct.c_void_p()
        st = self.lib.elph_ibzq_plan_create(wann.ctx, ct.byref(es), ct.byref(ps), PRECISION[precision],
                                            ct.byref(plan))
        wann._check(st, "ibzq plan")
        self.plan = plan
        self.stride = int(self.lib.elph_ibzq_plan_stride(plan))

    def run(self, istate_first=1, istate_last=None, out=None):
        if istate_last is None:
            istate_last = self.nwv_irred_q * self.nbr
        nst = istate_last - istate_first + 1
        counts = np.zeros(max(nst, 1), np.int64)
        if out is None:
            out = np.zeros((nst, self.stride))
        ptr, dev = _dev_ptr(out)
        st = self.lib.elph_ibzq_plan_run(self.plan, istate_first, istate_last, _p(counts, c_i64p), ptr, dev)
        self.wann._check(st, "ibzq run")
        return counts[:nst], out

    def timing(self):
        t = np.zeros(3)
        self.lib.elph_last_timing(self.wann.ctx, _p(t))
        return {"dominant_ms": t[0], "launches": int(t[1]), "all_ms": t[2]}

    def close(self):
        if getattr(self, "plan", None):
            self.lib.elph_ibzq_plan_destroy(self.plan)
            self.plan = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def write_g2_files(g2dir, kind, istate_first, counts, g2):
    """gk2.istate<i> ('k') / gq2.istate<i> ('q') in the reference's byte layout
    (src/interactions.f90:2337-2349, 1877-1889)."""
    lib = _lib.load()
    os.makedirs(g2dir, exist_ok=True)
    c, g = i64(counts), np.ascontiguousarray(g2, dtype=np.float64)
    st = lib.elph_write_g2_files(str(g2dir).encode(), kind.encode(), istate_first, len(c), _p(c, c_i64p), _p(g),
                                 g.shape[1])
    if st != 0:
        raise ElphError(st, f"writing g2 files into {g2dir} failed")


class G2Writer:
    """gk2/gq2.istate* files written on a host thread while the device
    computes the next states (elph_g2_writer_*): submit() takes the device
    rows a run just wrote and returns at once; the next run may write into
    the same device buffer (it waits on the device for the copy)."""

    def __init__(self, wann: Wannier, g2dir, kind, row_stride, max_states):
        self.wann, self.lib = wann, wann.lib
        os.makedirs(g2dir, exist_ok=True)
        h = ct.c_void_p()
        st = self.lib.elph_g2_writer_create(wann.ctx, str(g2dir).encode(), kind.encode(), int(row_stride),
                                            int(max_states), ct.byref(h))
        wann._check(st, "g2 writer")
        self.h = h

    def submit(self, istate_first, counts, g2_device):
        c = i64(counts)
        ptr, dev = _dev_ptr(g2_device)
        if not dev:
            raise ValueError("the rows must be device memory")
        st = self.lib.elph_g2_writer_submit(self.h, int(istate_first), len(c), _p(c, c_i64p), ptr)
        if st != 0:
            raise ElphError(st, f"g2 writer: {self.lib.elph_g2_writer_error(self.h).decode()}")

    def wait(self):
        st = self.lib.elph_g2_writer_wait(self.h)
        if st != 0:
            raise ElphError(st, f"g2 writer: {self.lib.elph_g2_writer_error(self.h).decode()}")

    def close(self):
        if getattr(self, "h", None):
            st = self.lib.elph_g2_writer_destroy(self.h)
            self.h = None
            if st != 0:
                raise ElphError(st, "g2 writer: a batch was not written")

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def read_g2_files(g2dir, kind, istate_first, nstates, row_stride):
    """The resume path of `read_gk2` / `read_gq2` (app/elphbolt.f90:185,212):
    (counts, rows) of states istate_first..istate_first+nstates-1 read from
    gk2/gq2.istate<i> files; counts -1 (row zero) where no file exists."""
    lib = _lib.load()
    counts = np.zeros(max(nstates, 1), np.int64)
    g = np.zeros((max(nstates, 1), row_stride))
    st = lib.elph_read_g2_files(str(g2dir).encode(), kind.encode(), istate_first, nstates, _p(counts, c_i64p), _p(g),
                                row_stride)
    if st != 0:
        raise ElphError(st, f"reading g2 files from {g2dir} failed")
    return counts[:nstates], g[:nstates]


def read_g2_file(path):
    """Reads one gk2/gq2 file back: (count, values) (the reader of
    src/interactions.f90:2202-2212 / 3065-3091 takes the first `count`)."""
    with open(path, "rb") as f:
        count = int(np.frombuffer(f.read(8), np.int64)[0])
        vals = np.frombuffer(f.read(), np.float64)
    return count, vals


def write_X_files(xdir, istate_first, counts, xplus, xminus, istate_el, istate_ph):
    """Xplus.istate<i> / Xminus.istate<i> in the reference's byte layout
    (src/interactions.f90:2356-2377) from host rows."""
    lib = _lib.load()
    os.makedirs(xdir, exist_ok=True)
    c = i64(counts)
    xp, xm = (np.ascontiguousarray(a, dtype=np.float64) for a in (xplus, xminus))
    se, sp = (np.ascontiguousarray(a, dtype=np.int64) for a in (istate_el, istate_ph))
    st = lib.elph_write_X_files(str(xdir).encode(), istate_first, len(c), _p(c, c_i64p), _p(xp), _p(xm),
                                _p(se, c_i64p), _p(sp, c_i64p), xp.shape[1])
    if st != 0:
        raise ElphError(st, f"writing X files into {xdir} failed")


def read_transition_probs_e(path):
    """read_transition_probs_e (src/interactions.f90:3065-3091): (N, TP,
    istate1, istate2)."""
    with open(path, "rb") as f:
        n = int(np.frombuffer(f.read(8), np.int64)[0])
        tp = np.frombuffer(f.read(8 * n), np.float64)
        s1 = np.frombuffer(f.read(8 * n), np.int64)
        s2 = np.frombuffer(f.read(8 * n), np.int64)
        rest = f.read()
    if rest:
        raise ValueError(f"{path}: {len(rest)} trailing bytes")
    return n, tp, s1, s2


def calculate_eph_interaction_ibzk(wann: Wannier, el, ph, key="g", mesh_ref=1, precision="fp64",
                                   istate_first=1, istate_last=None, g2dir=None):
    if key not in ("g",):
        raise ElphError(_lib.ELPH_EINVAL, "Invalid value of key in call to calculate_eph_interaction_ibzk. Exiting.")
    plan = IbzkPlan(wann, el, ph, mesh_ref=mesh_ref, precision=precision)
    counts, g2 = plan.run(istate_first, istate_last)
    plan.close()
    if g2dir is not None:
        write_g2_files(g2dir, "k", istate_first, counts, g2)
    return counts, g2


def calculate_eph_interaction_ibzq(wann: Wannier, el, ph, key="g", precision="fp64", istate_first=1,
                                   istate_last=None, g2dir=None):
    if key not in ("g",):
        raise ElphError(_lib.ELPH_EINVAL, "Invalid value of key in call to calculate_eph_interaction_ibzq. Exiting.")
    plan = IbzqPlan(wann, el, ph, precision=precision)
    counts, g2 = plan.run(istate_first, istate_last)
    plan.close()
    if g2dir is not None:
        write_g2_files(g2dir, "q", istate_first, counts, g2)
    return counts, g2


def calculate_gkRp(wann: Wannier, el):
    """g(k,Rp) for all IBZ k of `el` (calculate_gkRp, src/interactions.f90:1632-1671)."""
    k = np.stack([np.array(np.unravel_index(i - 1, tuple(el.wvmesh[::-1]))[::-1], dtype=np.float64) / el.wvmesh
                  for i in el.indexlist_irred])
    return wann.gkRp(k)


def calculate_gReq(wann: Wannier, ph):
    """g(Re,q) for all IBZ q of `ph` (calculate_gReq, src/interactions.f90:1603-1630)."""
    q = np.stack([np.array(np.unravel_index(i - 1, tuple(ph.wvmesh[::-1]))[::-1], dtype=np.float64) / ph.wvmesh
                  for i in ph.indexlist_irred])
    return wann.gReq(q)
