"""Serving of |g|^2 at arbitrary k- and q-lists through the library's request
server (include/elph_vertex.h, elph_server_*): the Wannier tensors of the
registered materials stay resident in HBM; requests are admitted first come
first served while their device data fits the server's pool, and a scheduler
thread runs them as work units of up to 64 k-points of one material, drawn
from one or more requests, on several streams at once.

A result holds its pool memory until `release` (or `fetch(..., release=True)`),
so a client that keeps results must size the pool for them.
"""
from __future__ import annotations

import ctypes as ct

import numpy as np

from . import _lib
from ._lib import ElphError, ElphServerStats, c_dblp

_PREC = {"fp64": _lib.ELPH_FP64, "mixed": _lib.ELPH_MIXED}


class Server:
    def __init__(self, precision: str = "mixed", pool_bytes: int = 16 << 30, nslots: int = 4, device: int = 0):
        self.lib = _lib.load()
        self.precision = precision
        h = ct.c_void_p()
        st = self.lib.elph_server_create(ct.byref(h), device, _PREC[precision], int(pool_bytes), int(nslots))
        if st != 0:
            raise ElphError(st, "elph_server_create failed")
        self.h = h
        self._mats = []   # Wannier objects must outlive the server
        self._shapes = []
        self._req_shape = {}

    def _err(self):
        m = self.lib.elph_server_error(self.h)
        return m.decode() if m else ""

    def add_material(self, w) -> int:
        mid = self.lib.elph_server_add_material(self.h, w.ctx)
        if mid < 0:
            raise ElphError(-mid, f"add_material: {self._err()}")
        self._mats.append(w)
        self._shapes.append((w.numwannbands, w.numbranches))
        return mid

    def submit(self, material: int, kvecs, qvecs) -> int:
        k = np.asfortranarray(np.atleast_2d(np.asarray(kvecs, np.float64)))
        q = np.asfortranarray(np.atleast_2d(np.asarray(qvecs, np.float64)))
        rid = self.lib.elph_server_submit(self.h, int(material), k.shape[0], k.ctypes.data_as(c_dblp),
                                          q.shape[0], q.ctypes.data_as(c_dblp))
        if rid < 0:
            raise ElphError(-rid, f"submit: {self._err()}")
        nb, nbr = self._shapes[material]
        self._req_shape[rid] = (k.shape[0], q.shape[0], nb, nb, nbr)
        return rid

    def wait(self, rid: int):
        """(status, latency in ms from submit to completion)."""
        lat = ct.c_double(0)
        st = self.lib.elph_server_wait(self.h, int(rid), ct.byref(lat))
        return st, lat.value

    def done(self, rid: int) -> bool:
        return self.lib.elph_server_test(self.h, int(rid)) == 1

    def fetch(self, rid: int, release: bool = False) -> np.ndarray:
        """|g|^2 (eV^2, float32) with shape (nk, nq, nb[m], nb[n], nbr[s])."""
        shape = self._req_shape[rid]
        out = np.empty(shape, np.float32)
        st = self.lib.elph_server_fetch(self.h, int(rid), out.ctypes.data_as(ct.POINTER(ct.c_float)))
        if st != 0:
            raise ElphError(st, f"fetch: {self._err()}")
        if release:
            self.release(rid)
        return out

    def release(self, rid: int):
        st = self.lib.elph_server_release(self.h, int(rid))
        if st != 0:
            raise ElphError(st, "release")
        self._req_shape.pop(rid, None)

    def stats(self) -> dict:
        s = ElphServerStats()
        self.lib.elph_server_stats(self.h, ct.byref(s))
        return {name: getattr(s, name) for name, _ in s._fields_}

    def close(self):
        if self.h:
            self.lib.elph_server_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
