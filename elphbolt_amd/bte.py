"""Electron RTA transport from the device e-ph rates: the dragless RTA branch
of the electron BTE (src/bte.f90:200-277) -- field terms (912-1011), RTA
responses and calculate_transport_coeff (src/bz_sums.f90:484-611) -- on the
device through the C ABI."""
from __future__ import annotations

import ctypes as ct

import numpy as np

from ._lib import c_i64p
from .interactions import _dev_ptr
from .wannier import Wannier, _p


class ElectronRTA:
    """Static inputs in the reference's layouts: ens (nk, nb) eV, vels
    (nk, nb, 3) km/s on indexlist (nk), nequiv (nk_ibz), ibz2fbz_map
    (maxeq, nk_ibz, 2), symmetrizers (3, 3, nk) or None, crotations
    (3, 3, nrot) or None, volume nm^3, mesh, spin degeneracy."""

    def __init__(self, wann: Wannier, ens, vels, indexlist, nequiv, ibz2fbz_map, volume, mesh, spindeg=2,
                 symmetrizers=None, crotations=None):
        self.wann, self.lib = wann, wann.lib
        ens = np.asfortranarray(ens, np.float64)
        vels = np.asfortranarray(vels, np.float64)
        self.nk, self.nb = ens.shape
        il = np.ascontiguousarray(indexlist, np.int64)
        neq = np.ascontiguousarray(nequiv, np.int64)
        m = np.asfortranarray(ibz2fbz_map, np.int64)
        self.nk_ibz = len(neq)
        sym = None if symmetrizers is None else np.asfortranarray(symmetrizers, np.float64)
        cr = None if crotations is None else np.asfortranarray(crotations, np.float64)
        nrot = 0 if cr is None else cr.shape[2]
        mesh = np.ascontiguousarray(mesh, np.int64)
        h = ct.c_void_p()
        st = self.lib.elph_el_transport_create(
            wann.ctx, self.nk, self.nb, _p(ens), _p(vels), _p(il, c_i64p), self.nk_ibz, m.shape[0],
            _p(neq, c_i64p), _p(m, c_i64p), None if sym is None else _p(sym), nrot,
            None if cr is None else _p(cr), float(volume), _p(mesh, c_i64p), int(spindeg), ct.byref(h))
        wann._check(st, "electron transport")
        self.h = h

    def rta(self, rates_ibz, T: float, chempot: float):
        """Transport tensors (nb, 3, 3) from the IBZ e-ph rates (device
        buffer, state (ik-1)*nb + m): kappa0 (W/m/K), sigmaS (A/m/K), sigma
        (1/Ohm/m), alphabyT (A/m/K)."""
        out = {k: np.zeros((self.nb, 3, 3), order="F") for k in ("kappa0", "sigmaS", "sigma", "alphabyT")}
        ptr, dev = _dev_ptr(rates_ibz)
        if not dev:
            raise ValueError("rates_ibz must be device memory")
        st = self.lib.elph_el_transport_rta(self.h, ptr, float(T), float(chempot), _p(out["kappa0"]),
                                            _p(out["sigmaS"]), _p(out["sigma"]), _p(out["alphabyT"]))
        self.wann._check(st, "electron RTA transport")
        return out

    @staticmethod
    def scalar(tensor):
        """trace(sum over bands) / 3, as the reference prints (src/bte.f90:293-300)."""
        return float(np.trace(np.asarray(tensor).sum(axis=0))) / 3.0

    def close(self):
        if getattr(self, "h", None):
            self.lib.elph_el_transport_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
