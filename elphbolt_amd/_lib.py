"""ctypes bindings of libelph_vertex.so (the C ABI declared in include/elph_vertex.h).

The library is loaded from the package directory; importing the product path
without it raises immediately (there is no CPU fallback).
"""
from __future__ import annotations

import ctypes as ct
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libelph_vertex.so")

c_i64p = ct.POINTER(ct.c_int64)
c_dblp = ct.POINTER(ct.c_double)

ELPH_OK, ELPH_EINVAL, ELPH_EDEVICE, ELPH_EIO, ELPH_ENEGPHONON, ELPH_EWINDOW, ELPH_ENOMEM, ELPH_EWEDGE = range(8)
ELPH_FP64, ELPH_MIXED = 0, 1


class ElphDims(ct.Structure):
    _fields_ = [("numwannbands", ct.c_int64), ("numbranches", ct.c_int64), ("nwsk", ct.c_int64),
                ("nwsq", ct.c_int64), ("nwsg", ct.c_int64), ("natoms", ct.c_int64),
                ("coarse_qmesh", ct.c_int64 * 3), ("lattvecs", ct.c_double * 9), ("polar", ct.c_int32),
                ("epsilon", ct.c_double * 9)]


class ElphElectrons(ct.Structure):
    _fields_ = [("wvmesh", ct.c_int64 * 3), ("enref", ct.c_double), ("fsthick", ct.c_double),
                ("nwv_irred", ct.c_int64), ("indexlist_irred", c_i64p), ("ens_irred", c_dblp),
                ("evecs_irred", c_dblp), ("nwv", ct.c_int64), ("indexlist", c_i64p), ("ens", c_dblp),
                ("evecs", c_dblp), ("nstates_inwindow", ct.c_int64)]


class ElphPhonons(ct.Structure):
    _fields_ = [("wvmesh", ct.c_int64 * 3), ("nwv", ct.c_int64), ("ens", c_dblp), ("evecs", c_dblp),
                ("nwv_irred", ct.c_int64), ("indexlist_irred", c_i64p)]


class ElphServerStats(ct.Structure):
    _fields_ = [("completed", ct.c_int64), ("failed", ct.c_int64), ("pending", ct.c_int64), ("active", ct.c_int64),
                ("vertices", ct.c_int64), ("units", ct.c_int64), ("pool_bytes", ct.c_int64),
                ("pool_used", ct.c_int64), ("max_slots_busy", ct.c_int64), ("busy_span_ms", ct.c_double)]


class ElphBatchClass(ct.Structure):
    _fields_ = [("precision", ct.c_int32), ("path", ct.c_int32), ("cb", ct.c_int32), ("rb", ct.c_int32),
                ("candidates", ct.c_int32), ("cols", ct.c_int64), ("kp", ct.c_int64), ("members", ct.c_int64),
                ("ms", ct.c_double)]


# exported symbol -> (restype, argtypes)
PROTOTYPES = {
    "elph_ctx_create": (ct.c_int, [ct.POINTER(ct.c_void_p), ct.c_int, ct.POINTER(ElphDims), c_i64p, c_i64p, c_i64p,
                                   c_i64p, c_i64p, c_i64p, c_dblp, c_dblp, c_dblp, c_dblp, c_dblp, c_dblp]),
    "elph_ctx_create_epw": (ct.c_int, [ct.POINTER(ct.c_void_p), ct.c_int, ct.c_char_p, ct.c_int64, c_dblp, c_dblp,
                                       c_dblp, c_i64p, ct.c_int32, c_dblp, c_dblp]),
    "elph_epw_read_dims": (ct.c_int, [ct.c_char_p, c_i64p]),
    "elph_epw_read": (ct.c_int, [ct.c_char_p, c_i64p, c_i64p, c_i64p, c_i64p, c_i64p, c_i64p, c_i64p, c_dblp, c_dblp,
                                 c_dblp]),
    "elph_ctx_destroy": (ct.c_int, [ct.c_void_p]),
    "elph_last_error": (ct.c_char_p, [ct.c_void_p]),
    "elph_ctx_dims": (ct.c_int, [ct.c_void_p, c_i64p]),
    "elph_el_wann": (ct.c_int, [ct.c_void_p, ct.c_int64, c_dblp, c_dblp, c_dblp, c_dblp, c_dblp]),
    "elph_ph_wann": (ct.c_int, [ct.c_void_p, ct.c_int64, c_dblp, c_dblp, c_dblp, c_dblp]),
    "elph_gkRp": (ct.c_int, [ct.c_void_p, ct.c_int64, c_dblp, c_dblp]),
    "elph_gReq": (ct.c_int, [ct.c_void_p, ct.c_int64, c_dblp, c_dblp]),
    "elph_g2": (ct.c_int, [ct.c_void_p, ct.c_int64, c_dblp, c_dblp, c_dblp, c_dblp, c_dblp, c_dblp, c_dblp,
                           ct.c_int64, c_i64p, ct.c_char, c_dblp]),
    "elph_ibzk_plan_create": (ct.c_int, [ct.c_void_p, ct.POINTER(ElphElectrons), ct.POINTER(ElphPhonons), ct.c_int64,
                                         ct.c_int32, ct.POINTER(ct.c_void_p)]),
    "elph_ibzk_plan_run": (ct.c_int, [ct.c_void_p, ct.c_int64, ct.c_int64, c_i64p, ct.c_void_p, ct.c_int32]),
    "elph_ibzk_plan_nprocs": (ct.c_int64, [ct.c_void_p]),
    "elph_ibzk_plan_run_noalloc": (ct.c_int, [ct.c_void_p, ct.c_int64, ct.c_int64, c_i64p, ct.c_void_p]),
    "elph_tetra_create": (ct.c_int, [ct.c_void_p, ct.c_int64, c_i64p, c_i64p, ct.c_int64, c_dblp,
                                     ct.POINTER(ct.c_void_p)]),
    "elph_triang_create": (ct.c_int, [ct.c_void_p, ct.c_int64, c_i64p, c_i64p, ct.c_int64, c_dblp,
                                      ct.POINTER(ct.c_void_p)]),
    "elph_tetra_tables": (ct.c_int, [ct.c_void_p, c_i64p, c_i64p, c_i64p, c_dblp]),
    "elph_tetra_delta": (ct.c_int, [ct.c_void_p, ct.c_int64, c_dblp, c_i64p, c_i64p, c_dblp]),
    "elph_tetra_destroy": (ct.c_int, [ct.c_void_p]),
    "elph_el_transport_create": (ct.c_int, [ct.c_void_p, ct.c_int64, ct.c_int64, c_dblp, c_dblp, c_i64p, ct.c_int64,
                                            ct.c_int64, c_i64p, c_i64p, c_dblp, ct.c_int64, c_dblp, ct.c_double,
                                            c_i64p, ct.c_int64, ct.POINTER(ct.c_void_p)]),
    "elph_el_transport_rta": (ct.c_int, [ct.c_void_p, ct.c_void_p, ct.c_double, ct.c_double, c_dblp, c_dblp, c_dblp,
                                         c_dblp]),
    "elph_el_transport_destroy": (ct.c_int, [ct.c_void_p]),
    "elph_ibzk_plan_X": (ct.c_int, [ct.c_void_p, ct.c_void_p, ct.c_double, ct.c_double, ct.c_int64, ct.c_int64,
                                    ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p]),
    "elph_ibzk_plan_destroy": (ct.c_int, [ct.c_void_p]),
    "elph_ibzk_g2": (ct.c_int, [ct.c_void_p, ct.POINTER(ElphElectrons), ct.POINTER(ElphPhonons), ct.c_int64, ct.c_int32,
                                ct.c_int64, ct.c_int64, c_i64p, c_dblp]),
    "elph_ibzq_plan_create": (ct.c_int, [ct.c_void_p, ct.POINTER(ElphElectrons), ct.POINTER(ElphPhonons), ct.c_int32,
                                         ct.POINTER(ct.c_void_p)]),
    "elph_ibzq_plan_run": (ct.c_int, [ct.c_void_p, ct.c_int64, ct.c_int64, c_i64p, ct.c_void_p, ct.c_int32]),
    "elph_ibzq_plan_stride": (ct.c_int64, [ct.c_void_p]),
    "elph_ibzq_plan_destroy": (ct.c_int, [ct.c_void_p]),
    "elph_write_X_files": (ct.c_int, [ct.c_char_p, ct.c_int64, ct.c_int64, c_i64p, c_dblp, c_dblp, c_i64p, c_i64p,
                                      ct.c_int64]),
    "elph_write_g2_files": (ct.c_int, [ct.c_char_p, ct.c_char, ct.c_int64, ct.c_int64, c_i64p, c_dblp, ct.c_int64]),
    "elph_g2_writer_create": (ct.c_int, [ct.c_void_p, ct.c_char_p, ct.c_char, ct.c_int64, ct.c_int64,
                                         ct.POINTER(ct.c_void_p)]),
    "elph_g2_writer_submit": (ct.c_int, [ct.c_void_p, ct.c_int64, ct.c_int64, c_i64p, ct.c_void_p]),
    "elph_g2_writer_wait": (ct.c_int, [ct.c_void_p]),
    "elph_g2_writer_error": (ct.c_char_p, [ct.c_void_p]),
    "elph_g2_writer_destroy": (ct.c_int, [ct.c_void_p]),
    "elph_read_g2_files": (ct.c_int, [ct.c_char_p, ct.c_char, ct.c_int64, ct.c_int64, c_i64p, c_dblp, ct.c_int64]),
    "elph_symmetry": (ct.c_int, [c_dblp, ct.c_int64, c_dblp, c_i64p, ct.c_double, c_i64p, c_i64p, c_i64p, c_dblp]),
    "elph_symmetry_cartesian": (ct.c_int, [c_dblp, ct.c_int64, c_dblp, c_i64p, ct.c_double, c_i64p, c_i64p, c_i64p,
                                           c_dblp, c_dblp]),
    "elph_symmetrizers": (ct.c_int, [c_i64p, ct.c_int64, c_dblp, c_dblp, ct.c_int64, c_i64p, c_dblp]),
    "elph_irred_wedge": (ct.c_int, [c_i64p, ct.c_int64, c_dblp, ct.c_int64, c_i64p, c_i64p, c_i64p, c_i64p, c_i64p]),
    "elph_calculate_electrons": (ct.c_int, [ct.c_void_p, ct.c_int64, c_dblp, c_i64p, ct.c_double, ct.c_double,
                                            ct.POINTER(ct.c_void_p), ct.POINTER(ElphElectrons)]),
    "elph_calculate_phonons": (ct.c_int, [ct.c_void_p, ct.c_int64, c_dblp, c_i64p, ct.POINTER(ct.c_void_p),
                                          ct.POINTER(ElphPhonons)]),
    "elph_state_free": (ct.c_int, [ct.c_void_p]),
    "elph_selftest_mfma": (ct.c_int, [ct.c_int, c_dblp]),
    "elph_device_alloc": (ct.c_int, [ct.c_void_p, ct.c_int64, ct.POINTER(ct.c_void_p)]),
    "elph_device_free": (ct.c_int, [ct.c_void_p, ct.c_void_p]),
    "elph_host_to_device": (ct.c_int, [ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_int64]),
    "elph_device_to_host": (ct.c_int, [ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_int64]),
    "elph_ctx_synchronize": (ct.c_int, [ct.c_void_p]),
    "elph_last_timing": (ct.c_int, [ct.c_void_p, c_dblp]),
    "elph_synth_vertex": (ct.c_int, [ct.c_void_p, ct.c_uint64]),
    "elph_device_meminfo": (ct.c_int, [ct.c_int, c_i64p, c_i64p]),
    "elph_device_copy": (ct.c_int, [ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_int64]),
    "elph_batch_create": (ct.c_int, [ct.POINTER(ct.c_void_p), ct.c_int, ct.c_int32]),
    "elph_batch_add": (ct.c_int, [ct.c_void_p, ct.c_void_p, ct.c_int64, ct.c_int64, ct.c_void_p, c_i64p]),
    "elph_batch_prepare": (ct.c_int, [ct.c_void_p, ct.c_int32]),
    "elph_batch_run": (ct.c_int, [ct.c_void_p, c_dblp]),
    "elph_batch_classes": (ct.c_int64, [ct.c_void_p, ct.POINTER(ElphBatchClass), ct.c_int64]),
    "elph_batch_error": (ct.c_char_p, [ct.c_void_p]),
    "elph_batch_destroy": (ct.c_int, [ct.c_void_p]),
    "elph_server_create": (ct.c_int, [ct.POINTER(ct.c_void_p), ct.c_int, ct.c_int32, ct.c_int64, ct.c_int32]),
    "elph_server_add_material": (ct.c_int, [ct.c_void_p, ct.c_void_p]),
    "elph_server_submit": (ct.c_int64, [ct.c_void_p, ct.c_int32, ct.c_int64, c_dblp, ct.c_int64, c_dblp]),
    "elph_server_wait": (ct.c_int, [ct.c_void_p, ct.c_int64, c_dblp]),
    "elph_server_test": (ct.c_int, [ct.c_void_p, ct.c_int64]),
    "elph_server_fetch": (ct.c_int, [ct.c_void_p, ct.c_int64, ct.POINTER(ct.c_float)]),
    "elph_server_result_device": (ct.c_int, [ct.c_void_p, ct.c_int64, ct.POINTER(ct.c_void_p), c_i64p]),
    "elph_server_release": (ct.c_int, [ct.c_void_p, ct.c_int64]),
    "elph_server_stats": (ct.c_int, [ct.c_void_p, ct.POINTER(ElphServerStats)]),
    "elph_server_error": (ct.c_char_p, [ct.c_void_p]),
    "elph_server_destroy": (ct.c_int, [ct.c_void_p]),
}

_lib = None


def load():
    """Load the in-tree library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with elphbolt_amd/csrc/build.sh "
                           "(or __graft_entry__.build()); there is no CPU fallback")
    lib = ct.CDLL(LIB_PATH)
    for name, (res, args) in PROTOTYPES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


class ElphError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[elph {code}] {msg}")
        self.code = code


def device_meminfo(device: int = 0):
    """(free, total) bytes of device memory."""
    f, t = ct.c_int64(0), ct.c_int64(0)
    load().elph_device_meminfo(device, ct.byref(f), ct.byref(t))
    return f.value, t.value
