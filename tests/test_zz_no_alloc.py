"""elph_ibzk_plan_run_noalloc (the allocation-free form of
elph_ibzk_plan_run, scratch from the library's per-GPU array) against
elph_ibzk_plan_run through the C ABI: two calls in a row, the second on a
smaller range, for the fused FP64 and mixed paths, the column-tiled general
path (whose Y chunks are cut to fit the array) and a material whose IBZ k
need several groups; the rows and counts equal bit for bit (same kernels)."""
import ctypes as ct

import numpy as np
import pytest

from elphbolt_amd import _lib, interactions, synth
from elphbolt_amd.wannier import Wannier
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _setup(mat):
    o = O.Wannier(mat)
    qr, nsr, _ = O.qrotations(mat, mat.wvmesh)
    el = O.Electrons(o, qr, nsr, mat.wvmesh, mat.enref, mat.fsthick)
    ph = O.Phonons(o, qr, nsr, mat.phmesh)
    return el, ph


def _noalloc(plan, first, last, buf):
    nst = last - first + 1
    counts = np.zeros(nst, np.int64)
    st = _lib.load().elph_ibzk_plan_run_noalloc(plan.plan, first, last, counts.ctypes.data_as(_lib.c_i64p),
                                                buf.ptr)
    plan.wann._check(st, "ibzk run noalloc")
    return counts


CASES = [
    ("fused_fp64", dict(nb=8, natoms=2, ws=3, wv=6, seed=91, bands_in_window=4), "fp64"),
    ("fused_mixed", dict(nb=8, natoms=2, ws=3, wv=6, seed=91, bands_in_window=4), "mixed"),
    ("general_fp64", dict(nb=12, natoms=3, ws=3, wv=4, seed=92), "fp64"),
    ("general_mixed", dict(nb=12, natoms=3, ws=3, wv=4, seed=92), "mixed"),
]


@pytest.mark.parametrize("name,kw,precision", CASES, ids=[c[0] for c in CASES])
def test_noalloc_equals_run(name, kw, precision):
    mat = synth.tiny(**kw)
    el, ph = _setup(mat)
    w = Wannier(mat)
    plan = interactions.IbzkPlan(w, el, ph, precision=precision)
    nst = el.nwv_irred * mat.nb
    print(f"noalloc {name}: {nst} states, nprocs {plan.nprocs}", flush=True)
    for first, last in ((1, nst), (3, min(nst, 3 + 2 * mat.nb))):
        n = last - first + 1
        a = w.device_buffer((n, plan.nprocs))
        b = w.device_buffer((n, plan.nprocs))
        c_ref, _ = plan.run(first, last, out=a)
        c_na = _noalloc(plan, first, last, b)
        assert np.array_equal(c_ref, c_na) and (c_ref >= 0).any()
        act = c_ref >= 0
        ra, rb = a.numpy()[act], b.numpy()[act]
        assert np.abs(ra).max() > 0
        assert np.array_equal(ra, rb)
        a.free()
        b.free()
    plan.close()


def test_noalloc_many_groups_at_config1_size():
    """8^3 WS cells (617), 24^3 mesh: one IBZ k needs ~10 MB of scratch, so a
    call over 40 IBZ k runs in several groups; the second call is smaller."""
    mat = synth.make_material("noalloc_ws8_k24", 4.36, [28.0855, 12.0107], 8, [8, 8, 8], [12, 12, 12], 0xE1B01799,
                              bands_in_window=4.0)
    el, ph = _setup(mat)
    w = Wannier(mat)
    for precision in ("fp64", "mixed"):
        plan = interactions.IbzkPlan(w, el, ph, precision=precision)
        for first, last in ((1, 40 * mat.nb), (9, 20 * mat.nb)):
            n = last - first + 1
            a = w.device_buffer((n, plan.nprocs))
            b = w.device_buffer((n, plan.nprocs))
            c_ref, _ = plan.run(first, last, out=a)
            c_na = _noalloc(plan, first, last, b)
            assert np.array_equal(c_ref, c_na)
            act = c_ref >= 0
            assert np.array_equal(a.numpy()[act], b.numpy()[act])
            print(f"  {precision} states {first}..{last}: {int(act.sum())} rows equal", flush=True)
            a.free()
            b.free()
        plan.close()
