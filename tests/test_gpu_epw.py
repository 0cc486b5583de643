"""A context read from EPW files (elph_ctx_create_epw, the drop-in for
read_EPW_Wannier, src/wannier.f90:289-428) equals one built from the same
arrays: el_wann, ph_wann, gkRp and an ibzk plan run give the same bits, for
the plain files and for the Fortran list-directed variants of the text."""
import numpy as np
import pytest

from elphbolt_amd import interactions, synth
from elphbolt_amd.wannier import Wannier
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("style", ["plain", "fortran"])
def test_epw_context_equals_array_context(tmp_path, style):
    mat = synth.tiny(nb=4, natoms=2, ws=3, wv=4, seed=82)
    synth.write_epw(mat, str(tmp_path), style=style)
    print(f"EPW context ({style})", flush=True)
    w_arr = Wannier(mat)
    w_epw = Wannier.from_epw(str(tmp_path), mat.natoms, mat.lattvecs, mat.basis, mat.masses, mat.coarse_qmesh)
    assert (w_epw.numwannbands, w_epw.numbranches, w_epw.nwsk) == (mat.nb, mat.nbr, mat.nwsk)
    rng = np.random.default_rng(3)
    k = rng.random((9, 3))
    for a, b in zip(w_arr.el_wann(k, velocities=True), w_epw.el_wann(k, velocities=True)):
        assert np.array_equal(a, b)
    for a, b in zip(w_arr.ph_wann(k)[:2], w_epw.ph_wann(k)[:2]):
        assert np.array_equal(a, b)
    assert np.array_equal(w_arr.gkRp(k[:3]), w_epw.gkRp(k[:3]))
    o = O.Wannier(mat)
    qr, nsr, _ = O.qrotations(mat, mat.wvmesh)
    el = O.Electrons(o, qr, nsr, mat.wvmesh, mat.enref, mat.fsthick)
    ph = O.Phonons(o, qr, nsr, mat.phmesh)
    last = min(el.nwv_irred * mat.nb, 16)
    ca, ga = interactions.IbzkPlan(w_arr, el, ph).run(1, last)
    cb, gb = interactions.IbzkPlan(w_epw, el, ph).run(1, last)
    assert np.array_equal(ca, cb) and (ca >= 0).any()
    assert np.array_equal(ga, gb)
