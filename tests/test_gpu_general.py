"""The general (column-tiled) ibzk path, used when numwannbands * numbranches
exceeds what the fused pair kernels hold (96 columns in FP64, 64 in mixed):
Y of a chunk of states through HBM, then the rotation / |g|^2 epilogue (the
row epilogue up to 8 bands and 6 branches, beyond that one wave per row,
which stages the row's Y and U_k' rows in LDS unless they would leave fewer
than two workgroups per CU: nb26_na4 reads them from memory).
Against the oracle: counts bit-exact, |g|^2 within 1e-12 (FP64) / 1e-5
(mixed) relative above 1e-6 * max, exact zeros equal."""
import numpy as np
import pytest

from elphbolt_amd import interactions, synth
from elphbolt_amd.wannier import Wannier
from oracle import oracle as O
from tests.parity_util import branch_group_sums, fine_phonon_energies

pytestmark = pytest.mark.gpu

FLOOR = 1e-6


def _rel(x, ref):
    x, ref = np.asarray(x).ravel(), np.asarray(ref).ravel()
    sel = np.abs(ref) > FLOOR * np.abs(ref).max()
    return float(np.max(np.abs(x[sel] - ref[sel]) / np.abs(ref[sel])))


def _state(mat):
    o = O.Wannier(mat)
    qr, nsr, _ = O.qrotations(mat, mat.wvmesh)
    el = O.Electrons(o, qr, nsr, mat.wvmesh, mat.enref, mat.fsthick)
    ph = O.Phonons(o, qr, nsr, mat.phmesh)
    return o, el, ph


CASES = [
    ("nb12_na3", dict(nb=12, natoms=3, ws=3, wv=4, seed=61), None, "fp64"),
    ("nb12_na3", dict(nb=12, natoms=3, ws=3, wv=4, seed=61), None, "mixed"),
    ("nb10_na4", dict(nb=10, natoms=4, ws=3, wv=4, seed=62), None, "fp64"),
    ("nb6_na4", dict(nb=6, natoms=4, ws=3, wv=4, seed=63), None, "mixed"),
    ("nb10_na4_polar", dict(nb=10, natoms=4, ws=3, wv=4, seed=64), "random", "fp64"),
    ("nb10_na4_polar", dict(nb=10, natoms=4, ws=3, wv=4, seed=64), "random", "mixed"),
    # beyond nb <= 8, nbr <= 6: one wave per row
    ("nb7_na5", dict(nb=7, natoms=5, ws=3, wv=4, seed=66), None, "fp64"),
    ("nb7_na5", dict(nb=7, natoms=5, ws=3, wv=4, seed=66), None, "mixed"),
    ("nb18_na2", dict(nb=18, natoms=2, ws=3, wv=4, seed=67), None, "fp64"),
    ("nb18_na2_polar", dict(nb=18, natoms=2, ws=3, wv=4, seed=68), "random", "mixed"),
    ("nb26_na4", dict(nb=26, natoms=4, ws=3, wv=4, seed=69), None, "mixed"),
]


@pytest.mark.parametrize("name,kw,polar,precision", CASES, ids=[f"{c[0]}-{c[3]}" for c in CASES])
def test_general_ibzk(name, kw, polar, precision):
    mat = synth.tiny(**kw)
    if polar:
        mat = synth.with_polar(mat, polar, seed=3)
    o, el, ph = _state(mat)
    last = min(el.nwv_irred * mat.nb, 30)
    print(f"general ibzk {name} {precision}: nb*nbr={mat.nb * mat.nbr} states 1..{last}", flush=True)
    c_o, g_o = O.ibzk_g2(o, el, ph, 1, 1, last)
    c_g, g_g = interactions.IbzkPlan(Wannier(mat), el, ph, precision=precision).run(1, last)
    assert np.array_equal(c_o, c_g)
    act = c_o >= 0
    assert act.any()
    assert np.array_equal(g_o[act] == 0.0, g_g[act] == 0.0)
    err = _rel(g_g[act], g_o[act])
    print(f"  max rel {err:.2e}", flush=True)
    assert err < (1e-12 if precision == "fp64" else 1e-5)


def test_general_mesh_ref2():
    mat = synth.tiny(nb=12, natoms=3, ws=3, wv=4, seed=65, mesh_ref=2)
    o, el, ph = _state(mat)
    last = min(el.nwv_irred * mat.nb, 12)
    c_o, g_o = O.ibzk_g2(o, el, ph, 2, 1, last)
    c_g, g_g = interactions.IbzkPlan(Wannier(mat), el, ph, mesh_ref=2, precision="fp64").run(1, last)
    assert np.array_equal(c_o, c_g)
    # sums over degenerate branch groups per (state, k', n): independent of
    # the device eigensolver's basis inside degenerate subspaces
    wf = fine_phonon_energies(o, mat.wvmesh)
    kirr = [O.demux_vector(int(i), mat.wvmesh, 0) for i in el.indexlist_irred]
    a = branch_group_sums(g_g, c_o, el, wf, 1, mat.nb, mat.nbr, kirr)
    b = branch_group_sums(g_o, c_o, el, wf, 1, mat.nb, mat.nbr, kirr)
    assert len(a) > (c_o[c_o >= 0].sum() // mat.nbr)
    assert _rel(a, b) < 1e-10
