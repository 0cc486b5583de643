"""GPU parity of the HIP path (through the C ABI) against the CPU oracle.

Both sides get the same seeded inputs; where the vertex depends on
eigenvectors, both get the same eigenvectors (|g|^2 is gauge invariant only up
to degenerate subspaces).  Tolerances: FP64 path 1e-12 relative on |g|^2
entries above 1e-6 * max|g|^2 (the floor below which cancellation over ~K
random-phase terms leaves FP64 rounding visible); index maps and counts exact.
"""
import numpy as np
import pytest

from elphbolt_amd import bz, interactions, synth
from elphbolt_amd.wannier import Wannier
from oracle import oracle as O
from tests.parity_util import branch_group_sums, fine_phonon_energies

pytestmark = pytest.mark.gpu

REL_FP64 = 1e-12
FLOOR = 1e-6


def _flush(msg):
    print(msg, flush=True)


def rel_err_above_floor(x, ref, floor=FLOOR):
    x, ref = np.asarray(x).ravel(), np.asarray(ref).ravel()
    m = np.abs(ref).max() if ref.size else 0.0
    sel = np.abs(ref) > floor * m
    if not sel.any():
        return 0.0
    return float(np.max(np.abs(x[sel] - ref[sel]) / np.abs(ref[sel])))


@pytest.fixture(scope="module")
def mats():
    return {
        "nb2": synth.tiny(nb=2, natoms=1, ws=3, wv=4, seed=11),
        "nb4": synth.tiny(nb=4, natoms=2, ws=3, wv=6, seed=12),
        "nb8": synth.tiny(nb=8, natoms=2, ws=3, wv=6, seed=13, bands_in_window=4),
    }


def test_mfma_layout():
    import ctypes as ct
    from elphbolt_amd import _lib
    err = ct.c_double(-1)
    assert _lib.load().elph_selftest_mfma(0, ct.byref(err)) == 0
    assert err.value < 1e-14


@pytest.mark.parametrize("key", ["nb2", "nb4", "nb8"])
def test_el_wann(mats, key):
    _flush(f"el_wann {key}")
    mat = mats[key]
    w, o = Wannier(mat), O.Wannier(mat)
    rng = np.random.default_rng(5)
    k = rng.random((37, 3))
    k[0] = 0.0
    e_g, v_g, vel_g = w.el_wann(k, velocities=True)
    e_o, v_o, vel_o = o.el_wann(k, velocities=True)
    scale = np.abs(e_o).max()
    assert np.max(np.abs(e_g - e_o)) < 1e-12 * scale
    # eigenvectors: projector equality for non-degenerate bands
    for ik in range(k.shape[0]):
        for b in range(mat.nb):
            gap = np.min(np.abs(np.delete(e_o[ik], b) - e_o[ik, b])) if mat.nb > 1 else 1.0
            if gap < 1e-6 * scale:
                continue
            ov = abs(np.vdot(v_o[ik, b], v_g[ik, b]))
            assert abs(ov - 1.0) < 1e-10
            assert abs(vel_g[ik, b] - vel_o[ik, b]).max() < 1e-8 * (1 + abs(vel_o).max())


@pytest.mark.parametrize("key", ["nb2", "nb4", "nb8"])
def test_ph_wann(mats, key):
    _flush(f"ph_wann {key}")
    mat = mats[key]
    w, o = Wannier(mat), O.Wannier(mat)
    rng = np.random.default_rng(6)
    q = rng.random((29, 3))
    q[0] = 0.0
    e_g, v_g, _ = w.ph_wann(q)
    e_o, v_o, _ = o.ph_wann(q)
    assert np.all(e_g[0, :3] == 0.0) and np.all(e_o[0, :3] == 0.0)
    assert np.max(np.abs(e_g - e_o)) < 1e-10 * np.abs(e_o).max()
    for iq in range(1, q.shape[0]):
        for s in range(mat.nbr):
            gap = np.min(np.abs(np.delete(e_o[iq], s) - e_o[iq, s]))
            if gap < 1e-5:
                continue
            assert abs(abs(np.vdot(v_o[iq, s], v_g[iq, s])) - 1.0) < 1e-8


@pytest.mark.parametrize("key", ["nb2", "nb4", "nb8"])
def test_gkRp_gReq(mats, key):
    _flush(f"gkRp/gReq {key}")
    mat = mats[key]
    w, o = Wannier(mat), O.Wannier(mat)
    rng = np.random.default_rng(7)
    ks = rng.random((5, 3))
    ks[0] = 0.0
    g = w.gkRp(ks)
    for i in range(ks.shape[0]):
        ref = o.gkRp(ks[i])
        assert np.max(np.abs(g[i] - ref)) < 1e-13 * np.abs(ref).max()
    gq = w.gReq(ks)
    for i in range(ks.shape[0]):
        ref = o.gReq(ks[i])
        assert np.max(np.abs(gq[i] - ref)) < 1e-13 * np.abs(ref).max()


@pytest.mark.parametrize("key", ["nb2", "nb4"])
def test_g2_list(mats, key):
    _flush(f"g2 list {key}")
    mat = mats[key]
    w, o = Wannier(mat), O.Wannier(mat)
    rng = np.random.default_rng(8)
    k, kp, q = rng.random(3), rng.random(3), rng.random(3)
    _, ek, _ = o.el_wann(np.stack([k, kp]))
    e_q, v_q, _ = o.ph_wann(q[None])
    gm = o.gkRp(k)
    for wannspace, gmix in (("ph", gm), ("el", o.gReq(q))):
        for m in range(mat.nb):
            for s in range(mat.nbr):
                ref = o.g2(k, q, ek[0, m], ek[1, (m + 1) % mat.nb], v_q[0, s], e_q[0, s], gmix, wannspace)
                got = w.g2(k, q, ek[0, m], ek[1, (m + 1) % mat.nb], v_q[0, s], e_q[0, s], gmix, wannspace)
                assert abs(got - ref) <= 1e-11 * abs(ref) + 1e-300


def _oracle_state(mat):
    o = O.Wannier(mat)
    qr, nsr, _ = O.qrotations(mat, mat.wvmesh)
    el = O.Electrons(o, qr, nsr, mat.wvmesh, mat.enref, mat.fsthick)
    ph = O.Phonons(o, qr, nsr, mat.phmesh)
    return o, el, ph


@pytest.mark.parametrize("key", ["nb2", "nb4", "nb8"])
def test_ibzk_fp64(mats, key):
    mat = mats[key]
    o, el, ph = _oracle_state(mat)
    nst = el.nwv_irred * mat.nb
    last = min(nst, 24)
    _flush(f"ibzk {key}: nwv={el.nwv} nirr={el.nwv_irred} states 1..{last}")
    c_o, g_o = O.ibzk_g2(o, el, ph, 1, 1, last)
    w = Wannier(mat)
    plan = interactions.IbzkPlan(w, el, ph, mesh_ref=1, precision="fp64")
    c_g, g_g = plan.run(1, last)
    assert np.array_equal(c_o, c_g)
    act = c_o >= 0
    assert act.any()
    assert np.all(c_o[act] == el.nstates_inwindow * mat.nbr)
    zero = g_o[act] == 0.0
    assert np.array_equal(zero, g_g[act] == 0.0)
    assert rel_err_above_floor(g_g[act], g_o[act]) < REL_FP64


def test_ibzk_mesh_ref2():
    mat = synth.tiny(nb=4, natoms=1, ws=3, wv=4, seed=21, mesh_ref=2)
    o, el, ph = _oracle_state(mat)
    last = min(el.nwv_irred * mat.nb, 12)
    _flush(f"ibzk mesh_ref=2: states 1..{last}")
    c_o, g_o = O.ibzk_g2(o, el, ph, 2, 1, last)
    w = Wannier(mat)
    plan = interactions.IbzkPlan(w, el, ph, mesh_ref=2, precision="fp64")
    c_g, g_g = plan.run(1, last)
    assert np.array_equal(c_o, c_g)
    # fine-mesh phonons come from the device eigensolver (its own basis inside
    # degenerate subspaces): compare sums over degenerate branch groups per
    # (state, k', n), which do not depend on that basis
    wf = fine_phonon_energies(o, mat.wvmesh)
    kirr = [O.demux_vector(int(i), mat.wvmesh, 0) for i in el.indexlist_irred]
    a = branch_group_sums(g_g, c_o, el, wf, 1, mat.nb, mat.nbr, kirr)
    b = branch_group_sums(g_o, c_o, el, wf, 1, mat.nb, mat.nbr, kirr)
    assert len(a) > (c_o[c_o >= 0].sum() // mat.nbr)      # groups finer than whole branch blocks
    assert rel_err_above_floor(a, b) < 1e-10


@pytest.mark.parametrize("key", ["nb2", "nb4", "nb8"])
def test_ibzq_fp64(mats, key):
    mat = mats[key]
    o, el, ph = _oracle_state(mat)
    last = min(ph.nwv_irred * mat.nbr, 18)
    _flush(f"ibzq {key}: states 1..{last}")
    c_o, g_o = O.ibzq_g2(o, el, ph, 1, last)
    w = Wannier(mat)
    plan = interactions.IbzqPlan(w, el, ph, precision="fp64")
    c_g, g_g = plan.run(1, last)
    assert np.array_equal(c_o, c_g)
    for t in range(len(c_o)):
        n = c_o[t]
        if n == 0:
            continue
        assert np.array_equal(g_o[t, :n] == 0.0, g_g[t, :n] == 0.0)
        assert rel_err_above_floor(g_g[t, :n], g_o[t, :n]) < REL_FP64


@pytest.mark.parametrize("key", ["nb4", "nb8"])
def test_setup_matches_oracle(mats, key):
    """calculate_electrons / calculate_phonons of the product vs the oracle:
    identical index lists and counts, energies to eigensolver precision."""
    mat = mats[key]
    _flush(f"setup {key}")
    o, el_o, ph_o = _oracle_state(mat)
    w = Wannier(mat)
    sym = bz.Symmetry(mat.lattvecs, mat.basis, mat.masses, mat.wvmesh)
    el = bz.Electrons(w, sym, mat.wvmesh, mat.enref, mat.fsthick)
    ph = bz.Phonons(w, sym, mat.phmesh)
    assert np.array_equal(el.indexlist, el_o.indexlist)
    assert np.array_equal(el.indexlist_irred, el_o.indexlist_irred)
    assert el.nstates_inwindow == el_o.nstates_inwindow
    assert np.max(np.abs(el.ens - el_o.ens)) < 1e-11
    assert np.max(np.abs(el.ens_irred - el_o.ens_irred)) < 1e-11
    assert np.array_equal(ph.indexlist_irred, ph_o.indexlist_irred)
    assert np.max(np.abs(ph.ens - ph_o.ens)) < 1e-10


def test_ibzk_product_setup_end_to_end(mats):
    """Product setup feeding the product vertex vs the oracle vertex fed by the
    same (product) eigenvectors."""
    mat = mats["nb4"]
    _flush("end-to-end nb4")
    w = Wannier(mat)
    sym = bz.Symmetry(mat.lattvecs, mat.basis, mat.masses, mat.wvmesh)
    el = bz.Electrons(w, sym, mat.wvmesh, mat.enref, mat.fsthick)
    ph = bz.Phonons(w, sym, mat.phmesh)
    last = min(el.nwv_irred * mat.nb, 16)
    c_g, g_g = interactions.IbzkPlan(w, el, ph).run(1, last)
    c_o, g_o = O.ibzk_g2(O.Wannier(mat), el, ph, 1, 1, last)
    assert np.array_equal(c_o, c_g)
    act = c_o >= 0
    assert rel_err_above_floor(g_g[act], g_o[act]) < REL_FP64
