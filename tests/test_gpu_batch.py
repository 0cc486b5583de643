"""Batched multi-material driver (elph_batch_*): materials of different band,
mode and WS counts (fused and general pair paths, a polar one) captured into
one hipGraph over several streams, tile shapes chosen by the autotuner.

Every replay (three, with eager library work and synchronisations between
them) is checked against the oracle (1e-12 FP64 / 1e-5 mixed relative above
1e-6 * max, counts exact) and bit for bit against an eager run of the same
plans with the shapes the autotuner chose."""
import numpy as np
import pytest

from elphbolt_amd import interactions, synth
from elphbolt_amd.wannier import Wannier
from oracle import oracle as O

pytestmark = pytest.mark.gpu

FLOOR = 1e-6


def _rel(x, ref):
    x, ref = np.asarray(x).ravel(), np.asarray(ref).ravel()
    sel = np.abs(ref) > FLOOR * np.abs(ref).max()
    return float(np.max(np.abs(x[sel] - ref[sel]) / np.abs(ref[sel])))


def _materials():
    return [synth.tiny(nb=2, natoms=1, ws=3, wv=4, seed=71),
            synth.tiny(nb=4, natoms=2, ws=3, wv=6, seed=72),
            synth.tiny(nb=12, natoms=3, ws=3, wv=4, seed=73),
            synth.with_polar(synth.tiny(nb=10, natoms=4, ws=3, wv=4, seed=74), "random", seed=4)]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batch_graph_vs_oracle_and_eager(precision):
    tol = 1e-12 if precision == "fp64" else 1e-5
    mats = _materials()
    setups = []
    for mat in mats:
        o = O.Wannier(mat)
        qr, nsr, _ = O.qrotations(mat, mat.wvmesh)
        el = O.Electrons(o, qr, nsr, mat.wvmesh, mat.enref, mat.fsthick)
        ph = O.Phonons(o, qr, nsr, mat.phmesh)
        last = min(el.nwv_irred * mat.nb, 20)
        c_o, g_o = O.ibzk_g2(o, el, ph, 1, 1, last)
        w = Wannier(mat)
        setups.append((mat, w, interactions.IbzkPlan(w, el, ph, precision=precision), last, c_o, g_o))
        print(f"batch {precision}: {mat.name} nb*nbr={mat.nb * mat.nbr} states 1..{last}", flush=True)
    b = interactions.Batch(nstreams=3)
    outs = [b.add(plan, 1, last) for (_, _, plan, last, _, _) in setups]
    b.prepare(autotune=True)
    classes = b.classes()
    print("classes", classes, flush=True)
    assert sum(c["members"] for c in classes) == len(mats)
    assert all(c["candidates"] >= 1 and c["ms"] > 0 for c in classes)
    # eager runs with the tuned shapes
    eager = []
    for (mat, w, plan, last, c_o, g_o) in setups:
        c_e, g_e = plan.run(1, last)
        assert np.array_equal(c_e, c_o)
        eager.append(g_e)
    for rep in range(3):
        ms = b.run()
        assert ms > 0
        for i, ((mat, w, plan, last, c_o, g_o), (counts, out)) in enumerate(zip(setups, outs)):
            assert np.array_equal(counts, c_o)
            act = c_o >= 0
            got = out.numpy()          # one row per in-window state
            assert got.shape[0] == act.sum()
            assert np.array_equal(got == 0.0, g_o[act] == 0.0)
            assert _rel(got, g_o[act]) < tol
            assert np.array_equal(got, eager[i][act])
        # eager work between replays
        setups[rep % len(setups)][2].run(1, 4)
        setups[0][1].synchronize()
