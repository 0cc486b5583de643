"""General ibzq path (numwannbands^2 beyond the fused kernel's 96 columns:
column tiles of Y in HBM, then the epilogue) against the oracle: counts exact,
|g|^2 within 1e-12 relative above 1e-6 * max, exact zeros equal; plain and
polar."""
import numpy as np
import pytest

from elphbolt_amd import interactions, synth
from elphbolt_amd.wannier import Wannier
from oracle import oracle as O

pytestmark = pytest.mark.gpu

CASES = [
    ("nb10_na1", lambda: synth.tiny(nb=10, natoms=1, ws=3, wv=4, seed=61, bands_in_window=4)),
    ("nb12_na2_polar", lambda: synth.with_polar(synth.tiny(nb=12, natoms=2, ws=3, wv=4, seed=62,
                                                           bands_in_window=5), "sic")),
]


def _rel(x, ref):
    sel = np.abs(ref) > 1e-6 * np.abs(ref).max()
    return float(np.max(np.abs(x[sel] - ref[sel]) / np.abs(ref[sel]))) if sel.any() else 0.0


@pytest.mark.parametrize("name,make", CASES, ids=[c[0] for c in CASES])
def test_ibzq_general(name, make):
    mat = make()
    o = O.Wannier(mat)
    qr, nsr, _ = O.qrotations(mat, mat.wvmesh)
    el = O.Electrons(o, qr, nsr, mat.wvmesh, mat.enref, mat.fsthick)
    ph = O.Phonons(o, qr, nsr, mat.phmesh)
    last = min(ph.nwv_irred * mat.nbr, 14)
    print(f"ibzq general {name}: nb^2={mat.nb ** 2}, states 1..{last}", flush=True)
    c_o, g_o = O.ibzq_g2(o, el, ph, 1, last)
    plan = interactions.IbzqPlan(Wannier(mat), el, ph, precision="fp64")
    c_g, g_g = plan.run(1, last)
    assert np.array_equal(c_o, c_g)
    assert c_o.max() > 0
    err = 0.0
    for t in range(len(c_o)):
        n = c_o[t]
        if n == 0:
            continue
        assert np.array_equal(g_o[t, :n] == 0.0, g_g[t, :n] == 0.0)
        err = max(err, _rel(g_g[t, :n], g_o[t, :n]))
    print(f"  max rel {err:.2e}", flush=True)
    assert err < 1e-12
