"""Fused ibzk pair kernels across band/branch shapes: the unrolled epilogue
for nb <= 8, nbr <= 6 (covered by the parity and mixed tests), the one for
nb <= 16, nbr <= 12 and the generic one (nb > 16), each in FP64 (1e-12) and
mixed precision (1e-5) against the oracle, relative above 1e-6 * max; counts
and exact zeros equal."""
import numpy as np
import pytest

from elphbolt_amd import interactions, synth
from elphbolt_amd.wannier import Wannier
from oracle import oracle as O

pytestmark = pytest.mark.gpu

FLOOR = 1e-6
TOL = {"fp64": 1e-12, "mixed": 1e-5}


def _rel(x, ref):
    x, ref = np.asarray(x).ravel(), np.asarray(ref).ravel()
    sel = np.abs(ref) > FLOOR * np.abs(ref).max()
    return float(np.max(np.abs(x[sel] - ref[sel]) / np.abs(ref[sel])))


# nb * nbr within the fused kernels' 64 (mixed) / 96 (FP64) columns
SHAPES = [
    ("nb2_na4", dict(nb=2, natoms=4, ws=3, wv=4, seed=81)),    # 24 columns, nbr 12
    ("nb10_na2", dict(nb=10, natoms=2, ws=3, wv=4, seed=82)),  # 60 columns, nb 10
    ("nb20_na1", dict(nb=20, natoms=1, ws=3, wv=4, seed=83)),  # 60 columns, nb 20
]


@pytest.fixture(scope="module")
def states():
    out = {}
    for name, kw in SHAPES:
        mat = synth.tiny(**kw)
        o = O.Wannier(mat)
        qr, nsr, _ = O.qrotations(mat, mat.wvmesh)
        el = O.Electrons(o, qr, nsr, mat.wvmesh, mat.enref, mat.fsthick)
        ph = O.Phonons(o, qr, nsr, mat.phmesh)
        last = min(el.nwv_irred * mat.nb, 30)
        c_o, g_o = O.ibzk_g2(o, el, ph, 1, 1, last)
        out[name] = (mat, el, ph, last, c_o, g_o)
    return out


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("name", [s[0] for s in SHAPES])
def test_fused_shape(states, name, precision):
    mat, el, ph, last, c_o, g_o = states[name]
    print(f"fused {name} {precision}: nb*nbr={mat.nb * mat.nbr} states 1..{last}", flush=True)
    c_g, g_g = interactions.IbzkPlan(Wannier(mat), el, ph, precision=precision).run(1, last)
    assert np.array_equal(c_o, c_g)
    act = c_o >= 0
    assert act.any()
    assert np.array_equal(g_o[act] == 0.0, g_g[act] == 0.0)
    err = _rel(g_g[act], g_o[act])
    print(f"  max rel {err:.2e}", flush=True)
    assert err < TOL[precision]
