"""The batched workload's materials under oracle parity: draws of the
multi-material generator at the benchmark's vertex cap (2 GB) and window
(~4 bands), spanning its size range -- 16 bands x 15 branches (the most
columns), 24 branches, 26 bands (column-tiled general path) and 857 WS cells
(fused path).  Each plan runs first on its own (its default path), then all
run together through one captured batch graph (elph_batch_*) with the tile
shapes the autotuner picks.  The first two IBZ k of each material are
checked against the oracle's calculate_eph_interaction_ibzk('g')
(src/interactions.f90:2095-2400) over two k' windows of the row (its start
and its middle): FP64 within 1e-12, mixed within 1e-5 relative above 1e-6 *
max, counts and zeros exact.

The electron and phonon states come from the product's setup (its parity
against the oracle is tested at the benchmark mesh in test_gpu_config1.py)
and feed both sides, so both use the same eigenvectors."""
import numpy as np
import pytest

from elphbolt_amd import bz, interactions, synth
from elphbolt_amd.wannier import Wannier
from oracle import oracle as O

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(900)]

FLOOR = 1e-6
BUDGET = 2e9
# nb16 nbr15 ws165 (most columns) | nb6 nbr24 ws279 | nb4 nbr9 ws857 (fused) | nb26 nbr6 ws165
PICKS = (45, 33, 49, 27)
KP_WINDOW = 8


@pytest.fixture(scope="module")
def batch_mats():
    out = []
    for i in PICKS:
        mat = synth.multi_material(i, budget_bytes=BUDGET, vertex="host", bands_in_window=4.0, pair_floor=True)
        w = Wannier(mat)
        sym = bz.Symmetry(mat.lattvecs, mat.basis, mat.masses, mat.wvmesh)
        el = bz.Electrons(w, sym, mat.wvmesh, mat.enref, mat.fsthick)
        ph = bz.Phonons(w, sym, mat.phmesh)
        last = 2 * mat.nb
        o = O.Wannier(mat)
        ref = [O.ibzk_g2_window(o, el, ph, 1, last, kp0, KP_WINDOW) for kp0 in (1, el.nwv // 2)]
        print(f"config 3 material {mat.name}: nb*nbr={mat.nb * mat.nbr} nws={mat.nwsk} nwv={el.nwv} "
              f"states 1..{last}, {int(sum((r[0] >= 0).sum() for r in ref[:1]))} in window", flush=True)
        out.append(dict(mat=mat, w=w, el=el, ph=ph, last=last, ref=ref))
    return out


def _worst(rows, counts, refs, plan):
    """Max relative error of compact rows (one per in-window state) against
    the oracle's k'-window rows; counts, zeros and row lengths checked."""
    act = counts >= 0
    assert rows.shape[0] == act.sum()
    assert np.all(counts[act] == plan.nprocs)
    row_of = np.cumsum(act) - 1
    pairs = []
    for c_o, g_o, off in refs:
        assert np.array_equal(c_o >= 0, act)
        for t in np.where(act)[0]:
            pairs.append((rows[row_of[t], off[t]:off[t] + c_o[t]], g_o[t, :c_o[t]]))
    gmax = max(float(np.abs(r).max()) for _, r in pairs)
    worst = 0.0
    for got, r in pairs:
        assert np.array_equal(got == 0.0, r == 0.0)
        sel = np.abs(r) > FLOOR * gmax
        if sel.any():
            worst = max(worst, float(np.max(np.abs(got[sel] - r[sel]) / np.abs(r[sel]))))
    return worst, len(pairs)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_config3_batch_vs_oracle(batch_mats, precision):
    tol = 1e-12 if precision == "fp64" else 1e-5
    b = interactions.Batch(nstreams=4)
    runs = []
    for m in batch_mats:
        plan = interactions.IbzkPlan(m["w"], m["el"], m["ph"], precision=precision)
        # the plan on its own, default path (fused where nb*nbr fits)
        c1, g1 = plan.run(1, m["last"])
        worst, nblk = _worst(g1[c1 >= 0], c1, m["ref"], plan)
        print(f"  {m['mat'].name} alone: {nblk} blocks, max rel {worst:.2e}", flush=True)
        assert worst < tol
        counts, out = b.add(plan, 1, m["last"])
        runs.append((m, plan, counts, out))
    b.prepare(autotune=True)
    cls = b.classes()
    paths = {c["path"] for c in cls}
    print(f"config 3 {precision}: {len(cls)} size classes, paths {sorted(paths)}", flush=True)
    assert sum(c["members"] for c in cls) == len(batch_mats)
    assert 1 in paths                # the column-tiled path is exercised
    for rep in range(2):
        assert b.run() > 0
        for m, plan, counts, out in runs:
            worst, nblk = _worst(out.numpy(), counts, m["ref"], plan)
            if rep == 0:
                print(f"  {m['mat'].name} in the batch: {nblk} blocks, max rel {worst:.2e}", flush=True)
            assert worst < tol
    b.close()
    for _, plan, _, _ in runs:
        plan.close()
