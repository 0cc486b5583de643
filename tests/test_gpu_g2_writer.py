"""The asynchronous gk2/gq2 writer (elph_g2_writer_*): batches of states run
into one device buffer, each batch handed to the writer right after its run
and the next run started at once into the same buffer; the files on disk
equal the rows of a synchronous run (ibzk 'k' files: count + the whole row;
ibzq 'q' files: count + count values; no file for skipped states), read back
through the resume reader.  A directory that cannot be written gives
ELPH_EIO from wait()."""
import os

import numpy as np
import pytest

from elphbolt_amd import interactions, synth
from elphbolt_amd._lib import ELPH_EIO, ElphError
from elphbolt_amd.wannier import Wannier
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup():
    mat = synth.tiny(nb=4, natoms=2, ws=3, wv=6, seed=12)
    o = O.Wannier(mat)
    qr, nsr, _ = O.qrotations(mat, mat.wvmesh)
    el = O.Electrons(o, qr, nsr, mat.wvmesh, mat.enref, mat.fsthick)
    ph = O.Phonons(o, qr, nsr, mat.phmesh)
    return mat, Wannier(mat), el, ph


def test_ibzk_files_while_running(setup, tmp_path):
    mat, w, el, ph = setup
    plan = interactions.IbzkPlan(w, el, ph)
    nst = el.nwv_irred * mat.nb
    c_ref, g_ref = plan.run(1, nst)
    batch = 24
    dev = w.device_buffer((batch, plan.nprocs))
    with interactions.G2Writer(w, tmp_path, "k", plan.nprocs, batch) as wr:
        for s0 in range(1, nst + 1, batch):
            s1 = min(nst, s0 + batch - 1)
            counts, _ = plan.run(s0, s1, out=dev)
            wr.submit(s0, counts, dev)
        wr.wait()
    c, g = interactions.read_g2_files(tmp_path, "k", 1, nst, plan.nprocs)
    assert np.array_equal(c, c_ref) and (c >= 0).any() and (c < 0).any()
    act = c >= 0
    assert np.array_equal(g[act], g_ref[act])
    assert sorted(os.listdir(tmp_path)) == sorted(f"gk2.istate{i + 1}" for i in np.where(act)[0])
    dev.free()
    plan.close()


def test_ibzq_files_while_running(setup, tmp_path):
    mat, w, el, ph = setup
    plan = interactions.IbzqPlan(w, el, ph)
    nst = min(ph.nwv_irred * mat.nbr, 30)
    c_ref, g_ref = plan.run(1, nst)
    batch = 12
    dev = w.device_buffer((batch, plan.stride))
    with interactions.G2Writer(w, tmp_path, "q", plan.stride, batch) as wr:
        for s0 in range(1, nst + 1, batch):
            s1 = min(nst, s0 + batch - 1)
            counts, _ = plan.run(s0, s1, out=dev)
            wr.submit(s0, counts, dev)
    c, g = interactions.read_g2_files(tmp_path, "q", 1, nst, plan.stride)
    assert np.array_equal(c, c_ref)
    for t in range(nst):
        assert np.array_equal(g[t, :c[t]], g_ref[t, :c[t]])
    dev.free()
    plan.close()


def test_unwritable_directory_is_an_error(setup, tmp_path):
    mat, w, el, ph = setup
    plan = interactions.IbzkPlan(w, el, ph)
    dev = w.device_buffer((8, plan.nprocs))
    counts, _ = plan.run(1, 8, out=dev)
    assert (counts >= 0).any()
    target = tmp_path / "gone"
    wr = interactions.G2Writer(w, target, "k", plan.nprocs, 8)
    os.rmdir(target)
    wr.submit(1, counts, dev)
    with pytest.raises(ElphError) as ei:
        wr.wait()
    assert ei.value.code == ELPH_EIO
    with pytest.raises(ElphError):
        wr.close()
    dev.free()
    plan.close()
