"""The Fortran driver (tests/fortran/elph_driver.f90 over
include/elph_vertex_mod.f90, built by __graft_entry__.build()) on the GPU, as
elphbolt's image 1 would run the path (app/elphbolt.f90:184-230): a context
from EPW files, calculate_electrons / calculate_phonons, the ibzk 'g' driver
in FP64 and the gk2.istate* files.  The files equal, bit for bit, the rows
the Python mirror gets from the same library calls, and agree with the
oracle (fed the same electron and phonon states) within 1e-12."""
import os
import subprocess

import numpy as np
import pytest

from elphbolt_amd import bz, interactions, synth
from elphbolt_amd.wannier import Wannier
from oracle import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "fortran", "elph_driver")


def test_fortran_driver_on_the_device(tmp_path):
    assert os.path.exists(EXE), "tests/fortran/elph_driver is not built (run __graft_entry__.build())"
    mat = synth.tiny(nb=4, natoms=2, ws=3, wv=4, seed=83)
    epw, g2dir = tmp_path / "epw", tmp_path / "g2"
    g2dir.mkdir()
    synth.write_epw(mat, str(epw), style="fortran")
    with open(epw / "crystal.txt", "w") as f:
        f.write(f"{mat.natoms}\n")
        f.write(" ".join(f"{x:.17e}" for x in np.asarray(mat.lattvecs).ravel(order="F")) + "\n")
        f.write(" ".join(f"{x:.17e}" for x in np.asarray(mat.basis).ravel(order="F")) + "\n")
        f.write(" ".join(f"{x:.17e}" for x in mat.masses) + "\n")
        f.write(" ".join(str(int(x)) for x in mat.coarse_qmesh) + "\n")
        f.write(" ".join(str(int(x)) for x in mat.wvmesh) + "\n")
        f.write(f"{mat.enref:.17e} {mat.fsthick:.17e}\n")
    print("Fortran driver, device mode", flush=True)
    r = subprocess.run([EXE, "device", str(epw), str(g2dir)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0 and "device ok" in r.stdout
    # the same through the Python mirror
    w = Wannier.from_epw(str(epw), mat.natoms, mat.lattvecs, mat.basis, mat.masses, mat.coarse_qmesh)
    sym = bz.Symmetry(mat.lattvecs, mat.basis, mat.masses, mat.wvmesh)
    el = bz.Electrons(w, sym, mat.wvmesh, mat.enref, mat.fsthick)
    ph = bz.Phonons(w, sym, mat.phmesh)
    nst = el.nwv_irred * mat.nb
    plan = interactions.IbzkPlan(w, el, ph)
    counts, rows = plan.run(1, nst)
    c_f, g_f = interactions.read_g2_files(str(g2dir), "k", 1, nst, plan.nprocs)
    assert np.array_equal(c_f, counts) and (counts >= 0).any()
    act = counts >= 0
    assert np.array_equal(g_f[act], rows[act])
    last = min(nst, 16)
    c_o, g_o = O.ibzk_g2(O.Wannier(mat), el, ph, 1, 1, last)
    assert np.array_equal(c_o, counts[:last])
    a = c_o >= 0
    ref, got = g_o[a].ravel(), g_f[:last][a].ravel()
    sel = np.abs(ref) > 1e-6 * np.abs(ref).max()
    assert np.max(np.abs(got[sel] - ref[sel]) / np.abs(ref[sel])) < 1e-12
