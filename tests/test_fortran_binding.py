"""The Fortran side of the boundary compiles and links: include/elph_vertex_mod.f90
(iso_c_binding interfaces of every entry point in include/elph_vertex.h) is
compiled with amdflang together with a plain Fortran driver program, linked
against libelph_vertex.so, and run.  The run exercises the host-only entry
points through the bind(C) types and value attributes: symmetry and IBZ
(48 operations, 413 of 24^3 points, as examples/Si/output_gcc.txt reports),
the EPW host reader on list-directed variants of the files, and the gk2 file
writer and its resume reader."""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from elphbolt_amd import _lib, synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
AMDFLANG = "/opt/rocm/bin/amdflang"


def _header_symbols():
    txt = open(os.path.join(ROOT, "include", "elph_vertex.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(elph_[A-Za-z0-9_]+)\s*\(", txt)))


def test_module_covers_the_header():
    txt = open(os.path.join(ROOT, "include", "elph_vertex_mod.f90")).read()
    bound = sorted(set(re.findall(r'bind\(C, name="(elph_[A-Za-z0-9_]+)"\)', txt)))
    assert bound == _header_symbols()


@pytest.mark.skipif(not os.path.exists(AMDFLANG), reason="amdflang is not installed")
def test_fortran_driver_host_entry_points(tmp_path):
    _lib.load()                                   # the library must be built
    build = tmp_path / "build"
    build.mkdir()
    subprocess.run([AMDFLANG, "-c", os.path.join(ROOT, "include", "elph_vertex_mod.f90"), "-o", "mod.o"],
                   cwd=build, check=True)
    exe = build / "drv"
    subprocess.run([AMDFLANG, os.path.join(ROOT, "tests", "fortran", "elph_driver.f90"), "mod.o", "-o", str(exe),
                    "-L" + os.path.join(ROOT, "elphbolt_amd"), "-lelph_vertex",
                    "-Wl,-rpath," + os.path.join(ROOT, "elphbolt_amd")], cwd=build, check=True)
    mat = synth.tiny(nb=4, natoms=2, ws=3, wv=4, seed=81)
    epw, g2 = tmp_path / "epw", tmp_path / "g2"
    g2.mkdir()
    synth.write_epw(mat, str(epw), style="fortran")
    out = subprocess.run([str(exe), "host", str(epw), str(g2)], capture_output=True, text=True, check=True).stdout
    lines = {ln.split()[0] + " " + ln.split()[1]: ln for ln in out.strip().splitlines() if len(ln.split()) > 1}
    assert "host ok" in out
    assert lines["symmetry nsymm"].split() == ["symmetry", "nsymm", "48", "nirr", "413", "nequiv_sum", "13824"]
    assert lines["epw dims"].split()[2:] == [str(x) for x in (mat.nb, mat.nwsk, mat.nbr, mat.nwsq, mat.nwsg)]
    f = lines["epw rcells_k_sum"].split()
    assert int(f[2]) == int(np.abs(mat.rcells_k).sum()) and int(f[4]) == int(mat.elwsdeg.sum())
    assert abs(float(f[6]) - float(np.abs(mat.Hwann).sum())) < 1e-12 * float(np.abs(mat.Hwann).sum())
    assert lines["g2files counts"].split()[2:] == ["4", "-1", "2"]
    assert (g2 / "gk2.istate7").stat().st_size == 8 + 8 * 4
    assert not (g2 / "gk2.istate8").exists()
    shutil.rmtree(build)
