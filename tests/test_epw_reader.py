"""read_EPW_Wannier (src/wannier.f90:289-428) through the library's host
reader: the files the generator writes in its plain form, and the same values
in other forms a Fortran list-directed read accepts (blanks inside the complex
parentheses, D and d exponents, a complex split over two records, comma
separators, header values continued onto the next record, trailing values
after the first item of a record), read back bit for bit; malformed and
truncated files are errors."""
import numpy as np
import pytest

from elphbolt_amd import synth
from elphbolt_amd._lib import ELPH_EIO, ElphError
from elphbolt_amd.wannier import read_epw

FIELDS = ("rcells_k", "rcells_q", "rcells_g", "elwsdeg", "phwsdeg", "gwsdeg", "Hwann", "Dphwann", "gwann")


@pytest.fixture(scope="module")
def mat():
    return synth.tiny(nb=4, natoms=2, ws=3, wv=4, seed=81)


@pytest.mark.parametrize("style", ["plain", "fortran"])
def test_epw_files_read_back_bitwise(mat, tmp_path, style):
    synth.write_epw(mat, str(tmp_path), ef=1.25, style=style)
    got = read_epw(str(tmp_path))
    for name in FIELDS:
        ref = np.asarray(getattr(mat, name))
        assert got[name].shape == ref.shape, name
        assert np.array_equal(got[name], ref), name


def test_fortran_forms_are_really_used(mat, tmp_path):
    synth.write_epw(mat, str(tmp_path), style="fortran")
    txt = open(tmp_path / "epwdata.fmt").read()
    assert "D+" in txt and "d" in txt and " ( " in txt and ",\n" in txt


@pytest.mark.parametrize("damage", ["short_gwann", "bad_complex", "short_cells", "missing_header"])
def test_malformed_files_are_errors(mat, tmp_path, damage):
    synth.write_epw(mat, str(tmp_path), style="fortran")
    if damage == "short_gwann":
        p = tmp_path / "epmatwp1"
        p.write_bytes(p.read_bytes()[:-16])
    elif damage == "bad_complex":
        p = tmp_path / "epwdata.fmt"
        lines = p.read_text().split("\n")
        lines[10] = " (1.0D0 ; 2.0D0)"
        p.write_text("\n".join(lines))
    elif damage == "short_cells":
        p = tmp_path / "rcells_q"
        p.write_text("\n".join(p.read_text().split("\n")[:-3]))
    else:
        (tmp_path / "epwdata.fmt").write_text("1.0\n")
    with pytest.raises(ElphError) as ei:
        read_epw(str(tmp_path))
    assert ei.value.code == ELPH_EIO
