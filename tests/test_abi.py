"""The C-ABI library loads (no GPU needed) and exports every entry point that
include/elph_vertex.h declares; the ctypes prototypes cover the same set."""
import ctypes as ct
import os
import re

import pytest

from elphbolt_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "elph_vertex.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(elph_[A-Za-z0-9_]+)\s*\(", txt)))


def test_header_symbols_exported():
    lib = _lib.load()
    names = _declared()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_prototypes_match_header():
    assert sorted(_lib.PROTOTYPES) == _declared()


def test_missing_library_fails_loudly(tmp_path, monkeypatch):
    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "nope.so"))
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        _lib.load()


def test_write_and_read_g2_files(tmp_path):
    import numpy as np

    from elphbolt_amd import interactions
    counts = np.array([3, -1, 2], np.int64)
    g2 = np.arange(12, dtype=np.float64).reshape(3, 4)
    interactions.write_g2_files(str(tmp_path), "k", 5, counts, g2)
    c, v = interactions.read_g2_file(str(tmp_path / "gk2.istate5"))
    assert c == 3 and np.array_equal(v, g2[0])        # 'k': count + whole nprocs buffer
    assert not (tmp_path / "gk2.istate6").exists()     # out-of-window state: no file
    interactions.write_g2_files(str(tmp_path), "q", 1, counts, g2)
    c, v = interactions.read_g2_file(str(tmp_path / "gq2.istate3"))
    assert c == 2 and np.array_equal(v, g2[2, :2])    # 'q': count + count values


def test_g2_files_round_trip_through_the_abi(tmp_path):
    """elph_read_g2_files (the read_gk2/read_gq2 resume) returns what
    elph_write_g2_files wrote; missing files are states the driver skipped;
    counts beyond the row stride, truncated files and a missing directory are
    errors, not silent truncation."""
    import numpy as np

    from elphbolt_amd import interactions
    from elphbolt_amd._lib import ELPH_EINVAL, ELPH_EIO, ElphError
    rng = np.random.default_rng(3)
    counts = np.array([5, -1, 0, 7], np.int64)
    g2 = rng.random((4, 7))
    for kind in ("k", "q"):
        interactions.write_g2_files(str(tmp_path), kind, 11, counts, g2)
        c, g = interactions.read_g2_files(str(tmp_path), kind, 11, 4, 7)
        assert np.array_equal(c, counts)
        for t in np.where(counts >= 0)[0]:
            assert np.array_equal(g[t, :counts[t]], g2[t, :counts[t]])
    with pytest.raises(ElphError) as ei:
        interactions.write_g2_files(str(tmp_path), "q", 1, np.array([8], np.int64), g2[:1])
    assert ei.value.code == ELPH_EINVAL
    # a truncated file (count says 5 values, 2 present)
    path = tmp_path / "gk2.istate11"
    path.write_bytes(path.read_bytes()[:8 + 16])
    with pytest.raises(ElphError) as ei:
        interactions.read_g2_files(str(tmp_path), "k", 11, 1, 7)
    assert ei.value.code == ELPH_EIO
    st = _lib.load().elph_write_g2_files(str(tmp_path / "no" / "such").encode(), b"k", 1, 1,
                                         counts.ctypes.data_as(_lib.c_i64p), g2.ctypes.data_as(_lib.c_dblp), 7)
    assert st == ELPH_EIO


def test_symmetry_host_path():
    """elph_symmetry / elph_irred_wedge are host code: same counts as the
    reference's published values."""
    import numpy as np

    from elphbolt_amd import bz, synth
    L = synth.fcc_lattice(5.43)
    basis = np.array([[0, 0.25], [0, 0.25], [0, 0.25]], float)
    s = bz.Symmetry(L, basis, [28.0, 28.0], [24, 24, 24])
    assert s.nsymm == 48
    irr, neq = bz.irred_wedge([24, 24, 24], s)
    assert len(irr) == 413 and neq.sum() == 24 ** 3
    s2 = bz.Symmetry(L, basis, [28.0, 12.0], [6, 6, 6])
    assert s2.nsymm == 24
    assert len(bz.irred_wedge([6, 6, 6], s2)[0]) == 16


def test_write_and_read_X_files(tmp_path):
    """Xplus/Xminus.istate<i>: int64 count, count X, count istate_el, count
    istate_ph (src/interactions.f90:2356-2377), as read_transition_probs_e
    reads them (3065-3091); no file for a state outside the window."""
    import numpy as np

    from elphbolt_amd import interactions
    counts = np.array([3, -1, 0, 4], np.int64)
    rng = np.random.default_rng(0)
    xp, xm = rng.random((4, 5)), rng.random((4, 5))
    se, sp = rng.integers(1, 99, (4, 5)), -rng.integers(1, 99, (4, 5))
    interactions.write_X_files(str(tmp_path), 7, counts, xp, xm, se, sp)
    for t, c in enumerate(counts):
        for name, x in (("Xplus", xp), ("Xminus", xm)):
            path = tmp_path / f"{name}.istate{7 + t}"
            if c < 0:
                assert not path.exists()
                continue
            n, tp, s1, s2 = interactions.read_transition_probs_e(str(path))
            assert n == c and np.array_equal(tp, x[t, :c])
            assert np.array_equal(s1, se[t, :c]) and np.array_equal(s2, sp[t, :c])
            assert path.stat().st_size == 8 + 24 * c
