"""Cartesian symmetry operations (src/symmetry.f90:131-137) and the
symmetrizers of wave-vector dependent vectors (src/electron.f90:462-476) that
the field terms are multiplied by (src/bte.f90:242-246, 262-266), in the
oracle and in the product (host code).  Pins: the Cartesian operations of the
fcc point groups are orthogonal, integer-valued in the Cartesian frame of an
fcc lattice, and closed under multiplication; a symmetrizer is the mean of the
little group's operations, hence an orthogonal projector onto the vectors the
little group leaves invariant: zero at Gamma (Oh and Td leave no vector
invariant), the identity at a point with no symmetry but the identity, and the
projector onto the line for a point on the Gamma-L line; product and oracle
agree."""
import numpy as np
import pytest

from elphbolt_amd import bz, synth
from oracle import oracle as O


@pytest.fixture(scope="module", params=["si_Oh", "sic_Td"])
def case(request):
    m = synth.tiny(nb=2, natoms=2, ws=3, wv=6, seed=5)
    if request.param == "si_Oh":
        m.masses = np.array([28.0855, 28.0855])
    else:
        m.masses = np.array([28.0855, 12.0107])
    mesh = np.array([6, 6, 6])
    qr, nsr, ns = O.qrotations(m, mesh)
    co, cr = O.cartesian_rotations(m, mesh)
    return request.param, m, mesh, qr, nsr, ns, co, cr


def test_cartesian_operations(case):
    name, m, mesh, qr, nsr, ns, co, cr = case
    assert ns == (48 if name == "si_Oh" else 24)
    assert cr.shape[2] == nsr == 48
    for C in np.moveaxis(cr, 2, 0):
        assert np.allclose(C.T @ C, np.eye(3), atol=1e-12)
        assert np.allclose(C, np.rint(C), atol=1e-12)      # cubic axes: entries 0, +-1
    ints = {tuple(np.rint(C).astype(int).ravel()) for C in np.moveaxis(cr, 2, 0)}
    assert len(ints) == nsr
    for A in np.moveaxis(cr, 2, 0):
        for B in np.moveaxis(cr, 2, 0)[:8]:
            assert tuple(np.rint(A @ B).astype(int).ravel()) in ints
    # the first nsymm filtered operations are the original ones
    assert np.allclose(cr[:, :, :ns], co)


def test_symmetrizers_are_little_group_projectors(case):
    name, m, mesh, qr, nsr, ns, co, cr = case
    S = O.symmetrizers(mesh, qr, nsr, ns, co)                 # whole mesh, (3, 3, 216)
    for i in range(S.shape[2]):
        P = S[:, :, i]
        assert np.allclose(P @ P, P, atol=1e-12) and np.allclose(P, P.T, atol=1e-12)
    assert np.allclose(S[:, :, 0], 0.0)                       # Gamma
    rec = 2 * np.pi * np.linalg.inv(np.asarray(m.lattvecs)).T
    # a point on Gamma-L: k = t (b1 + b2 + b3)
    idx = O.mux_vector(np.array([1, 1, 1]), mesh, 0)
    P = S[:, :, idx - 1]
    kc = rec @ np.array([1.0, 1.0, 1.0])
    kc /= np.linalg.norm(kc)
    assert np.allclose(P, np.outer(kc, kc), atol=1e-12)
    # rank of each symmetrizer = dimension of the invariant subspace (0..3)
    ranks = {int(round(np.trace(S[:, :, i]))) for i in range(S.shape[2])}
    assert ranks <= {0, 1, 2, 3} and 0 in ranks and 1 in ranks


def test_identity_for_a_point_without_symmetry():
    m = synth.tiny(nb=2, natoms=2, ws=3, wv=6, seed=5)
    m.masses = np.array([28.0855, 12.0107])
    mesh = np.array([13, 13, 13])
    qr, nsr, ns = O.qrotations(m, mesh)
    co, _ = O.cartesian_rotations(m, mesh)
    idx = O.mux_vector(np.array([1, 2, 5]), mesh, 0)
    S = O.symmetrizers(mesh, qr, nsr, ns, co, indexlist=np.array([idx]))
    assert np.allclose(S[:, :, 0], np.eye(3), atol=1e-14)


def test_product_equals_oracle(case):
    name, m, mesh, qr, nsr, ns, co, cr = case
    sym = bz.Symmetry(m.lattvecs, m.basis, m.masses, mesh)
    assert (sym.nsymm, sym.nsymm_rot) == (ns, nsr)
    assert np.allclose(sym.crotations_orig, co, atol=1e-13)
    assert np.allclose(sym.crotations, cr, atol=1e-13)
    rng = np.random.default_rng(1)
    il = np.sort(rng.choice(np.arange(1, 217), 60, replace=False))
    a = bz.symmetrizers(mesh, sym, il)
    b = O.symmetrizers(mesh, qr, nsr, ns, co, indexlist=il)
    assert np.allclose(a, b, atol=1e-13)
