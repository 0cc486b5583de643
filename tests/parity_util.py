"""Helpers of the GPU parity tests (not a test module)."""
import numpy as np

from oracle import oracle as O


def branch_group_sums(rows, counts, el, ph_oracle_w, first, nb, nbr, kint_irr, tol=1e-9):
    """Sums of |g|^2 over groups of degenerate phonon branches, per (state,
    k', in-window n), for ibzk rows (k' ascending, n, s).  ph_oracle_w(q_int)
    gives the branch energies (eV) at the electron-mesh point q_int; branches
    whose energies differ by less than `tol` form one group (the sum over a
    degenerate group does not depend on the eigenvector basis inside it).
    Returns a flat array in a fixed order."""
    mesh = np.asarray(el.wvmesh)
    inw = np.abs(np.asarray(el.ens) - el.enref) <= el.fsthick
    kp = np.array([O.demux_vector(int(i), mesh, 0) for i in el.indexlist])
    out = []
    for t, c in enumerate(counts):
        if c < 0:
            continue
        _, ik = O.demux_state(first + t, nb)
        q_all = (kp - np.asarray(kint_irr[ik - 1])) % mesh
        pos = 0
        for r in range(el.nwv):
            w = ph_oracle_w(tuple(q_all[r]))
            lab = np.zeros(nbr, int)
            order = np.argsort(w)
            for a, b in zip(order[:-1], order[1:]):
                lab[b] = lab[a] + (0 if abs(w[b] - w[a]) < tol else 1)
            for n in range(nb):
                if not inw[r, n]:
                    continue
                blk = rows[t, pos:pos + nbr]
                out.extend(np.bincount(lab, weights=blk, minlength=nbr)[:lab.max() + 1])
                pos += nbr
        assert pos == c
    return np.asarray(out)


def fine_phonon_energies(o, mesh):
    """Energies of ph_wann at every point of the electron mesh, keyed by its
    integer coordinates (the phonons a mesh_ref > 1 run interpolates)."""
    mesh = np.asarray(mesh)
    pts = np.array([O.demux_vector(i, mesh, 0) for i in range(1, int(np.prod(mesh)) + 1)])
    w, _, _ = o.ph_wann(pts / mesh)
    table = {tuple(p): w[i] for i, p in enumerate(pts)}
    return lambda q: table[tuple(q)]
