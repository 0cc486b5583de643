"""elphbolt_amd: MI355X-native electron-phonon vertex path of elphbolt.

Public pieces: `Wannier` (src/wannier.f90 type Wannier), `bz.Symmetry`,
`bz.Electrons`, `bz.Phonons` (the state the drivers read), and the drivers
`interactions.calculate_eph_interaction_ibzk/ibzq`.  All compute runs in
libelph_vertex.so on the GPU; `synth` builds seeded EPW-format inputs.
"""
from . import constants  # noqa: F401

__all__ = ["constants", "synth", "wannier", "bz", "interactions"]
