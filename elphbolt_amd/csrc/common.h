// Shared declarations of the elph_vertex library (host side).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/elph_vertex.h"

namespace elph {

// Physical constants of src/params.f90:27-57.
constexpr double kQE = 1.602176634e-19;
constexpr double kME = 9.1093837015e-31;
constexpr double kAMU = 1.66053906660e-27;
constexpr double kHBAR = 1.05457172647e-22;
constexpr double kRyd2eV = 0.5 * 27.2116;
constexpr double kRyd2amu = 2.0 * kME / kAMU;
constexpr double kBohr2nm = 0.052917721092;
constexpr double kRyd2radTHz = 0.5 * 27.2116 * kQE / kHBAR;
constexpr double kG2Unit = kRyd2eV * kRyd2eV * kRyd2eV * kRyd2amu;  // src/wannier.f90:923

struct Error {
    int code;
    std::string msg;
};

// Fortran nint (round half away from zero) and modulo (sign of divisor).
inline int64_t fnint(double x) { return (int64_t)llround(x); }
inline int64_t fmodulo(int64_t a, int64_t p) {
    int64_t r = a % p;
    return (r != 0 && ((r < 0) != (p < 0))) ? r + p : r;
}

// Index algebra of src/misc.f90:906-990 (mux/demux of mesh points and states).
inline int64_t mux_vector0(const int64_t v[3], const int64_t mesh[3]) {
    return (v[2] * mesh[1] + v[1]) * mesh[0] + v[0] + 1;
}
inline void demux_vector0(int64_t i, int64_t v[3], const int64_t mesh[3]) {
    int64_t aux = (i - 1) / mesh[0];
    v[0] = (i - 1) % mesh[0];
    v[2] = aux / mesh[1];
    v[1] = aux % mesh[1];
}
inline void demux_state(int64_t m, int64_t nbands, int64_t *iband, int64_t *ik) {
    *iband = fmodulo(m - 1, nbands) + 1;
    *ik = (m - 1) / nbands + 1;
}
int64_t binsearch(const int64_t *array, int64_t n, int64_t e);  // src/misc.f90:738-763

// Crystal geometry (src/crystal.f90:282-293).
void reciprocal_lattice(const double L[9], double rec[9], double *volume);

// Symmetry (src/symmetry.f90) and irreducible wedge.
int find_point_group(const double L[9], const double *basis, const int64_t *types, int64_t natoms, double symprec,
                     std::vector<int> &rot);
// qrot (and, when crot is given, the Cartesian operations in the same order):
// the operations and their time-reversed negatives that are compatible with
// the mesh, duplicates removed (src/symmetry.f90:131-205).
int64_t q_rotations(const double L[9], const std::vector<int> &rot, const int64_t mesh[3], std::vector<double> &qrot,
                    std::vector<double> *crot = nullptr);
std::vector<double> cartesian_rotations(const double L[9], const std::vector<int> &rot);
void symmetrizers(const int64_t mesh[3], int64_t nsymm, const double *qrot, const double *crot_orig, int64_t nk,
                  const int64_t *indexlist, double *out);
int64_t irred_wedge(const int64_t mesh[3], int64_t nsr, const double *qrot, const int64_t *indexlist, int64_t nlist,
                    int64_t *irr, int64_t *nequiv, int64_t *ibz2fbz);

// EPW reader (src/wannier.f90:289-428).
struct EpwData {
    int64_t nb = 0, nwsk = 0, nbr = 0, nwsq = 0, nwsg = 0;
    std::vector<double> Hwann, Dphwann, gwann;  // interleaved complex, Fortran layouts
    std::vector<int64_t> rcells_k, rcells_q, rcells_g, elwsdeg, phwsdeg, gwsdeg;
};
int read_epw(const std::string &dir, EpwData &out, std::string &err);

// WS wrapping vectors of src/wannier.f90:532-582.
int ws_wrapping_vectors(const double lvec[9], const double vl[3], double eps, std::vector<int> &wrappers);

}  // namespace elph
