// Host-side pieces of the vertex path: index search, crystal geometry,
// symmetry operations and the irreducible wedge, the EPW reader, WS wrapping
// vectors and the g2 file writer.
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <unordered_set>

#include "common.h"

namespace elph {

// src/misc.f90:738-763: bisection on a sorted list; the midpoint rule picks
// index 3 of [-105,0,105,105,105] for 105 (test/test_misc.f90:110-116).
int64_t binsearch(const int64_t *array, int64_t n, int64_t e) {
    int64_t a = 1, b = n, m = (b + a) / 2;
    int64_t mid = array[m - 1];
    while (mid != e) {
        if (e > mid) a = m + 1;
        else if (e < mid) b = m - 1;
        if (a > b) return -1;
        m = (b + a) / 2;
        mid = array[m - 1];
    }
    return m;
}

void reciprocal_lattice(const double L[9], double rec[9], double *volume) {
    for (int i = 0; i < 3; ++i) {
        int j = (i + 1) % 3, k = (j + 1) % 3;
        const double *a = &L[3 * j], *b = &L[3 * k];
        rec[3 * i + 0] = a[1] * b[2] - a[2] * b[1];
        rec[3 * i + 1] = a[2] * b[0] - a[0] * b[2];
        rec[3 * i + 2] = a[0] * b[1] - a[1] * b[0];
    }
    double vol = std::fabs(L[0] * rec[0] + L[1] * rec[1] + L[2] * rec[2]);
    double vbz = 2.0 * M_PI / vol;
    for (int i = 0; i < 9; ++i) rec[i] *= vbz;
    *volume = vol;
}

static void matvec3(const double m[9], const double v[3], double out[3]) {
    for (int r = 0; r < 3; ++r) out[r] = m[r] * v[0] + m[r + 3] * v[1] + m[r + 6] * v[2];
}

// Operations x -> R x + t (fractional coordinates) that keep the lattice metric
// and carry every atom onto an atom of its species.  spglib's dataset
// (src/symmetry.f90:97) lists the same operations for these structures; only
// the set matters downstream because the irreducible wedge keeps the first
// mesh point of each orbit.  Rotation entries of a lattice-preserving matrix
// in a reduced basis of a cubic/fcc/bcc lattice lie in {-1,0,1}.
int find_point_group(const double L[9], const double *basis, const int64_t *types, int64_t natoms, double symprec,
                     std::vector<int> &rot) {
    double G[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0;
            for (int c = 0; c < 3; ++c) s += L[3 * i + c] * L[3 * j + c];
            G[i + 3 * j] = s;
        }
    const double gscale = (G[0] + G[4] + G[8]) / 3.0;
    rot.clear();
    int e[9];
    for (int code = 0; code < 19683; ++code) {
        int c = code;
        for (int i = 0; i < 9; ++i) { e[i] = c % 3 - 1; c /= 3; }
        const int det = e[0] * (e[4] * e[8] - e[7] * e[5]) - e[3] * (e[1] * e[8] - e[7] * e[2]) +
                        e[6] * (e[1] * e[5] - e[4] * e[2]);
        if (det != 1 && det != -1) continue;
        bool ok = true;
        for (int i = 0; i < 3 && ok; ++i)
            for (int j = 0; j < 3 && ok; ++j) {
                double s = 0;
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b) s += e[a + 3 * i] * G[a + 3 * b] * e[b + 3 * j];
                if (std::fabs(s - G[i + 3 * j]) > 1e-8 * gscale) ok = false;
            }
        if (!ok) continue;
        bool found = false;
        for (int64_t j0 = 0; j0 < natoms && !found; ++j0) {
            if (types[j0] != types[0]) continue;
            double t[3];
            for (int r = 0; r < 3; ++r)
                t[r] = basis[3 * j0 + r] - (e[r] * basis[0] + e[r + 3] * basis[1] + e[r + 6] * basis[2]);
            bool all = true;
            for (int64_t a = 0; a < natoms && all; ++a) {
                double x[3];
                for (int r = 0; r < 3; ++r)
                    x[r] = e[r] * basis[3 * a] + e[r + 3] * basis[3 * a + 1] + e[r + 6] * basis[3 * a + 2] + t[r];
                bool hit = false;
                for (int64_t b = 0; b < natoms && !hit; ++b) {
                    if (types[b] != types[a]) continue;
                    double d[3], dc[3];
                    for (int r = 0; r < 3; ++r) { d[r] = x[r] - basis[3 * b + r]; d[r] -= std::nearbyint(d[r]); }
                    matvec3(L, d, dc);
                    if (std::sqrt(dc[0] * dc[0] + dc[1] * dc[1] + dc[2] * dc[2]) < symprec) hit = true;
                }
                all = hit;
            }
            found = all;
        }
        if (!found) continue;
        rot.insert(rot.end(), e, e + 9);
    }
    return (int)(rot.size() / 9);
}

static void inv3(const double a[9], double r[9]) {
    double det = a[0] * (a[4] * a[8] - a[7] * a[5]) - a[3] * (a[1] * a[8] - a[7] * a[2]) +
                 a[6] * (a[1] * a[5] - a[4] * a[2]);
    r[0] = (a[4] * a[8] - a[5] * a[7]) / det; r[3] = (a[6] * a[5] - a[3] * a[8]) / det;
    r[6] = (a[3] * a[7] - a[6] * a[4]) / det; r[1] = (a[7] * a[2] - a[1] * a[8]) / det;
    r[4] = (a[0] * a[8] - a[6] * a[2]) / det; r[7] = (a[6] * a[1] - a[0] * a[7]) / det;
    r[2] = (a[1] * a[5] - a[4] * a[2]) / det; r[5] = (a[3] * a[2] - a[0] * a[5]) / det;
    r[8] = (a[0] * a[4] - a[3] * a[1]) / det;
}

// Image of the 0-based mesh point ijk under qrot: the muxed index, or -1 when
// it is off the mesh (src/symmetry.f90:214-229, 277-286).
static int64_t star_image(const double *qr, const int64_t ijk[3], const int64_t mesh[3]) {
    double f[3] = {(double)ijk[0] / (double)mesh[0], (double)ijk[1] / (double)mesh[1], (double)ijk[2] / (double)mesh[2]};
    double v[3];
    matvec3(qr, f, v);
    int64_t iv[3];
    double d2 = 0;
    for (int c = 0; c < 3; ++c) {
        v[c] = (double)mesh[c] * v[c];
        iv[c] = fnint(v[c]);
        double d = std::fabs(v[c] - (double)iv[c]);
        d2 += d * d;
    }
    if (std::sqrt(d2) >= 1e-2) return -1;
    int64_t m[3] = {fmodulo(iv[0], mesh[0]), fmodulo(iv[1], mesh[1]), fmodulo(iv[2], mesh[2])};
    return mux_vector0(m, mesh);
}

// src/symmetry.f90:134-211: q-space rotations G R G^-1 (G = L^T L), time
// reversal appended, mesh-incompatible rotations and duplicates of the
// real-space matrices dropped.
int64_t q_rotations(const double L[9], const std::vector<int> &rot, const int64_t mesh[3], std::vector<double> &qrot,
                    std::vector<double> *crot) {
    const int64_t ns = (int64_t)(rot.size() / 9), n2 = 2 * ns;
    double G[9], Gi[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0;
            for (int c = 0; c < 3; ++c) s += L[3 * i + c] * L[3 * j + c];
            G[i + 3 * j] = s;
        }
    inv3(G, Gi);
    std::vector<double> qr(9 * n2);
    std::vector<int> rr(9 * n2);
    for (int64_t s = 0; s < ns; ++s) {
        double GR[9];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                double v = 0;
                for (int a = 0; a < 3; ++a) v += G[i + 3 * a] * rot[9 * s + a + 3 * j];
                GR[i + 3 * j] = v;
            }
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                double v = 0;
                for (int a = 0; a < 3; ++a) v += GR[i + 3 * a] * Gi[a + 3 * j];
                qr[9 * s + i + 3 * j] = v;
                qr[9 * (s + ns) + i + 3 * j] = -v;
            }
        for (int i = 0; i < 9; ++i) { rr[9 * s + i] = rot[9 * s + i]; rr[9 * (s + ns) + i] = -rot[9 * s + i]; }
    }
    std::vector<char> valid(n2, 1);
    const int64_t nq = mesh[0] * mesh[1] * mesh[2];
    for (int64_t s = 0; s < n2; ++s)
        for (int64_t i = 1; i <= nq && valid[s]; ++i) {
            int64_t ijk[3];
            demux_vector0(i, ijk, mesh);
            if (star_image(&qr[9 * s], ijk, mesh) < 0) valid[s] = 0;
        }
    for (int64_t s = 1; s < n2; ++s)
        for (int64_t t = 0; t < s; ++t) {
            if (!valid[t]) continue;
            if (std::memcmp(&rr[9 * s], &rr[9 * t], 9 * sizeof(int)) == 0) { valid[s] = 0; break; }
        }
    qrot.clear();
    for (int64_t s = 0; s < n2; ++s)
        if (valid[s]) qrot.insert(qrot.end(), &qr[9 * s], &qr[9 * s] + 9);
    if (crot) {
        // the Cartesian operations in the same order (src/symmetry.f90:131-137,
        // 154-159, 191-205): L W L^-1 for the first half, its negative for the
        // time-reversed half
        const std::vector<double> c = cartesian_rotations(L, rot);
        crot->clear();
        for (int64_t s = 0; s < n2; ++s) {
            if (!valid[s]) continue;
            for (int i = 0; i < 9; ++i) crot->push_back(s < ns ? c[9 * s + i] : -c[9 * (s - ns) + i]);
        }
    }
    return (int64_t)(qrot.size() / 9);
}

// src/symmetry.f90:131-137: the Cartesian form C = L W L^-1 of each operation
// W (fractional coordinates, x' = W x + w), column-major 3x3 blocks.
std::vector<double> cartesian_rotations(const double L[9], const std::vector<int> &rot) {
    const int64_t ns = (int64_t)(rot.size() / 9);
    double Li[9];
    inv3(L, Li);
    std::vector<double> out(9 * ns);
    for (int64_t s = 0; s < ns; ++s) {
        double LW[9];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                double v = 0;
                for (int a = 0; a < 3; ++a) v += L[i + 3 * a] * rot[9 * s + a + 3 * j];
                LW[i + 3 * j] = v;
            }
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                double v = 0;
                for (int a = 0; a < 3; ++a) v += LW[i + 3 * a] * Li[a + 3 * j];
                out[9 * s + i + 3 * j] = v;
            }
    }
    return out;
}

// src/electron.f90:462-476 (and phonon.f90:245-264): the symmetrizer of list
// point i is the mean of the Cartesian operations crot_orig[jj], jj < nsymm,
// whose reciprocal operation qrot[jj] maps the point onto itself (ShengBTE
// style; the reference pairs the first nsymm entries of the filtered
// reciprocal list with the unfiltered Cartesian list, as done here).  Points
// whose only such operation is one keep it undivided.
void symmetrizers(const int64_t mesh[3], int64_t nsymm, const double *qrot, const double *crot_orig, int64_t nk,
                  const int64_t *indexlist, double *out) {
    for (int64_t i = 0; i < nk; ++i) {
        const int64_t ii = indexlist ? indexlist[i] : i + 1;
        int64_t ijk[3];
        demux_vector0(ii, ijk, mesh);
        double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        int64_t kk = 0;
        for (int64_t jj = 0; jj < nsymm; ++jj) {
            if (star_image(&qrot[9 * jj], ijk, mesh) != ii) continue;
            for (int e = 0; e < 9; ++e) acc[e] += crot_orig[9 * jj + e];
            ++kk;
        }
        for (int e = 0; e < 9; ++e) out[9 * i + e] = kk > 1 ? acc[e] / (double)kk : acc[e];
    }
}

// src/symmetry.f90:303-451 find_irred_wedge.  A point starts a new orbit when
// no earlier orbit reached it; membership is a flag array over the mesh instead
// of the reference's distributed running-list scan (same result).
int64_t irred_wedge(const int64_t mesh[3], int64_t nsr, const double *qrot, const int64_t *indexlist, int64_t nlist,
                    int64_t *irr, int64_t *nequiv, int64_t *ibz2fbz) {
    const int64_t nmesh = mesh[0] * mesh[1] * mesh[2];
    std::vector<char> seen(nmesh + 1, 0);
    std::vector<int64_t> images(nsr);
    int64_t nirr = 0, counter = 0;
    for (int64_t i = 0; i < nlist; ++i) {
        const int64_t imux = indexlist ? indexlist[i] : i + 1;
        if (seen[imux]) continue;
        int64_t ijk[3];
        demux_vector0(imux, ijk, mesh);
        irr[nirr] = imux;
        int64_t neq = 0;
        for (int64_t s = 0; s < nsr; ++s) {
            const int64_t image = star_image(&qrot[9 * s], ijk, mesh);
            bool present = false;
            for (int64_t l = 0; l < neq; ++l)
                if (images[l] == image) { present = true; break; }
            if (present) continue;
            images[neq] = image;
            if (ibz2fbz) {
                ibz2fbz[neq + nsr * (nirr + nlist * 0)] = s + 1;
                ibz2fbz[neq + nsr * (nirr + nlist * 1)] = image;
            }
            neq++;
            if (image >= 1 && image <= nmesh) seen[image] = 1;
        }
        nequiv[nirr] = neq;
        counter += neq;
        nirr++;
    }
    if (counter != nlist) return -1;
    return nirr;
}

// src/wannier.f90:532-582: lattice vectors S with |lvec (frac(vl) + S')| minimal
// (within eps) over a 7^3 search box; S = S' - int(vl) (truncation toward zero).
int ws_wrapping_vectors(const double lvec[9], const double vl[3], double eps, std::vector<int> &wrappers) {
    int iv0[3];
    double wl[3];
    for (int c = 0; c < 3; ++c) { iv0[c] = (int)vl[c]; wl[c] = vl[c] - (double)iv0[c]; }
    double d = 1.7976931348623157e308;
    int n = 0;
    double norms[343];
    int idx = 0;
    for (int k = -3; k <= 3; ++k)
        for (int j = -3; j <= 3; ++j)
            for (int i = -3; i <= 3; ++i, ++idx) {
                double v[3] = {wl[0] + i, wl[1] + j, wl[2] + k}, wc[3];
                matvec3(lvec, v, wc);
                double nrm = std::sqrt(wc[0] * wc[0] + wc[1] * wc[1] + wc[2] * wc[2]);
                norms[idx] = nrm;
                if (std::fabs(nrm - d) < eps) n = n + 1;
                else if (nrm < d) { n = 1; d = nrm; }
            }
    wrappers.clear();
    idx = 0;
    for (int k = -3; k <= 3; ++k)
        for (int j = -3; j <= 3; ++j)
            for (int i = -3; i <= 3; ++i, ++idx)
                if (std::fabs(norms[idx] - d) < eps) {
                    wrappers.push_back(i - iv0[0]);
                    wrappers.push_back(j - iv0[1]);
                    wrappers.push_back(k - iv0[2]);
                }
    return (int)(wrappers.size() / 3);
}

// Fortran list-directed input (the `read(u, *)` statements of
// src/wannier.f90:313-340, 379-427): every statement starts at a new record
// and skips what is left of its last record; values are separated by blanks
// or a comma, may continue onto the next records, and a slash ends the
// statement; reals take E, D or Q exponents (or a signed exponent with no
// letter); a complex is "(re, im)" with blanks allowed around both parts and
// the pair allowed to continue onto the next record.
class ListDirected {
  public:
    explicit ListDirected(std::istream &in) {
        std::string line;
        while (std::getline(in, line)) {
            if (!line.empty() && line.back() == '\r') line.pop_back();
            recs_.push_back(line);
        }
    }
    // starts the next statement: the rest of the current record is skipped
    void statement() {
        if (started_) ++rec_;
        started_ = true;
        pos_ = 0;
        ended_ = false;
        first_ = true;
    }
    bool real(double &x) {
        std::string t;
        if (!item(t)) return false;
        return to_real(t, x);
    }
    bool integer(int64_t &x) {
        std::string t;
        if (!item(t)) return false;
        char *e = nullptr;
        x = std::strtoll(t.c_str(), &e, 10);
        return e && *e == '\0' && !t.empty();
    }
    bool complex(double &re, double &im) {
        if (!skip_separators()) return false;
        if (cur() != '(') return false;
        ++pos_;
        std::string a, b;
        if (!blanks_across() || !number(a) || !blanks_across() || cur() != ',') return false;
        ++pos_;
        if (!blanks_across() || !number(b) || !blanks_across() || cur() != ')') return false;
        ++pos_;
        first_ = false;
        return to_real(a, re) && to_real(b, im);
    }

  private:
    std::vector<std::string> recs_;
    size_t rec_ = 0, pos_ = 0;
    bool started_ = false, ended_ = false, first_ = true;  // first_: no value read yet in this statement

    char cur() const { return rec_ < recs_.size() && pos_ < recs_[rec_].size() ? recs_[rec_][pos_] : '\0'; }
    bool eor() const { return rec_ >= recs_.size() || pos_ >= recs_[rec_].size(); }
    static bool blank(char c) { return c == ' ' || c == '\t'; }
    bool next_record() {
        if (rec_ + 1 >= recs_.size()) return false;
        ++rec_;
        pos_ = 0;
        return true;
    }
    // blanks within a record and record ends (inside a complex constant)
    bool blanks_across() {
        for (;;) {
            while (!eor() && blank(cur())) ++pos_;
            if (!eor()) return true;
            if (!next_record()) return false;
        }
    }
    // moves to the start of the next value: blanks, record ends and at most
    // one comma after the previous value
    bool skip_separators() {
        if (ended_) return false;
        bool comma = false;
        for (;;) {
            while (!eor() && blank(cur())) ++pos_;
            if (eor()) {  // a record end separates like a blank
                if (!next_record()) return false;
                continue;
            }
            const char c = cur();
            if (c == '/') { ended_ = true; return false; }
            if (c == ',') {
                if (comma || first_) return false;  // null values are not accepted here
                comma = true;
                ++pos_;
                continue;
            }
            return true;
        }
    }
    bool number(std::string &t) {
        t.clear();
        while (!eor()) {
            const char c = cur();
            if (blank(c) || c == ',' || c == ')' || c == '/' || c == '(') break;
            t.push_back(c);
            ++pos_;
        }
        return !t.empty();
    }
    bool item(std::string &t) {
        if (!skip_separators()) return false;
        if (!number(t)) return false;
        first_ = false;
        return true;
    }
    static bool to_real(const std::string &t, double &x) {
        std::string u;
        bool expo = false, digits = false;
        for (size_t i = 0; i < t.size(); ++i) {
            char c = t[i];
            if (c == 'd' || c == 'D' || c == 'q' || c == 'Q' || c == 'e' || c == 'E') {
                c = 'E';
                expo = true;
            } else if ((c == '+' || c == '-') && i > 0 && !expo && digits) {
                u.push_back('E');  // "1.0+05": exponent without a letter
                expo = true;
            } else if (std::isdigit((unsigned char)c)) {
                digits = true;
            }
            u.push_back(c);
        }
        char *e = nullptr;
        x = std::strtod(u.c_str(), &e);
        return digits && e && *e == '\0';
    }
};

int read_epw(const std::string &dir, EpwData &o, std::string &err) {
    std::ifstream f(dir + "/epwdata.fmt");
    if (!f) { err = "cannot open " + dir + "/epwdata.fmt"; return ELPH_EIO; }
    ListDirected in(f);
    double ef;
    in.statement();
    if (!in.real(ef)) { err = "bad Fermi energy record in epwdata.fmt"; return ELPH_EIO; }
    in.statement();
    if (!in.integer(o.nb) || !in.integer(o.nwsk) || !in.integer(o.nbr) || !in.integer(o.nwsq) ||
        !in.integer(o.nwsg) || o.nb <= 0 || o.nbr <= 0 || o.nwsk <= 0 || o.nwsq <= 0 || o.nwsg <= 0) {
        err = "bad header in epwdata.fmt";
        return ELPH_EIO;
    }
    const int64_t ndummy = (o.nbr / 3 + 1) * 9;  // Born charges + epsilon, read and ignored
    in.statement();
    for (int64_t i = 0; i < ndummy; ++i) {
        double x;
        if (!in.real(x)) { err = "bad Born charge / epsilon records in epwdata.fmt"; return ELPH_EIO; }
    }
    auto read_block = [&](std::vector<double> &dst, int64_t nws, int64_t n) -> bool {
        dst.assign(2 * nws * n * n, 0.0);
        for (int64_t ib = 0; ib < n; ++ib)
            for (int64_t jb = 0; jb < n; ++jb)
                for (int64_t iuc = 0; iuc < nws; ++iuc) {
                    double re, im;
                    in.statement();
                    if (!in.complex(re, im)) return false;
                    const int64_t off = iuc + nws * (ib + n * jb);
                    dst[2 * off] = re;
                    dst[2 * off + 1] = im;
                }
        return true;
    };
    if (!read_block(o.Hwann, o.nwsk, o.nb)) { err = "bad Hwann records in epwdata.fmt"; return ELPH_EIO; }
    if (!read_block(o.Dphwann, o.nwsq, o.nbr)) { err = "bad Dphwann records in epwdata.fmt"; return ELPH_EIO; }
    const int64_t ng = o.nb * o.nb * o.nwsk * o.nbr * o.nwsg;
    o.gwann.resize(2 * ng);
    FILE *fb = std::fopen((dir + "/epmatwp1").c_str(), "rb");
    if (!fb) { err = "cannot open " + dir + "/epmatwp1"; return ELPH_EIO; }
    size_t got = std::fread(o.gwann.data(), sizeof(double), 2 * ng, fb);
    std::fclose(fb);
    if ((int64_t)got != 2 * ng) { err = "epmatwp1 is shorter than nb*nb*nwsk*nbr*nwsg complex values"; return ELPH_EIO; }
    auto read_cells = [&](const char *tag, int64_t n, std::vector<int64_t> &cells, std::vector<int64_t> &deg) -> bool {
        std::ifstream fc(dir + "/rcells_" + tag), fd(dir + "/wsdeg_" + tag);
        if (!fc || !fd) return false;
        ListDirected lc(fc), ld(fd);
        cells.assign(3 * n, 0);
        deg.assign(n, 0);
        for (int64_t i = 0; i < n; ++i) {  // (nws,3) column-major, one statement per cell
            lc.statement();
            ld.statement();
            if (!lc.integer(cells[i]) || !lc.integer(cells[i + n]) || !lc.integer(cells[i + 2 * n])) return false;
            if (!ld.integer(deg[i])) return false;
        }
        return true;
    };
    if (!read_cells("k", o.nwsk, o.rcells_k, o.elwsdeg)) { err = "bad rcells_k/wsdeg_k"; return ELPH_EIO; }
    if (!read_cells("q", o.nwsq, o.rcells_q, o.phwsdeg)) { err = "bad rcells_q/wsdeg_q"; return ELPH_EIO; }
    if (!read_cells("g", o.nwsg, o.rcells_g, o.gwsdeg)) { err = "bad rcells_g/wsdeg_g"; return ELPH_EIO; }
    return ELPH_OK;
}

}  // namespace elph

extern "C" int elph_epw_read_dims(const char *dir, int64_t dims[5]) {
    if (!dir || !dims) return ELPH_EINVAL;
    elph::EpwData e;
    std::string err;
    const int st = elph::read_epw(dir, e, err);
    if (st != ELPH_OK) {
        std::fprintf(stderr, "elph_epw_read_dims: %s\n", err.c_str());
        return st;
    }
    const int64_t d[5] = {e.nb, e.nwsk, e.nbr, e.nwsq, e.nwsg};
    std::memcpy(dims, d, sizeof(d));
    return ELPH_OK;
}

extern "C" int elph_epw_read(const char *dir, const int64_t dims[5], int64_t *rcells_k, int64_t *rcells_q,
                             int64_t *rcells_g, int64_t *elwsdeg, int64_t *phwsdeg, int64_t *gwsdeg, double *Hwann,
                             double *Dphwann, double *gwann) {
    if (!dir || !dims) return ELPH_EINVAL;
    elph::EpwData e;
    std::string err;
    const int st = elph::read_epw(dir, e, err);
    if (st != ELPH_OK) {
        std::fprintf(stderr, "elph_epw_read: %s\n", err.c_str());
        return st;
    }
    if (dims[0] != e.nb || dims[1] != e.nwsk || dims[2] != e.nbr || dims[3] != e.nwsq || dims[4] != e.nwsg)
        return ELPH_EINVAL;
    auto put = [](auto *dst, const auto &v) {
        if (dst) std::memcpy(dst, v.data(), v.size() * sizeof(v[0]));
    };
    put(rcells_k, e.rcells_k);
    put(rcells_q, e.rcells_q);
    put(rcells_g, e.rcells_g);
    put(elwsdeg, e.elwsdeg);
    put(phwsdeg, e.phwsdeg);
    put(gwsdeg, e.gwsdeg);
    put(Hwann, e.Hwann);
    put(Dphwann, e.Dphwann);
    put(gwann, e.gwann);
    return ELPH_OK;
}

extern "C" int elph_write_g2_files(const char *dir, char kind, int64_t istate_first, int64_t nstates,
                                   const int64_t *counts, const double *g2, int64_t row_stride) {
    if (!dir || !counts || !g2 || nstates < 0 || row_stride < 0 || (kind != 'k' && kind != 'q')) return ELPH_EINVAL;
    for (int64_t t = 0; t < nstates; ++t) {
        if (counts[t] < 0) continue;
        const int64_t c = counts[t];
        if (c > row_stride) return ELPH_EINVAL;
        char name[64];
        std::snprintf(name, sizeof(name), "%s.istate%lld", kind == 'k' ? "gk2" : "gq2",
                      (long long)(istate_first + t));
        FILE *f = std::fopen((std::string(dir) + "/" + name).c_str(), "wb");
        if (!f) return ELPH_EIO;
        // gk2: the whole row buffer (src/interactions.f90:2344-2348); gq2: count values (1884-1888)
        const int64_t nwrite = kind == 'k' ? row_stride : c;
        bool ok = std::fwrite(&c, sizeof(int64_t), 1, f) == 1;
        ok = ok && std::fwrite(g2 + t * row_stride, sizeof(double), nwrite, f) == (size_t)nwrite;
        if (std::fclose(f) != 0 || !ok) return ELPH_EIO;
    }
    return ELPH_OK;
}

extern "C" int elph_read_g2_files(const char *dir, char kind, int64_t istate_first, int64_t nstates, int64_t *counts,
                                  double *g2, int64_t row_stride) {
    if (!dir || !counts || !g2 || nstates < 0 || row_stride < 0 || (kind != 'k' && kind != 'q')) return ELPH_EINVAL;
    for (int64_t t = 0; t < nstates; ++t) {
        char name[64];
        std::snprintf(name, sizeof(name), "%s.istate%lld", kind == 'k' ? "gk2" : "gq2",
                      (long long)(istate_first + t));
        FILE *f = std::fopen((std::string(dir) + "/" + name).c_str(), "rb");
        if (!f) {
            // the drivers write no file for a state outside the window
            counts[t] = -1;
            continue;
        }
        int64_t c = -1;
        bool ok = std::fread(&c, sizeof(int64_t), 1, f) == 1 && c >= 0 && c <= row_stride;
        ok = ok && std::fread(g2 + t * row_stride, sizeof(double), c, f) == (size_t)c;
        std::fclose(f);
        if (!ok) return ELPH_EIO;
        counts[t] = c;
    }
    return ELPH_OK;
}

extern "C" int elph_write_X_files(const char *dir, int64_t istate_first, int64_t nstates, const int64_t *counts,
                                  const double *xplus, const double *xminus, const int64_t *istate_el,
                                  const int64_t *istate_ph, int64_t row_stride) {
    if (!dir || !counts || !xplus || !xminus || !istate_el || !istate_ph || nstates < 0) return ELPH_EINVAL;
    for (int64_t t = 0; t < nstates; ++t) {
        if (counts[t] < 0) continue;
        const int64_t c = counts[t];
        if (c > row_stride) return ELPH_EINVAL;
        for (int pm = 0; pm < 2; ++pm) {
            char name[64];
            std::snprintf(name, sizeof(name), "%s.istate%lld", pm == 0 ? "Xplus" : "Xminus",
                          (long long)(istate_first + t));
            FILE *f = std::fopen((std::string(dir) + "/" + name).c_str(), "wb");
            if (!f) return ELPH_EIO;
            const size_t o = (size_t)t * row_stride;
            bool ok = std::fwrite(&c, sizeof(int64_t), 1, f) == 1;
            ok = ok && std::fwrite((pm == 0 ? xplus : xminus) + o, sizeof(double), c, f) == (size_t)c;
            ok = ok && std::fwrite(istate_el + o, sizeof(int64_t), c, f) == (size_t)c;
            ok = ok && std::fwrite(istate_ph + o, sizeof(int64_t), c, f) == (size_t)c;
            if (std::fclose(f) != 0 || !ok) return ELPH_EIO;
        }
    }
    return ELPH_OK;
}

extern "C" int elph_symmetry(const double *lattvecs, int64_t natoms, const double *basis, const int64_t *atomtypes,
                             double symprec, const int64_t *mesh, int64_t *nsymm, int64_t *nsymm_rot,
                             double *qrotations) {
    std::vector<int> rot;
    int ns = elph::find_point_group(lattvecs, basis, atomtypes, natoms, symprec, rot);
    if (ns <= 0) return ELPH_EINVAL;
    std::vector<double> qr;
    int64_t nsr = elph::q_rotations(lattvecs, rot, mesh, qr);
    *nsymm = ns;
    *nsymm_rot = nsr;
    std::memcpy(qrotations, qr.data(), sizeof(double) * qr.size());
    return ELPH_OK;
}

extern "C" int elph_symmetry_cartesian(const double *lattvecs, int64_t natoms, const double *basis,
                                       const int64_t *atomtypes, double symprec, const int64_t *mesh, int64_t *nsymm,
                                       int64_t *nsymm_rot, double *crotations_orig, double *crotations) {
    std::vector<int> rot;
    int ns = elph::find_point_group(lattvecs, basis, atomtypes, natoms, symprec, rot);
    if (ns <= 0) return ELPH_EINVAL;
    std::vector<double> qr, cr;
    int64_t nsr = elph::q_rotations(lattvecs, rot, mesh, qr, &cr);
    const std::vector<double> co = elph::cartesian_rotations(lattvecs, rot);
    *nsymm = ns;
    *nsymm_rot = nsr;
    if (crotations_orig) std::memcpy(crotations_orig, co.data(), sizeof(double) * co.size());
    if (crotations) std::memcpy(crotations, cr.data(), sizeof(double) * cr.size());
    return ELPH_OK;
}

extern "C" int elph_symmetrizers(const int64_t *mesh, int64_t nsymm, const double *qrotations,
                                 const double *crotations_orig, int64_t nk, const int64_t *indexlist,
                                 double *symmetrizers) {
    if (!mesh || !qrotations || !crotations_orig || !symmetrizers || nk < 0 || nsymm <= 0) return ELPH_EINVAL;
    elph::symmetrizers(mesh, nsymm, qrotations, crotations_orig, nk, indexlist, symmetrizers);
    return ELPH_OK;
}

extern "C" int elph_irred_wedge(const int64_t *mesh, int64_t nsymm_rot, const double *qrotations, int64_t nlist,
                                const int64_t *indexlist, int64_t *nirr, int64_t *indexlist_irred, int64_t *nequiv,
                                int64_t *ibz2fbz) {
    int64_t n = elph::irred_wedge(mesh, nsymm_rot, qrotations, indexlist, nlist, indexlist_irred, nequiv, ibz2fbz);
    if (n < 0) return ELPH_EWEDGE;
    *nirr = n;
    return ELPH_OK;
}
