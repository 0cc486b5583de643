// Electron and phonon state the vertex drivers consume:
// calculate_electrons (src/electron.f90:321-595, steps 1-11) and the Wannier
// branch of calculate_phonons (src/phonon.f90:159-310).  Band structures come
// from the device el_wann / ph_wann; mesh bookkeeping runs on the host.
#include <cmath>
#include <cstring>
#include <vector>

#include "common.h"

using namespace elph;

struct elph_state {
    std::vector<int64_t> indexlist_irred, indexlist;
    std::vector<double> ens_irred, ens, evecs_irred, evecs;
};

namespace {

std::vector<double> mesh_kvecs(const int64_t *mesh, const int64_t *list, int64_t n) {
    std::vector<double> k(3 * n);
    for (int64_t i = 0; i < n; ++i) {
        int64_t v[3];
        demux_vector0(list ? list[i] : i + 1, v, mesh);
        for (int d = 0; d < 3; ++d) k[i + n * d] = (double)v[d] / (double)mesh[d];
    }
    return k;
}

}  // namespace

extern "C" int elph_calculate_electrons(elph_ctx *ctx, int64_t nsr, const double *qrot, const int64_t *mesh,
                                        double enref, double fsthick, elph_state **state, elph_electrons *view) {
    int64_t dims[6];
    if (elph_ctx_dims(ctx, dims) != ELPH_OK || !state || !view) return ELPH_EINVAL;
    const int64_t nb = dims[0];
    auto *s = new elph_state();
    auto fail = [&](int code) { delete s; return code; };
    const int64_t nwv = mesh[0] * mesh[1] * mesh[2];
    // 2. IBZ of the full mesh
    std::vector<int64_t> irr(nwv), neq(nwv), map(nsr * nwv * 2);
    const int64_t nirr = irred_wedge(mesh, nsr, qrot, nullptr, nwv, irr.data(), neq.data(), map.data());
    if (nirr < 0) return fail(ELPH_EWEDGE);
    // 3. IBZ energies
    std::vector<double> kir = mesh_kvecs(mesh, irr.data(), nirr), eir(nirr * nb);
    int st = elph_el_wann(ctx, nirr, kir.data(), eir.data(), nullptr, nullptr, nullptr);
    if (st != ELPH_OK) return fail(st);
    // 4. FBZ energies copied from the IBZ over each orbit
    std::vector<double> efbz(nwv * nb);
    for (int64_t i = 0; i < nirr; ++i)
        for (int64_t l = 0; l < neq[i]; ++l) {
            const int64_t il = map[l + nsr * (i + nwv * 1)];
            for (int64_t b = 0; b < nb; ++b) efbz[(il - 1) + nwv * b] = eir[i + nirr * b];
        }
    // 5-6. energy window (any band within fsthick), ascending index list
    for (int64_t ik = 0; ik < nwv; ++ik) {
        bool in = false;
        for (int64_t b = 0; b < nb; ++b)
            if (std::fabs(efbz[ik + nwv * b] - enref) <= fsthick) in = true;
        if (in) s->indexlist.push_back(ik + 1);
    }
    const int64_t cnt = (int64_t)s->indexlist.size();
    if (cnt == 0) return fail(ELPH_EWINDOW);
    s->ens.resize(cnt * nb);
    for (int64_t i = 0; i < cnt; ++i)
        for (int64_t b = 0; b < nb; ++b) s->ens[i + cnt * b] = efbz[(s->indexlist[i] - 1) + nwv * b];
    // 7. FBZ-block eigenvectors computed directly
    std::vector<double> kw = mesh_kvecs(mesh, s->indexlist.data(), cnt), etmp(cnt * nb);
    s->evecs.resize(2 * cnt * nb * nb);
    st = elph_el_wann(ctx, cnt, kw.data(), etmp.data(), nullptr, s->evecs.data(), nullptr);
    if (st != ELPH_OK) return fail(st);
    // 8. IBZ of the blocks
    std::vector<int64_t> irr2(cnt), neq2(cnt), map2(nsr * cnt * 2);
    const int64_t nirr2 = irred_wedge(mesh, nsr, qrot, s->indexlist.data(), cnt, irr2.data(), neq2.data(), map2.data());
    if (nirr2 < 0) return fail(ELPH_EWEDGE);
    s->indexlist_irred.assign(irr2.begin(), irr2.begin() + nirr2);
    // 9. IBZ-block energies and eigenvectors
    std::vector<double> kir2 = mesh_kvecs(mesh, s->indexlist_irred.data(), nirr2);
    s->ens_irred.resize(nirr2 * nb);
    s->evecs_irred.resize(2 * nirr2 * nb * nb);
    st = elph_el_wann(ctx, nirr2, kir2.data(), s->ens_irred.data(), nullptr, s->evecs_irred.data(), nullptr);
    if (st != ELPH_OK) return fail(st);
    // 10. in-window FBZ-block states
    int64_t nstates = 0;
    for (int64_t i = 0; i < cnt * nb; ++i)
        if (std::fabs(s->ens[i] - enref) <= fsthick) nstates++;
    // 11. FBZ-block energies from the IBZ blocks
    for (int64_t i = 0; i < nirr2; ++i)
        for (int64_t l = 0; l < neq2[i]; ++l) {
            const int64_t il = map2[l + nsr * (i + cnt * 1)];
            const int64_t aux = binsearch(s->indexlist.data(), cnt, il);
            if (aux > 0)
                for (int64_t b = 0; b < nb; ++b) s->ens[(aux - 1) + cnt * b] = s->ens_irred[i + nirr2 * b];
        }
    std::memset(view, 0, sizeof(*view));
    for (int d = 0; d < 3; ++d) view->wvmesh[d] = mesh[d];
    view->enref = enref;
    view->fsthick = fsthick;
    view->nwv_irred = nirr2;
    view->indexlist_irred = s->indexlist_irred.data();
    view->ens_irred = s->ens_irred.data();
    view->evecs_irred = s->evecs_irred.data();
    view->nwv = cnt;
    view->indexlist = s->indexlist.data();
    view->ens = s->ens.data();
    view->evecs = s->evecs.data();
    view->nstates_inwindow = nstates;
    *state = s;
    return ELPH_OK;
}

extern "C" int elph_calculate_phonons(elph_ctx *ctx, int64_t nsr, const double *qrot, const int64_t *mesh,
                                      elph_state **state, elph_phonons *view) {
    int64_t dims[6];
    if (elph_ctx_dims(ctx, dims) != ELPH_OK || !state || !view) return ELPH_EINVAL;
    const int64_t nbr = dims[1];
    auto *s = new elph_state();
    const int64_t nq = mesh[0] * mesh[1] * mesh[2];
    std::vector<double> qv = mesh_kvecs(mesh, nullptr, nq);
    s->ens.resize(nq * nbr);
    s->evecs.resize(2 * nq * nbr * nbr);
    int st = elph_ph_wann(ctx, nq, qv.data(), s->ens.data(), s->evecs.data(), nullptr);
    if (st != ELPH_OK) { delete s; return st; }
    std::vector<int64_t> irr(nq), neq(nq), map(nsr * nq * 2);
    const int64_t nirr = irred_wedge(mesh, nsr, qrot, nullptr, nq, irr.data(), neq.data(), map.data());
    if (nirr < 0) { delete s; return ELPH_EWEDGE; }
    for (int64_t i = 0; i < nirr; ++i) {
        const int64_t ii = irr[i];
        for (int64_t l = 0; l < neq[i]; ++l) {
            const int64_t il = map[l + nsr * (i + nq * 1)];
            for (int64_t b = 0; b < nbr; ++b) s->ens[(il - 1) + nq * b] = s->ens[(ii - 1) + nq * b];
        }
    }
    s->indexlist_irred.assign(irr.begin(), irr.begin() + nirr);
    std::memset(view, 0, sizeof(*view));
    for (int d = 0; d < 3; ++d) view->wvmesh[d] = mesh[d];
    view->nwv = nq;
    view->ens = s->ens.data();
    view->evecs = s->evecs.data();
    view->nwv_irred = nirr;
    view->indexlist_irred = s->indexlist_irred.data();
    *state = s;
    return ELPH_OK;
}

extern "C" int elph_state_free(elph_state *s) {
    delete s;
    return ELPH_OK;
}
