// tangent.cpp - forward-mode derivatives of the FP64 oracle (the reference gradient of the adjoint).
//
// TEST INFRASTRUCTURE ONLY (see oracle.h).  oracle_impl.c is compiled here a third time with REAL =
// orc::Dual: every arithmetic operation of the FP64 oracle runs on values and ORC_ND directional
// derivatives at once.  Decisions (cutoffs, cell indices, neighbour lists, bond detection) are taken
// on the values exactly as in the FP64 build, so the derivatives are those of the piecewise-smooth
// FP64 map along the FP64 trajectory, and the values are bitwise the FP64 oracle's.
#include "dual.hpp"

#define ORC_DUAL
#include "oracle_impl.c"

#include <vector>

namespace {

using orc::Dual;
using orc::ND;

void seed_scales(const orc_params *p, const double *pdir) {
    for (int j = 0; j < 4; ++j) {
        g_scale[j] = Dual(p->parScale[j]);
        if (pdir)
            for (int k = 0; k < ND; ++k) g_scale[j].d[k] = pdir[k * 4 + j];
    }
}

// Dual positions (x, y, z, w) with position directions rdir[k][3*n+d].
std::vector<Dual> dual_pos(const double *pos4, const double *rdir, int n) {
    std::vector<Dual> r((size_t)4 * n);
    for (int i = 0; i < n; ++i)
        for (int d = 0; d < 4; ++d) {
            r[4 * i + d] = Dual(pos4[4 * i + d]);
            if (rdir && d < 3)
                for (int k = 0; k < ND; ++k) r[4 * i + d].d[k] = rdir[(size_t)k * 3 * n + 3 * i + d];
        }
    return r;
}

}  // namespace

extern "C" {

int orc_tangent_window(const orc_params *p, double *pos4, double *vel, double *acc, int nsteps, int stepCount0,
                       int rebuildEvery, const double *pdir, const double *rdir, const double *vdir, double *J,
                       double *dJ) {
    const int n = p->nMol;
    seed_scales(p, pdir);
    std::vector<Dual> r = dual_pos(pos4, rdir, n), v((size_t)3 * n), a((size_t)3 * n);
    for (int i = 0; i < 3 * n; ++i) {
        v[i] = Dual(vel[i]);
        if (vdir)
            for (int k = 0; k < ND; ++k) v[i].d[k] = vdir[(size_t)k * 3 * n + i];
    }
    Dual Jd[2];
    const int err = orc_window_dual(p, r.data(), v.data(), a.data(), nsteps, stepCount0, rebuildEvery, Jd);
    for (int i = 0; i < 4 * n; ++i) pos4[i] = r[i].v;
    for (int i = 0; i < 3 * n; ++i) {
        vel[i] = v[i].v;
        acc[i] = a[i].v;
    }
    for (int q = 0; q < 2; ++q) {
        J[q] = Jd[q].v;
        for (int k = 0; k < ND; ++k) dJ[q * ND + k] = Jd[q].d[k];
    }
    return err;
}

void orc_tangent_forces(const orc_params *p, const double *pos4, const int *NN, const int *NBL, const double *pdir,
                        const double *rdir, double *acc, double *dacc, double *adh, double *dadh) {
    const int n = p->nMol, nMe = p->nMolMe;
    seed_scales(p, pdir);
    std::vector<Dual> r = dual_pos(pos4, rdir, n), rho((size_t)(nMe ? nMe : 1)), a((size_t)3 * n);
    if (nMe) orc_rho_dual(p, r.data(), NN, NBL, rho.data());
    orc_forces_dual(p, r.data(), NN, NBL, rho.data(), a.data(), nullptr);
    const Dual e = orc_adhesion_dual(p, r.data(), NN, NBL);
    for (int i = 0; i < 3 * n; ++i) {
        acc[i] = a[i].v;
        for (int k = 0; k < ND; ++k) dacc[(size_t)k * 3 * n + i] = a[i].d[k];
    }
    *adh = e.v;
    for (int k = 0; k < ND; ++k) dadh[k] = e.d[k];
}

}  // extern "C"
