/*
 * CPU oracle for the NEMOLite2D time step -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the serial reference model
 * (original/nemolite2d.f90), used by the test-suite and by bench.py's CPU
 * baseline as the checker.  Nothing in the product library links or calls
 * it.  Every array keeps the reference's Fortran shape and lower bounds
 * (column-major, first index fastest) and every expression keeps the
 * reference's operation order, so that -O2 -ffp-contract=off builds give the
 * same IEEE double results as the reference built with flang at the same
 * flags.
 *
 * How it is pinned (tests/test_oracle.py):
 *   - golden end-of-run ua/va checksums (E16.8) and decimated go2d_*.dat
 *     snapshots from the reference binary (tests/golden/, made by
 *     tests/golden/make_golden.py from oracle/_ref);
 *   - closed forms: rest state stays at rest, the first step's tide value,
 *     zero edges, uniform-field interpolation, mass budget of continuity,
 *     Flather row given the inner row.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "nemo2d_oracle.h"

/* Physical constants, original/nemolite2d.f90:15-18. */
static const double PI = 3.1415926535897932;
static const double GRAV = 9.80665;
static const double OMEGA = 7.292116e-05;
#define D2R (PI / 180.0)

typedef struct {
    double *p;
    int lb1, ub1, lb2, ub2;
} farr;

struct o2d {
    int jpi, jpj;
    o2d_params prm;
    double tide_amp, tide_omega;  /* original/nemolite2d.f90:629-630 */
    int *pt;             /* (0:jpi+1, 0:jpj+1) */
    farr e1t, e2t, e1u, e2u, e1f, e2f, e1v, e2v, e12t, e12u, e12v;
    farr gphiu, gphiv, gphif, xt, yt, ht, hu, hv, hf;
    farr sshb, sshb_u, sshb_v, sshn, sshn_u, sshn_v, ssha, ssha_u, ssha_v;
    farr un, vn, ua, va;
    /* sin(gphi*d2r) per U/V point, filled for the duration of o2d_run (the
     * latitudes do not change within a run); NULL: evaluate per point */
    double *sin_u, *sin_v;
};

/* Fortran-style element access with explicit lower bounds. */
#define F(a, i, j) ((a).p[((i) - (a).lb1) + (long)((j) - (a).lb2) * ((a).ub1 - (a).lb1 + 1)])
#define PT(m, i, j) ((m)->pt[(i) + (long)(j) * ((m)->jpi + 2)])

static int falloc(farr *a, int lb1, int ub1, int lb2, int ub2) {
    a->lb1 = lb1; a->ub1 = ub1; a->lb2 = lb2; a->ub2 = ub2;
    long n = (long)(ub1 - lb1 + 1) * (ub2 - lb2 + 1);
    a->p = (double *)calloc(n > 0 ? n : 1, sizeof(double));
    return a->p != NULL;
}

/* allocation: original/nemolite2d.f90:285-320 (all arrays zeroed). */
o2d *o2d_create(int jpi, int jpj) {
    if (jpi < 1 || jpj < 1) return NULL;
    o2d *m = (o2d *)calloc(1, sizeof(o2d));
    if (!m) return NULL;
    m->jpi = jpi; m->jpj = jpj;
    m->tide_amp = 0.2;
    m->tide_omega = 2.0 * 3.14159 / (12.42 * 3600.0);
    int ok = 1;
    ok &= falloc(&m->e1t, 1, jpi, 1, jpj);   ok &= falloc(&m->e2t, 1, jpi, 1, jpj);
    ok &= falloc(&m->e1u, 0, jpi, 1, jpj);   ok &= falloc(&m->e2u, 0, jpi, 1, jpj);
    ok &= falloc(&m->e1f, 0, jpi, 0, jpj);   ok &= falloc(&m->e2f, 0, jpi, 0, jpj);
    ok &= falloc(&m->e1v, 1, jpi, 0, jpj);   ok &= falloc(&m->e2v, 1, jpi, 0, jpj);
    ok &= falloc(&m->e12t, 1, jpi, 1, jpj);  ok &= falloc(&m->e12u, 0, jpi, 1, jpj);
    ok &= falloc(&m->e12v, 1, jpi, 0, jpj);
    ok &= falloc(&m->gphiu, 0, jpi, 1, jpj); ok &= falloc(&m->gphiv, 1, jpi, 0, jpj);
    ok &= falloc(&m->gphif, 0, jpi, 0, jpj);
    ok &= falloc(&m->xt, 1, jpi, 1, jpj);    ok &= falloc(&m->yt, 1, jpi, 1, jpj);
    ok &= falloc(&m->ht, 1, jpi, 1, jpj);    ok &= falloc(&m->hu, 0, jpi, 1, jpj);
    ok &= falloc(&m->hv, 1, jpi, 0, jpj);    ok &= falloc(&m->hf, 0, jpi, 0, jpj);
    ok &= falloc(&m->sshb, 1, jpi, 1, jpj);  ok &= falloc(&m->sshb_u, 0, jpi, 1, jpj);
    ok &= falloc(&m->sshb_v, 1, jpi, 0, jpj);
    ok &= falloc(&m->sshn, 1, jpi, 1, jpj);  ok &= falloc(&m->sshn_u, 0, jpi, 1, jpj);
    ok &= falloc(&m->sshn_v, 1, jpi, 0, jpj);
    ok &= falloc(&m->ssha, 1, jpi, 1, jpj);  ok &= falloc(&m->ssha_u, 0, jpi, 1, jpj);
    ok &= falloc(&m->ssha_v, 1, jpi, 0, jpj);
    ok &= falloc(&m->un, 0, jpi, 1, jpj);    ok &= falloc(&m->vn, 1, jpi, 0, jpj);
    ok &= falloc(&m->ua, 0, jpi, 1, jpj);    ok &= falloc(&m->va, 1, jpi, 0, jpj);
    m->pt = (int *)calloc((size_t)(jpi + 2) * (jpj + 2), sizeof(int));
    if (!ok || !m->pt) { o2d_destroy(m); return NULL; }
    return m;
}

void o2d_destroy(o2d *m) {
    if (!m) return;
    farr *all[] = {&m->e1t, &m->e2t, &m->e1u, &m->e2u, &m->e1f, &m->e2f, &m->e1v, &m->e2v,
                   &m->e12t, &m->e12u, &m->e12v, &m->gphiu, &m->gphiv, &m->gphif, &m->xt,
                   &m->yt, &m->ht, &m->hu, &m->hv, &m->hf, &m->sshb, &m->sshb_u, &m->sshb_v,
                   &m->sshn, &m->sshn_u, &m->sshn_v, &m->ssha, &m->ssha_u, &m->ssha_v,
                   &m->un, &m->vn, &m->ua, &m->va};
    for (size_t k = 0; k < sizeof(all) / sizeof(all[0]); ++k) free(all[k]->p);
    free(m->pt);
    free(m);
}

void o2d_set_params(o2d *m, const o2d_params *p) { m->prm = *p; }

int o2d_jpi(const o2d *m) { return m->jpi; }
int o2d_jpj(const o2d *m) { return m->jpj; }

static farr *lookup(o2d *m, const char *name) {
#define L(x) if (!strcmp(name, #x)) return &m->x;
    L(e1t) L(e2t) L(e1u) L(e2u) L(e1f) L(e2f) L(e1v) L(e2v) L(e12t) L(e12u) L(e12v)
    L(gphiu) L(gphiv) L(gphif) L(xt) L(yt) L(ht) L(hu) L(hv) L(hf)
    L(sshb) L(sshb_u) L(sshb_v) L(sshn) L(sshn_u) L(sshn_v) L(ssha) L(ssha_u) L(ssha_v)
    L(un) L(vn) L(ua) L(va)
#undef L
    return NULL;
}

double *o2d_array(o2d *m, const char *name, int bounds[4]) {
    farr *a = lookup(m, name);
    if (!a) return NULL;
    bounds[0] = a->lb1; bounds[1] = a->ub1; bounds[2] = a->lb2; bounds[3] = a->ub2;
    return a->p;
}

int *o2d_pt(o2d *m) { return m->pt; }

/* grid, case jphgr_msh=1: original/nemolite2d.f90:180-279. */
void o2d_grid_default(o2d *m) {
    const int jpi = m->jpi, jpj = m->jpj;
    const double dx = m->prm.dx, dy = m->prm.dy, dep = m->prm.dep_const;
    for (int jj = 0; jj <= jpj + 1; ++jj)
        for (int ji = 0; ji <= jpi + 1; ++ji) PT(m, ji, jj) = 1;
    for (int jj = 0; jj <= jpj + 1; ++jj) { PT(m, 0, jj) = 0; PT(m, jpi + 1, jj) = 0; }
    for (int ji = 0; ji <= jpi + 1; ++ji) PT(m, ji, jpj + 1) = 0;
    for (int ji = 1; ji <= jpi; ++ji) PT(m, ji, 0) = -1;

    for (int jj = 1; jj <= jpj; ++jj)
        for (int ji = 1; ji <= jpi; ++ji) { F(m->e1t, ji, jj) = dx; F(m->e2t, ji, jj) = dy; }
    for (int jj = 1; jj <= jpj; ++jj)
        for (int ji = 0; ji <= jpi; ++ji) {
            F(m->e1u, ji, jj) = dx; F(m->e2u, ji, jj) = dy; F(m->gphiu, ji, jj) = 50.0;
        }
    for (int jj = 0; jj <= jpj; ++jj)
        for (int ji = 1; ji <= jpi; ++ji) {
            F(m->e1v, ji, jj) = dx; F(m->e2v, ji, jj) = dy; F(m->gphiv, ji, jj) = 50.0;
        }
    for (int jj = 0; jj <= jpj; ++jj)
        for (int ji = 0; ji <= jpi; ++ji) {
            F(m->e1f, ji, jj) = dx; F(m->e2f, ji, jj) = dy; F(m->gphif, ji, jj) = 50.0;
        }
    /* cumulative coordinates, original/nemolite2d.f90:224-233 */
    for (int jj = 1; jj <= jpj; ++jj) F(m->xt, 1, jj) = 0.0 + 0.5 * F(m->e1t, 1, jj);
    for (int ji = 1; ji <= jpi; ++ji) F(m->yt, ji, 1) = 0.0 + 0.5 * F(m->e2t, ji, 1);
    for (int ji = 2; ji <= jpi; ++ji)
        for (int jj = 1; jj <= jpj; ++jj) F(m->xt, ji, jj) = F(m->xt, ji - 1, jj) + dx;
    for (int jj = 2; jj <= jpj; ++jj)
        for (int ji = 1; ji <= jpi; ++ji) F(m->yt, ji, jj) = F(m->yt, ji, jj - 1) + dy;

    for (int jj = 1; jj <= jpj; ++jj)
        for (int ji = 1; ji <= jpi; ++ji) F(m->ht, ji, jj) = dep;
    for (int jj = 1; jj <= jpj; ++jj)
        for (int ji = 0; ji <= jpi; ++ji) F(m->hu, ji, jj) = dep;
    for (int jj = 0; jj <= jpj; ++jj)
        for (int ji = 1; ji <= jpi; ++ji) F(m->hv, ji, jj) = dep;
    o2d_grid_areas(m);
}

/* cell areas, original/nemolite2d.f90:262-279. */
void o2d_grid_areas(o2d *m) {
    const int jpi = m->jpi, jpj = m->jpj;
    for (int jj = 1; jj <= jpj; ++jj)
        for (int ji = 1; ji <= jpi; ++ji)
            F(m->e12t, ji, jj) = F(m->e1t, ji, jj) * F(m->e2t, ji, jj);
    for (int jj = 1; jj <= jpj; ++jj)
        for (int ji = 0; ji <= jpi; ++ji)
            F(m->e12u, ji, jj) = F(m->e1u, ji, jj) * F(m->e2u, ji, jj);
    for (int jj = 0; jj <= jpj; ++jj)
        for (int ji = 1; ji <= jpi; ++ji)
            F(m->e12v, ji, jj) = F(m->e1v, ji, jj) * F(m->e2v, ji, jj);
}

/* initialisation, original/nemolite2d.f90:324-371: sshn is taken as set
 * (zero after create), sshn_u/sshn_v interpolated from it, un = vn = 0. */
void o2d_initialise(o2d *m) {
    const int jpi = m->jpi, jpj = m->jpj;
    for (int ji = 0; ji <= jpi; ++ji)
        for (int jj = 1; jj <= jpj; ++jj) {
            int i1 = ji + 1 < jpi ? ji + 1 : jpi;
            int i2 = ji > 1 ? ji : 1;
            double r = F(m->e12t, i1, jj) * F(m->sshn, i1, jj) + F(m->e12t, i2, jj) * F(m->sshn, i2, jj);
            F(m->sshn_u, ji, jj) = 0.5 * r / F(m->e12u, ji, jj);
        }
    for (int ji = 1; ji <= jpi; ++ji)
        for (int jj = 0; jj <= jpj; ++jj) {
            int j1 = jj + 1 < jpj ? jj + 1 : jpj;
            int j2 = jj > 1 ? jj : 1;
            double r = F(m->e12t, ji, j1) * F(m->sshn, ji, j1) + F(m->e12t, ji, j2) * F(m->sshn, ji, j2);
            F(m->sshn_v, ji, jj) = 0.5 * r / F(m->e12v, ji, jj);
        }
    for (int jj = 1; jj <= jpj; ++jj)
        for (int ji = 0; ji <= jpi; ++ji) F(m->un, ji, jj) = 0.0;
    for (int jj = 0; jj <= jpj; ++jj)
        for (int ji = 1; ji <= jpi; ++ji) F(m->vn, ji, jj) = 0.0;
}

/* continuity, original/nemolite2d.f90:390-410. */
void o2d_continuity(o2d *m) {
    const int jpi = m->jpi, jpj = m->jpj;
    const double rdt = m->prm.rdt;
    for (int jj = 1; jj <= jpj; ++jj)
        for (int ji = 1; ji <= jpi; ++ji) {
            double r1 = (F(m->sshn_u, ji, jj) + F(m->hu, ji, jj)) * F(m->un, ji, jj);
            double r2 = (F(m->sshn_u, ji - 1, jj) + F(m->hu, ji - 1, jj)) * F(m->un, ji - 1, jj);
            double r3 = (F(m->sshn_v, ji, jj) + F(m->hv, ji, jj)) * F(m->vn, ji, jj);
            double r4 = (F(m->sshn_v, ji, jj - 1) + F(m->hv, ji, jj - 1)) * F(m->vn, ji, jj - 1);
            F(m->ssha, ji, jj) = F(m->sshn, ji, jj) + (r2 - r1 + r4 - r3) * rdt / F(m->e12t, ji, jj);
        }
}

/* u half of momentum, original/nemolite2d.f90:430-517. */
void o2d_momentum_u(o2d *m) {
    const int jpi = m->jpi, jpj = m->jpj;
    const double rdt = m->prm.rdt, cbfr = m->prm.cbfr, visc = m->prm.visc;
    for (int jj = 1; jj <= jpj; ++jj)
        for (int ji = 1; ji <= jpi - 1; ++ji) {
            if (PT(m, ji, jj) + PT(m, ji + 1, jj) <= 0) continue;
            if (PT(m, ji, jj) <= 0 || PT(m, ji + 1, jj) <= 0) continue;

            double u_e = 0.5 * (F(m->un, ji, jj) + F(m->un, ji + 1, jj)) * F(m->e2t, ji + 1, jj);
            double depe = F(m->ht, ji + 1, jj) + F(m->sshn, ji + 1, jj);
            double u_w = 0.5 * (F(m->un, ji, jj) + F(m->un, ji - 1, jj)) * F(m->e2t, ji, jj);
            double depw = F(m->ht, ji, jj) + F(m->sshn, ji, jj);
            double v_sc = 0.5 * (F(m->vn, ji, jj - 1) + F(m->vn, ji + 1, jj - 1));
            double v_s = 0.5 * v_sc * (F(m->e1v, ji, jj - 1) + F(m->e1v, ji + 1, jj - 1));
            double deps = 0.5 * (F(m->hv, ji, jj - 1) + F(m->sshn_v, ji, jj - 1) +
                                 F(m->hv, ji + 1, jj - 1) + F(m->sshn_v, ji + 1, jj - 1));
            double v_nc = 0.5 * (F(m->vn, ji, jj) + F(m->vn, ji + 1, jj));
            double v_n = 0.5 * v_nc * (F(m->e1v, ji, jj) + F(m->e1v, ji + 1, jj));
            double depn = 0.5 * (F(m->hv, ji, jj) + F(m->sshn_v, ji, jj) +
                                 F(m->hv, ji + 1, jj) + F(m->sshn_v, ji + 1, jj));

            double uu_w = (0.5 - copysign(0.5, u_w)) * F(m->un, ji, jj) +
                          (0.5 + copysign(0.5, u_w)) * F(m->un, ji - 1, jj);
            double uu_e = (0.5 + copysign(0.5, u_e)) * F(m->un, ji, jj) +
                          (0.5 - copysign(0.5, u_e)) * F(m->un, ji + 1, jj);
            double uu_s, uu_n;
            int wall_s = PT(m, ji, jj - 1) <= 0 || PT(m, ji + 1, jj - 1) <= 0;
            int wall_n = PT(m, ji, jj + 1) <= 0 || PT(m, ji + 1, jj + 1) <= 0;
            if (wall_s)
                uu_s = (0.5 - copysign(0.5, v_s)) * F(m->un, ji, jj);
            else
                uu_s = (0.5 - copysign(0.5, v_s)) * F(m->un, ji, jj) +
                       (0.5 + copysign(0.5, v_s)) * F(m->un, ji, jj - 1);
            if (wall_n)
                uu_n = (0.5 + copysign(0.5, v_n)) * F(m->un, ji, jj);
            else
                uu_n = (0.5 + copysign(0.5, v_n)) * F(m->un, ji, jj) +
                       (0.5 - copysign(0.5, v_n)) * F(m->un, ji, jj + 1);
            double adv = uu_w * u_w * depw - uu_e * u_e * depe + uu_s * v_s * deps - uu_n * v_n * depn;

            double dudx_e = (F(m->un, ji + 1, jj) - F(m->un, ji, jj)) / F(m->e1t, ji + 1, jj) *
                            (F(m->ht, ji + 1, jj) + F(m->sshn, ji + 1, jj));
            double dudx_w = (F(m->un, ji, jj) - F(m->un, ji - 1, jj)) / F(m->e1t, ji, jj) *
                            (F(m->ht, ji, jj) + F(m->sshn, ji, jj));
            double dudy_s, dudy_n;
            if (wall_s)
                dudy_s = 0.0;
            else
                dudy_s = (F(m->un, ji, jj) - F(m->un, ji, jj - 1)) /
                         (F(m->e2u, ji, jj) + F(m->e2u, ji, jj - 1)) *
                         (F(m->hu, ji, jj) + F(m->sshn_u, ji, jj) + F(m->hu, ji, jj - 1) + F(m->sshn_u, ji, jj - 1));
            if (wall_n)
                dudy_n = 0.0;
            else
                dudy_n = (F(m->un, ji, jj + 1) - F(m->un, ji, jj)) /
                         (F(m->e2u, ji, jj) + F(m->e2u, ji, jj + 1)) *
                         (F(m->hu, ji, jj) + F(m->sshn_u, ji, jj) + F(m->hu, ji, jj + 1) + F(m->sshn_u, ji, jj + 1));
            double vis = (dudx_e - dudx_w) * F(m->e2u, ji, jj) + (dudy_n - dudy_s) * F(m->e1u, ji, jj) * 0.5;
            vis = visc * vis;

            const double sphi = m->sin_u ? m->sin_u[(ji - m->gphiu.lb1) + (long)(jj - m->gphiu.lb2) * (m->jpi + 1)]
                                         : sin(F(m->gphiu, ji, jj) * D2R);
            double cor = 0.5 * (2.0 * OMEGA * sphi * (v_sc + v_nc)) *
                         F(m->e12u, ji, jj) * (F(m->hu, ji, jj) + F(m->sshn_u, ji, jj));
            double hpg = -GRAV * (F(m->hu, ji, jj) + F(m->sshn_u, ji, jj)) * F(m->e2u, ji, jj) *
                         (F(m->sshn, ji + 1, jj) - F(m->sshn, ji, jj));
            F(m->ua, ji, jj) = (F(m->un, ji, jj) * (F(m->hu, ji, jj) + F(m->sshn_u, ji, jj)) +
                                rdt * (adv + vis + cor + hpg) / F(m->e12u, ji, jj)) /
                               (F(m->hu, ji, jj) + F(m->ssha_u, ji, jj)) / (1.0 + cbfr * rdt);
        }
}

/* v half of momentum, original/nemolite2d.f90:519-609. */
void o2d_momentum_v(o2d *m) {
    const int jpi = m->jpi, jpj = m->jpj;
    const double rdt = m->prm.rdt, cbfr = m->prm.cbfr, visc = m->prm.visc;
    for (int jj = 1; jj <= jpj - 1; ++jj)
        for (int ji = 1; ji <= jpi; ++ji) {
            /* the first test looks east, not north (original/nemolite2d.f90:523) */
            if (PT(m, ji, jj) + PT(m, ji + 1, jj) <= 0) continue;
            if (PT(m, ji, jj) <= 0 || PT(m, ji, jj + 1) <= 0) continue;

            double v_n = 0.5 * (F(m->vn, ji, jj) + F(m->vn, ji, jj + 1)) * F(m->e1t, ji, jj + 1);
            double depn = F(m->ht, ji, jj + 1) + F(m->sshn, ji, jj + 1);
            double v_s = 0.5 * (F(m->vn, ji, jj) + F(m->vn, ji, jj - 1)) * F(m->e1t, ji, jj);
            double deps = F(m->ht, ji, jj) + F(m->sshn, ji, jj);
            double u_wc = 0.5 * (F(m->un, ji - 1, jj) + F(m->un, ji - 1, jj + 1));
            double u_w = 0.5 * u_wc * (F(m->e2u, ji - 1, jj) + F(m->e2u, ji - 1, jj + 1));
            double depw = 0.50 * (F(m->hu, ji - 1, jj) + F(m->sshn_u, ji - 1, jj) +
                                  F(m->hu, ji - 1, jj + 1) + F(m->sshn_u, ji - 1, jj + 1));
            double u_ec = 0.5 * (F(m->un, ji, jj) + F(m->un, ji, jj + 1));
            double u_e = 0.5 * u_ec * (F(m->e2u, ji, jj) + F(m->e2u, ji, jj + 1));
            double depe = 0.50 * (F(m->hu, ji, jj) + F(m->sshn_u, ji, jj) +
                                  F(m->hu, ji, jj + 1) + F(m->sshn_u, ji, jj + 1));

            double vv_s = (0.5 - copysign(0.5, v_s)) * F(m->vn, ji, jj) +
                          (0.5 + copysign(0.5, v_s)) * F(m->vn, ji, jj - 1);
            double vv_n = (0.5 + copysign(0.5, v_n)) * F(m->vn, ji, jj) +
                          (0.5 - copysign(0.5, v_n)) * F(m->vn, ji, jj + 1);
            int wall_w = PT(m, ji - 1, jj) <= 0 || PT(m, ji - 1, jj + 1) <= 0;
            int wall_e = PT(m, ji + 1, jj) <= 0 || PT(m, ji + 1, jj + 1) <= 0;
            double vv_w, vv_e;
            if (wall_w)
                vv_w = (0.5 - copysign(0.5, u_w)) * F(m->vn, ji, jj);
            else
                vv_w = (0.5 - copysign(0.5, u_w)) * F(m->vn, ji, jj) +
                       (0.5 + copysign(0.5, u_w)) * F(m->vn, ji - 1, jj);
            if (wall_e)
                vv_e = (0.5 + copysign(0.5, u_e)) * F(m->vn, ji, jj);
            else
                vv_e = (0.5 + copysign(0.5, u_e)) * F(m->vn, ji, jj) +
                       (0.5 - copysign(0.5, u_e)) * F(m->vn, ji + 1, jj);
            double adv = vv_w * u_w * depw - vv_e * u_e * depe + vv_s * v_s * deps - vv_n * v_n * depn;

            double dvdy_n = (F(m->vn, ji, jj + 1) - F(m->vn, ji, jj)) / F(m->e2t, ji, jj + 1) *
                            (F(m->ht, ji, jj + 1) + F(m->sshn, ji, jj + 1));
            double dvdy_s = (F(m->vn, ji, jj) - F(m->vn, ji, jj - 1)) / F(m->e2t, ji, jj) *
                            (F(m->ht, ji, jj) + F(m->sshn, ji, jj));
            double dvdx_w, dvdx_e;
            if (wall_w)
                dvdx_w = 0.0;
            else
                dvdx_w = (F(m->vn, ji, jj) - F(m->vn, ji - 1, jj)) /
                         (F(m->e1v, ji, jj) + F(m->e1v, ji - 1, jj)) *
                         (F(m->hv, ji, jj) + F(m->sshn_v, ji, jj) + F(m->hv, ji - 1, jj) + F(m->sshn_v, ji - 1, jj));
            if (wall_e)
                dvdx_e = 0.0;
            else
                dvdx_e = (F(m->vn, ji + 1, jj) - F(m->vn, ji, jj)) /
                         (F(m->e1v, ji, jj) + F(m->e1v, ji + 1, jj)) *
                         (F(m->hv, ji, jj) + F(m->sshn_v, ji, jj) + F(m->hv, ji + 1, jj) + F(m->sshn_v, ji + 1, jj));
            double vis = (dvdy_n - dvdy_s) * F(m->e1v, ji, jj) + (dvdx_e - dvdx_w) * F(m->e2v, ji, jj) * 0.5;
            vis = visc * vis;

            const double sphi = m->sin_v ? m->sin_v[(ji - m->gphiv.lb1) + (long)(jj - m->gphiv.lb2) * m->jpi]
                                         : sin(F(m->gphiv, ji, jj) * D2R);
            double cor = -0.5 * (2.0 * OMEGA * sphi * (u_ec + u_wc)) *
                         F(m->e12v, ji, jj) * (F(m->hv, ji, jj) + F(m->sshn_v, ji, jj));
            double hpg = -GRAV * (F(m->hv, ji, jj) + F(m->sshn_v, ji, jj)) * F(m->e1v, ji, jj) *
                         (F(m->sshn, ji, jj + 1) - F(m->sshn, ji, jj));
            F(m->va, ji, jj) = (F(m->vn, ji, jj) * (F(m->hv, ji, jj) + F(m->sshn_v, ji, jj)) +
                                rdt * (adv + vis + cor + hpg) / F(m->e12v, ji, jj)) /
                               (F(m->hv, ji, jj) + F(m->ssha_v, ji, jj)) / (1.0 + cbfr * rdt);
        }
}

/* Tide value of the clamped-ssh open boundary, original/nemolite2d.f90:629-635. */
double o2d_tide(double rtime) {
    const double amp_tide = 0.2;
    const double omega_tide = 2.0 * 3.14159 / (12.42 * 3600.0);
    return amp_tide * sin(omega_tide * rtime);
}

/* Tide amplitude (m) and period (hours); the defaults are the serial
 * model's.  Other values are for ensemble members (no reference
 * counterpart). */
void o2d_set_tide(o2d *m, double amp, double period_hours) {
    m->tide_amp = amp;
    m->tide_omega = 2.0 * 3.14159 / (period_hours * 3600.0);
}

/* bc_ssh, original/nemolite2d.f90:628-645. */
void o2d_bc_ssh(o2d *m, double rtime) {
    const int jpi = m->jpi, jpj = m->jpj;
    const double tide = m->tide_amp * sin(m->tide_omega * rtime);
    for (int jj = 1; jj <= jpj; ++jj)
        for (int ji = 1; ji <= jpi; ++ji) {
            if (PT(m, ji, jj) <= 0) continue;
            if (PT(m, ji, jj - 1) < 0 || PT(m, ji, jj + 1) < 0 || PT(m, ji + 1, jj) < 0 ||
                PT(m, ji - 1, jj) < 0)
                F(m->ssha, ji, jj) = tide;
        }
}

/* bc_solid_u, original/nemolite2d.f90:648-652. */
void o2d_bc_solid_u(o2d *m) {
    for (int jj = 1; jj <= m->jpj; ++jj)
        for (int ji = 0; ji <= m->jpi; ++ji)
            if (PT(m, ji, jj) * PT(m, ji + 1, jj) == 0) F(m->ua, ji, jj) = 0.0;
}

/* bc_solid_v, original/nemolite2d.f90:656-660. */
void o2d_bc_solid_v(o2d *m) {
    for (int jj = 0; jj <= m->jpj; ++jj)
        for (int ji = 1; ji <= m->jpi; ++ji)
            if (PT(m, ji, jj) * PT(m, ji, jj + 1) == 0) F(m->va, ji, jj) = 0.0;
}

/* bc_flather_u, original/nemolite2d.f90:669-680 (sequential in ji). */
void o2d_bc_flather_u(o2d *m) {
    for (int jj = 1; jj <= m->jpj; ++jj)
        for (int ji = 0; ji <= m->jpi; ++ji) {
            if (PT(m, ji, jj) + PT(m, ji + 1, jj) <= -1) continue;
            int jiu;
            if (PT(m, ji, jj) < 0) jiu = ji + 1;
            else if (PT(m, ji + 1, jj) < 0) jiu = ji - 1;
            else continue;
            F(m->ua, ji, jj) = F(m->ua, jiu, jj) +
                               sqrt(GRAV / F(m->hu, ji, jj)) * (F(m->sshn_u, ji, jj) - F(m->sshn_u, jiu, jj));
        }
}

/* bc_flather_v, original/nemolite2d.f90:684-695 (sequential in jj). */
void o2d_bc_flather_v(o2d *m) {
    for (int jj = 0; jj <= m->jpj; ++jj)
        for (int ji = 1; ji <= m->jpi; ++ji) {
            if (PT(m, ji, jj) + PT(m, ji, jj + 1) <= -1) continue;
            int jjv;
            if (PT(m, ji, jj) < 0) jjv = jj + 1;
            else if (PT(m, ji, jj + 1) < 0) jjv = jj - 1;
            else continue;
            F(m->va, ji, jj) = F(m->va, ji, jjv) +
                               sqrt(GRAV / F(m->hv, ji, jj)) * (F(m->sshn_v, ji, jj) - F(m->sshn_v, ji, jjv));
        }
}

/* next, original/nemolite2d.f90:704-769. */
void o2d_next(o2d *m) {
    const int jpi = m->jpi, jpj = m->jpj;
    for (int jj = 1; jj <= jpj; ++jj)
        for (int ji = 0; ji <= jpi; ++ji) F(m->un, ji, jj) = F(m->ua, ji, jj);
    for (int jj = 0; jj <= jpj; ++jj)
        for (int ji = 1; ji <= jpi; ++ji) F(m->vn, ji, jj) = F(m->va, ji, jj);
    for (int jj = 1; jj <= jpj; ++jj)
        for (int ji = 1; ji <= jpi; ++ji) F(m->sshn, ji, jj) = F(m->ssha, ji, jj);
    for (int jj = 1; jj <= jpj; ++jj)
        for (int ji = 0; ji <= jpi; ++ji) {
            int a = PT(m, ji, jj), b = PT(m, ji + 1, jj);
            if (a + b <= 0) continue;
            if (a * b > 0) {
                double r = F(m->e12t, ji, jj) * F(m->sshn, ji, jj) + F(m->e12t, ji + 1, jj) * F(m->sshn, ji + 1, jj);
                F(m->sshn_u, ji, jj) = 0.5 * r / F(m->e12u, ji, jj);
            } else if (a <= 0) {
                F(m->sshn_u, ji, jj) = F(m->sshn, ji + 1, jj);
            } else if (b <= 0) {
                F(m->sshn_u, ji, jj) = F(m->sshn, ji, jj);
            }
        }
    for (int jj = 0; jj <= jpj; ++jj)
        for (int ji = 1; ji <= jpi; ++ji) {
            int a = PT(m, ji, jj), b = PT(m, ji, jj + 1);
            if (a + b <= 0) continue;
            if (a * b > 0) {
                double r = F(m->e12t, ji, jj) * F(m->sshn, ji, jj) + F(m->e12t, ji, jj + 1) * F(m->sshn, ji, jj + 1);
                F(m->sshn_v, ji, jj) = 0.5 * r / F(m->e12v, ji, jj);
            } else if (a <= 0) {
                F(m->sshn_v, ji, jj) = F(m->sshn, ji, jj + 1);
            } else if (b <= 0) {
                F(m->sshn_v, ji, jj) = F(m->sshn, ji, jj);
            }
        }
}

/* bc, original/nemolite2d.f90:617-700, in its order. */
void o2d_bc(o2d *m, double rtime) {
    o2d_bc_ssh(m, rtime);
    o2d_bc_solid_u(m);
    o2d_bc_solid_v(m);
    o2d_bc_flather_u(m);
    o2d_bc_flather_v(m);
}

/* step, original/nemolite2d.f90:375-386 (without the output call). */
void o2d_step(o2d *m, int istp) {
    double rtime = (double)istp * m->prm.rdt;
    o2d_continuity(m);
    o2d_momentum_u(m);
    o2d_momentum_v(m);
    o2d_bc(m, rtime);
    o2d_next(m);
}

/* Neumaier-compensated sum of |a(1:jpi,1:jpj)|, columns in Fortran order.
 * The reference's checksum is SUM(ABS(section)) (field_mod.f90:935-956);
 * compensation makes the result independent of summation order to about
 * one ulp. */
static double abs_sum(const farr *a, int jpi, int jpj) {
    double s = 0.0, c = 0.0;
    for (int jj = 1; jj <= jpj; ++jj)
        for (int ji = 1; ji <= jpi; ++ji) {
            double x = fabs(F(*a, ji, jj));
            double t = s + x;
            if (fabs(s) >= x) c += (s - t) + x;
            else c += (x - t) + s;
            s = t;
        }
    return s + c;
}

void o2d_checksums(o2d *m, double out[3]) {
    out[0] = abs_sum(&m->sshn, m->jpi, m->jpj);
    out[1] = abs_sum(&m->un, m->jpi, m->jpj);
    out[2] = abs_sum(&m->vn, m->jpi, m->jpj);
}

/* Run nsteps from nit000, writing the per-step checksums (3 per step) to
 * sums (may be NULL).  Returns the wall time of the step loop in seconds. */
#include <time.h>
double o2d_run(o2d *m, int nit000, int nitend, double *sums) {
    struct timespec t0, t1;
    /* the same sin values the per-point evaluation gives, computed once */
    const long nu = (long)(m->jpi + 1) * m->jpj, nv = (long)m->jpi * (m->jpj + 1);
    m->sin_u = (double *)malloc(nu * sizeof(double));
    m->sin_v = (double *)malloc(nv * sizeof(double));
    if (m->sin_u && m->sin_v) {
        for (long k = 0; k < nu; ++k) m->sin_u[k] = sin(m->gphiu.p[k] * D2R);
        for (long k = 0; k < nv; ++k) m->sin_v[k] = sin(m->gphiv.p[k] * D2R);
    } else {
        free(m->sin_u); free(m->sin_v);
        m->sin_u = m->sin_v = NULL;
    }
    clock_gettime(CLOCK_MONOTONIC, &t0);
    for (int istp = nit000; istp <= nitend; ++istp) {
        o2d_step(m, istp);
        if (sums) o2d_checksums(m, sums + 3 * (long)(istp - nit000));
    }
    clock_gettime(CLOCK_MONOTONIC, &t1);
    free(m->sin_u); free(m->sin_v);
    m->sin_u = m->sin_v = NULL;
    return (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
}
