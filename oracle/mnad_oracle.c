/*
 * mnad_oracle.c -- FP64 CPU restatement of the MNAD-MRTM hot path.
 * TEST INFRASTRUCTURE ONLY (the checker, never the product); see mnad_oracle.h.
 *
 * Parity pinned by tests/test_oracle_*.py: golden fixtures of an independent
 * sparse-matrix transliteration, plane-wave consistency, linear-field and
 * row-sum invariants, CPML closed forms, symmetry, linearity, reversibility.
 */
#include "mnad_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define IDX(g, j, i) ((size_t)(j) * (size_t)(g)->nx + (size_t)(i))

/* Value at (j,i), zero outside the padded grid: the *_gen.m index ranges
 * (e.g. j = 2:nz in Auu_gen.m:26) simply omit neighbours beyond the grid. */
static inline double at(const mo_grid *g, const double *F, int j, int i)
{
    if (j < 0 || j >= g->nz || i < 0 || i >= g->nx) return 0.0;
    return F[IDX(g, j, i)];
}

int mo_snap_schedule(int nt, double vmin, double vmax,
                     int *nf, int *T, int *K, int *ncol)
{
    /* fwd_PDE_exact_MNAD.m:46-47: nf = floor(nt/(2*3*max(v)/min(v))),
     * T = floor(nt/nf/2); :53 UU has 2*nf+1 columns, grown if needed. */
    int nf_ = (int)floor((double)nt / (2.0 * 3.0 * vmax / vmin));
    if (nf_ <= 0) return -1;
    int T_ = (int)floor((double)nt / (double)nf_ / 2.0);
    if (T_ <= 0) return -1;
    int K_ = nt / T_;
    int nc = 2 * nf_ + 1;
    if (K_ > nc) nc = K_;
    if (nf) *nf = nf_;
    if (T) *T = T_;
    if (K) *K = K_;
    if (ncol) *ncol = nc;
    return 0;
}

void mo_cpml_coeffs(const mo_grid *g, const double *v, double R, int side,
                    double *c1, double *c2, double *c3)
{
    /* C_cof.m:4-65.  Profile coordinate x_b = [D:-h:h, 0..., h:h:D]
     * (C_cof.m:6,38); sigma = 3 log(1/R)/(2 D^3) * v * x^2 on the side's
     * strip only (V_xl(:,d_pml+1:nx) = 0 etc.). */
    const int npml = g->npml;
    const int nz = g->nz, nx = g->nx;
    const double h = (side < 2) ? g->dx : g->dz;
    const double D = npml * h;
    const double k = (3.0 * log(1.0 / R)) / (2.0 * D * D * D);
    const double dt = g->dt;
    for (int j = 0; j < nz; ++j) {
        for (int i = 0; i < nx; ++i) {
            double x = 0.0;
            int on = 0;
            switch (side) {
            case 0: if (i < npml) { on = 1; x = (npml - i) * h; } break;
            case 1: if (i >= nx - npml) { on = 1; x = (i - (nx - npml) + 1) * h; } break;
            case 2: if (j < npml) { on = 1; x = (npml - j) * h; } break;
            case 3: if (j >= nz - npml) { on = 1; x = (j - (nz - npml) + 1) * h; } break;
            }
            double sig0 = on ? k * v[IDX(g, j, i)] : 0.0;
            double sig = sig0 * (x * x);
            double dsig = 2.0 * sig0 * x;
            double den = 1.0 + sig * (dt / 2.0);
            c1[IDX(g, j, i)] = (1.0 - sig * (dt / 2.0)) / den;
            c2[IDX(g, j, i)] = (dsig * dt) / den;
            c3[IDX(g, j, i)] = (sig * dt) / den;
        }
    }
}

/* Literal operator entries at one output point (*_gen.m), 30 values:
 * they depend only on v at the output point, so they are computed once. */
enum {
    C_UU_Z, C_UU_X, C_UU_C, C_UU_D,               /* Auu_gen.m:30,37,44,65 */
    C_UV_1, C_UV_2, C_UV_E, C_UV_7, C_UV_8,       /* Auv_gen.m:121-170 */
    C_UW_1, C_UW_2, C_UW_E, C_UW_7, C_UW_8,       /* Auw_gen.m:30-77 */
    C_VV_Z, C_VV_X, C_VV_C, C_VV_D,               /* Avv_gen.m:30-84 */
    C_VU_M, C_VU_P, C_VU_R,                       /* Avu_gen.m:30-64 */
    C_VW_S,                                       /* Avw_gen.m:30-48 */
    C_WW_Z, C_WW_X, C_WW_C, C_WW_D,               /* Aww_gen.m:32-86 */
    C_WU_M, C_WU_P, C_WU_R,                       /* Awu_gen.m:30-66 */
    C_WV_S,                                       /* Awv_gen.m:32-50 */
    NCOEF
};

static void coef_point(const mo_grid *g, double vv, double *c)
{
    const double dx = g->dx, dz = g->dz, dt = g->dt;
    const double A = vv * vv * (dt * dt);            /* v(j,i).^2*dt^2 */
    const double B = pow(vv, 4) * pow(dt, 4);        /* v(j,i).^4*dt^4 */
    const double dx2 = dx * dx, dz2 = dz * dz;

    /* Auu: (+-1,0), (0,+-1), (0,0), corners */
    c[C_UU_Z] = (24.0 / 11.0) * A / dz2 - B * (15.0 / (11.0 * dz2 * dz2) + 1.0 / (dx2 * dz2));
    c[C_UU_X] = (24.0 / 11.0) * A / dx2 - B * (15.0 / (11.0 * dx2 * dx2) + 1.0 / (dx2 * dz2));
    c[C_UU_C] = 2.0 - (48.0 / 11.0) * A * (1.0 / dx2 + 1.0 / dz2)
              + (30.0 / 11.0) * B * (1.0 / (dx2 * dx2) + 1.0 / (dz2 * dz2) + 11.0 / (15.0 * dx2 * dz2));
    c[C_UU_D] = B / (2.0 * dx2 * dz2);
    /* Auv: V at (0,-1) v1, (0,+1) v2, (-1,-1) -e, (+1,-1) -e, (-1,+1) +e,
     * (+1,+1) +e, (0,-2) v7, (0,+2) v8 */
    {
        double q = B * (47.0 / (66.0 * dx2 * dx) + 1.0 / (6.0 * dx * dz2));
        c[C_UV_1] = (-20.0 / 33.0) * A / dx + q;
        c[C_UV_2] = (20.0 / 33.0) * A / dx - q;
        c[C_UV_E] = B / (12.0 * dx * dz2);
        c[C_UV_7] = (1.0 / 132.0) * A / dx - (1.0 / 66.0) * B / (dx2 * dx);
        c[C_UV_8] = (-1.0 / 132.0) * A / dx + (1.0 / 66.0) * B / (dx2 * dx);
    }
    /* Auw: W at (+1,0) v1, (-1,0) v2, (-1,-1) -e, (+1,-1) +e, (-1,+1) -e,
     * (+1,+1) +e, (+2,0) v7, (-2,0) v8 */
    {
        double q = B * (47.0 / (66.0 * dz2 * dz) + 1.0 / (6.0 * dz * dx2));
        c[C_UW_1] = (20.0 / 33.0) * A / dz - q;
        c[C_UW_2] = (-20.0 / 33.0) * A / dz + q;
        c[C_UW_E] = B / (12.0 * dz * dx2);
        c[C_UW_7] = (-1.0 / 132.0) * A / dz + (1.0 / 66.0) * B / (dz2 * dz);
        c[C_UW_8] = (1.0 / 132.0) * A / dz - (1.0 / 66.0) * B / (dz2 * dz);
    }
    /* Avv */
    c[C_VV_Z] = A / dz2 - 2.0 * B / (dx2 * dz2);
    c[C_VV_X] = -(3.0 / 2.0) * A / dx2 + B * (5.0 / (2.0 * dx2 * dx2) + 1.0 / (2.0 * dx2 * dz2));
    c[C_VV_C] = 2.0 - A * (12.0 / dx2 + 2.0 / dz2) + B * (10.0 / (dx2 * dx2) + 4.0 / (dx2 * dz2));
    c[C_VV_D] = -B / (4.0 * dx2 * dz2);
    /* Avu: U at (0,-1) -P+Q, (0,+1) P-Q, (+-1,-1) +R, (+-1,+1) -R */
    {
        double P = B * (15.0 / (2.0 * pow(dx, 5)) - 1.0 / (dx * dz2 * dz2) + 5.0 / (2.0 * dx2 * dx * dz2));
        double Q = A * (15.0 / (2.0 * dx2 * dx) - 1.0 / (dx * dz2));
        c[C_VU_M] = -P + Q;
        c[C_VU_P] = P - Q;
        c[C_VU_R] = B * (5.0 / (4.0 * dx2 * dx * dz2) - 1.0 / (2.0 * dx * dz2 * dz2)) + A / (2.0 * dx * dz2);
    }
    /* Avw: W at (-1,-1) S, (+1,-1) -S, (-1,+1) -S, (+1,+1) S */
    c[C_VW_S] = B / (4.0 * dz2 * dz * dx) - A / (4.0 * dz * dx);
    /* Aww */
    c[C_WW_Z] = -(3.0 / 2.0) * A / dz2 + B * (5.0 / (2.0 * dz2 * dz2) + 1.0 / (2.0 * dx2 * dz2));
    c[C_WW_X] = A / dx2 - 2.0 * B / (dz2 * dx2);
    c[C_WW_C] = 2.0 - A * (12.0 / dz2 + 2.0 / dx2) + B * (10.0 / (dz2 * dz2) + 4.0 / (dx2 * dz2));
    c[C_WW_D] = -B / (4.0 * dx2 * dz2);
    /* Awu: U at (-1,0) -P+Q, (+1,0) P-Q, (-1,+-1) +R, (+1,+-1) -R */
    {
        double P = B * (15.0 / (2.0 * pow(dz, 5)) - 1.0 / (dz * dx2 * dx2) + 5.0 / (2.0 * dz2 * dz * dx2));
        double Q = A * (15.0 / (2.0 * dz2 * dz) - 1.0 / (dz * dx2));
        c[C_WU_M] = -P + Q;
        c[C_WU_P] = P - Q;
        c[C_WU_R] = B * (5.0 / (4.0 * dz2 * dz * dx2) - 1.0 / (2.0 * dz * dx2 * dx2)) + A * (1.0 / (2.0 * dz * dx2));
    }
    /* Awv: V at (-1,-1) S, (+1,-1) -S, (-1,+1) -S, (+1,+1) S */
    c[C_WV_S] = B / (4.0 * dx2 * dx * dz) - A / (4.0 * dz * dx);
}

static double *coef_table(const mo_grid *g, const double *v)
{
    size_t n = (size_t)g->nz * g->nx;
    double *c = malloc(n * NCOEF * sizeof(double));
    if (!c) return NULL;
    for (size_t p = 0; p < n; ++p) coef_point(g, v[p], c + p * NCOEF);
    return c;
}

/* out = [Auu Auv Auw; Avu Avv Avw; Awu Awv Aww] (U;V;W) at one point. */
static void apply_point(const mo_grid *g, const double *c,
                        const double *U, const double *V, const double *W,
                        int j, int i, double *ou, double *ov, double *ow)
{
#define Uu(oz, ox) at(g, U, j + (oz), i + (ox))
#define Vv(oz, ox) at(g, V, j + (oz), i + (ox))
#define Ww(oz, ox) at(g, W, j + (oz), i + (ox))
    double nu = c[C_UU_C] * Uu(0, 0)
              + c[C_UU_Z] * (Uu(-1, 0) + Uu(1, 0))
              + c[C_UU_X] * (Uu(0, -1) + Uu(0, 1))
              + c[C_UU_D] * (Uu(-1, -1) + Uu(1, -1) + Uu(-1, 1) + Uu(1, 1));
    nu += c[C_UV_1] * Vv(0, -1) + c[C_UV_2] * Vv(0, 1)
        - c[C_UV_E] * Vv(-1, -1) - c[C_UV_E] * Vv(1, -1)
        + c[C_UV_E] * Vv(-1, 1) + c[C_UV_E] * Vv(1, 1)
        + c[C_UV_7] * Vv(0, -2) + c[C_UV_8] * Vv(0, 2);
    nu += c[C_UW_1] * Ww(1, 0) + c[C_UW_2] * Ww(-1, 0)
        - c[C_UW_E] * Ww(-1, -1) + c[C_UW_E] * Ww(1, -1)
        - c[C_UW_E] * Ww(-1, 1) + c[C_UW_E] * Ww(1, 1)
        + c[C_UW_7] * Ww(2, 0) + c[C_UW_8] * Ww(-2, 0);

    double nv = c[C_VV_C] * Vv(0, 0)
              + c[C_VV_Z] * (Vv(-1, 0) + Vv(1, 0))
              + c[C_VV_X] * (Vv(0, -1) + Vv(0, 1))
              + c[C_VV_D] * (Vv(-1, -1) + Vv(1, -1) + Vv(-1, 1) + Vv(1, 1));
    nv += c[C_VU_M] * Uu(0, -1) + c[C_VU_P] * Uu(0, 1)
        + c[C_VU_R] * (Uu(-1, -1) + Uu(1, -1)) - c[C_VU_R] * (Uu(-1, 1) + Uu(1, 1));
    nv += c[C_VW_S] * (Ww(-1, -1) - Ww(1, -1) - Ww(-1, 1) + Ww(1, 1));

    double nw = c[C_WW_C] * Ww(0, 0)
              + c[C_WW_Z] * (Ww(-1, 0) + Ww(1, 0))
              + c[C_WW_X] * (Ww(0, -1) + Ww(0, 1))
              + c[C_WW_D] * (Ww(-1, -1) + Ww(1, -1) + Ww(-1, 1) + Ww(1, 1));
    nw += c[C_WU_M] * Uu(-1, 0) + c[C_WU_P] * Uu(1, 0)
        + c[C_WU_R] * (Uu(-1, -1) + Uu(-1, 1)) - c[C_WU_R] * (Uu(1, -1) + Uu(1, 1));
    nw += c[C_WV_S] * (Vv(-1, -1) - Vv(1, -1) - Vv(-1, 1) + Vv(1, 1));
#undef Uu
#undef Vv
#undef Ww
    *ou = nu;
    *ov = nv;
    *ow = nw;
}

static void apply_all(const mo_grid *g, const double *ctab,
                      const double *U, const double *V, const double *W,
                      double *oU, double *oV, double *oW)
{
    for (int j = 0; j < g->nz; ++j)
        for (int i = 0; i < g->nx; ++i) {
            size_t p = IDX(g, j, i);
            apply_point(g, ctab + p * NCOEF, U, V, W, j, i, &oU[p], &oV[p], &oW[p]);
        }
}

void mo_apply_A(const mo_grid *g, const double *v,
                const double *U, const double *V, const double *W,
                double *oU, double *oV, double *oW)
{
    double *ctab = coef_table(g, v);
    if (!ctab) return;
    apply_all(g, ctab, U, V, W, oU, oV, oW);
    free(ctab);
}

/* ------------------------------------------------------------------ CPML */

typedef struct {
    /* psi[f][lev]: field f in {U,V,W}, level 1..3 (index 0..2), per direction.
     * Left/right (top/bottom) memories never overlap (nx >= 2 npml), so one
     * array per (field, level, direction) holds both sides. */
    double *px[3][3];
    double *pz[3][3];
    double *c1x, *c2x, *c3x;   /* left strip + right strip merged */
    double *c1z, *c2z, *c3z;
    double *sx;                /* +1 on the left strip, -1 on the right */
    double *sz;                /* +1 on the top strip, -1 on the bottom */
} cpml_state;

static int cpml_init(cpml_state *c, const mo_grid *g, const double *v)
{
    size_t n = (size_t)g->nz * g->nx;
    memset(c, 0, sizeof(*c));
    for (int f = 0; f < 3; ++f)
        for (int l = 0; l < 3; ++l) {
            c->px[f][l] = calloc(n, sizeof(double));
            c->pz[f][l] = calloc(n, sizeof(double));
            if (!c->px[f][l] || !c->pz[f][l]) return -1;
        }
    double *t1 = malloc(n * sizeof(double)), *t2 = malloc(n * sizeof(double)),
           *t3 = malloc(n * sizeof(double));
    c->c1x = malloc(n * sizeof(double)); c->c2x = malloc(n * sizeof(double));
    c->c3x = malloc(n * sizeof(double)); c->c1z = malloc(n * sizeof(double));
    c->c2z = malloc(n * sizeof(double)); c->c3z = malloc(n * sizeof(double));
    c->sx = malloc(n * sizeof(double)); c->sz = malloc(n * sizeof(double));
    if (!t1 || !t2 || !t3 || !c->c1x || !c->c2x || !c->c3x || !c->c1z ||
        !c->c2z || !c->c3z || !c->sx || !c->sz) return -1;
    const double R = 1e-4;  /* fwd_PDE_exact_MNAD.m:88, Born_adj_MNAD.m:81 */
    /* x: left then right */
    mo_cpml_coeffs(g, v, R, 0, c->c1x, c->c2x, c->c3x);
    mo_cpml_coeffs(g, v, R, 1, t1, t2, t3);
    for (int j = 0; j < g->nz; ++j)
        for (int i = 0; i < g->nx; ++i) {
            size_t p = IDX(g, j, i);
            c->sx[p] = 0.0;
            if (i < g->npml) c->sx[p] = 1.0;
            if (i >= g->nx - g->npml) {
                c->sx[p] = -1.0;
                c->c1x[p] = t1[p]; c->c2x[p] = t2[p]; c->c3x[p] = t3[p];
            }
        }
    mo_cpml_coeffs(g, v, R, 2, c->c1z, c->c2z, c->c3z);
    mo_cpml_coeffs(g, v, R, 3, t1, t2, t3);
    for (int j = 0; j < g->nz; ++j)
        for (int i = 0; i < g->nx; ++i) {
            size_t p = IDX(g, j, i);
            c->sz[p] = 0.0;
            if (j < g->npml) c->sz[p] = 1.0;
            if (j >= g->nz - g->npml) {
                c->sz[p] = -1.0;
                c->c1z[p] = t1[p]; c->c2z[p] = t2[p]; c->c3z[p] = t3[p];
            }
        }
    free(t1); free(t2); free(t3);
    return 0;
}

static void cpml_free(cpml_state *c)
{
    for (int f = 0; f < 3; ++f)
        for (int l = 0; l < 3; ++l) { free(c->px[f][l]); free(c->pz[f][l]); }
    free(c->c1x); free(c->c2x); free(c->c3x);
    free(c->c1z); free(c->c2z); free(c->c3z);
    free(c->sx); free(c->sz);
}

/* One CPML memory triple at point p (Ux123_gen.m:9-15 / Uz123_gen.m:28-34):
 * D = F(-1) - F(+1) on the left/top side and its negative on the right/bottom
 * side (sgn), L = F(-1) - 2F + F(+1). Returns the new psi1. */
static inline double psi_point(double *p1, double *p2, double *p3, size_t p,
                               double sgn, double Fd, double F0, double Fa,
                               double c1, double c2, double c3, double h)
{
    double D = sgn * (Fd - Fa);
    double L = Fd - 2.0 * F0 + Fa;
    p3[p] = c1 * p3[p] + (c2 / (2.0 * h)) * D;
    p2[p] = c1 * p2[p] + (c2 / h) * D + c3 * (L / (h * h) - p3[p]);
    p1[p] = c1 * p1[p] + (c2 / (2.0 * h)) * D + c3 * (2.0 * L / (h * h) - p2[p]);
    return p1[p];
}

/* Ux123_gen.m / Uz123_gen.m for one field F; writes F1_x = psi1_l + psi1_r
 * and F1_z = psi1_t + psi1_b.  Off the strips c1 = 1 and c2 = c3 = 0
 * (C_cof.m:48,61,80,93 zero sigma there), so the memories and the outputs
 * stay exactly zero off the strips and only strip points are visited. */
static void cpml_update(const mo_grid *g, cpml_state *c, int f, const double *F,
                        double *out_x, double *out_z)
{
    const int nz = g->nz, nx = g->nx, np = g->npml;
    for (int j = 0; j < nz; ++j) {
        for (int side = 0; side < 2; ++side) {
            int i0 = side ? nx - np : 0, i1 = side ? nx : np;
            for (int i = i0; i < i1; ++i) {
                size_t p = IDX(g, j, i);
                out_x[p] = psi_point(c->px[f][0], c->px[f][1], c->px[f][2], p, c->sx[p],
                                     at(g, F, j, i - 1), F[p], at(g, F, j, i + 1),
                                     c->c1x[p], c->c2x[p], c->c3x[p], g->dx);
            }
        }
    }
    for (int side = 0; side < 2; ++side) {
        int j0 = side ? nz - np : 0, j1 = side ? nz : np;
        for (int j = j0; j < j1; ++j)
            for (int i = 0; i < nx; ++i) {
                size_t p = IDX(g, j, i);
                out_z[p] = psi_point(c->pz[f][0], c->pz[f][1], c->pz[f][2], p, c->sz[p],
                                     at(g, F, j - 1, i), F[p], at(g, F, j + 1, i),
                                     c->c1z[p], c->c2z[p], c->c3z[p], g->dz);
            }
    }
}

/* ---------------------------------------------------------- propagators */

typedef struct {
    double *pU, *pV, *pW, *U, *V, *W, *nU, *nV, *nW;
    double *a;                      /* V2dt2 */
    double *ctab;                   /* operator entries per point */
    double *ux, *uz, *vx, *vz, *wx, *wz;
    cpml_state cp;
} prop_state;

static void prop_free(prop_state *s)
{
    free(s->pU); free(s->pV); free(s->pW);
    free(s->U); free(s->V); free(s->W);
    free(s->nU); free(s->nV); free(s->nW);
    free(s->a);
    free(s->ctab);
    free(s->ux); free(s->uz); free(s->vx); free(s->vz); free(s->wx); free(s->wz);
    cpml_free(&s->cp);
}

static int prop_init(prop_state *s, const mo_grid *g, const double *v, int a_form)
{
    size_t n = (size_t)g->nz * g->nx;
    memset(s, 0, sizeof(*s));
    double **arrs[] = {&s->pU, &s->pV, &s->pW, &s->U, &s->V, &s->W, &s->nU, &s->nV,
                       &s->nW, &s->a, &s->ux, &s->uz, &s->vx, &s->vz, &s->wx, &s->wz};
    for (size_t k = 0; k < sizeof(arrs) / sizeof(arrs[0]); ++k) {
        *arrs[k] = calloc(n, sizeof(double));
        if (!*arrs[k]) return -1;
    }
    for (size_t p = 0; p < n; ++p) {
        if (a_form == 0) s->a[p] = v[p] * v[p] * (g->dt * g->dt);   /* fwd :39 */
        else { double t = v[p] * g->dt; s->a[p] = t * t; }          /* adj :36 */
    }
    s->ctab = coef_table(g, v);
    if (!s->ctab) return -1;
    return cpml_init(&s->cp, g, v);
}

/* The update of fwd_PDE_exact_MNAD.m:114-116 (and Born_adj_MNAD.m:104-106
 * with U_r read as U) from already-updated CPML outputs. */
static void prop_update(const mo_grid *g, prop_state *s)
{
    size_t n = (size_t)g->nz * g->nx;
    apply_all(g, s->ctab, s->U, s->V, s->W, s->nU, s->nV, s->nW);
    for (size_t p = 0; p < n; ++p) {
        s->nU[p] += -s->pU[p] - s->a[p] * (s->ux[p] + s->uz[p]);
        s->nV[p] += -s->pV[p] - s->a[p] * (s->vx[p] + s->vz[p]);
        s->nW[p] += -s->pW[p] - s->a[p] * (s->wx[p] + s->wz[p]);
    }
}

static void prop_rotate(prop_state *s)
{
    double *t;
    t = s->pU; s->pU = s->U; s->U = s->nU; s->nU = t;
    t = s->pV; s->pV = s->V; s->V = s->nV; s->nV = t;
    t = s->pW; s->pW = s->W; s->W = s->nW; s->nW = t;
}

static int check_grid(const mo_grid *g)
{
    if (g->nz < 1 || g->nx < 1 || g->npml < 0) return -1;
    if (g->nx < 2 * g->npml || g->nz < 2 * g->npml) return -1;
    if (!(g->dx > 0) || !(g->dz > 0) || !(g->dt > 0)) return -1;
    return 0;
}

static void minmax(const double *v, size_t n, double *mn, double *mx)
{
    double a = v[0], b = v[0];
    for (size_t p = 1; p < n; ++p) { if (v[p] < a) a = v[p]; if (v[p] > b) b = v[p]; }
    *mn = a; *mx = b;
}

int mo_forward(const mo_grid *g, const double *v, int nt, const double *w,
               int zs, int xs, int nr, const int *zr, const int *xr,
               int snap, double *gather, double *Uo, double *Vo, double *Wo,
               double *UU, double *illum)
{
    if (check_grid(g) || nt < 1 || !w) return -1;
    if (zs < 0 || zs >= g->nz || xs < 0 || xs >= g->nx) return -2;
    for (int r = 0; r < nr; ++r)
        if (zr[r] < 0 || zr[r] >= g->nz || xr[r] < 0 || xr[r] >= g->nx) return -2;
    size_t n = (size_t)g->nz * g->nx;
    const int np = g->npml, nzi = g->nz - 2 * np, nxi = g->nx - 2 * np;
    const size_t ni = (size_t)nzi * nxi;
    int T = 0, K = 0;
    if (snap) {
        double vmin, vmax;
        minmax(v, n, &vmin, &vmax);
        if (mo_snap_schedule(nt, vmin, vmax, NULL, &T, &K, NULL)) return -3;
        if (illum) memset(illum, 0, ni * sizeof(double));
    }
    prop_state s;
    if (prop_init(&s, g, v, 0)) { prop_free(&s); return -4; }
    const double dt2 = g->dt * g->dt;
    int ttt = 0;
    for (int tt = 1; tt <= nt; ++tt) {
        /* :95-108 CPML memories from the current (pre-injection) fields */
        cpml_update(g, &s.cp, 0, s.U, s.ux, s.uz);
        cpml_update(g, &s.cp, 1, s.V, s.vx, s.vz);
        cpml_update(g, &s.cp, 2, s.W, s.wx, s.wz);
        /* :111 source into the current U */
        s.U[IDX(g, zs, xs)] += s.a[IDX(g, zs, xs)] * w[tt - 1];
        /* :114-116 */
        prop_update(g, &s);
        /* :119-123 receivers sample the injected current U */
        if (gather)
            for (int r = 0; r < nr; ++r)
                gather[(size_t)(tt - 1) * nr + r] = s.U[IDX(g, zr[r], xr[r])];
        /* :126-143 */
        if (snap && tt % T == 0) {
            double *col = UU ? UU + (size_t)ttt * ni : NULL;
            for (int j = 0; j < nzi; ++j)
                for (int i = 0; i < nxi; ++i) {
                    size_t p = IDX(g, j + np, i + np);
                    double d2u;
                    if (tt == 1) d2u = s.nU[p] / dt2;
                    else if (tt == 2) d2u = (s.nU[p] - 2.0 * s.U[p]) / dt2;
                    else d2u = (s.nU[p] - 2.0 * s.U[p] + s.pU[p]) / dt2;
                    size_t q = (size_t)j * nxi + i;
                    if (col) col[q] = d2u;
                    if (illum) illum[q] += d2u * d2u;
                }
            ++ttt;
        }
        prop_rotate(&s);   /* :146-153 */
    }
    if (snap && illum)
        for (size_t q = 0; q < ni; ++q) illum[q] *= T;   /* :162 */
    if (Uo) memcpy(Uo, s.U, n * sizeof(double));
    if (Vo) memcpy(Vo, s.V, n * sizeof(double));
    if (Wo) memcpy(Wo, s.W, n * sizeof(double));
    prop_free(&s);
    return 0;
}

int mo_adjoint(const mo_grid *g, const double *v, int nt,
               int nr, const int *zr, const int *xr, const double *data,
               const double *UU, double *Im)
{
    if (check_grid(g) || nt < 1) return -1;
    for (int r = 0; r < nr; ++r)
        if (zr[r] < 0 || zr[r] >= g->nz || xr[r] < 0 || xr[r] >= g->nx) return -2;
    size_t n = (size_t)g->nz * g->nx;
    const int np = g->npml, nzi = g->nz - 2 * np, nxi = g->nx - 2 * np;
    const size_t ni = (size_t)nzi * nxi;
    double vmin, vmax;
    minmax(v, n, &vmin, &vmax);
    int T, K, ncol;
    if (mo_snap_schedule(nt, vmin, vmax, NULL, &T, &K, &ncol)) return -3;  /* :41-42 */
    prop_state s;
    if (prop_init(&s, g, v, 1)) { prop_free(&s); return -4; }
    memset(Im, 0, ni * sizeof(double));
    int ttt = 1;
    for (int tt = nt; tt >= 1; --tt) {
        cpml_update(g, &s.cp, 0, s.U, s.ux, s.uz);     /* :88-101 */
        cpml_update(g, &s.cp, 1, s.V, s.vx, s.vz);
        cpml_update(g, &s.cp, 2, s.W, s.wx, s.wz);
        prop_update(g, &s);                          /* :104-106, U_r := U */
        for (int r = 0; r < nr; ++r) {                  /* :109-111 */
            size_t p = IDX(g, zr[r], xr[r]);
            s.U[p] += s.a[p] * data[(size_t)(tt - 1) * nr + r];
        }
        prop_rotate(&s);                                /* :116-123 */
        if (tt % T == 0) {                              /* :127-134 */
            int col = ncol - ttt + 1;                   /* 1-based UU(:,end-ttt+1) */
            if (col >= 1 && col <= K && UU) {           /* columns > K are zeros */
                const double *us = UU + (size_t)(col - 1) * ni;
                for (int j = 0; j < nzi; ++j)
                    for (int i = 0; i < nxi; ++i) {
                        size_t q = (size_t)j * nxi + i;
                        Im[q] += us[q] * s.U[IDX(g, j + np, i + np)];
                    }
            }
            ++ttt;
        }
    }
    for (size_t q = 0; q < ni; ++q) Im[q] *= T;         /* :138 */
    prop_free(&s);
    return 0;
}

int mo_steps(const mo_grid *g, const double *v, int nsteps,
             double *pU, double *pV, double *pW,
             double *cU, double *cV, double *cW)
{
    if (check_grid(g)) return -1;
    size_t n = (size_t)g->nz * g->nx;
    prop_state s;
    if (prop_init(&s, g, v, 0)) { prop_free(&s); return -4; }
    memcpy(s.pU, pU, n * sizeof(double)); memcpy(s.pV, pV, n * sizeof(double));
    memcpy(s.pW, pW, n * sizeof(double)); memcpy(s.U, cU, n * sizeof(double));
    memcpy(s.V, cV, n * sizeof(double)); memcpy(s.W, cW, n * sizeof(double));
    for (int k = 0; k < nsteps; ++k) {
        cpml_update(g, &s.cp, 0, s.U, s.ux, s.uz);
        cpml_update(g, &s.cp, 1, s.V, s.vx, s.vz);
        cpml_update(g, &s.cp, 2, s.W, s.wx, s.wz);
        prop_update(g, &s);
        prop_rotate(&s);
    }
    memcpy(pU, s.pU, n * sizeof(double)); memcpy(pV, s.pV, n * sizeof(double));
    memcpy(pW, s.pW, n * sizeof(double)); memcpy(cU, s.U, n * sizeof(double));
    memcpy(cV, s.V, n * sizeof(double)); memcpy(cW, s.W, n * sizeof(double));
    prop_free(&s);
    return 0;
}
