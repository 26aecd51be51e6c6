/*
 * mnad_oracle.h -- FP64 CPU restatement of the MNAD-MRTM hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product library (libmnad) never links or calls it.
 *
 * It restates, in stencil form and double precision, the MATLAB reference:
 *   SRC/MNAD/{Auu,Auv,Auw,Avu,Avv,Avw,Awu,Awv,Aww}_gen.m  (operator entries)
 *   SRC/MNAD/Ux123_gen.m, SRC/MNAD/Uz123_gen.m              (CPML memories)
 *   SRC/MNAD/C_cof.m                                        (CPML coefficients)
 *   SRC/MNAD/fwd_PDE_exact_MNAD.m                           (forward shot)
 *   SRC/migration/Born_adj_MNAD.m                           (RTM adjoint + imaging)
 *
 * Layout: every grid array is row-major [nz][nx] (x fastest), rows = z.
 * Indices are 0-based.  Interior arrays ([nzi][nxi], nzi = nz - 2 npml,
 * nxi = nx - 2 npml) are row-major too.  Gathers are [nt][nr] (receiver
 * fastest).  Snapshot matrices are [K][nzi*nxi] (snapshot-major).
 *
 * Parity pins (see tests/): golden fixtures from an independent sparse-matrix
 * transliteration (tests/golden/make_golden.py), closed-form plane-wave
 * consistency, static linear-field invariants, Auu row sums, CPML closed forms,
 * mirror symmetry, linearity and time reversibility.
 */
#ifndef MNAD_ORACLE_H
#define MNAD_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int nz, nx;      /* padded grid size (rows = z, columns = x) */
    int npml;        /* CPML thickness in cells (d_pml) */
    double dx, dz, dt;
} mo_grid;

/* Snapshot schedule of fwd_PDE_exact_MNAD.m:46-47,53 and the column count
 * Born_adj_MNAD.m:128 indexes (MATLAB grows UU when K > 2 nf + 1).
 * Returns 0, or -1 when nf == 0 (T would be Inf: the reference images NaN). */
int mo_snap_schedule(int nt, double vmin, double vmax,
                     int *nf, int *T, int *K, int *ncol);

/* C_cof.m restated for one side profile: c1, c2, c3 at every grid point.
 * side: 0 = x-left, 1 = x-right, 2 = z-top, 3 = z-bottom.  Arrays [nz][nx]. */
void mo_cpml_coeffs(const mo_grid *g, const double *v, double R, int side,
                    double *c1, double *c2, double *c3);

/* The 9 operators applied to (U,V,W) (fwd_PDE_exact_MNAD.m:114-116 without
 * the -p and -a*psi terms): out_u = Auu U + Auv V + Auw W, etc.
 * a_form: 0 -> a = v.^2*dt^2 (forward, :39), 1 -> (v*dt)^2 (adjoint, :36);
 * operator entries always use v(j,i).^2*dt^2 and v(j,i).^4*dt^4 as *_gen.m. */
void mo_apply_A(const mo_grid *g, const double *v,
                const double *U, const double *V, const double *W,
                double *oU, double *oV, double *oW);

/* Forward shot (fwd_PDE_exact_MNAD.m).
 * zs, xs: 0-based padded source indices; zr, xr: 0-based receivers.
 * gather: [nt][nr] (may be NULL).  U,V,W: final fields [nz][nx] (may be NULL).
 * snap: 0/1 (indx).  When 1, UU gets [K][nzi*nxi] d2u snapshots and illum
 * gets T*sum(d2u^2) [nzi*nxi] (fwd_PDE_exact_MNAD.m:126-143,162).
 * Returns 0, or a negative code on invalid input. */
int mo_forward(const mo_grid *g, const double *v, int nt, const double *w,
               int zs, int xs, int nr, const int *zr, const int *xr,
               int snap, double *gather, double *U, double *V, double *W,
               double *UU, double *illum);

/* RTM adjoint (Born_adj_MNAD.m) with the reading U_r := U.
 * data: [nt][nr] residual.  UU: [K][nzi*nxi] as written by mo_forward on the
 * same v.  Im: [nzi*nxi], already multiplied by T (Born_adj_MNAD.m:138). */
int mo_adjoint(const mo_grid *g, const double *v, int nt,
               int nr, const int *zr, const int *xr, const double *data,
               const double *UU, double *Im);

/* Same as mo_forward but starting from a given state and running nsteps of
 * the bare update (no source, no receivers): used by the time-reversibility
 * and invariant tests.  p*: previous level, c*: current level; on return
 * c* hold the newest level and p* the one before. */
int mo_steps(const mo_grid *g, const double *v, int nsteps,
             double *pU, double *pV, double *pW,
             double *cU, double *cV, double *cW);

#ifdef __cplusplus
}
#endif
#endif
