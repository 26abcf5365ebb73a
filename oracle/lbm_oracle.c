/*
 * CPU oracle (TEST INFRASTRUCTURE ONLY) -- see lbm_oracle.h.
 *
 * Every routine restates one reference routine; the citation on each names
 * the NREL/marbles file:line it follows.  Floating-point expressions keep the
 * reference's operation order exactly (C and C++ share the usual arithmetic
 * conversions and left-to-right association), and the file is built with
 * -ffp-contract=off so no multiply-add is fused.  AMReX macro expansions are
 * restated textually: AMREX_D_TERM(a,b,c) is `a b c` with no added
 * parentheses, which matters in get_temperature (vector overload) below.
 */
#include "lbm_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* Stencil (Source/Stencil.H:41-167, Constants.H:8-66)                        */
/* ------------------------------------------------------------------------ */

/* Stencil.H:49-79 */
static const int EV[ORC_NQ][3] = {
    {0, 0, 0},
    {1, 0, 0},   {-1, 0, 0},  {0, 1, 0},   {0, -1, 0},  {0, 0, 1},
    {0, 0, -1},  {1, 1, 0},   {-1, -1, 0}, {1, -1, 0},  {-1, 1, 0},
    {1, 0, 1},   {-1, 0, -1}, {1, 0, -1},  {-1, 0, 1},  {0, 1, 1},
    {0, -1, -1}, {0, 1, -1},  {0, -1, 1},  {1, 1, 1},   {-1, -1, -1},
    {1, -1, 1},  {-1, 1, -1}, {1, -1, -1}, {-1, 1, 1},  {1, 1, -1},
    {-1, -1, 1}};

/* Stencil.H:125-134 */
static const int BOUNCE[ORC_NQ] = {0,  2,  1,  4,  3,  6,  5,  8,  7,
                                   10, 9,  12, 11, 14, 13, 16, 15, 18,
                                   17, 20, 19, 22, 21, 24, 23, 26, 25};

/* Stencil.H:136-167: reflection of one velocity component (slip walls). */
static int BOUNCE_DIR[3][ORC_NQ];

static double W[ORC_NQ];
static const double THETA0 = 1.0 / 3.0; /* Stencil.H:42 */

/* Constants.H:40-58 */
static const double AIR_MOLAR_MASS = 28.96;
#define SMALL_NUM (2.220446049250313080847263336181640625e-16 * 1e10)

static int g_stencil_ready = 0;
static int g_threads = 1;

static void stencil_init(void)
{
    if (g_stencil_ready) return;
    /* Stencil.H:81-90 */
    const double WSC1 = 2.0 / 27.0;
    const double WFCC1 = 1.0 / 54.0;
    const double WBCC1 = 1.0 / 216.0;
    const double WO = 1.0 - (6.0 * (WSC1) + 12.0 * (WFCC1) + 8.0 * (WBCC1));
    for (int q = 0; q < ORC_NQ; q++) {
        int s = abs(EV[q][0]) + abs(EV[q][1]) + abs(EV[q][2]);
        W[q] = s == 0 ? WO : s == 1 ? WSC1 : s == 2 ? WFCC1 : WBCC1;
    }
    for (int d = 0; d < 3; d++)
        for (int q = 0; q < ORC_NQ; q++) {
            int r[3] = {EV[q][0], EV[q][1], EV[q][2]};
            r[d] = -r[d];
            for (int p = 0; p < ORC_NQ; p++)
                if (EV[p][0] == r[0] && EV[p][1] == r[1] && EV[p][2] == r[2])
                    BOUNCE_DIR[d][q] = p;
        }
    g_stencil_ready = 1;
}

void orc_stencil(int *evs, int *bounce, double *w)
{
    stencil_init();
    for (int q = 0; q < ORC_NQ; q++) {
        for (int d = 0; d < 3; d++) evs[3 * q + d] = EV[q][d];
        bounce[q] = BOUNCE[q];
        w[q] = W[q];
    }
}

void orc_bounce_dir(int d, int *out)
{
    stencil_init();
    for (int q = 0; q < ORC_NQ; q++) out[q] = BOUNCE_DIR[d][q];
}

void orc_set_threads(int n) { g_threads = n < 1 ? 1 : n; }
int orc_get_threads(void) { return g_threads; }

/* ------------------------------------------------------------------------ */
/* Lattice math (Source/Utilities.H:11-350)                                   */
/* ------------------------------------------------------------------------ */

/* Utilities.H:11-40 (product-form equilibrium, no correction) */
static inline double set_equilibrium_value(
    double rho, const double vel[3], double r_temperature, const int ev[3])
{
    const double pxx = vel[0] * vel[0] + r_temperature;
    const double pyy = vel[1] * vel[1] + r_temperature;
    const double phix =
        ev[0] * 0.5 * vel[0] + abs(ev[0]) * (1.50 * pxx - 1.0) - pxx + 1.0;
    const double phiy =
        ev[1] * 0.5 * vel[1] + abs(ev[1]) * (1.50 * pyy - 1.0) - pyy + 1.0;
    const double pzz = vel[2] * vel[2] + r_temperature;
    const double phiz =
        ev[2] * 0.5 * vel[2] + abs(ev[2]) * (1.50 * pzz - 1.0) - pzz + 1.0;
    return rho * phix * phiy * phiz;
}

/* Utilities.H:42-63 (product-form equilibrium, extended pressure tensor) */
static inline double set_extended_equilibrium_value(
    double rho, const double vel[3], double pxx, double pyy, double pzz,
    const int ev[3])
{
    const double phix =
        ev[0] * 0.5 * vel[0] + abs(ev[0]) * (1.50 * pxx - 1.0) - pxx + 1.0;
    const double phiy =
        ev[1] * 0.5 * vel[1] + abs(ev[1]) * (1.50 * pyy - 1.0) - pyy + 1.0;
    const double phiz =
        ev[2] * 0.5 * vel[2] + abs(ev[2]) * (1.50 * pzz - 1.0) - pzz + 1.0;
    return rho * phix * phiy * phiz;
}

/* Utilities.H:65-163, called everywhere with frame_velocity = 0, s = 1 and
 * theta0 = Stencil::THETA0; the general form is kept. */
static inline double set_extended_grad_expansion_generic(
    double rho, const double momentum[3], const double flux[6], double wt,
    const int ev[3], double theta0, const double frame[3], double s)
{
    const double jx = momentum[0];
    const double jy = momentum[1];
    const double jz = momentum[2];
    const double ux = frame[0];
    const double uy = frame[1];
    const double uz = frame[2];
    const double pxx = flux[0], pyy = flux[1], pzz = flux[2];
    const double pxy = flux[3];
    const double pxz = flux[4];
    const double pyz = flux[5];

    const double stheta0 = s * theta0;
    const double one_bystheta0 = 1.0 / stheta0;
    const double a1x = ((jx)-rho * ux) * one_bystheta0,
                 a1y = ((jy)-rho * uy) * one_bystheta0,
                 a1z = ((jz)-rho * uz) * one_bystheta0;

    const double
        a2xx = ((pxx)-rho * s * stheta0 - rho * ux * ux - ux * ((jx)-rho * ux) -
                ux * ((jx)-rho * ux)) *
               one_bystheta0 * one_bystheta0,
        a2yy = ((pyy)-rho * s * stheta0 - rho * uy * uy - uy * ((jy)-rho * uy) -
                uy * ((jy)-rho * uy)) *
               one_bystheta0 * one_bystheta0,
        a2zz = ((pzz)-rho * s * stheta0 - rho * uz * uz - uz * ((jz)-rho * uz) -
                uz * ((jz)-rho * uz)) *
               one_bystheta0 * one_bystheta0;

    const double a2xy = ((pxy)-0 - rho * ux * uy - ux * ((jy)-rho * uy) -
                         uy * ((jx)-rho * ux)) *
                        one_bystheta0 * one_bystheta0;
    const double a2xz = ((pxz)-0 - rho * ux * uz - ux * ((jz)-rho * uz) -
                         uz * ((jx)-rho * ux)) *
                        one_bystheta0 * one_bystheta0;
    const double a2yz = ((pyz)-0 - rho * uy * uz - uy * ((jz)-rho * uz) -
                         uz * ((jy)-rho * uy)) *
                        one_bystheta0 * one_bystheta0;

    /* AMREX_D_TERM(a1x * ev[0], +a1y * ev[1], +a1z * ev[2]) */
    double f = rho + a1x * ev[0] + a1y * ev[1] + a1z * ev[2];

    /* Utilities.H:149-158, the D_TERM expands to one left-to-right sum */
    f += 0.5 * ((ev[0] * ev[0] - theta0) * a2xx
                + (ev[1] * ev[1] - theta0) * a2yy + 2.0 * (ev[0] * ev[1] - 0) * a2xy
                + (ev[2] * ev[2] - theta0) * a2zz +
                2.0 * (ev[0] * ev[2] - 0) * a2xz +
                2.0 * (ev[1] * ev[2] - 0) * a2yz);

    f *= wt;
    return f;
}

/* Utilities.H:176-185 (RealVect overload): the D_TERM sum is not
 * parenthesised, so it associates with the 2*Cv*T term. */
static inline double get_energy_vec(double T, double rho, const double vel[3], double Cv)
{
    return rho * (2.0 * Cv * T + vel[0] * vel[0] + vel[1] * vel[1] + vel[2] * vel[2]);
}

/* Utilities.H:187-196 (scalar overload, parenthesised |u|^2) */
static inline double get_temperature_scalar(
    double two_rho_e, double rho, double ux, double uy, double uz, double Cv)
{
    return (0.50 / Cv) * ((two_rho_e / rho) - (ux * ux + uy * uy + uz * uz));
}

/* Utilities.H:198-207 (RealVect overload).  AMREX_D_TERM(vel[0]*vel[0],
 * +vel[1]*vel[1], +vel[2]*vel[2]) expands to `vel[0]*vel[0] +vel[1]*vel[1]
 * +vel[2]*vel[2]` after a minus sign, so only u_x^2 is subtracted and
 * u_y^2, u_z^2 are added.  The reference computes exactly this. */
static inline double get_temperature_vec(
    double two_rho_e, double rho, const double vel[3], double Cv)
{
    return (0.50 / Cv) *
           ((two_rho_e / rho) - vel[0] * vel[0] + vel[1] * vel[1] + vel[2] * vel[2]);
}

/* Utilities.H:245-277 (RealVect overload) */
static inline void get_equilibrium_moments_vec(
    double rho, const double vel[3], double total_energy, double Cv, double R,
    double qEq[3], double *RxxEq, double *RyyEq, double *RzzEq, double *RxyEq,
    double *RxzEq, double *RyzEq)
{
    double energy = total_energy / (2.0 * rho);
    double temperature = get_temperature_vec(total_energy, rho, vel, Cv);
    double p = rho * R * temperature;
    double h = energy + (p / rho);

    qEq[0] = 2.0 * rho * vel[0] * h;
    qEq[1] = 2.0 * rho * vel[1] * h;

    *RxxEq = 2.0 * rho * vel[0] * vel[0] * (h + (p / rho)) + 2.0 * p * h;
    *RyyEq = 2.0 * rho * vel[1] * vel[1] * (h + (p / rho)) + 2.0 * p * h;
    *RxyEq = 2.0 * rho * vel[0] * vel[1] * (h + (p / rho)) + 0;
    qEq[2] = 2.0 * rho * vel[2] * h;
    *RzzEq = 2.0 * rho * vel[2] * vel[2] * (h + (p / rho)) + 2.0 * p * h;
    *RxzEq = 2.0 * rho * vel[0] * vel[2] * (h + (p / rho)) + 0;
    *RyzEq = 2.0 * rho * vel[1] * vel[2] * (h + (p / rho)) + 0;
}

/* Exposed wrappers for the tests. */
double orc_set_equilibrium_value(double rho, const double vel[3], double rt, int q)
{
    stencil_init();
    return set_equilibrium_value(rho, vel, rt, EV[q]);
}
double orc_set_extended_equilibrium_value(
    double rho, const double vel[3], double pxx, double pyy, double pzz, int q)
{
    stencil_init();
    return set_extended_equilibrium_value(rho, vel, pxx, pyy, pzz, EV[q]);
}
double orc_grad_expansion(double rho, const double mom[3], const double flux[6], int q)
{
    stencil_init();
    const double zero[3] = {0.0, 0.0, 0.0};
    return set_extended_grad_expansion_generic(
        rho, mom, flux, W[q], EV[q], THETA0, zero, 1.0);
}
void orc_get_equilibrium_moments(
    double rho, const double vel[3], double total_energy, double cv, double R,
    double qeq[3], double req[6])
{
    get_equilibrium_moments_vec(
        rho, vel, total_energy, cv, R, qeq, &req[0], &req[1], &req[2], &req[3],
        &req[4], &req[5]);
}
double orc_get_energy_vec(double T, double rho, const double vel[3], double cv)
{
    return get_energy_vec(T, rho, vel, cv);
}
double orc_get_temperature_scalar(
    double two_rho_e, double rho, double ux, double uy, double uz, double cv)
{
    return get_temperature_scalar(two_rho_e, rho, ux, uy, uz, cv);
}
double orc_get_temperature_vec(double two_rho_e, double rho, const double vel[3], double cv)
{
    return get_temperature_vec(two_rho_e, rho, vel, cv);
}

/* Utilities.H:315-350 for one coarse value (ratio 2, masked_val = -1). */
double orc_masked_avgdown8(const double v[8])
{
    double c = 0;
    double vol = 0;
    for (int n = 0; n < 8; n++) {
        const double fval = v[n];
        if (fabs(fval - (-1.0)) > SMALL_NUM) {
            c += fval;
            vol += 1.0;
        }
    }
    return vol > 0.0 ? c / vol : NAN; /* NAN = "coarse value left unchanged" */
}

/* ------------------------------------------------------------------------ */
/* Level storage                                                              */
/* ------------------------------------------------------------------------ */

typedef struct {
    int lo[3], hi[3], n[3];
    int64_t ncell;
} fabdim;

struct orc_level {
    int dom_lo[3], dom_hi[3], nd[3], periodic[3];
    double prob_lo[3], prob_hi[3], dx[3], inv_dx[3];
    int nbox, ng;
    int (*vlo)[3], (*vhi)[3];
    fabdim *fdg; /* ng ghosts: f, g, is_fluid */
    fabdim *fd1; /* 1 ghost: macrodata (LBM.H:229) */
    fabdim *fd0; /* 0 ghosts: eq, eq_g, derived (LBM.H:231-232) */
    double **f, **g, **tmp, **md, **dq, **eq, **eqg;
    int32_t **isf;
    int32_t *owner; /* domain cell -> owning box, -1 if uncovered */
    orc_phys ph;
    orc_bc bc;
};

static inline int64_t fidx(const fabdim *fd, int c, int i, int j, int k)
{
    return (((int64_t)c * fd->n[2] + (k - fd->lo[2])) * fd->n[1] + (j - fd->lo[1])) *
               fd->n[0] +
           (i - fd->lo[0]);
}

static inline int fab_contains(const fabdim *fd, int i, int j, int k)
{
    return i >= fd->lo[0] && i <= fd->hi[0] && j >= fd->lo[1] && j <= fd->hi[1] &&
           k >= fd->lo[2] && k <= fd->hi[2];
}

static inline int valid_contains(const orc_level *L, int b, int i, int j, int k)
{
    return i >= L->vlo[b][0] && i <= L->vhi[b][0] && j >= L->vlo[b][1] &&
           j <= L->vhi[b][1] && k >= L->vlo[b][2] && k <= L->vhi[b][2];
}

static inline int dom_contains(const orc_level *L, int i, int j, int k)
{
    return i >= L->dom_lo[0] && i <= L->dom_hi[0] && j >= L->dom_lo[1] &&
           j <= L->dom_hi[1] && k >= L->dom_lo[2] && k <= L->dom_hi[2];
}

static void fabdim_set(fabdim *fd, const int lo[3], const int hi[3], int ng)
{
    fd->ncell = 1;
    for (int d = 0; d < 3; d++) {
        fd->lo[d] = lo[d] - ng;
        fd->hi[d] = hi[d] + ng;
        fd->n[d] = fd->hi[d] - fd->lo[d] + 1;
        fd->ncell *= fd->n[d];
    }
}

static double *dalloc(int64_t n, double v)
{
    double *p = (double *)malloc(sizeof(double) * (size_t)n);
    if (!p) {
        fprintf(stderr, "oracle: out of memory (%lld doubles)\n", (long long)n);
        abort();
    }
    for (int64_t i = 0; i < n; i++) p[i] = v;
    return p;
}

orc_level *orc_create(
    const int dom_lo[3], const int dom_hi[3], const int periodic[3],
    const double prob_lo[3], const double prob_hi[3], int nbox,
    const int *box_lo, const int *box_hi, int ng)
{
    stencil_init();
    orc_level *L = (orc_level *)calloc(1, sizeof(orc_level));
    for (int d = 0; d < 3; d++) {
        L->dom_lo[d] = dom_lo[d];
        L->dom_hi[d] = dom_hi[d];
        L->nd[d] = dom_hi[d] - dom_lo[d] + 1;
        L->periodic[d] = periodic[d];
        L->prob_lo[d] = prob_lo[d];
        L->prob_hi[d] = prob_hi[d];
        /* AMReX Geometry: CellSize = (prob_hi - prob_lo) / n_cell,
         * InvCellSize = n_cell / (prob_hi - prob_lo) */
        L->dx[d] = (prob_hi[d] - prob_lo[d]) / L->nd[d];
        L->inv_dx[d] = L->nd[d] / (prob_hi[d] - prob_lo[d]);
    }
    L->nbox = nbox;
    L->ng = ng;
    L->vlo = malloc(sizeof(int[3]) * nbox);
    L->vhi = malloc(sizeof(int[3]) * nbox);
    L->fdg = calloc(nbox, sizeof(fabdim));
    L->fd1 = calloc(nbox, sizeof(fabdim));
    L->fd0 = calloc(nbox, sizeof(fabdim));
    L->f = calloc(nbox, sizeof(double *));
    L->g = calloc(nbox, sizeof(double *));
    L->tmp = calloc(nbox, sizeof(double *));
    L->md = calloc(nbox, sizeof(double *));
    L->dq = calloc(nbox, sizeof(double *));
    L->eq = calloc(nbox, sizeof(double *));
    L->eqg = calloc(nbox, sizeof(double *));
    L->isf = calloc(nbox, sizeof(int32_t *));
    int64_t ndom = (int64_t)L->nd[0] * L->nd[1] * L->nd[2];
    L->owner = malloc(sizeof(int32_t) * (size_t)ndom);
    for (int64_t n = 0; n < ndom; n++) L->owner[n] = -1;
    for (int b = 0; b < nbox; b++) {
        for (int d = 0; d < 3; d++) {
            L->vlo[b][d] = box_lo[3 * b + d];
            L->vhi[b][d] = box_hi[3 * b + d];
        }
        fabdim_set(&L->fdg[b], L->vlo[b], L->vhi[b], ng);
        fabdim_set(&L->fd1[b], L->vlo[b], L->vhi[b], 1);
        fabdim_set(&L->fd0[b], L->vlo[b], L->vhi[b], 0);
        /* MultiFab data is uninitialised in the reference; zero here. */
        L->f[b] = dalloc(ORC_NQ * L->fdg[b].ncell, 0.0);
        L->g[b] = dalloc(ORC_NQ * L->fdg[b].ncell, 0.0);
        L->tmp[b] = dalloc(ORC_NQ * L->fdg[b].ncell, 0.0);
        L->md[b] = dalloc(ORC_NMD * L->fd1[b].ncell, 0.0);
        L->dq[b] = dalloc(ORC_NDQ * L->fd0[b].ncell, 0.0);
        L->eq[b] = dalloc(ORC_NQ * L->fd0[b].ncell, 0.0);
        L->eqg[b] = dalloc(ORC_NQ * L->fd0[b].ncell, 0.0);
        L->isf[b] = calloc((size_t)(2 * L->fdg[b].ncell), sizeof(int32_t));
        for (int k = L->vlo[b][2]; k <= L->vhi[b][2]; k++)
            for (int j = L->vlo[b][1]; j <= L->vhi[b][1]; j++)
                for (int i = L->vlo[b][0]; i <= L->vhi[b][0]; i++) {
                    int64_t o = ((int64_t)(k - dom_lo[2]) * L->nd[1] + (j - dom_lo[1])) *
                                    L->nd[0] +
                                (i - dom_lo[0]);
                    L->owner[o] = b;
                }
    }
    /* defaults: LBM.H:272-318 */
    L->ph.nu = 1.0;
    L->ph.alpha = 1.0;
    L->ph.adiabatic_exponent = 5.0 / 3.0;
    L->ph.m_bar = AIR_MOLAR_MASS;
    L->ph.dt = 1.0;
    L->ph.mesh_speed = 1.0;
    for (int n = 0; n < 6; n++) L->bc.bc_type[n] = ORC_BC_PERIODIC;
    L->bc.velbc_kind = 0;
    L->bc.vel_dir = 1;
    L->bc.mach_ref = 0.005;
    L->bc.density = 1.0;
    L->bc.temperature = 1.0 / 3.0;
    L->bc.adiabatic_exponent = 5.0 / 3.0;
    L->bc.m_bar = AIR_MOLAR_MASS;
    return L;
}

void orc_destroy(orc_level *L)
{
    if (!L) return;
    for (int b = 0; b < L->nbox; b++) {
        free(L->f[b]);
        free(L->g[b]);
        free(L->tmp[b]);
        free(L->md[b]);
        free(L->dq[b]);
        free(L->eq[b]);
        free(L->eqg[b]);
        free(L->isf[b]);
    }
    free(L->f);
    free(L->g);
    free(L->tmp);
    free(L->md);
    free(L->dq);
    free(L->eq);
    free(L->eqg);
    free(L->isf);
    free(L->owner);
    free(L->vlo);
    free(L->vhi);
    free(L->fdg);
    free(L->fd1);
    free(L->fd0);
    free(L);
}

void orc_set_phys(orc_level *L, const orc_phys *p) { L->ph = *p; }
void orc_set_bc(orc_level *L, const orc_bc *bc) { L->bc = *bc; }
int orc_nbox(const orc_level *L) { return L->nbox; }

int64_t orc_fab_size(const orc_level *L, int kind, int b)
{
    switch (kind) {
    case ORC_FAB_F:
    case ORC_FAB_G: return ORC_NQ * L->fdg[b].ncell;
    case ORC_FAB_MD: return ORC_NMD * L->fd1[b].ncell;
    case ORC_FAB_DQ: return ORC_NDQ * L->fd0[b].ncell;
    case ORC_FAB_ISFLUID: return 2 * L->fdg[b].ncell;
    }
    return 0;
}

static void *fab_ptr(const orc_level *L, int kind, int b)
{
    switch (kind) {
    case ORC_FAB_F: return L->f[b];
    case ORC_FAB_G: return L->g[b];
    case ORC_FAB_MD: return L->md[b];
    case ORC_FAB_DQ: return L->dq[b];
    case ORC_FAB_ISFLUID: return L->isf[b];
    }
    return NULL;
}

void orc_get_fab(const orc_level *L, int kind, int b, void *out)
{
    size_t es = kind == ORC_FAB_ISFLUID ? sizeof(int32_t) : sizeof(double);
    memcpy(out, fab_ptr(L, kind, b), es * (size_t)orc_fab_size(L, kind, b));
}

void orc_set_fab(orc_level *L, int kind, int b, const void *in)
{
    size_t es = kind == ORC_FAB_ISFLUID ? sizeof(int32_t) : sizeof(double);
    memcpy(fab_ptr(L, kind, b), in, es * (size_t)orc_fab_size(L, kind, b));
}

static inline double gas_constant(const orc_level *L)
{
    /* LBM.cpp:639 specific_gas_constant = m_R_u / m_m_bar */
    return AIR_MOLAR_MASS / L->ph.m_bar;
}

/* ------------------------------------------------------------------------ */
/* FillBoundary: copy valid data into ghost cells across boxes and periodic  */
/* images (AMReX FabArray::FillBoundary semantics; used at LBM.cpp:603,805,   */
/* 905,1209,1234,1261).  Ghost cells outside a non-periodic domain face and  */
/* ghost cells over uncovered index space are left untouched.                */
/* ------------------------------------------------------------------------ */

/* The image of a ghost cell one period away (AMReX's periodic shifts span
 * one period); a ghost more than one period outside a short periodic
 * direction (ng > n) has none and is left as it is. */
static int periodic_image(const orc_level *L, const int iv[3], int s[3])
{
    for (int d = 0; d < 3; d++) {
        s[d] = iv[d];
        if (s[d] < L->dom_lo[d]) {
            if (!L->periodic[d]) return 0;
            s[d] += L->nd[d];
        } else if (s[d] > L->dom_hi[d]) {
            if (!L->periodic[d]) return 0;
            s[d] -= L->nd[d];
        }
        if (s[d] < L->dom_lo[d] || s[d] > L->dom_hi[d]) return 0;
    }
    return 1;
}

static inline int owner_of(const orc_level *L, const int s[3])
{
    int64_t o = ((int64_t)(s[2] - L->dom_lo[2]) * L->nd[1] + (s[1] - L->dom_lo[1])) *
                    L->nd[0] +
                (s[0] - L->dom_lo[0]);
    return L->owner[o];
}

static void fill_boundary_generic(
    orc_level *L, void **arr, const fabdim *fds, int ncomp, int is_int)
{
    for (int b = 0; b < L->nbox; b++) {
        const fabdim *fd = &fds[b];
#pragma omp parallel for num_threads(g_threads) schedule(static)
        for (int k = fd->lo[2]; k <= fd->hi[2]; k++)
            for (int j = fd->lo[1]; j <= fd->hi[1]; j++)
                for (int i = fd->lo[0]; i <= fd->hi[0]; i++) {
                    if (valid_contains(L, b, i, j, k)) continue;
                    const int iv[3] = {i, j, k};
                    int s[3];
                    if (!periodic_image(L, iv, s)) continue;
                    const int ob = owner_of(L, s);
                    if (ob < 0) continue;
                    const fabdim *sd = &fds[ob];
                    for (int c = 0; c < ncomp; c++) {
                        if (is_int)
                            ((int32_t *)arr[b])[fidx(fd, c, i, j, k)] =
                                ((int32_t *)arr[ob])[fidx(sd, c, s[0], s[1], s[2])];
                        else
                            ((double *)arr[b])[fidx(fd, c, i, j, k)] =
                                ((double *)arr[ob])[fidx(sd, c, s[0], s[1], s[2])];
                    }
                }
    }
}

void orc_fill_boundary(orc_level *L, int kind)
{
    switch (kind) {
    case ORC_FAB_F: fill_boundary_generic(L, (void **)L->f, L->fdg, ORC_NQ, 0); break;
    case ORC_FAB_G: fill_boundary_generic(L, (void **)L->g, L->fdg, ORC_NQ, 0); break;
    case ORC_FAB_MD: fill_boundary_generic(L, (void **)L->md, L->fd1, ORC_NMD, 0); break;
    case ORC_FAB_ISFLUID:
        fill_boundary_generic(L, (void **)L->isf, L->fdg, 2, 1);
        break;
    }
}

/* ------------------------------------------------------------------------ */
/* Physical boundary fill: BCFill::operator() (Source/BC.H:344-471) applied   */
/* to every fab cell outside the domain in a non-periodic direction, in       */
/* (k,j,i) order.  The Constant velocity operator is VelocityBC.H:235-281     */
/* with its constructor VelocityBC.cpp:6-21.                                  */
/* ------------------------------------------------------------------------ */

static void velbc_op(
    const orc_level *L, const int iv[3], const orc_bc *bc, double *rho, double vel[3], double *R, double *T,
    double *gamma)
{
    if (bc->velbc_kind == 2) {
        /* Channel: VelocityBC.cpp:23-37 and VelocityBC.H:90-129 */
        const double speed_of_sound_ref =
            sqrt(bc->adiabatic_exponent * (AIR_MOLAR_MASS / bc->m_bar) * bc->temperature);
        const double u_ref = bc->mach_ref * speed_of_sound_ref;
        *rho = bc->density;
        const double c1 = (double)(iv[1] * (L->dom_hi[1] - iv[1]));
        const double c2 = (double)(iv[2] * (L->dom_hi[2] - iv[2]));
        vel[0] = 16.0 * u_ref * c1 * c2 / pow((double)(L->dom_hi[1] + 1), 4.0);
        *R = (AIR_MOLAR_MASS / bc->m_bar);
        *T = bc->temperature;
        *gamma = bc->adiabatic_exponent;
    } else if (bc->velbc_kind == 3) {
        /* Parabolic: VelocityBC.cpp:39-54 and VelocityBC.H:139-192 */
        const double speed_of_sound_ref =
            sqrt(bc->adiabatic_exponent * (AIR_MOLAR_MASS / bc->m_bar) * bc->temperature);
        const double um = bc->mach_ref * speed_of_sound_ref;
        const int nd = bc->normal_dir, td = bc->vel_dir;
        *rho = bc->density;
        const double height = L->prob_hi[nd] - L->prob_lo[nd];
        const double x = L->prob_lo[nd] + (iv[nd] + 0.5) * L->dx[nd];
        vel[td] = 4.0 * um * x * (height - x) / (height * height);
        *R = (AIR_MOLAR_MASS / bc->m_bar);
        *T = bc->temperature;
        *gamma = bc->adiabatic_exponent;
    } else if (bc->velbc_kind == 1) {
        /* VelocityBC.cpp:6-21 and VelocityBC.H:253-271 */
        const double speed_of_sound_ref =
            sqrt(bc->adiabatic_exponent * (AIR_MOLAR_MASS / bc->m_bar) * bc->temperature);
        const double u0 = bc->mach_ref * speed_of_sound_ref;
        *rho = bc->density;
        vel[bc->vel_dir] = u0;
        *R = (AIR_MOLAR_MASS / bc->m_bar);
        *T = bc->temperature;
        *gamma = bc->adiabatic_exponent;
    }
    /* NoOp (VelocityBC.H:206-233) leaves every argument untouched. */
}

static void bcfill_cell(orc_level *L, int b, double *data, const int iv[3], int is_energy)
{
    const fabdim *fd = &L->fdg[b];
    const int normal_dirs[2] = {1, -1};
    for (int idir = 0; idir < 3; ++idir) {
        for (int lohi = 0; lohi < 2; lohi++) {
            const int normal_dir = normal_dirs[lohi];
            const int bc_idx = idir + lohi * 3;
            if (((lohi == 0) && (iv[idir] < L->dom_lo[idir])) ||
                ((lohi == 1) && (iv[idir] > L->dom_hi[idir]))) {
                for (int q = 0; q < ORC_NQ; q++) {
                    const int *ev = EV[q];
                    const int ivn[3] = {iv[0] + ev[0], iv[1] + ev[1], iv[2] + ev[2]};
                    const int64_t o = fidx(fd, q, iv[0], iv[1], iv[2]);
                    if (valid_contains(L, b, ivn[0], ivn[1], ivn[2])) {
                        const int t = L->bc.bc_type[bc_idx];
                        if (t == ORC_BC_NOSLIP) {
                            data[o] = data[fidx(fd, BOUNCE[q], ivn[0], ivn[1], ivn[2])];
                        } else if (t == ORC_BC_VELOCITY) {
                            double vel_bc[3] = {0.0, 0.0, 0.0};
                            double R = 1.0, temperature = 1.0 / 3.0;
                            double gamma = 5.0 / 3.0;
                            double rho_bc = 0.0;
                            velbc_op(L, iv, &L->bc, &rho_bc, vel_bc, &R, &temperature, &gamma);
                            if (is_energy) {
                                /* BC.H:141-189 velocity_bc_g */
                                const double cv = R / (gamma - 1.0);
                                const double two_rho_e =
                                    get_energy_vec(temperature, rho_bc, vel_bc, cv);
                                double heat_flux[3] = {0.0, 0.0, 0.0};
                                double fl[6];
                                get_equilibrium_moments_vec(
                                    rho_bc, vel_bc, two_rho_e, cv, R, heat_flux, &fl[0],
                                    &fl[1], &fl[2], &fl[3], &fl[4], &fl[5]);
                                const double zero[3] = {0.0, 0.0, 0.0};
                                data[o] = set_extended_grad_expansion_generic(
                                    two_rho_e, heat_flux, fl, W[q], ev, THETA0, zero, 1.0);
                            } else {
                                /* BC.H:191-218 velocity_bc_f */
                                data[o] = set_equilibrium_value(
                                    rho_bc, vel_bc, R * temperature, ev);
                            }
                        } else if (t == ORC_BC_OUTFLOW) {
                            /* BC.H:323-342 */
                            int s[3] = {iv[0], iv[1], iv[2]};
                            s[idir] += 1 * normal_dir;
                            data[o] = data[fidx(fd, q, s[0], s[1], s[2])];
                        } else if (t == ORC_BC_SLIPX) {
                            data[o] = data[fidx(fd, BOUNCE_DIR[0][q], ivn[0], ivn[1], ivn[2])];
                        } else if (t == ORC_BC_SLIPY) {
                            data[o] = data[fidx(fd, BOUNCE_DIR[1][q], ivn[0], ivn[1], ivn[2])];
                        } else if (t == ORC_BC_SLIPZ) {
                            data[o] = data[fidx(fd, BOUNCE_DIR[2][q], ivn[0], ivn[1], ivn[2])];
                        }
                        else if (t == ORC_BC_PRESSURE) {
                            /* BCFill's PRESSURE case (BC.H:421-452): the
                             * boundary state starts at rho 1, R 1, T 1/3,
                             * gamma 5/3 and the velocity operator overrides
                             * it; its velocity is not used */
                            double vel_op[3] = {0.0, 0.0, 0.0};
                            double R = 1.0, temperature = 1.0 / 3.0;
                            double gamma = 5.0 / 3.0;
                            double rho_bc = 1.0;
                            velbc_op(L, iv, &L->bc, &rho_bc, vel_op, &R, &temperature, &gamma);
                            if (is_energy) {
                                /* pressure_bc_g (BC.H:273-321) with
                                 * rho_bc_out = 1 as the caller passes it */
                                const double rho_bc_out = 1.0;
                                double vel_bc[3] = {0.0, 0.0, 0.0};
                                vel_bc[idir] = normal_dir * (1.0 - (rho_bc_out) / rho_bc);
                                const double cv = R / (gamma - 1.0);
                                const double two_rho_e = get_energy_vec(temperature, rho_bc, vel_bc, cv);
                                double heat_flux[3] = {0.0, 0.0, 0.0};
                                double fl[6];
                                get_equilibrium_moments_vec(
                                    rho_bc, vel_bc, two_rho_e, cv, R, heat_flux, &fl[0], &fl[1], &fl[2],
                                    &fl[3], &fl[4], &fl[5]);
                                const double zero[3] = {0.0, 0.0, 0.0};
                                data[o] = set_extended_grad_expansion_generic(
                                    two_rho_e, heat_flux, fl, W[q], ev, THETA0, zero, 1.0);
                            } else {
                                /* pressure_bc_f (BC.H:222-271): the density
                                 * leaving the fluid cell ivn through the face
                                 * (twice the outgoing, once the tangential
                                 * populations it receives) sets the normal
                                 * velocity of the ghost's equilibrium */
                                double rho_out = 0.0, rho_tan = 0.0;
                                for (int qq = 0; qq < ORC_NQ; qq++) {
                                    const int *evn_out = EV[qq];
                                    const int bounce_qq = BOUNCE[qq];
                                    const int *evn_in = EV[bounce_qq];
                                    const int64_t r = fidx(fd, bounce_qq, ivn[0] + evn_out[0],
                                                           ivn[1] + evn_out[1], ivn[2] + evn_out[2]);
                                    if (evn_in[idir] == -normal_dir) {
                                        rho_out += 2.0 * data[r];
                                    } else if (evn_in[idir] == 0) {
                                        rho_tan += 1.0 * data[r];
                                    }
                                }
                                double vel_bc[3] = {0.0, 0.0, 0.0};
                                vel_bc[idir] = normal_dir * (1.0 - (rho_out + rho_tan) / rho_bc);
                                data[o] = set_equilibrium_value(rho_bc, vel_bc, R * temperature, ev);
                            }
                        }
                        /* PERIODIC: nothing here */
                    } else {
                        data[o] = -1.0; /* BC.H:463-466 */
                    }
                }
            }
        }
    }
}

/* PhysBCFunct's application of BCFill (FillPatchOps.H:113,153) to every
 * fab cell beyond a non-periodic face.  The reference runs the functor on
 * all those cells at once (a GPU kernel: an OUTFLOW ghost that copies
 * another ghost races with that ghost's own fill); here the cells are
 * taken in order of their distance beyond the domain (the sum over
 * non-periodic directions of how many cells out they lie), so that the
 * cell an OUTFLOW ghost copies (one step closer) is final before it is
 * read: an edge ghost takes the value its face neighbour's boundary fill
 * gives.  Within one distance the fills are independent. */
void orc_physbc(orc_level *L, int lattice)
{
    double **arr = lattice == 0 ? L->f : L->g;
    const int maxdist = 3 * L->ng;
    for (int dist = 1; dist <= maxdist; dist++)
        for (int b = 0; b < L->nbox; b++) {
            const fabdim *fd = &L->fdg[b];
            for (int k = fd->lo[2]; k <= fd->hi[2]; k++)
                for (int j = fd->lo[1]; j <= fd->hi[1]; j++)
                    for (int i = fd->lo[0]; i <= fd->hi[0]; i++) {
                        const int iv[3] = {i, j, k};
                        int out = 0;
                        for (int d = 0; d < 3; d++) {
                            if (L->periodic[d]) continue;
                            if (iv[d] < L->dom_lo[d]) out += L->dom_lo[d] - iv[d];
                            if (iv[d] > L->dom_hi[d]) out += iv[d] - L->dom_hi[d];
                        }
                        if (out == dist) bcfill_cell(L, b, arr[b], iv, lattice == 1);
                    }
        }
}

/* FillPatchOps::fillpatch for level 0 (FillPatchOps.H:75-132): the bounce
 * prefill (92-110), then FillPatchSingleLevel = FillBoundary + physbc. */
/* FillPatchOps.H:92-108: domain-outside ghosts take the opposite population
 * of their in-domain neighbour (overwritten afterwards by physbc). */
static void bounce_prefill(orc_level *L, double **arr)
{
    for (int b = 0; b < L->nbox; b++) {
        const fabdim *fd = &L->fdg[b];
        double *f = arr[b];
#pragma omp parallel for num_threads(g_threads) schedule(static)
        for (int k = fd->lo[2]; k <= fd->hi[2]; k++)
            for (int j = fd->lo[1]; j <= fd->hi[1]; j++)
                for (int i = fd->lo[0]; i <= fd->hi[0]; i++) {
                    if (dom_contains(L, i, j, k)) continue;
                    for (int q = 0; q < ORC_NQ; q++) {
                        const int a = i + EV[q][0], c = j + EV[q][1], e = k + EV[q][2];
                        if (dom_contains(L, a, c, e) && fab_contains(fd, a, c, e))
                            f[fidx(fd, q, i, j, k)] = f[fidx(fd, BOUNCE[q], a, c, e)];
                    }
                }
    }
}

void orc_fillpatch(orc_level *L, int lattice)
{
    bounce_prefill(L, lattice == 0 ? L->f : L->g);
    orc_fill_boundary(L, lattice == 0 ? ORC_FAB_F : ORC_FAB_G);
    orc_physbc(L, lattice);
}

/* ------------------------------------------------------------------------ */
/* LBM::stream (Source/LBM.cpp:558-604): push with halfway bounce-back into  */
/* a temporary pre-filled with -1, copy back, FillBoundary.                  */
/* ------------------------------------------------------------------------ */
void orc_stream(orc_level *L, int lattice)
{
    double **arr = lattice == 0 ? L->f : L->g;
    for (int b = 0; b < L->nbox; b++) {
        const fabdim *fd = &L->fdg[b];
        const double *f = arr[b];
        double *fs = L->tmp[b];
        const int32_t *isf = L->isf[b];
        const int64_t n = ORC_NQ * fd->ncell;
#pragma omp parallel for num_threads(g_threads) schedule(static)
        for (int64_t m = 0; m < n; m++) fs[m] = -1.0;
#pragma omp parallel for num_threads(g_threads) schedule(static)
        for (int k = fd->lo[2]; k <= fd->hi[2]; k++)
            for (int j = fd->lo[1]; j <= fd->hi[1]; j++)
                for (int i = fd->lo[0]; i <= fd->hi[0]; i++) {
                    if (isf[fidx(fd, 0, i, j, k)] != 1) continue;
                    for (int q = 0; q < ORC_NQ; q++) {
                        const int a = i + EV[q][0], c = j + EV[q][1], e = k + EV[q][2];
                        if (!fab_contains(fd, a, c, e)) continue;
                        if (isf[fidx(fd, 0, a, c, e)] != 0)
                            fs[fidx(fd, q, a, c, e)] = f[fidx(fd, q, i, j, k)];
                        else
                            fs[fidx(fd, BOUNCE[q], i, j, k)] = f[fidx(fd, q, i, j, k)];
                    }
                }
        /* MultiFab::Copy(fs, f_star) over all cells incl. ghosts == swap */
        arr[b] = fs;
        L->tmp[b] = (double *)f;
    }
    orc_fill_boundary(L, lattice == 0 ? ORC_FAB_F : ORC_FAB_G);
}

/* ------------------------------------------------------------------------ */
/* LBM::f_to_macrodata (Source/LBM.cpp:810-906)                              */
/* ------------------------------------------------------------------------ */
void orc_f_to_macrodata(orc_level *L)
{
    const double l_mesh_speed = L->ph.mesh_speed;
    const double specific_gas_constant = gas_constant(L);
    const double cv = specific_gas_constant / (L->ph.adiabatic_exponent - 1.0);
    for (int b = 0; b < L->nbox; b++) {
        const fabdim *fd = &L->fdg[b];
        const fabdim *m1 = &L->fd1[b];
        const double *f = L->f[b], *g = L->g[b];
        double *md = L->md[b];
        const int32_t *isf = L->isf[b];
#pragma omp parallel for num_threads(g_threads) schedule(static)
        for (int k = m1->lo[2]; k <= m1->hi[2]; k++)
            for (int j = m1->lo[1]; j <= m1->hi[1]; j++)
                for (int i = m1->lo[0]; i <= m1->hi[0]; i++) {
                    if (isf[fidx(fd, 0, i, j, k)] != 1) continue;
                    double rho = 0.0, u = 0.0, v = 0.0, w = 0.0;
                    double pxx = 0.0, pyy = 0.0, pzz = 0.0, pxy = 0.0, pxz = 0.0,
                           pyz = 0.0;
                    double two_rho_e = 0.0, qx = 0.0, qy = 0.0, qz = 0.0;
                    for (int q = 0; q < ORC_NQ; q++) {
                        const double fq = f[fidx(fd, q, i, j, k)];
                        const double gq = g[fidx(fd, q, i, j, k)];
                        const int *ev = EV[q];
                        rho += fq;
                        u += ev[0] * fq, v += ev[1] * fq, w += ev[2] * fq;
                        pxx += ev[0] * ev[0] * fq;
                        pyy += ev[1] * ev[1] * fq;
                        pxy += ev[0] * ev[1] * fq;
                        pzz += ev[2] * ev[2] * fq;
                        pxz += ev[0] * ev[2] * fq;
                        pyz += ev[1] * ev[2] * fq;
                        two_rho_e += gq;
                        qx += ev[0] * gq, qy += ev[1] * gq, qz += ev[2] * gq;
                    }
                    u *= l_mesh_speed / rho, v *= l_mesh_speed / rho,
                        w *= l_mesh_speed / rho;
#define MD(c) md[fidx(m1, (c), i, j, k)]
                    MD(0) = rho;
                    MD(1) = u;
                    MD(2) = v;
                    MD(3) = w;
                    MD(4) = sqrt(u * u + v * v + w * w);
                    MD(9) = pxx;
                    MD(10) = pyy;
                    MD(11) = pzz;
                    MD(12) = pxy;
                    MD(13) = pxz;
                    MD(14) = pyz;
                    MD(5) = two_rho_e;
                    MD(15) = qx;
                    MD(16) = qy;
                    MD(17) = qz;
                    const double temperature =
                        get_temperature_scalar(two_rho_e, rho, u, v, w, cv);
                    MD(18) = temperature;
                    MD(6) = rho * u *
                            ((1.0 - 3.0 * specific_gas_constant * temperature) - u * u);
                    MD(7) = rho * v *
                            ((1.0 - 3.0 * specific_gas_constant * temperature) - v * v);
                    MD(8) = rho * w *
                            ((1.0 - 3.0 * specific_gas_constant * temperature) - w * w);
#undef MD
                }
    }
    orc_fill_boundary(L, ORC_FAB_MD);
}

/* ------------------------------------------------------------------------ */
/* gradient (Utilities.H:279-312) and LBM::compute_q_corrections              */
/* (LBM.cpp:959-991).  The one-sided test uses the domain box, not the       */
/* periodicity.                                                               */
/* ------------------------------------------------------------------------ */
static inline double gradient(
    const orc_level *L, int b, int dir, int comp, const int iv[3])
{
    const fabdim *fd = &L->fdg[b];
    const fabdim *m1 = &L->fd1[b];
    const int32_t *isf = L->isf[b];
    const double *md = L->md[b];
    int ivp[3] = {iv[0], iv[1], iv[2]};
    int ivm[3] = {iv[0], iv[1], iv[2]};
    ivp[dir] += 1;
    ivm[dir] -= 1;
    double vp = 0.0, vc = 0.0, vm = 0.0;
    const int p_bad = !dom_contains(L, ivp[0], ivp[1], ivp[2]) ||
                      isf[fidx(fd, 0, ivp[0], ivp[1], ivp[2])] != 1;
    const int m_bad = !dom_contains(L, ivm[0], ivm[1], ivm[2]) ||
                      isf[fidx(fd, 0, ivm[0], ivm[1], ivm[2])] != 1;
    if (p_bad && m_bad) {
        vp = 0.0;
        vc = 0.0;
        vm = 0.0;
    } else if (p_bad) {
        vp = 0.0;
        vc = md[fidx(m1, comp, iv[0], iv[1], iv[2])];
        vm = -md[fidx(m1, comp, ivm[0], ivm[1], ivm[2])];
    } else if (m_bad) {
        vp = md[fidx(m1, comp, ivp[0], ivp[1], ivp[2])];
        vc = -md[fidx(m1, comp, iv[0], iv[1], iv[2])];
        vm = 0.0;
    } else {
        vp = 0.5 * md[fidx(m1, comp, ivp[0], ivp[1], ivp[2])];
        vc = 0.0;
        vm = -0.5 * md[fidx(m1, comp, ivm[0], ivm[1], ivm[2])];
    }
    return (vp + vc + vm) * L->inv_dx[dir];
}

/* the vorticity LBM::compute_derived (LBM.cpp:927-950) computes at one
 * cell from the level's macrodata; 0 where the cell is not fluid */
static void vorticity_at(const orc_level *L, int b, int fluid, const int iv[3], double w[4])
{
    w[0] = w[1] = w[2] = w[3] = 0.0;
    if (!fluid) return;
    const double vx = gradient(L, b, 0, 2, iv);
    const double wx = gradient(L, b, 0, 3, iv);
    const double uy = gradient(L, b, 1, 1, iv);
    const double wy = gradient(L, b, 1, 3, iv);
    const double uz = gradient(L, b, 2, 1, iv);
    const double vz = gradient(L, b, 2, 2, iv);
    w[0] = wy - vz;
    w[1] = uz - wx;
    w[2] = vx - uy;
    w[3] = sqrt((wy - vz) * (wy - vz) + (uz - wx) * (uz - wx) + (vx - uy) * (vx - uy));
}

/* LBM::plot_file_mf (LBM.cpp:1629-1669): per box, the valid cells'
 * variables [19 macrodata, (27 f, 27 g if streaming), (vort_x, vort_y,
 * vort_z, vort_mag, dQCorrX..Z if derived), is_fluid, eb_boundary],
 * component-major (x fastest), boxes back to back.  The macrodata and
 * dQCorr are those of the last collide (or of the initial state), the
 * vorticity is compute_derived's of those macrodata; the level is not
 * modified. */
void orc_plot_vars(const orc_level *L, int streaming, int derived, double *out)
{
    const int nc = 19 + (streaming ? 2 * ORC_NQ : 0) + (derived ? 7 : 0) + 2;
    int64_t off = 0;
    for (int b = 0; b < L->nbox; b++) {
        const fabdim *fd = &L->fdg[b];
        const fabdim *m1 = &L->fd1[b];
        const fabdim *d0 = &L->fd0[b];
        const int nx = d0->hi[0] - d0->lo[0] + 1, ny = d0->hi[1] - d0->lo[1] + 1,
                  nz = d0->hi[2] - d0->lo[2] + 1;
        const int64_t nv = (int64_t)nx * ny * nz;
        for (int k = d0->lo[2]; k <= d0->hi[2]; k++)
            for (int j = d0->lo[1]; j <= d0->hi[1]; j++)
                for (int i = d0->lo[0]; i <= d0->hi[0]; i++) {
                    const int64_t c = ((int64_t)(k - d0->lo[2]) * ny + (j - d0->lo[1])) * nx + (i - d0->lo[0]);
                    int n = 0;
                    for (int m = 0; m < 19; m++) out[off + (n++) * nv + c] = L->md[b][fidx(m1, m, i, j, k)];
                    if (streaming) {
                        for (int q = 0; q < ORC_NQ; q++) out[off + (n++) * nv + c] = L->f[b][fidx(fd, q, i, j, k)];
                        for (int q = 0; q < ORC_NQ; q++) out[off + (n++) * nv + c] = L->g[b][fidx(fd, q, i, j, k)];
                    }
                    if (derived) {
                        double w[4];
                        const int iv[3] = {i, j, k};
                        vorticity_at(L, b, L->isf[b][fidx(fd, 0, i, j, k)] == 1, iv, w);
                        for (int m = 0; m < 4; m++) out[off + (n++) * nv + c] = w[m];
                        for (int m = 0; m < 3; m++) out[off + (n++) * nv + c] = L->dq[b][fidx(d0, m, i, j, k)];
                    }
                    for (int m = 0; m < 2; m++) out[off + (n++) * nv + c] = (double)L->isf[b][fidx(fd, m, i, j, k)];
                }
        off += nc * nv;
    }
}

/* ------------------------------------------------------------------------ */
/* LBM::compute_derived (Source/LBM.cpp:909-955): vorticity of the velocity  */
/* of the current populations (macrodata recomputed from them first).        */
/* out: per box, the valid cells' 4 components [vort_x, vort_y, vort_z,      */
/* vort_mag] component-major (x fastest), boxes back to back; 0 in solids.   */
/* ------------------------------------------------------------------------ */
void orc_vorticity(orc_level *L, double *out)
{
    orc_f_to_macrodata(L);
    int64_t off = 0;
    for (int b = 0; b < L->nbox; b++) {
        const fabdim *fd = &L->fdg[b];
        const fabdim *d0 = &L->fd0[b];
        const int32_t *isf = L->isf[b];
        const int nx = d0->hi[0] - d0->lo[0] + 1, ny = d0->hi[1] - d0->lo[1] + 1,
                  nz = d0->hi[2] - d0->lo[2] + 1;
        const int64_t nv = (int64_t)nx * ny * nz;
#pragma omp parallel for num_threads(g_threads) schedule(static)
        for (int k = d0->lo[2]; k <= d0->hi[2]; k++)
            for (int j = d0->lo[1]; j <= d0->hi[1]; j++)
                for (int i = d0->lo[0]; i <= d0->hi[0]; i++) {
                    const int64_t c = ((int64_t)(k - d0->lo[2]) * ny + (j - d0->lo[1])) * nx + (i - d0->lo[0]);
                    double w[4];
                    const int iv[3] = {i, j, k};
                    vorticity_at(L, b, isf[fidx(fd, 0, i, j, k)] == 1, iv, w);
                    for (int n = 0; n < 4; n++) out[off + n * nv + c] = w[n];
                }
        off += 4 * nv;
    }
}

void orc_compute_q_corrections(orc_level *L)
{
    for (int b = 0; b < L->nbox; b++) {
        const fabdim *fd = &L->fdg[b];
        const fabdim *d0 = &L->fd0[b];
        double *dq = L->dq[b];
        const int32_t *isf = L->isf[b];
#pragma omp parallel for num_threads(g_threads) schedule(static)
        for (int k = d0->lo[2]; k <= d0->hi[2]; k++)
            for (int j = d0->lo[1]; j <= d0->hi[1]; j++)
                for (int i = d0->lo[0]; i <= d0->hi[0]; i++) {
                    if (isf[fidx(fd, 0, i, j, k)] != 1) continue;
                    const int iv[3] = {i, j, k};
                    dq[fidx(d0, 0, i, j, k)] = gradient(L, b, 0, 6, iv);
                    dq[fidx(d0, 1, i, j, k)] = gradient(L, b, 1, 7, iv);
                    dq[fidx(d0, 2, i, j, k)] = gradient(L, b, 2, 8, iv);
                }
    }
}

/* ------------------------------------------------------------------------ */
/* LBM::macrodata_to_equilibrium (Source/LBM.cpp:621-762)                     */
/* ------------------------------------------------------------------------ */
void orc_macrodata_to_equilibrium(orc_level *L)
{
    const double specific_gas_constant = gas_constant(L);
    const double cv = specific_gas_constant / (L->ph.adiabatic_exponent - 1.0);
    const double nu = L->ph.nu;
    const double dt = L->ph.dt;
    const double alpha = L->ph.alpha;
    const double theta0 = THETA0;
    const double zero_vec[3] = {0.0, 0.0, 0.0};
    for (int b = 0; b < L->nbox; b++) {
        const fabdim *fd = &L->fdg[b];
        const fabdim *m1 = &L->fd1[b];
        const fabdim *d0 = &L->fd0[b];
        const double *md = L->md[b];
        const double *dq = L->dq[b];
        double *eq = L->eq[b], *eqg = L->eqg[b];
        const int32_t *isf = L->isf[b];
#pragma omp parallel for num_threads(g_threads) schedule(static)
        for (int k = d0->lo[2]; k <= d0->hi[2]; k++)
            for (int j = d0->lo[1]; j <= d0->hi[1]; j++)
                for (int i = d0->lo[0]; i <= d0->hi[0]; i++) {
                    if (isf[fidx(fd, 0, i, j, k)] != 1) continue;
#define MD(c) md[fidx(m1, (c), i, j, k)]
                    const double rho = MD(0);
                    const double vel[3] = {MD(1), MD(2), MD(3)};
                    const double two_rho_e = MD(5);
                    const double temperature = MD(18);
                    const double dqx = dq[fidx(d0, 0, i, j, k)];
                    const double dqy = dq[fidx(d0, 1, i, j, k)];
                    const double dqz = dq[fidx(d0, 2, i, j, k)];
                    for (int q = 0; q < ORC_NQ; q++) {
                        const double wt = W[q];
                        const int *ev = EV[q];
                        const double omega =
                            1.0 / (nu / (specific_gas_constant * temperature * dt) + 0.5);
                        const double omega_one =
                            1.0 /
                            (alpha / (specific_gas_constant * temperature * dt) + 0.5);
                        const double omega_one_by_omega = omega_one / omega;
                        const double omega_corr = (2.0 - omega) / (2.0 * omega * rho);

                        const double pxx_ext = vel[0] * vel[0] +
                                               specific_gas_constant * temperature +
                                               dt * (omega_corr)*dqx;
                        const double pyy_ext = vel[1] * vel[1] +
                                               specific_gas_constant * temperature +
                                               dt * (omega_corr)*dqy;
                        const double pzz_ext = vel[2] * vel[2] +
                                               specific_gas_constant * temperature +
                                               dt * (omega_corr)*dqz;

                        eq[fidx(d0, q, i, j, k)] = set_extended_equilibrium_value(
                            rho, vel, pxx_ext, pyy_ext, pzz_ext, ev);

                        double qx_eq = 0.0, qy_eq = 0.0, qz_eq = 0.0;
                        double rxx_eq = 0.0, ryy_eq = 0.0, rzz_eq = 0.0, rxy_eq = 0.0,
                               rxz_eq = 0.0, ryz_eq = 0.0;
                        double heat_flux[3] = {0.0, 0.0, 0.0};
                        get_equilibrium_moments_vec(
                            rho, vel, two_rho_e, cv, specific_gas_constant, heat_flux,
                            &rxx_eq, &ryy_eq, &rzz_eq, &rxy_eq, &rxz_eq, &ryz_eq);
                        qx_eq = heat_flux[0];
                        qy_eq = heat_flux[1];
                        qz_eq = heat_flux[2];

                        const double pxx = MD(9);
                        const double pyy = MD(10);
                        const double pzz = MD(11);
                        const double pxy = MD(12);
                        const double pxz = MD(13);
                        const double pyz = MD(14);
                        const double qx = MD(15);
                        const double qy = MD(16);
                        const double qz = MD(17);

                        qx_eq *= omega_one_by_omega;
                        qy_eq *= omega_one_by_omega;
                        qz_eq *= omega_one_by_omega;

                        qx_eq += (1.0 - omega_one_by_omega) *
                                 (qx - 2.0 * vel[0] * pxx - 2.0 * vel[1] * pxy -
                                  2.0 * vel[2] * pxz - vel[0] * dt * dqx);
                        qy_eq += (1.0 - omega_one_by_omega) *
                                 (qy - 2.0 * vel[0] * pxy - 2.0 * vel[1] * pyy -
                                  2.0 * vel[2] * pyz - vel[1] * dt * dqy);
                        qz_eq += (1.0 - omega_one_by_omega) *
                                 (qz - 2.0 * vel[0] * pxz - 2.0 * vel[1] * pyz -
                                  2.0 * vel[2] * pzz - vel[2] * dt * dqz);

                        const double heat_flux_mrt[3] = {qx_eq, qy_eq, qz_eq};
                        const double flux_of_heat_flux[6] = {rxx_eq, ryy_eq, rzz_eq,
                                                             rxy_eq, rxz_eq, ryz_eq};
                        eqg[fidx(d0, q, i, j, k)] = set_extended_grad_expansion_generic(
                            two_rho_e, heat_flux_mrt, flux_of_heat_flux, wt, ev, theta0,
                            zero_vec, 1.0);
                    }
#undef MD
                }
    }
}

/* ------------------------------------------------------------------------ */
/* LBM::relax_f_to_equilibrium (Source/LBM.cpp:765-807)                       */
/* ------------------------------------------------------------------------ */
void orc_relax_f_to_equilibrium(orc_level *L)
{
    const double specific_gas_constant = gas_constant(L);
    const double nu = L->ph.nu;
    const double dt = L->ph.dt;
    for (int b = 0; b < L->nbox; b++) {
        const fabdim *fd = &L->fdg[b];
        const fabdim *m1 = &L->fd1[b];
        const fabdim *d0 = &L->fd0[b];
        double *f = L->f[b], *g = L->g[b];
        const double *eq = L->eq[b], *eqg = L->eqg[b], *md = L->md[b];
        const int32_t *isf = L->isf[b];
#pragma omp parallel for num_threads(g_threads) schedule(static)
        for (int k = d0->lo[2]; k <= d0->hi[2]; k++)
            for (int j = d0->lo[1]; j <= d0->hi[1]; j++)
                for (int i = d0->lo[0]; i <= d0->hi[0]; i++) {
                    if (isf[fidx(fd, 0, i, j, k)] != 1) continue;
                    const double temperature = md[fidx(m1, 18, i, j, k)];
                    const double omega =
                        1.0 / (nu / (specific_gas_constant * temperature * dt) + 0.5);
                    for (int q = 0; q < ORC_NQ; q++) {
                        const int64_t o = fidx(fd, q, i, j, k);
                        const int64_t e = fidx(d0, q, i, j, k);
                        f[o] += omega * (eq[e] - f[o]);
                        g[o] += omega * (eqg[e] - g[o]);
                    }
                }
    }
    orc_fill_boundary(L, ORC_FAB_F);
    orc_fill_boundary(L, ORC_FAB_G);
}

/* LBM::collide (Source/LBM.cpp:607-618) */
void orc_collide(orc_level *L)
{
    orc_f_to_macrodata(L);
    orc_compute_q_corrections(L);
    orc_macrodata_to_equilibrium(L);
    orc_relax_f_to_equilibrium(L);
}

/* Single-level LBM::evolve (LBM.cpp:398-448) without post_time_step
 * (vorticity and EB forces do not feed back into the state). */
void orc_evolve(orc_level *L, int nsteps)
{
    for (int s = 0; s < nsteps; s++) {
        orc_fillpatch(L, 0);
        orc_fillpatch(L, 1);
        orc_stream(L, 0);
        orc_stream(L, 1);
        orc_collide(L);
    }
}

void orc_conservation(const orc_level *L, double out[5])
{
    for (int n = 0; n < 5; n++) out[n] = 0.0;
    for (int b = 0; b < L->nbox; b++) {
        const fabdim *fd = &L->fdg[b];
        for (int k = L->vlo[b][2]; k <= L->vhi[b][2]; k++)
            for (int j = L->vlo[b][1]; j <= L->vhi[b][1]; j++)
                for (int i = L->vlo[b][0]; i <= L->vhi[b][0]; i++) {
                    if (L->isf[b][fidx(fd, 0, i, j, k)] != 1) continue;
                    for (int q = 0; q < ORC_NQ; q++) {
                        const double fq = L->f[b][fidx(fd, q, i, j, k)];
                        out[0] += fq;
                        out[1] += EV[q][0] * fq;
                        out[2] += EV[q][1] * fq;
                        out[3] += EV[q][2] * fq;
                        out[4] += L->g[b][fidx(fd, q, i, j, k)];
                    }
                }
    }
}

/* ------------------------------------------------------------------------ */
/* Initialisation                                                             */
/* ------------------------------------------------------------------------ */

/* EB sphere classification (AMReX EB2 SphereIF via EB.cpp:36).  A cell is
 * covered when all 8 of its corner nodes lie in the body; |x-c|^2 == r^2
 * counts as body.  Ties and EB2's small-cell fixes: parity unpinned. */
int orc_sphere_covered(
    const double node_lo[3], const double dx[3], const double center[3], double radius,
    int has_fluid_inside)
{
    const double r2 = radius * radius;
    for (int a = 0; a < 2; a++)
        for (int c = 0; c < 2; c++)
            for (int e = 0; e < 2; e++) {
                const double x = node_lo[0] + a * dx[0] - center[0];
                const double y = node_lo[1] + c * dx[1] - center[1];
                const double z = node_lo[2] + e * dx[2] - center[2];
                const double d2 = x * x + y * y + z * z;
                const int in_body = has_fluid_inside ? (d2 >= r2) : (d2 <= r2);
                if (!in_body) return 0;
            }
    return 1;
}

/* EB cylinder classification (AMReX EB2 CylinderIF, geom_type "cylinder"
 * via EB.cpp:28-37; inputs eb2.cylinder_radius, _height, _direction,
 * _center, _has_fluid_inside as in single_cylinder.inp:38-43).  The body is
 * the set of points whose distance from the axis through `center` along
 * `direction` is at most `radius` and whose distance from `center` along
 * the axis is at most height / 2; a cell is covered when all 8 corner nodes
 * lie in it (nodes on the surface count as body).  Ties: parity unpinned,
 * as for the sphere. */
int orc_cylinder_covered(
    const double node_lo[3], const double dx[3], const double center[3], double radius,
    double height, int direction, int has_fluid_inside)
{
    const double r2 = radius * radius, hh = 0.5 * height;
    const int u = direction == 0 ? 1 : 0, v = direction == 2 ? 1 : 2;
    for (int a = 0; a < 2; a++)
        for (int c = 0; c < 2; c++)
            for (int e = 0; e < 2; e++) {
                const double p[3] = {node_lo[0] + a * dx[0] - center[0],
                                     node_lo[1] + c * dx[1] - center[1],
                                     node_lo[2] + e * dx[2] - center[2]};
                const double d2 = p[u] * p[u] + p[v] * p[v];
                const double h = fabs(p[direction]);
                const int inside = d2 <= r2 && h <= hh;
                const int in_body = has_fluid_inside ? (d2 >= r2 || h >= hh) : inside;
                if (!in_body) return 0;
            }
    return 1;
}

/* LBM::initialize_is_fluid (LBM.cpp:1213-1262) given comp0 for every fab
 * cell: FillBoundary, EB boundary ring (comp1), FillBoundary. */
static void finish_is_fluid(orc_level *L)
{
    orc_fill_boundary(L, ORC_FAB_ISFLUID);
    for (int b = 0; b < L->nbox; b++) {
        const fabdim *fd = &L->fdg[b];
        int32_t *isf = L->isf[b];
        const int ngm = L->ng - 1;
        for (int k = L->vlo[b][2] - ngm; k <= L->vhi[b][2] + ngm; k++)
            for (int j = L->vlo[b][1] - ngm; j <= L->vhi[b][1] + ngm; j++)
                for (int i = L->vlo[b][0] - ngm; i <= L->vhi[b][0] + ngm; i++) {
                    int all_covered = 1;
                    const int iv[3] = {i, j, k};
                    for (int idir = 0; idir < 3; idir++) {
                        int m[3] = {iv[0], iv[1], iv[2]}, p[3] = {iv[0], iv[1], iv[2]};
                        m[idir] -= 1;
                        p[idir] += 1;
                        all_covered &= (isf[fidx(fd, 0, m[0], m[1], m[2])] == 0) &&
                                       (isf[fidx(fd, 0, p[0], p[1], p[2])] == 0);
                    }
                    if ((all_covered) || (isf[fidx(fd, 0, i, j, k)] == 1))
                        isf[fidx(fd, 1, i, j, k)] = 0;
                    else
                        isf[fidx(fd, 1, i, j, k)] = 1;
                }
    }
    orc_fill_boundary(L, ORC_FAB_ISFLUID);
}

void orc_init_flags_all_regular(orc_level *L)
{
    for (int b = 0; b < L->nbox; b++) {
        const fabdim *fd = &L->fdg[b];
        for (int64_t n = 0; n < 2 * fd->ncell; n++) L->isf[b][n] = 0;
        for (int64_t n = 0; n < fd->ncell; n++) L->isf[b][n] = 1;
    }
    finish_is_fluid(L);
}

void orc_init_flags_sphere(
    orc_level *L, const double center[3], double radius, int has_fluid_inside)
{
    for (int b = 0; b < L->nbox; b++) {
        const fabdim *fd = &L->fdg[b];
        for (int64_t n = 0; n < 2 * fd->ncell; n++) L->isf[b][n] = 0;
#pragma omp parallel for num_threads(g_threads) schedule(static)
        for (int k = fd->lo[2]; k <= fd->hi[2]; k++)
            for (int j = fd->lo[1]; j <= fd->hi[1]; j++)
                for (int i = fd->lo[0]; i <= fd->hi[0]; i++) {
                    const double node[3] = {L->prob_lo[0] + i * L->dx[0],
                                            L->prob_lo[1] + j * L->dx[1],
                                            L->prob_lo[2] + k * L->dx[2]};
                    L->isf[b][fidx(fd, 0, i, j, k)] =
                        orc_sphere_covered(node, L->dx, center, radius, has_fluid_inside)
                            ? 0
                            : 1;
                }
    }
    finish_is_fluid(L);
}

void orc_init_flags_cylinder(
    orc_level *L, const double center[3], double radius, double height, int direction,
    int has_fluid_inside)
{
    for (int b = 0; b < L->nbox; b++) {
        const fabdim *fd = &L->fdg[b];
        for (int64_t n = 0; n < 2 * fd->ncell; n++) L->isf[b][n] = 0;
#pragma omp parallel for num_threads(g_threads) schedule(static)
        for (int k = fd->lo[2]; k <= fd->hi[2]; k++)
            for (int j = fd->lo[1]; j <= fd->hi[1]; j++)
                for (int i = fd->lo[0]; i <= fd->hi[0]; i++) {
                    const double node[3] = {L->prob_lo[0] + i * L->dx[0],
                                            L->prob_lo[1] + j * L->dx[1],
                                            L->prob_lo[2] + k * L->dx[2]};
                    L->isf[b][fidx(fd, 0, i, j, k)] =
                        orc_cylinder_covered(node, L->dx, center, radius, height, direction,
                                             has_fluid_inside)
                            ? 0
                            : 1;
                }
    }
    finish_is_fluid(L);
}

/* solid_mask over the domain (x fastest); cells outside the domain are
 * fluid.  Used for synthetic geometries in tests. */
void orc_init_flags_from_mask(orc_level *L, const uint8_t *solid_mask)
{
    for (int b = 0; b < L->nbox; b++) {
        const fabdim *fd = &L->fdg[b];
        for (int64_t n = 0; n < 2 * fd->ncell; n++) L->isf[b][n] = 0;
        for (int k = fd->lo[2]; k <= fd->hi[2]; k++)
            for (int j = fd->lo[1]; j <= fd->hi[1]; j++)
                for (int i = fd->lo[0]; i <= fd->hi[0]; i++) {
                    const int iv[3] = {i, j, k};
                    int s[3];
                    int solid = 0;
                    if (periodic_image(L, iv, s)) {
                        int64_t o =
                            ((int64_t)(s[2] - L->dom_lo[2]) * L->nd[1] + (s[1] - L->dom_lo[1])) *
                                L->nd[0] +
                            (s[0] - L->dom_lo[0]);
                        solid = solid_mask[o] != 0;
                    }
                    L->isf[b][fidx(fd, 0, i, j, k)] = solid ? 0 : 1;
                }
    }
    finish_is_fluid(L);
}

/* ic::Initializer::initialize (IC.H:460-525) for one cell's state. */
static void init_cell(
    orc_level *L, int b, int i, int j, int k, double rho, const double vel[3],
    double temperature, double specific_gas_constant, double gamma)
{
    const fabdim *fd = &L->fdg[b];
    const double zero_vec[3] = {0.0, 0.0, 0.0};
    const double theta0 = THETA0;
    for (int q = 0; q < ORC_NQ; q++) {
        const double wt = W[q];
        const int *ev = EV[q];
        L->f[b][fidx(fd, q, i, j, k)] =
            set_equilibrium_value(rho, vel, specific_gas_constant * temperature, ev);
        const double cv = specific_gas_constant / (gamma - 1.0);
        const double two_rho_e = get_energy_vec(temperature, rho, vel, cv);
        double heat_flux[3] = {0.0, 0.0, 0.0};
        double fl[6];
        get_equilibrium_moments_vec(
            rho, vel, two_rho_e, cv, specific_gas_constant, heat_flux, &fl[0], &fl[1],
            &fl[2], &fl[3], &fl[4], &fl[5]);
        L->g[b][fidx(fd, q, i, j, k)] = set_extended_grad_expansion_generic(
            two_rho_e, heat_flux, fl, wt, ev, theta0, zero_vec, 1.0);
    }
}

/* ic::TaylorGreen::DeviceOp::thermal_initialize (IC.H:119-153) with the
 * constructor IC.cpp:34-45: L = 1/pi, T = 1/3 and gamma = 5/3 are fixed. */
void orc_ic_taylorgreen(orc_level *L, double rho0, double v0, const double omega[3])
{
    const double Lw = 1.0 / M_PI;
    for (int b = 0; b < L->nbox; b++) {
        const fabdim *fd = &L->fdg[b];
#pragma omp parallel for num_threads(g_threads) schedule(static)
        for (int k = fd->lo[2]; k <= fd->hi[2]; k++)
            for (int j = fd->lo[1]; j <= fd->hi[1]; j++)
                for (int i = fd->lo[0]; i <= fd->hi[0]; i++) {
                    const double x = L->prob_lo[0] + (i + 0.5) * L->dx[0];
                    const double y = L->prob_lo[1] + (j + 0.5) * L->dx[1];
                    const double z = L->prob_lo[2] + (k + 0.5) * L->dx[2];
                    const double omega0 = omega[0];
                    const double omega1 = omega[1];
                    const double omega2 = omega[2];
                    const double rho = rho0 + rho0 * v0 * v0 / 16.0 *
                                                  (cos(2.0 * omega0 * x / Lw) +
                                                   cos(2.0 * omega1 * y / Lw)) *
                                                  (cos(2.0 * omega2 * z / Lw) + 2.0);
                    double vel[3];
                    vel[0] = v0 * sin(omega0 * x / Lw) * cos(omega1 * y / Lw) *
                             cos(omega2 * z / Lw);
                    vel[1] = -v0 * cos(omega0 * x / Lw) * sin(omega1 * y / Lw) *
                             cos(omega2 * z / Lw);
                    vel[2] = 0.0;
                    const double temperature = 1.0 / 3.0;
                    const double R = AIR_MOLAR_MASS / AIR_MOLAR_MASS;
                    const double gamma = 5.0 / 3.0;
                    init_cell(L, b, i, j, k, rho, vel, temperature, R, gamma);
                }
    }
}

/* ic::Constant (IC.H:13-69, IC.cpp:4-32): the velocity is always
 * mach_components * speed_of_sound_ref. */
void orc_ic_constant(
    orc_level *L, double density, const double mach[3], double temperature,
    double adiabatic_exponent, double m_bar)
{
    const double speed_of_sound_ref =
        sqrt(adiabatic_exponent * (AIR_MOLAR_MASS / m_bar) * temperature);
    double vel[3];
    for (int n = 0; n < 3; n++) vel[n] = mach[n] * speed_of_sound_ref;
    const double R = AIR_MOLAR_MASS / m_bar;
    for (int b = 0; b < L->nbox; b++) {
        const fabdim *fd = &L->fdg[b];
#pragma omp parallel for num_threads(g_threads) schedule(static)
        for (int k = fd->lo[2]; k <= fd->hi[2]; k++)
            for (int j = fd->lo[1]; j <= fd->hi[1]; j++)
                for (int i = fd->lo[0]; i <= fd->hi[0]; i++)
                    init_cell(L, b, i, j, k, density, vel, temperature, R,
                              adiabatic_exponent);
    }
}

/* Remainder of MakeNewLevelFromScratch (LBM.cpp:1186-1198) after the IC:
 * fill_f_inside_eb (1278-1298), FillBoundary f,g, zeroed macro fields,
 * f_to_macrodata, equilibria (with dQ = 0), dQ. */
void orc_init_finish(orc_level *L)
{
    for (int b = 0; b < L->nbox; b++) {
        const fabdim *fd = &L->fdg[b];
#pragma omp parallel for num_threads(g_threads) schedule(static)
        for (int k = fd->lo[2]; k <= fd->hi[2]; k++)
            for (int j = fd->lo[1]; j <= fd->hi[1]; j++)
                for (int i = fd->lo[0]; i <= fd->hi[0]; i++) {
                    if (L->isf[b][fidx(fd, 0, i, j, k)] != 0) continue;
                    for (int q = 0; q < ORC_NQ; q++) {
                        L->f[b][fidx(fd, q, i, j, k)] = 0.0;
                        L->g[b][fidx(fd, q, i, j, k)] = 0.0;
                    }
                }
    }
    orc_fill_boundary(L, ORC_FAB_F);
    orc_fill_boundary(L, ORC_FAB_G);
    for (int b = 0; b < L->nbox; b++) {
        memset(L->md[b], 0, sizeof(double) * (size_t)(ORC_NMD * L->fd1[b].ncell));
        memset(L->dq[b], 0, sizeof(double) * (size_t)(ORC_NDQ * L->fd0[b].ncell));
        memset(L->eq[b], 0, sizeof(double) * (size_t)(ORC_NQ * L->fd0[b].ncell));
        memset(L->eqg[b], 0, sizeof(double) * (size_t)(ORC_NQ * L->fd0[b].ncell));
    }
    orc_f_to_macrodata(L);
    orc_macrodata_to_equilibrium(L);
    orc_compute_q_corrections(L);
}

/* ic::Initializer::initialize (IC.H:460-525) from per-cell fields. */
void orc_ic_fields(
    orc_level *L, const double *rho, const double *u, const double *v, const double *w,
    const double *T, double adiabatic_exponent, double m_bar)
{
    const double R = AIR_MOLAR_MASS / m_bar;
    int64_t base = 0;
    for (int b = 0; b < L->nbox; b++) {
        const fabdim *fd = &L->fdg[b];
#pragma omp parallel for num_threads(g_threads) schedule(static)
        for (int k = fd->lo[2]; k <= fd->hi[2]; k++)
            for (int j = fd->lo[1]; j <= fd->hi[1]; j++)
                for (int i = fd->lo[0]; i <= fd->hi[0]; i++) {
                    const int64_t c = base + fidx(fd, 0, i, j, k);
                    const double vel[3] = {u[c], v[c], w[c]};
                    init_cell(L, b, i, j, k, rho[c], vel, T[c], R, adiabatic_exponent);
                }
        base += fd->ncell;
    }
}

/* LBM::compute_eb_forces (LBM.cpp:994-1044), one level with no finer level
 * (mask = 0): over valid cells with is_fluid comp1 == 1 (solid cells with a
 * fluid face neighbour), F += 2 c_q f_q(x - c_q) is_fluid(x - c_q, 0).
 * Summed in (box, k, j, i, q) order. */
void orc_eb_force(const orc_level *L, double out[3])
{
    out[0] = out[1] = out[2] = 0.0;
    for (int b = 0; b < L->nbox; b++) {
        const fabdim *fd = &L->fdg[b];
        const double *f = L->f[b];
        const int32_t *isf = L->isf[b];
        for (int k = L->vlo[b][2]; k <= L->vhi[b][2]; k++)
            for (int j = L->vlo[b][1]; j <= L->vhi[b][1]; j++)
                for (int i = L->vlo[b][0]; i <= L->vhi[b][0]; i++) {
                    if (isf[fidx(fd, 1, i, j, k)] != 1) continue;
                    double fs[3] = {0.0, 0.0, 0.0};
                    for (int q = 0; q < ORC_NQ; q++) {
                        const int *ev = EV[q];
                        const int *eb = EV[BOUNCE[q]];
                        const int a = i + eb[0], c = j + eb[1], e = k + eb[2];
                        for (int idir = 0; idir < 3; idir++)
                            fs[idir] += 2.0 * ev[idir] * f[fidx(fd, q, a, c, e)] *
                                        isf[fidx(fd, 0, a, c, e)];
                    }
                    out[0] += fs[0];
                    out[1] += fs[1];
                    out[2] += fs[2];
                }
    }
}

/* ------------------------------------------------------------------------ */
/* Two levels, refinement ratio 2.                                           */
/* ------------------------------------------------------------------------ */

static void amr_check(const orc_level *crse, const orc_level *fine)
{
    for (int d = 0; d < 3; d++) {
        if (fine->nd[d] != 2 * crse->nd[d] || fine->periodic[d] != crse->periodic[d]) {
            fprintf(stderr, "oracle: fine level is not the coarse domain refined by 2\n");
            abort();
        }
    }
    if (fine->ng < 2) {
        fprintf(stderr, "oracle: the fine level needs at least 2 ghost layers\n");
        abort();
    }
    for (int b = 0; b < fine->nbox; b++)
        for (int d = 0; d < 3; d++)
            if ((fine->vlo[b][d] & 1) || !(fine->vhi[b][d] & 1)) {
                fprintf(stderr, "oracle: fine box %d is not coarse-aligned\n", b);
                abort();
            }
}

/* average_down_with_ghosts (Utilities.cpp:5-28), crse_ng = 1, ratio 2.
 * Step 1 (cfine.ParallelCopy(crse)): every cfine cell starts as the coarse
 * value of its periodic image (the reference copies without periodicity
 * and leaves out-of-domain cfine ghosts uninitialised; they only matter
 * where all 8 children are masked).  Step 2: masked_avgdown in place.
 * Step 3 (crse.ParallelCopy(cfine, periodic)): cfine boxes are copied back
 * in box order, ghosts included, so the last box wins on overlaps. */
void orc_amr_coalesce(const orc_level *fine, orc_level *crse, int lattice)
{
    amr_check(crse, fine);
    double **fa = lattice == 0 ? fine->f : fine->g;
    double **ca = lattice == 0 ? crse->f : crse->g;
    fabdim *cf = calloc((size_t)fine->nbox, sizeof(fabdim));
    double **cv = calloc((size_t)fine->nbox, sizeof(double *));
    for (int b = 0; b < fine->nbox; b++) {
        int clo[3], chi[3];
        for (int d = 0; d < 3; d++) {
            clo[d] = fine->vlo[b][d] / 2;
            chi[d] = fine->vhi[b][d] / 2;
        }
        fabdim_set(&cf[b], clo, chi, 1);
        cv[b] = dalloc(ORC_NQ * cf[b].ncell, 0.0);
        const fabdim *cd = &cf[b];
        const fabdim *fd = &fine->fdg[b];
        const double *fv = fa[b];
        double *out = cv[b];
#pragma omp parallel for num_threads(g_threads) schedule(static)
        for (int k = cd->lo[2]; k <= cd->hi[2]; k++)
            for (int j = cd->lo[1]; j <= cd->hi[1]; j++)
                for (int i = cd->lo[0]; i <= cd->hi[0]; i++) {
                    const int iv[3] = {i, j, k};
                    int s[3];
                    const int has_image = periodic_image(crse, iv, s);
                    const int ob = has_image ? owner_of(crse, s) : -1;
                    for (int n = 0; n < ORC_NQ; n++) {
                        /* cfine.ParallelCopy(crse, ..., IntVect(0), ng) */
                        out[fidx(cd, n, i, j, k)] =
                            ob >= 0 ? ca[ob][fidx(&crse->fdg[ob], n, s[0], s[1], s[2])] : 0.0;
                        /* masked_avgdown (Utilities.H:315-350) */
                        double c = 0;
                        double vol = 0;
                        for (int kref = 0; kref < 2; ++kref)
                            for (int jref = 0; jref < 2; ++jref)
                                for (int iref = 0; iref < 2; ++iref) {
                                    const double fval = fv[fidx(
                                        fd, n, 2 * i + iref, 2 * j + jref, 2 * k + kref)];
                                    if (fabs(fval - (-1.0)) > SMALL_NUM) {
                                        c += fval;
                                        vol += 1.0;
                                    }
                                }
                        if (vol > 0.0) out[fidx(cd, n, i, j, k)] = c / vol;
                    }
                }
    }
    /* crse.ParallelCopy(cfine, ..., ng, IntVect(0), periodicity) */
    for (int b = 0; b < fine->nbox; b++) {
        const fabdim *cd = &cf[b];
        for (int k = cd->lo[2]; k <= cd->hi[2]; k++)
            for (int j = cd->lo[1]; j <= cd->hi[1]; j++)
                for (int i = cd->lo[0]; i <= cd->hi[0]; i++) {
                    const int iv[3] = {i, j, k};
                    int s[3];
                    if (!periodic_image(crse, iv, s)) continue;
                    const int ob = owner_of(crse, s);
                    if (ob < 0) continue;
                    for (int n = 0; n < ORC_NQ; n++)
                        ca[ob][fidx(&crse->fdg[ob], n, s[0], s[1], s[2])] =
                            cv[b][fidx(cd, n, i, j, k)];
                }
        free(cv[b]);
    }
    free(cv);
    free(cf);
}

/* Conservative linear interpolation with monotonised-central limited
 * slopes: child = c + s_x o_x + s_y o_y + s_z o_z, o = -1/4 (low child) or
 * +1/4 (high child) in coarse-cell units, so the mean of the 8 children is c
 * and linear coarse data is reproduced exactly.  Slope per direction:
 * s = sign(dc) min(|dc|, 2|dl|, 2|dr|) with dc = (c+ - c-)/2, dl = c - c-,
 * dr = c+ - c, and 0 where dl and dr differ in sign (local extremum) or a
 * neighbour is outside the domain. */
/* Coarse-to-fine interpolation of FillPatchTwoLevels / InterpFromCoarseLevel
 * with m_mapper = &amrex::cell_cons_interp (FillPatchOps.H:124-128,174-177,
 * 190): AMReX's CellConservativeLinear with do_linear_limiting = false, whose
 * published steps (AMReX MFInterp: mf_compute_slopes_{x,y,z},
 * mf_cell_cons_lin_interp_mcslope, mf_cell_cons_lin_interp) are restated
 * here for refinement ratio 2.  AMReX itself is the un-vendored submodule
 * Submodules/AMReX (commit not recorded), so the steps follow its published
 * source as described, not a copy of it:
 *  - per direction, the centred difference dc, replaced at an ext_dir face
 *    (every non-periodic face, LBM.cpp:108-150) by the one-sided second-order
 *    form through the ghost value (4-point form, or a 3-point fallback when
 *    the cell two further in is not available);
 *  - MC limiting: s = sign(dc) min(|dc|, min(|df|, |db|)) with df, db twice
 *    the forward and backward differences, 0 when they differ in sign;
 *  - one factor alpha <= 1 per coarse cell so that no child leaves the
 *    min/max of the 3x3x3 coarse neighbourhood: the largest child deviation
 *    dumax = sum_d |s_d| (r-1)/(2r) is held within umax - u and u - umin;
 *  - child = u + xoff s_x alpha + yoff s_y alpha + zoff s_z alpha with
 *    offsets -+1/4.
 * u[27]: the neighbourhood, u[(oz+1)*9 + (oy+1)*3 + (ox+1)]; u2[2d] the
 * value at P + 2 e_d (read at a low face), u2[2d+1] at P - 2 e_d (high
 * face); flags: bit d low ext_dir face in d, bit 3+d high face, bit 6+d
 * u2[2d] available, bit 9+d u2[2d+1] available.  slope[d] receives
 * s_d alpha. */
void orc_amr_slopes(const double u[27], const double u2[6], int flags, double slope[3])
{
    const double uc = u[13];
    double s[3];
    for (int d = 0; d < 3; d++) {
        const int st = d == 0 ? 1 : (d == 1 ? 3 : 9);
        const double um = u[13 - st], up = u[13 + st];
        double dc = 0.5 * (up - um);
        if (flags & (1 << d)) {
            if (flags & (1 << (6 + d)))
                dc = -16. / 15. * um + 0.5 * uc + 2. / 3. * up - 0.1 * u2[2 * d];
            else
                dc = 0.25 * (up + 5. * uc - 6. * um);
        }
        if (flags & (1 << (3 + d))) {
            if (flags & (1 << (9 + d)))
                dc = 16. / 15. * up - 0.5 * uc - 2. / 3. * um + 0.1 * u2[2 * d + 1];
            else
                dc = -0.25 * (um + 5. * uc - 6. * up);
        }
        const double df = 2.0 * (up - uc);
        const double db = 2.0 * (uc - um);
        double sl = (df * db >= 0.0) ? fmin(fabs(df), fabs(db)) : 0.0;
        s[d] = copysign(1.0, dc) * fmin(sl, fabs(dc));
    }
    double alpha = 1.0;
    if (s[0] != 0.0 || s[1] != 0.0 || s[2] != 0.0) {
        const double dumax = fabs(s[0]) * 1.0 / 4.0 + fabs(s[1]) * 1.0 / 4.0 + fabs(s[2]) * 1.0 / 4.0;
        double umax = uc, umin = uc;
        for (int n = 0; n < 27; n++) {
            umin = fmin(umin, u[n]);
            umax = fmax(umax, u[n]);
        }
        if (dumax * alpha > (umax - uc)) alpha = (umax - uc) / dumax;
        if (dumax * alpha > (uc - umin)) alpha = (uc - umin) / dumax;
    }
    for (int d = 0; d < 3; d++) slope[d] = s[d] * alpha;
}

double orc_amr_interp(const double u[27], const double u2[6], int flags, const int hi_child[3])
{
    double sl[3];
    orc_amr_slopes(u, u2, flags, sl);
    const double ox = hi_child[0] ? 0.25 : -0.25;
    const double oy = hi_child[1] ? 0.25 : -0.25;
    const double oz = hi_child[2] ? 0.25 : -0.25;
    return u[13] + ox * sl[0] + oy * sl[1] + oz * sl[2];
}

/* Coarse value of component q at cell c as the interpolation's coarse
 * patch holds it: the valid value of the cell or of its periodic image;
 * beyond a non-periodic face, -1.  There the patch (the coarsened fine
 * region grown by one cell) is filled by BCFill, whose `inside` box is the
 * patch shrunk by the level's 3 ghost cells (BC.H:375-380, m_ng =
 * m_f[0].nGrowVect(), LBM.cpp:1394); a ghost cell's neighbour iv + c_q lies
 * at most one cell inside the domain, never inside that box, so every
 * component gets the -1 dummy (BC.H:463-466). */
static double crse_patch_value(const orc_level *crse, double *const *ca, int q, const int c[3])
{
    int s[3];
    if (!periodic_image(crse, c, s)) return -1.0;
    const int ob = owner_of(crse, s);
    return ca[ob][fidx(&crse->fdg[ob], q, s[0], s[1], s[2])];
}

/* the neighbourhood, face flags and extra values of coarse parent P */
static void crse_neighbourhood(const orc_level *crse, double *const *ca, int q, const int P[3], double u[27],
                               double u2[6], int *flags)
{
    int fl = 0;
    for (int d = 0; d < 3; d++) {
        if (crse->periodic[d]) continue;
        if (P[d] == crse->dom_lo[d]) fl |= 1 << d;
        if (P[d] == crse->dom_hi[d]) fl |= 1 << (3 + d);
        if (P[d] + 2 <= crse->dom_hi[d]) fl |= 1 << (6 + d);
        if (P[d] - 2 >= crse->dom_lo[d]) fl |= 1 << (9 + d);
    }
    for (int oz = -1; oz <= 1; oz++)
        for (int oy = -1; oy <= 1; oy++)
            for (int ox = -1; ox <= 1; ox++) {
                const int c[3] = {P[0] + ox, P[1] + oy, P[2] + oz};
                u[(oz + 1) * 9 + (oy + 1) * 3 + (ox + 1)] = crse_patch_value(crse, ca, q, c);
            }
    for (int d = 0; d < 3; d++) {
        int a[3] = {P[0], P[1], P[2]}, b[3] = {P[0], P[1], P[2]};
        a[d] += 2;
        b[d] -= 2;
        u2[2 * d] = (fl & (1 << d)) && (fl & (1 << (6 + d))) ? crse_patch_value(crse, ca, q, a) : 0.0;
        u2[2 * d + 1] = (fl & (1 << (3 + d))) && (fl & (1 << (9 + d))) ? crse_patch_value(crse, ca, q, b) : 0.0;
    }
    *flags = fl;
}

void orc_amr_explode(const orc_level *crse, orc_level *fine, int lattice)
{
    amr_check(crse, fine);
    double **fa = lattice == 0 ? fine->f : fine->g;
    double **ca = lattice == 0 ? crse->f : crse->g;
    bounce_prefill(fine, fa);
    orc_fill_boundary(fine, lattice == 0 ? ORC_FAB_F : ORC_FAB_G);
    for (int b = 0; b < fine->nbox; b++) {
        const fabdim *fd = &fine->fdg[b];
#pragma omp parallel for num_threads(g_threads) schedule(static)
        for (int k = fd->lo[2]; k <= fd->hi[2]; k++)
            for (int j = fd->lo[1]; j <= fd->hi[1]; j++)
                for (int i = fd->lo[0]; i <= fd->hi[0]; i++) {
                    if (valid_contains(fine, b, i, j, k)) continue;
                    const int iv[3] = {i, j, k};
                    int s[3];
                    if (!periodic_image(fine, iv, s)) continue; /* physbc */
                    if (owner_of(fine, s) >= 0) continue;       /* FillBoundary */
                    const int P[3] = {s[0] / 2, s[1] / 2, s[2] / 2};
                    if (owner_of(crse, P) < 0) {
                        fprintf(stderr, "oracle: coarse level does not cover its domain\n");
                        abort();
                    }
                    const int hi_child[3] = {s[0] & 1, s[1] & 1, s[2] & 1};
                    for (int q = 0; q < ORC_NQ; q++) {
                        double u[27], u2[6];
                        int fl;
                        crse_neighbourhood(crse, ca, q, P, u, u2, &fl);
                        fa[b][fidx(fd, q, i, j, k)] = orc_amr_interp(u, u2, fl, hi_child);
                    }
                }
    }
    orc_physbc(fine, lattice);
}

/* LBM::RemakeLevel (LBM.cpp:1302-1365) for the fine level: FillPatchTwoLevels
 * (FillPatchOps.H:118-129) of the new level's valid cells -- the old fine
 * level's value where it has the cell, else the coarse interpolation of
 * orc_amr_explode -- then fill_f_inside_eb (0 in every solid fab cell).
 * The new level's flags are set. */
void orc_amr_regrid_fill(const orc_level *crse, const orc_level *old_fine, orc_level *fine)
{
    amr_check(crse, fine);
    for (int lattice = 0; lattice < 2; lattice++) {
        double **fa = lattice == 0 ? fine->f : fine->g;
        double *const *oa = lattice == 0 ? old_fine->f : old_fine->g;
        double *const *ca = lattice == 0 ? crse->f : crse->g;
        for (int b = 0; b < fine->nbox; b++) {
            const fabdim *fd = &fine->fdg[b];
            for (int k = fine->vlo[b][2]; k <= fine->vhi[b][2]; k++)
                for (int j = fine->vlo[b][1]; j <= fine->vhi[b][1]; j++)
                    for (int i = fine->vlo[b][0]; i <= fine->vhi[b][0]; i++) {
                        const int s[3] = {i, j, k};
                        const int ob = owner_of(old_fine, s);
                        if (ob >= 0) {
                            const fabdim *od = &old_fine->fdg[ob];
                            for (int q = 0; q < ORC_NQ; q++)
                                fa[b][fidx(fd, q, i, j, k)] = oa[ob][fidx(od, q, i, j, k)];
                            continue;
                        }
                        const int P[3] = {i / 2, j / 2, k / 2};
                        const int hi_child[3] = {s[0] & 1, s[1] & 1, s[2] & 1};
                        for (int q = 0; q < ORC_NQ; q++) {
                            double u[27], u2[6];
                            int fl;
                            crse_neighbourhood(crse, ca, q, P, u, u2, &fl);
                            fa[b][fidx(fd, q, i, j, k)] = orc_amr_interp(u, u2, fl, hi_child);
                        }
                    }
            const int32_t *isf = fine->isf[b];
            for (int k = fd->lo[2]; k <= fd->hi[2]; k++)
                for (int j = fd->lo[1]; j <= fd->hi[1]; j++)
                    for (int i = fd->lo[0]; i <= fd->hi[0]; i++)
                        if (isf[fidx(fd, 0, i, j, k)] == 0)
                            for (int q = 0; q < ORC_NQ; q++) fa[b][fidx(fd, q, i, j, k)] = 0.0;
        }
    }
}

/* LBM::evolve -> time_step(0) -> [fillpatch(1); 2 x (physbc(1); time_step(1)
 * -> advance(1))] -> advance(0) with average_down_to (LBM.cpp:416-419,
 * 498-512, 523-544). */
void orc_amr_evolve(orc_level *crse, orc_level *fine, int nsteps)
{
    amr_check(crse, fine);
    for (int s = 0; s < nsteps; s++) {
        orc_fillpatch(crse, 0);
        orc_fillpatch(crse, 1);
        orc_amr_explode(crse, fine, 0);
        orc_amr_explode(crse, fine, 1);
        for (int sub = 0; sub < 2; sub++) {
            orc_physbc(fine, 0);
            orc_physbc(fine, 1);
            orc_stream(fine, 0);
            orc_stream(fine, 1);
            orc_collide(fine);
        }
        orc_stream(crse, 0);
        orc_stream(crse, 1);
        orc_amr_coalesce(fine, crse, 0);
        orc_amr_coalesce(fine, crse, 1);
        orc_collide(crse);
    }
}
