/*
 * CPU oracle for the MARBLES D3Q27 thermal lattice-Boltzmann hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * solver's per-step algorithm (NREL/marbles, Source/LBM.cpp, Utilities.H,
 * BC.H, FillPatchOps.H, IC.H, Stencil.H), written so that the GPU engine in
 * marbles_amd/ can be checked against it.  Only tests/, the smoke check and
 * the CPU-baseline leg of bench.py may load it.  The product path never
 * links, loads or calls anything in this directory.
 *
 * Storage follows the reference's AMReX FAB convention: every box owns a
 * fab covering its valid region grown by `ng` ghost cells, component-major
 * ("SoA"): data[((c*nz + k)*ny + j)*nx + i] with (i,j,k) counted from the
 * fab's low corner.
 *
 * Parity pins for each part are listed in tests/test_oracle_*.py.  Parts that
 * no reference output can pin (AMReX internals: EB2 sphere classification
 * ties, coarse/fine interpolation, ParallelCopy overlap order) are marked
 * "parity unpinned" where they are defined below.
 */
#ifndef MARBLES_LBM_ORACLE_H
#define MARBLES_LBM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_NQ 27
#define ORC_NMD 19
#define ORC_NDQ 3

/* Boundary types, Source/BC.H:13-20. */
#define ORC_BC_PERIODIC 0
#define ORC_BC_NOSLIP 1
#define ORC_BC_VELOCITY 2
#define ORC_BC_PRESSURE 3
#define ORC_BC_OUTFLOW 5
#define ORC_BC_SLIPX 6
#define ORC_BC_SLIPY 7
#define ORC_BC_SLIPZ 8

/* What orc_get_fab / orc_set_fab address. */
#define ORC_FAB_F 0
#define ORC_FAB_G 1
#define ORC_FAB_MD 2
#define ORC_FAB_DQ 3
#define ORC_FAB_ISFLUID 4 /* int32, 2 comps */

typedef struct orc_level orc_level;

/* Physical parameters of one level (Source/LBM.H:272-318, LBM.cpp:274-298). */
typedef struct {
    double nu;                 /* lbm.nu */
    double alpha;              /* lbm.alpha (defaults to nu) */
    double adiabatic_exponent; /* lbm.adiabatic_exponent */
    double m_bar;              /* lbm.mean_molecular_mass */
    double dt;                 /* m_dts[lev] */
    double mesh_speed;         /* dx_outer / dt_outer */
} orc_phys;

/* Physical-boundary description (Source/BC.H:22-63, VelocityBC.H:235-281). */
typedef struct {
    int bc_type[6];   /* m_bc_type: index idir + lohi*3 */
    int velbc_kind;   /* 0 = "noop", 1 = "constant", 2 = "channel", 3 = "parabolic" */
    int vel_dir;      /* velocity_bc_constant.dir (reference default 1); parabolic: tangential_dir */
    double mach_ref;  /* velocity_bc_<kind>.Mach_ref */
    double density;   /* velocity_bc_<kind>.initial_density */
    double temperature;
    double adiabatic_exponent;
    double m_bar;
    int normal_dir;   /* velocity_bc_parabolic.normal_dir (reference default 0) */
} orc_bc;

/* ---- level lifetime ---------------------------------------------------- */
orc_level *orc_create(
    const int dom_lo[3], const int dom_hi[3], const int periodic[3],
    const double prob_lo[3], const double prob_hi[3], int nbox,
    const int *box_lo, const int *box_hi, int ng);
void orc_destroy(orc_level *L);
void orc_set_phys(orc_level *L, const orc_phys *p);
void orc_set_bc(orc_level *L, const orc_bc *bc);
int orc_nbox(const orc_level *L);
/* number of doubles (or int32) in one box's fab of the given kind */
int64_t orc_fab_size(const orc_level *L, int kind, int box);
void orc_get_fab(const orc_level *L, int kind, int box, void *out);
void orc_set_fab(orc_level *L, int kind, int box, const void *in);

/* ---- initialisation (LBM.cpp:1148-1199, 1201-1298; IC.H:460-525) -------- */
void orc_init_flags_all_regular(orc_level *L);
void orc_init_flags_sphere(
    orc_level *L, const double center[3], double radius, int has_fluid_inside);
void orc_init_flags_cylinder(
    orc_level *L, const double center[3], double radius, double height, int direction,
    int has_fluid_inside);
void orc_init_flags_from_mask(orc_level *L, const uint8_t *solid_mask);
void orc_ic_taylorgreen(orc_level *L, double rho0, double v0, const double omega[3]);
void orc_ic_constant(
    orc_level *L, double density, const double mach[3], double temperature,
    double adiabatic_exponent, double m_bar);
/* ic::Initializer::initialize (IC.H:460-525) from per-cell macroscopic
 * fields given for every fab cell of every box (box order, x fastest):
 * f = f_eq^prod(rho, u, R T), g = Grad equilibrium. */
void orc_ic_fields(
    orc_level *L, const double *rho, const double *u, const double *v, const double *w,
    const double *T, double adiabatic_exponent, double m_bar);
/* fill_f_inside_eb, FillBoundary, f_to_macrodata, equilibria, dQ */
void orc_init_finish(orc_level *L);

/* ---- the per-step pieces ---------------------------------------------- */
void orc_fill_boundary(orc_level *L, int kind);      /* AMReX FillBoundary */
void orc_fillpatch(orc_level *L, int lattice);      /* FillPatchOps::fillpatch, lev 0 */
void orc_physbc(orc_level *L, int lattice);         /* FillPatchOps::physbc */
void orc_stream(orc_level *L, int lattice);         /* LBM::stream */
void orc_f_to_macrodata(orc_level *L);
void orc_compute_q_corrections(orc_level *L);
void orc_macrodata_to_equilibrium(orc_level *L);
void orc_relax_f_to_equilibrium(orc_level *L);
void orc_collide(orc_level *L);                     /* LBM::collide */
/* single-level LBM::evolve: nsteps x (fillpatch f,g; stream f,g; collide) */
void orc_evolve(orc_level *L, int nsteps);
/* conservation sums over valid fluid cells: mass, x/y/z momentum, 2rhoE */
void orc_conservation(const orc_level *L, double out[5]);
/* LBM::compute_eb_forces (LBM.cpp:994-1044) on one level: momentum-exchange
 * force on the embedded boundary from the current (post-collision) f */
void orc_eb_force(const orc_level *L, double out[3]);
/* LBM::compute_derived (LBM.cpp:909-955): vorticity (x, y, z, magnitude) of
 * the current populations' velocity over every box's valid cells (per box:
 * 4 components, component-major, x fastest); recomputes the macrodata. */
void orc_vorticity(orc_level *L, double *out);
/* LBM::plot_file_mf (LBM.cpp:1629-1669) of the level's current data: per
 * box, the valid cells' [19 macrodata, (f, g), (vorticity, dQCorr),
 * is_fluid, eb_boundary] component-major, boxes back to back. */
void orc_plot_vars(const orc_level *L, int streaming, int derived, double *out);

/* ---- two levels, refinement ratio 2 -------------------------------------
 * The coarse level must cover its domain; the fine level's domain is the
 * coarse one refined by 2, its boxes are coarse-aligned (even lo, odd hi)
 * and it has ng >= 2 ghost layers. */
/* average_down_with_ghosts (Utilities.cpp:5-28) with ng = 1, ratio 2:
 * coarse cells under the fine boxes and one coarse ring around each become
 * the masked mean of their 8 fine children (masked_avgdown, Utilities.H:
 * 315-350).  Where the rings of several fine boxes overlap, the box latest
 * in the box list wins (AMReX leaves the order of overlapping ParallelCopy
 * sources unspecified). */
void orc_amr_coalesce(const orc_level *fine, orc_level *crse, int lattice);
/* FillPatchOps::fillpatch, two-level branch (FillPatchOps.H:118-129):
 * fine ghost cells covered by fine boxes are copied (FillBoundary), fine
 * ghost cells over uncovered domain are interpolated from the coarse level
 * (own conservative, MC-limited linear scheme in place of AMReX
 * cell_cons_interp), ghost cells beyond non-periodic faces get physbc. */
void orc_amr_explode(const orc_level *crse, orc_level *fine, int lattice);
/* LBM::RemakeLevel's fill of a new fine level (valid cells from the old
 * fine level where it has them, else the coarse interpolation; 0 in solid
 * cells); the new level's flags must be set. */
void orc_amr_regrid_fill(const orc_level *crse, const orc_level *old_fine, orc_level *fine);
/* the interpolation alone (AMReX cell_cons_interp, ratio 2) for one coarse
 * parent: u[27] its 3x3x3 neighbourhood (u[(oz+1)*9+(oy+1)*3+(ox+1)]; a
 * cell beyond a non-periodic face holds -1), u2[2d] / u2[2d+1] the values
 * at P + 2e_d / P - 2e_d, flags bit d / 3+d: P at the low / high ext_dir
 * face in d, bit 6+d / 9+d: u2[2d] / u2[2d+1] available.  orc_amr_slopes
 * gives the limited slopes (MC slopes times the min/max factor alpha);
 * orc_amr_interp the child (hi_child[d] = 0 low, 1 high). */
void orc_amr_slopes(const double u[27], const double u2[6], int flags, double slope[3]);
double orc_amr_interp(const double u[27], const double u2[6], int flags, const int hi_child[3]);
/* nsteps coarse steps of LBM::evolve/time_step/advance (LBM.cpp:398-544)
 * for a two-level hierarchy; the fine level advances two substeps per
 * coarse step with its own phys (dt halved). */
void orc_amr_evolve(orc_level *crse, orc_level *fine, int nsteps);

/* ---- per-cell math, exposed so tests can pin it directly ---------------- */
void orc_stencil(int *evs /*27*3*/, int *bounce /*27*/, double *w /*27*/);
/* slip-wall reflection table of direction d (bounce_Xdir/Ydir/Zdir) */
void orc_bounce_dir(int d, int *out /*27*/);
double orc_set_equilibrium_value(double rho, const double vel[3], double rt, int q);
double orc_set_extended_equilibrium_value(
    double rho, const double vel[3], double pxx, double pyy, double pzz, int q);
double orc_grad_expansion(
    double rho, const double mom[3], const double flux[6], int q);
void orc_get_equilibrium_moments(
    double rho, const double vel[3], double total_energy, double cv, double R,
    double qeq[3], double req[6]);
double orc_get_energy_vec(double T, double rho, const double vel[3], double cv);
double orc_get_temperature_scalar(
    double two_rho_e, double rho, double ux, double uy, double uz, double cv);
double orc_get_temperature_vec(double two_rho_e, double rho, const double vel[3], double cv);
double orc_masked_avgdown8(const double v[8]);
int orc_sphere_covered(
    const double node_lo[3], const double dx[3], const double center[3],
    double radius, int has_fluid_inside);
int orc_cylinder_covered(
    const double node_lo[3], const double dx[3], const double center[3],
    double radius, double height, int direction, int has_fluid_inside);

/* OpenMP threads used by the oracle (1 = serial). */
void orc_set_threads(int n);
int orc_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
