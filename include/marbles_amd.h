/*
 * marbles_amd: C ABI of the MI355X (gfx950) lattice-Boltzmann engine.
 *
 * The engine replaces the per-level hot path of NREL/marbles: the D3Q27
 * thermal collide (f and the energy populations g), the pull-stream with
 * halfway bounce-back, the ghost-cell fill, and the coarse/fine coalescence.
 * Each entry point names the reference interface it replaces (paths are in
 * the NREL/marbles tree).
 *
 * Conventions
 *  - Plain C: opaque handles, plain pointers and sizes.  Every function
 *    returns 0 on success or a negative MRB_E* code; mrb_last_error()
 *    returns the message of the last failure on the calling thread.
 *  - `stream` is a hipStream_t passed as void* (NULL = default stream).
 *    Calls are asynchronous on that stream unless stated otherwise and are
 *    not thread-safe per handle.
 *  - Host arrays use the reference's FAB convention: for each box (in the
 *    order given at creation) a fab covering the valid box grown by `ng`
 *    cells, component-major: data[((c*nz + k)*ny + j)*nx + i].
 *  - No entry point allocates device memory except mrb_level_create (and
 *    mrb_level_destroy, which frees it).
 */
#ifndef MARBLES_AMD_H
#define MARBLES_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MRB_OK 0
#define MRB_E_ARG -1      /* invalid argument */
#define MRB_E_HIP -2      /* HIP runtime error */
#define MRB_E_UNSUPPORTED -3
#define MRB_E_STATE -4    /* call not valid for this level's layout/state */

/* Storage layouts of a level. */
#define MRB_LAYOUT_FAB 0     /* reference layout: boxes with ng ghost cells */
#define MRB_LAYOUT_COMPACT 1 /* one box == the domain, no ghosts; fused path */

/* Population storage precision. */
#define MRB_STORAGE_F64 0         /* FP64 */
#define MRB_STORAGE_F32_SHIFTED 1 /* FP32 of f - f_eq(rho0, 0, T0) */
#define MRB_STORAGE_F16_SCALED 2  /* FP16 of (f - shift) * 2^e, e per (box, lattice);
                                     compact layout, boxes = z-slabs of 8 planes */

/* Boundary types, identical values to Source/BC.H:13-20. */
#define MRB_BC_PERIODIC 0
#define MRB_BC_NOSLIP 1
#define MRB_BC_VELOCITY 2
#define MRB_BC_PRESSURE 3 /* pressure_bc_f/g (BC.H:222-321); mrb_set_bc requires the face's
                             tangential axes periodic or NOSLIP on both sides, and two
                             cells or more along its own axis */
#define MRB_BC_OUTFLOW 5
#define MRB_BC_SLIPX 6
#define MRB_BC_SLIPY 7
#define MRB_BC_SLIPZ 8

/* Which array mrb_upload / mrb_download address. */
#define MRB_ARR_F 0
#define MRB_ARR_G 1
#define MRB_ARR_ISFLUID 2 /* int32, 2 comps (is_fluid, eb_boundary) */

typedef struct mrb_level mrb_level;

/* Level geometry and storage (Source/LBM.cpp:1148-1199 defines the level
 * data: f, g with 27 comps and 3 ghosts, is_fluid with 2 comps). */
typedef struct {
    int ncell[3];       /* domain size in cells; the domain is [0, ncell-1] */
    int periodic[3];    /* geometry.is_periodic */
    double prob_lo[3];  /* geometry.prob_lo */
    double prob_hi[3];  /* geometry.prob_hi */
    int nbox;           /* number of boxes */
    const int *box_lo;  /* nbox*3 valid-box low corners */
    const int *box_hi;  /* nbox*3 valid-box high corners (inclusive) */
    int ng;             /* ghost layers (MRB_LAYOUT_FAB); 0 for COMPACT */
    int layout;         /* MRB_LAYOUT_* */
    int nlattice;       /* 2 = f and g (thermal model) */
    int storage;        /* MRB_STORAGE_* */
} mrb_level_desc;

/* Physics of one level (Source/LBM.cpp:274-298, 639-644, 1047-1077). */
typedef struct {
    double nu;                 /* lbm.nu */
    double alpha;              /* lbm.alpha */
    double adiabatic_exponent; /* lbm.adiabatic_exponent */
    double m_bar;              /* lbm.mean_molecular_mass */
    double dt;                 /* level time step (1 on level 0) */
    double mesh_speed;         /* lbm.dx_outer / lbm.dt_outer */
} mrb_phys;

/* Physical boundaries (Source/BC.H:22-63, 474-504; the "constant" velocity
 * operator Source/VelocityBC.H:235-281 with VelocityBC.cpp:6-21). */
typedef struct {
    int bc_type[6];     /* m_bc_type: index idir + lohi*3 */
    int velbc_kind;     /* 0 = "noop", 1 = "constant", 2 = "channel" (VelocityBC.H:90-129),
                           3 = "parabolic" (VelocityBC.H:139-192) */
    int vel_dir;        /* velocity_bc_constant.dir; parabolic: tangential_dir */
    double mach_ref;    /* velocity_bc_<kind>.Mach_ref */
    double density;     /* velocity_bc_<kind>.initial_density */
    double temperature; /* velocity_bc_<kind>.initial_temperature */
    double adiabatic_exponent;
    double m_bar;
    int normal_dir;     /* velocity_bc_parabolic.normal_dir */
} mrb_bc;

/* ---- lifetime ----------------------------------------------------------
 * Replaces the MultiFab definitions of LBM::MakeNewLevelFromScratch
 * (Source/LBM.cpp:1159-1178).  Allocates A/B population buffers, flags and
 * scratch.  Synchronous. */
int mrb_level_create(const mrb_level_desc *desc, mrb_level **out);
int mrb_level_destroy(mrb_level *lvl);
int mrb_set_phys(mrb_level *lvl, const mrb_phys *phys);
int mrb_set_bc(mrb_level *lvl, const mrb_bc *bc);
/* Reference shift state used by the reduced-precision storages:
 * rho0, T0 (f_eq and g_eq at rest).  Defaults: 1, 1/3.  Set it to the
 * flow's reference density and temperature before the initial state: the
 * stored values are the differences to it, and their rounding error scales
 * with their size.  On a reduced-precision level with populations set, a
 * call (or an mrb_set_phys) that would change the shift returns
 * MRB_E_STATE. */
int mrb_set_shift_state(mrb_level *lvl, double rho0, double T0);
/* Element count of one box's host fab for `arr`. */
int64_t mrb_fab_size(const mrb_level *lvl, int arr, int box);

/* ---- data movement (synchronous) ----------------------------------------
 * Host arrays hold all boxes' fabs back to back (box order of creation). */
int mrb_upload(mrb_level *lvl, int arr, const void *host);
int mrb_download(const mrb_level *lvl, int arr, void *host);

/* ---- flags (Source/LBM.cpp:1213-1262; EB sphere through Source/EB.cpp:5-38)
 * After mrb_upload(MRB_ARR_ISFLUID) or mrb_compute_flags_sphere /
 * mrb_compute_flags_cylinder the level derives its stream masks.  A cell is
 * solid when all 8 of its corner nodes lie in the body (a node on the
 * surface counts as body).  Synchronous. */
int mrb_compute_flags_sphere(
    mrb_level *lvl, const double center[3], double radius, int has_fluid_inside);
/* EB2 geom_type "cylinder" (Source/EB.cpp:28-37, the eb2.cylinder_radius,
 * _height, _direction, _center, _has_fluid_inside inputs of
 * Tests/test_files/single_cylinder/single_cylinder.inp:38-43): a cylinder
 * along axis `direction` (0, 1, 2) of the given radius, centred at
 * `center` and `height` long. */
int mrb_compute_flags_cylinder(
    mrb_level *lvl, const double center[3], double radius, double height, int direction,
    int has_fluid_inside);
int mrb_finalize_flags(mrb_level *lvl);

/* ---- initial state (Source/IC.H:460-525, IC.cpp:4-45; LBM.cpp:1278-1298)
 * Host-evaluated (FP64 libm) and uploaded: f = product equilibrium,
 * g = Grad equilibrium, then zero inside the EB.  Synchronous. */
int mrb_init_taylorgreen(mrb_level *lvl, double rho0, double v0, const double omega[3]);
int mrb_init_constant(
    mrb_level *lvl, double density, const double mach[3], double temperature,
    double adiabatic_exponent, double m_bar);
/* Initial state from per-cell macroscopic fields, the data a user
 * ic::InitializerBase plugin (Source/IC.H:427-531) computes per cell: rho,
 * velocity (u, v, w) and temperature for every fab cell of every box (box
 * order, x fastest; mrb_fab_size(lvl, MRB_ARR_ISFLUID, b) / 2 cells per
 * box), then f = f_eq(rho, u, R T), g = Grad equilibrium, zero inside the EB
 * (IC.H:460-525, LBM.cpp:1278-1298).  Synchronous. */
int mrb_init_fields(
    mrb_level *lvl, const double *rho, const double *u, const double *v, const double *w,
    const double *T, double adiabatic_exponent, double m_bar);

/* ---- reference-layout stages (MRB_LAYOUT_FAB) ----------------------------
 * FillPatchOps::fillpatch at level 0 (Source/FillPatchOps.H:75-132):
 * FillBoundary across boxes and periodic images, then the physical
 * boundary functor BCFill::operator() (Source/BC.H:344-471). */
int mrb_fill_ghosts(mrb_level *lvl, int lattice, void *stream);
/* FillPatchOps::physbc (Source/FillPatchOps.H:142-159) only. */
int mrb_physbc(mrb_level *lvl, int lattice, void *stream);
/* LBM::stream (Source/LBM.cpp:558-604): push-equivalent pull over every
 * fab cell with halfway bounce-back and the -1 sentinel, then
 * FillBoundary.  Bit-exact with the reference. */
int mrb_stream(mrb_level *lvl, int lattice, void *stream);
/* LBM::collide (Source/LBM.cpp:607-618): f_to_macrodata (810-906),
 * compute_q_corrections (959-991), macrodata_to_equilibrium (621-762),
 * relax_f_to_equilibrium (765-807) and its FillBoundary. */
int mrb_collide(mrb_level *lvl, void *stream);

/* ---- hot path (MRB_LAYOUT_COMPACT) ---------------------------------------
 * One full single-level step (LBM::evolve body, Source/LBM.cpp:416-420 for
 * one level): ghost fill, stream f and g, collide -- fused, with periodic
 * wrap and physical boundaries resolved inside the kernel.  Swaps A/B.
 * Steps of one level are ordered by the caller (one stream, or events):
 * besides the A/B buffers they share the level's Q-correction array and,
 * for the single-pass sweeps, its ring of published planes and its launch
 * count. */
int mrb_stream_collide(mrb_level *lvl, void *stream);
/* nsteps fused steps. */
int mrb_advance(mrb_level *lvl, int nsteps, void *stream);
/* nsteps fused steps with HIP events recorded on `stream` around each kernel
 * of the step; kernel_ms[k] receives the summed duration of kernel slot k
 * (slots are named by mrb_hot_kernel_name, separated by '+'), *nslots the
 * number of slots (at most 4).  Synchronous.  Compact layout only. */
int mrb_advance_timed(mrb_level *lvl, int nsteps, void *stream, double *kernel_ms, int *nslots);
/* Kernel variants of the compact-layout step.
 *   0  two-pass step (Q-correction pass + collide pass) on 64x4-cell blocks,
 *      the compiler's register budget;
 *   1  the same with a 3-waves-per-SIMD budget;
 *   2  k_onepass: one pass that recomputes the neighbours' Q_corr;
 *   3  the two-pass step on 64x8-cell blocks;
 *   4  the two-pass step on 64x2-cell blocks;
 *   5  scaled FP16 storage only, no PRESSURE faces: the collide pass in
 *      FP32 arithmetic (FP64 sums of rho, E and the equilibria; the mass
 *      and energy residuals go to the rest direction);
 *   6  the two-pass step on 128x2-cell blocks;
 *   7  the two-pass step on 256x1-cell blocks;
 *   8  variant 5 on 256x1-cell blocks;
 *   9  shifted FP32 storage only, no PRESSURE faces: the collide pass in
 *      FP32 arithmetic on the stored deviations from the rest state (FP64
 *      moments and equilibrium coefficients; mass and energy residuals on
 *      the rest direction, momentum residuals on the axis pairs),
 *      64x2-cell blocks;
 *  10  variant 9 on 256x1-cell blocks;
 *  11  k_sweep: FP64 or shifted FP32 storage with rows of at most 256
 *      cells, one pass as a z-sweep with one resident block per x-row
 *      (MRB_E_UNSUPPORTED when the kernel's occupancy does not let every
 *      row be resident at once);
 *  12  k_sweep_x64: the same z-sweep on single-wave blocks of 64 cells of
 *      a row (any row length; MRB_E_UNSUPPORTED when the blocks do not fit
 *      the GPU at once);
 *  13  k_sweep_f32: shifted FP32 storage only, no PRESSURE faces: the
 *      z-sweep of variant 11 with the collide of variant 9;
 *  14+ the single-pass LDS z-sweep with different tile shapes.
 * The FP64-arithmetic variants give the same populations bit for bit; 5
 * and 8 give the same as each other, and 9, 10 and 13.  A level
 * starts on variant 7 (FP64 storage), 9 (shifted FP32) or 5 (scaled FP16).
 * While a level has a PRESSURE face, variants 5 and 9 are replaced by 0 and
 * 4, 8 and 10 by 7, 13 by 11 (mrb_set_bc), and restored when the faces
 * change back.
 * A variant that does not support the level returns MRB_E_UNSUPPORTED from
 * the step. */
int mrb_hot_kernel_count(void);
const char *mrb_hot_kernel_variant_name(int index);
int mrb_set_hot_kernel(mrb_level *lvl, int index);
/* Autotuner: times every variant (`reps` launches each) on the level's
 * current state without advancing it -- each trial reads the current
 * buffers and writes the spare ones -- and selects the fastest.  Compact
 * layout: the kernel variants above (a variant that does not support the
 * level reports -1; on reduced-precision levels only the variants with
 * the current variant's collide arithmetic, FP32 for 5, 8, 9 and 10 and
 * FP64 for the others, are tried, and the FP16 saturation count is left
 * as it was).
 * Fab layout: the persistent-block count of the batched
 * two-pass step over the box work-list.  ms_out[i] (may be NULL, else
 * mrb_autotune_count(lvl) entries) receives variant i's mean time and *best
 * its index.  Synchronous. */
int mrb_autotune(mrb_level *lvl, int reps, void *stream, double *ms_out, int *best);
int mrb_autotune_count(const mrb_level *lvl);
const char *mrb_autotune_name(const mrb_level *lvl, int index);
/* Stream only (debug/parity; LBM::stream for f and g on the compact
 * layout, ghosts resolved inline).  Swaps A/B. */
int mrb_stream_compact(mrb_level *lvl, void *stream);

/* ---- reductions ------------------------------------------------------------
 * FP64 sums over valid fluid cells: mass, x/y/z momentum, 2*rho*E.
 * Synchronous (waits for `stream`). */
int mrb_reduce_conservation(const mrb_level *lvl, double out[5], void *stream);
/* The same five sums without any allocation: no counterpart in the
 * reference; for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops).  The per-block partial sums go to a
 * static device array of the library (5 x 4096 doubles, one copy per GPU,
 * held until the process ends; the conservation kernel strides over any
 * level size with at most 4096 blocks, so no level exceeds it) and are
 * added on the device in the order mrb_reduce_conservation adds them, so
 * the sums are bit-identical.  out_device: 5 doubles of device memory on
 * the current device (the level's).  Asynchronous on `stream`; nothing is
 * allocated, freed or waited for.  Calls that overlap on one GPU share the
 * array and race: order them on one stream. */
int mrb_reduce_conservation_noalloc(const mrb_level *lvl, double *out_device, void *stream);
/* LBM::compute_eb_forces (Source/LBM.cpp:994-1044) on one level (no finer
 * level masking): momentum-exchange force on the embedded boundary,
 * F = sum over solid cells with a fluid face neighbour and over q of
 * 2 c_q f_q(x - c_q) is_fluid(x - c_q), from the current populations.  FP64
 * sums; synchronous.  Compact layout: sources beyond a non-periodic face
 * contribute nothing. */
int mrb_eb_force(const mrb_level *lvl, double out[3], void *stream);

/* ---- two levels, refinement ratio 2 (MRB_LAYOUT_FAB, FP64) ---------------
 * The coarse level covers its domain; the fine level's domain is the coarse
 * one refined by 2 (same physical extent and periodicity), its boxes are
 * coarse-aligned (even lo, odd hi) and it has ng >= 2.  mrb_amr_link builds
 * the transfer maps once (host work proportional to the fab cells) and keeps
 * them with the fine level, which must not outlive the coarse one.
 * Relinking replaces the maps (after a regrid). */
int mrb_amr_link(mrb_level *crse, mrb_level *fine);
/* LBM::average_down_to (Source/LBM.cpp:1571-1584) ->
 * average_down_with_ghosts (Source/Utilities.cpp:5-28, Utilities.H:353-358)
 * with crse_ng = 1: coarse cells under the fine boxes and one coarse ring
 * around each become the masked mean of their 8 fine children
 * (masked_avgdown, Utilities.H:315-350; unchanged when all 8 are -1).
 * Where rings of several fine boxes overlap, the box latest in the fine box
 * list wins.  Both lattices; asynchronous on `stream`. */
int mrb_coalesce(const mrb_level *fine, mrb_level *crse, void *stream);
/* FillPatchOps::fillpatch on a fine level (Source/FillPatchOps.H:66-130,
 * FillPatchTwoLevels branch): FillBoundary among fine boxes (every ghost
 * layer), fine ghost cells over uncovered domain interpolated from the
 * coarse level (AMReX cell_cons_interp, FillPatchOps.H:190, restated:
 * monotonised-central slopes, one-sided at the non-periodic faces the
 * reference marks ext_dir, one factor per coarse cell keeping the children
 * within the 3x3x3 neighbourhood's min/max), then the physical boundary
 * functor.  Asynchronous. */
int mrb_explode(const mrb_level *crse, mrb_level *fine, int lattice, void *stream);
/* nsteps coarse steps of LBM::evolve -> time_step -> advance
 * (Source/LBM.cpp:398-544) on a linked pair: coarse ghost fill; explode
 * f, g; twice (fine physbc, fine stream f, g, fine collide); coarse
 * stream f, g; coalesce; coarse collide.  The fine level's dt must be half
 * the coarse one (compute_dt, LBM.cpp:1047-1077).  Inside the step the fine
 * FillBoundary refreshes the first ghost layer and the deeper ghosts that
 * a step kernel reads (next to uncovered ghosts or non-periodic faces);
 * other deeper fine ghosts keep older values until mrb_fill_ghosts or
 * mrb_explode fills every layer.  Valid cells are unaffected.
 * Asynchronous. */
int mrb_amr_advance(mrb_level *crse, mrb_level *fine, int nsteps, void *stream);
/* Regrid (AmrCore::regrid -> LBM::RemakeLevel, Source/LBM.cpp:1302-1365):
 * mrb_cluster_tags turns refinement tags on the coarse domain (one byte per
 * coarse cell, x fastest; e.g. from mrb_tag_cells) into fine boxes (ratio
 * 2): the tags grown by n_error_buf cells, coarsened to blocks of
 * blocking_factor fine cells, clustered by Berger-Rigoutsos to grid_eff,
 * chopped to max_grid_size (amr.n_error_buf, amr.blocking_factor,
 * amr.max_grid_size, amr.grid_eff).  AMReX's own clustering is not part of
 * the reference, so the boxes follow the published method, not AMReX's
 * tie-breaking.  *nbox, and with non-null arrays (3 ints per box) the
 * boxes.  Host only. */
int mrb_cluster_tags(const uint8_t *tags, const int ncell[3], int n_error_buf, int blocking_factor,
                     int max_grid_size, double grid_eff, int *nbox, int *box_lo, int *box_hi, int max_boxes);
/* The new fine level's populations (FillPatchTwoLevels in RemakeLevel,
 * FillPatchOps.H:118-129): each valid cell from the old fine level where
 * the old level has it, else the coarse interpolation of the explosion;
 * 0 in solid cells (fill_f_inside_eb).  new_fine: created with the new
 * boxes, flags set; afterwards mrb_amr_link(crse, new_fine) and the old
 * level can be destroyed.  Synchronous. */
int mrb_amr_regrid_fill(mrb_level *crse, const mrb_level *old_fine, mrb_level *new_fine);
/* Same as mrb_amr_advance, with each coarse step replayed from a hipGraph
 * captured on first use (one graph per coarse buffer parity; recaptured
 * after mrb_set_phys, mrb_set_bc, a stage-level mrb_stream or a relink).
 * Results are identical to mrb_amr_advance.  Asynchronous. */
int mrb_amr_advance_graph(mrb_level *crse, mrb_level *fine, int nsteps, void *stream);

/* ---- isothermal variant with partial bounce-back, and its adjoint ---------
 * No reference counterpart as a whole: the f update of LBM::collide
 * (Source/LBM.cpp:621-807) with the temperature held at T0 (g not stored,
 * Q-correction kept), LBM::stream's pull (LBM.cpp:558-604) with a
 * constant-velocity inlet at x-lo (BC.H:395-419) and an outflow at x-hi
 * (BC.H:435-447), y and z periodic, followed by the Walsh-Burwinkle-Saar
 * partial bounce-back f_q <- (1 - n_s) f_q + n_s f*_qbar with a per-cell
 * solid fraction n_s in [0, 1] (n_s = 0: the reference's fluid update).
 * Loss J = mean outlet-plane u_x + 1e-3 sum rho |u|^2 / 2 of the final
 * state.  Arrays (n_s: N doubles; u0, du0: 3N, component-major) are device
 * pointers, cell = (z*ny + y)*nx + x. */
typedef struct mrb_iso mrb_iso;
int mrb_iso_create(const int ncell[3], double T0, double rho0, double u_in, mrb_iso **out);
int mrb_iso_destroy(mrb_iso *iso);
/* nsteps steps from f = f_eq(rho0, u0(x), R T0); *J receives the loss. */
int mrb_iso_forward(
    mrb_iso *iso, const double *ns, double tau, const double *u0, int nsteps, double *J, void *stream);
/* The loss and its exact gradient with respect to tau, n_s (per cell) and
 * u0 (per cell, 3 comps) by the hand-written adjoint of the step, with
 * checkpoints every `every` steps and the states between checkpoints
 * recomputed: mrb_iso_record followed by mrb_iso_reverse.  every = 0:
 * ceil(sqrt(nsteps)) (fewest buffers); every < 0: the longest segment whose
 * states fit half of the free device memory, the whole run (no
 * recomputation) when they fit, never shorter than ceil(sqrt(nsteps)).
 * Every interval gives the same gradient, bit for bit.  Synchronous. */
int mrb_iso_gradient(
    mrb_iso *iso, const double *ns, double tau, const double *u0, int nsteps, int every, double *J,
    double *dtau, double *dns, double *du0, void *stream);
/* The forward sweep of mrb_iso_gradient alone: nsteps steps keeping the
 * checkpoints; *J receives the loss.  ns and u0 must stay allocated and
 * unchanged until mrb_iso_reverse; mrb_iso_forward in between discards the
 * recording.  Synchronous. */
int mrb_iso_record(
    mrb_iso *iso, const double *ns, double tau, const double *u0, int nsteps, int every, double *J,
    void *stream);
/* The reverse sweep of the last mrb_iso_record (recomputing each earlier
 * segment from its checkpoint): the gradients as in mrb_iso_gradient.  Uses
 * the recording up.  Synchronous. */
int mrb_iso_reverse(mrb_iso *iso, double *dtau, double *dns, double *du0, void *stream);
/* The checkpoint interval of the last mrb_iso_record (after the every <= 0
 * rules). */
int mrb_iso_record_interval(const mrb_iso *iso, int *every);
/* Final populations of the last mrb_iso_forward (27 N doubles, host). */
int mrb_iso_state(const mrb_iso *iso, double *host_f);

/* ---- checkpoint files ----------------------------------------------------
 * LBM::write_checkpoint_file (Source/LBM.cpp:1684-1750): <dir>/Header
 * ("Checkpoint file for LBM", finest level, per level step, dt and time,
 * then the BoxArrays) and per level the f and g MultiFabs in AMReX's VisMF
 * format (<dir>/Level_<l>/f_00_H, f_00_D_00000, g_00_H, g_00_D_00000):
 * fab levels with their ghost cells, compact levels as one box without
 * ghosts.  Synchronous. */
int mrb_write_checkpoint(const char *dir, mrb_level *const *levels, int nlevels, const int *steps,
                         const double *dts, const double *times);
/* LBM::read_checkpoint_file (LBM.cpp:1752-1800): the level count, and with
 * non-null arrays (max_levels entries) each level's step, dt and time. */
int mrb_checkpoint_header(const char *dir, int *nlevels, int *steps, double *dts, double *times,
                          int max_levels);
/* The BoxArray of one level of a checkpoint: *nbox, and with non-null
 * arrays (3 ints per box, max_boxes boxes) the boxes' low and high cells. */
int mrb_checkpoint_boxes(const char *dir, int level, int *nbox, int *box_lo, int *box_hi, int max_boxes);
/* LBM::read_checkpoint_file (LBM.cpp:1800-1860): the f and g populations of
 * every level into levels created with the checkpoint's boxes (fab layout,
 * the same ghost width) or domain (compact), whose flags are already set;
 * solid cells get 0 (fill_f_inside_eb, LBM.cpp:1278-1298) and the ghost
 * cells are refilled before the next step.  Synchronous. */
int mrb_read_checkpoint(const char *dir, mrb_level *const *levels, int nlevels);

/* ---- forces file ---------------------------------------------------------
 * LBM::open_forces_file / output_forces_file (Source/LBM.cpp:1925-1969)
 * for the EB force of mrb_eb_force: open writes the header ("time", "fx",
 * "fy", "fz" in 24-character columns) unless the file exists and
 * initialize is 0 (a restart appends); append writes the time and the
 * three components, width 24, 16 significant digits.  Host only. */
int mrb_forces_file_open(const char *path, int initialize);
int mrb_forces_file_append(const char *path, double time, const double force[3]);

/* ---- plotfile ------------------------------------------------------------
 * LBM::write_plot_file (Source/LBM.cpp:1629-1682) for one level:
 * the variables of LBM::plot_file_mf in its order -- the 19 macrodata
 * components (rho, vel_x, vel_y, vel_z, vel_mag, two_rho_e, QCorrX..Z,
 * pxx, pyy, pzz, pxy, pxz, pyz, qx, qy, qz, temperature) of the last step's
 * post-stream state (what m_macrodata holds after the collide; the stored
 * state before any step), with MRB_PLOT_STREAMING the 27 f and 27 g
 * populations, with MRB_PLOT_DERIVED vort_x, vort_y, vort_z, vort_mag
 * (compute_derived) and dQCorrX..Z (compute_q_corrections) of those
 * macrodata, then is_fluid and eb_boundary -- written as AMReX's
 * plotfile (<dir>/Header "HyperCLaw-V1.1", <dir>/Level_0/Cell_H and
 * Cell_D_00000).  Synchronous. */
#define MRB_PLOT_STREAMING 1 /* lbm.save_streaming */
#define MRB_PLOT_DERIVED 2   /* lbm.save_derived */
int mrb_write_plotfile(const char *dir, mrb_level *lvl, double time, int step, int flags);
/* amrex::WriteMultiLevelPlotfile of LBM::write_plot_file for levels 0 ..
 * nlevels-1 (refinement ratio 2; a two-level hierarchy linked by
 * mrb_amr_link after mrb_amr_advance, or one level): time = the coarse
 * level's, steps = each level's step count; <dir>/Level_<l>/Cell. */
int mrb_write_plotfile_levels(const char *dir, mrb_level *const *levels, int nlevels, double time, const int *steps,
                              int flags);
/* The plotfile's variable count for these flags. */
int mrb_plot_nvars(int flags);
/* The plotfile's fields into host memory: per box, mrb_plot_nvars(flags)
 * variables over the box's valid cells (x fastest), boxes back to back.
 * On a fab level the macrodata are those of the last stage collide, of the
 * state before the last batched step pulled (with the coalesced fine values
 * on the coarse level of a hierarchy, and the ring-1 ghosts of its fine
 * level), or of the stored state before any step.  Synchronous. */
int mrb_plot_fields(mrb_level *lvl, double *out, int flags);

/* Density, velocity (x, y, z) and temperature of the current state per
 * cell into the device array out (5 N doubles, field-major; zeros in solid
 * cells), as LBM::f_to_macrodata (Source/LBM.cpp:810-906) computes them.
 * Compact layout, every storage.  Asynchronous. */
int mrb_macro_fields(const mrb_level *lvl, double *out, void *stream);

/* LBM::compute_derived (Source/LBM.cpp:909-955): vorticity x, y, z and
 * magnitude of the current state's velocity into the device array out
 * (4 N doubles, field-major; 0 in solid cells), with the reference's
 * gradient() (Utilities.H:279-312: central differences, one-sided next to a
 * domain face or a solid cell).  The velocity is that of the stored
 * (post-collision) populations, equal to the last collide's post-stream
 * velocity up to rounding.  Compact layout.  Asynchronous. */
int mrb_vorticity(mrb_level *lvl, double *out, void *stream);

/* Refinement tags for LBM::ErrorEst (LBM.cpp:1597-1611), the criteria of
 * tagging.<name>.value_greater / value_less / adjacent_difference_greater
 * (LBM.cpp:351-377) evaluated on a device field of N doubles (for example a
 * component of mrb_macro_fields or mrb_vorticity): tags[cell] = 1 where
 * field > value, field < value, or the largest |difference| to a face
 * neighbour > value (neighbours wrap on periodic axes and are skipped beyond
 * non-periodic faces).  Compact layout.  Asynchronous. */
#define MRB_TAG_GREATER 0
#define MRB_TAG_LESS 1
#define MRB_TAG_ADJACENT_DIFFERENCE 2
/* LBM::ErrorEst (LBM.cpp:1597-1611) on any level: the plot variable named
 * `field` (a macrodata or derived name of LBM.cpp:14-60, e.g. "rho",
 * "vel_mag", "vort_mag") of the current state, evaluated on the device,
 * and the criterion of mrb_tag_cells over the domain; host_tags: one byte
 * per domain cell (x fastest), the input of mrb_cluster_tags.  The level's
 * boxes must cover the domain (the levels ErrorEst tags do).  Synchronous. */
int mrb_tag_level(mrb_level *lvl, const char *field, int criterion, double value, uint8_t *host_tags);
int mrb_tag_cells(const mrb_level *lvl, const double *field, int criterion, double value, uint8_t *tags,
                  void *stream);

/* ---- diagnostics ------------------------------------------------------------ */
const char *mrb_last_error(void);
/* lbm::stencil::Stencil (Source/Stencil.H:41-169): the 27 lattice
 * velocities (q-major, 3 ints each), weights and opposite directions in
 * the reference's order. */
int mrb_stencil(int evs[81], double weights[27], int bounce_dirs[27]);
/* Name of the kernel variant used by mrb_stream_collide. */
const char *mrb_hot_kernel_name(const mrb_level *lvl);
/* Device pointers of the current population buffers (for zero-copy use by
 * the Python/torch wrappers).  lattice 0 = f, 1 = g.  Fab layout: each
 * box's FAB, components contiguous.  Compact layout: f_q and g_q of a cell
 * side by side, element (q*N + cell)*2 + lattice of the f pointer (the g
 * pointer is the f pointer plus one element). */
int mrb_population_ptr(const mrb_level *lvl, int lattice, void **dev_ptr, int64_t *bytes);
/* FP16 storage: number of block writes (256 cells, one lattice) since the
 * level was created in which some value exceeded the FP16 range at its
 * box's scale and was clamped (the per-box scale leaves a factor 16 of
 * headroom; 0 for the other storages).  No reference counterpart. */
int mrb_storage_saturations(const mrb_level *lvl, int64_t *count);
/* The rest-state values subtracted by the reduced-precision storages, per
 * direction (zero for FP64), for callers that read the raw buffers of
 * mrb_population_ptr.  No reference counterpart. */
int mrb_storage_shift(const mrb_level *lvl, int lattice, double out[27]);
/* Streaming device-to-device copy of `bytes` (a multiple of 16, 16-byte
 * aligned buffers) by a hand-written read + write kernel on `blocks` blocks
 * of 256 threads (0: 1024), enqueued on `stream`: the measured HBM ceiling
 * a step kernel's bandwidth is compared with.  No reference counterpart. */
int mrb_stream_copy(void *dst, const void *src, int64_t bytes, int blocks, void *stream);

#ifdef __cplusplus
}
#endif
#endif
