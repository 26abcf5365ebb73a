// Internal definitions of the marbles_amd engine (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "lbm_math.h"
#include "marbles_amd.h"

namespace mrb {

// One box of a level as the kernels see it.  A box's fab is the valid box
// grown by ng; every per-box array stores its fab component-major starting
// at base*ncomp (base = number of fab cells of all earlier boxes).
struct BoxDev {
    int lo[3];  // fab low corner (valid low corner - ng)
    int n[3];   // fab extents
    int vlo[3], vhi[3];
    int64_t base;
    int64_t ncell;
};

// FillBoundary copy: fab cell (dst_box, dst_local) <- (src_box, src_local).
struct GhostPair {
    int32_t dst_box, src_box;
    int64_t dst_local, src_local;
};

// Velocity-inlet ghost values and boundary types.
// Velocity-inflow profile of the Channel and Parabolic velocity operators
// (VelocityBC.H:90-192): the ghost cell's velocity from its index.
struct VelProfile {
    int kind;     // 0: constant values (fval/gval); 2: channel; 3: parabolic
    int nd, td;   // parabolic: normal and tangential directions
    int domhi[3];
    double vref;  // channel u_ref, parabolic um (Mach_ref * speed of sound)
    double rho, R, T, gamma;
    double plo, phi, dxn;  // parabolic: prob_lo, prob_hi and dx along nd
};

__host__ __device__ inline void profile_velocity(const VelProfile &P, const int iv[3], double vel[3])
{
    vel[0] = vel[1] = vel[2] = 0.0;
    if (P.kind == 2) {
        const double c1 = double(iv[1] * (P.domhi[1] - iv[1]));
        const double c2 = double(iv[2] * (P.domhi[2] - iv[2]));
        const double n = double(P.domhi[1] + 1);
        vel[0] = 16.0 * P.vref * c1 * c2 / (((n * n) * n) * n);
    } else if (P.kind == 3) {
        const double height = P.phi - P.plo;
        const double x = P.plo + (iv[P.nd] + 0.5) * P.dxn;
        vel[P.td] = 4.0 * P.vref * x * (height - x) / (height * height);
    }
}

struct BCDev {
    int type[6];
    double fval[NQ];
    double gval[NQ];
    VelProfile prof;
    // compact layout, profile kinds: per face, the f and g ghost values over
    // the face's transverse coordinates t1 < t2 (each from -1 to n), entry
    // ((t1 + 1) * tabw + t2 + 1) * 2NQ + [q | NQ + q]
    const double *tab[6];
    int tabw[6];
    // PRESSURE faces (BC.H:222-321), all on one axis pdir (-1: none): the
    // boundary density and R T of the f ghost equilibrium, and per
    // direction q the g ghost value of the face q enters through
    // (c_q[pdir] = +1: the low face, -1: the high face)
    int pdir;
    double prho, prt;
    double pgval[NQ];
};

// Geometry of the compact (single box, no ghost) layout.
struct CompactGeom {
    int n[3];
    int periodic[3];
    int64_t N;  // cells
};

// Bits of the per-cell stream mask of the compact layout.
//   bit q (0..26): the pull source x - c_q is fluid (a ghost beyond a
//                  physical boundary counts as fluid); bit 0 = own fluid.
//   kMaskGeneral:  some pull source needs a periodic wrap or is a ghost.
constexpr uint32_t kMaskAllFluid = (1u << NQ) - 1u;
constexpr uint32_t kMaskGeneral = 1u << 27;

enum class Storage { F64 = 0, F32S = 1, F16S = 2 };

// Batched fab-layout passes: a tile is kFabTile consecutive cells (x
// fastest) of one box's valid region, grown by 0 or 1 cells.
constexpr int kFabTile = 256;
struct FabTile {
    int32_t box;
    int32_t first;
};

// Two-level transfer maps (built by mrb_amr_link, held by the fine level).
// coarse cell <- masked mean of its 2x2x2 fine children
struct CoalesceEntry {
    int64_t crse_local;  // coarse fab cell (local index in crse_box's fab)
    int64_t fine_local;  // child (0,0,0) (local index in fine_box's fab)
    int32_t crse_box;
    int32_t fine_box;
};
// fine ghost cell over uncovered domain <- interpolation from its parent
struct ExplodeEntry {
    int64_t fine_local;
    int64_t crse_local;  // the parent
    int32_t fine_box;
    int32_t crse_box;
    int32_t bits;    // bit d: high child in direction d
    int32_t parent;  // index of the parent in the level's ExplodeParent list
};

// A coarse parent of interpolated fine cells: its slopes are computed once
// (k_explode_slopes) and read by its children.  flags (as the oracle's
// orc_amr_slopes): bit d / 3+d the parent lies at the low / high face of a
// non-periodic direction d (ext_dir: neighbours beyond it read -1, the slope
// is one-sided); bit 6+d / 9+d the cell P + 2e_d / P - 2e_d lies inside the
// domain (the 4-point one-sided form, else the 3-point one).
struct ExplodeParent {
    int64_t crse_local;
    int32_t crse_box;
    int32_t flags;
};

// Single-pass z-sweep kernel shapes: rows per tile and z planes per chunk.
struct FusedVariant {
    const char *name;
    int ty;
    int zc;
    int hold;  // 1: keep the look-ahead plane's populations in registers
};
extern const FusedVariant kFusedVariants[];
extern const int kNumFusedVariants;

// FP16 storage with a per-box scale: stored value = (population - shift) *
// 2^e, with e per (scale box, lattice).  A scale box of the compact layout
// is a slab of kScaleSlab z-planes over the whole x-y extent, so that the
// cells a wave pulls from (one z-plane and its two neighbours) use at most
// three boxes' scales, read as wave-uniform values.  e is chosen from the
// largest |population - shift| m of the box's fluid cells in the previous
// state so that m * 2^e < 2^kF16Target: the FP16 maximum (65504) then
// leaves a factor 16 of headroom for growth within one step; larger values
// are clamped to the FP16 range, and each (block, lattice) write that
// clamped is counted.  Solid cells hold 0 (downloads return the stream's
// -1 sentinel there).
// The largest |population - shift| a box is taken to hold is at least
// 2^-20: the FP32-arithmetic collide writes rounding noise of about
// 2^-24 of a population (up to ~6e-8) into a box exactly at the shift
// state, which a larger exponent would push past the FP16 range.
constexpr int kScaleSlab = 8;
constexpr int kF16Target = 12;
constexpr float kF16MinMax = 9.5367431640625e-07f;  // 2^-20
__host__ __device__ inline int f16_exponent(float m)
{
    if (!(m > kF16MinMax)) m = kF16MinMax;
    int k;
    (void)frexpf(m, &k);  // m < 2^k
    const int e = kF16Target - k;
    return e < -24 ? -24 : (e > 40 ? 40 : e);
}
int f16_pack(mrb_level *L, int lattice, const double *src, int64_t src_cs, hipStream_t s);
int f16_unpack(const mrb_level *L, int lattice, double *dst, hipStream_t s);
int compact_macro(const mrb_level *L, double *out, hipStream_t s);
int compact_vorticity(mrb_level *L, double *out, hipStream_t s);
// pull of cell loc (fluid) in box B: f*_q = fluid(x - c_q) ? f_q(x - c_q) : f_qbar(x)
__device__ __forceinline__ void pull_fab(
    const double *__restrict__ f, const double *__restrict__ g, const BoxDev &B, int64_t loc,
    uint32_t m, double fv[NQ], double gv[NQ])
{
    const double *fb = f + B.base * NQ;
    const double *gb = g + B.base * NQ;
    const int64_t sy = B.n[0], sz = int64_t(B.n[0]) * B.n[1];
    static_for<NQ>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        constexpr int cx = ev(q, 0), cy = ev(q, 1), cz = ev(q, 2);
        constexpr int qb = kBounce[q];
        const int64_t src = loc - (cx + cy * sy + cz * sz);
        const int64_t e = ((m >> q) & 1u) ? int64_t(q) * B.ncell + src : int64_t(qb) * B.ncell + loc;
        fv[q] = fb[e];
        gv[q] = gb[e];
    });
}

// plotfile fields of a compact level: the 19 macrodata components of the
// last step's post-stream state (or of the stored state before any step)
// into md, and from them vorticity (4) and dQCorr (3) into der
int compact_plot_fields(mrb_level *L, double *md, double *der, hipStream_t s);
// fab layout: the plotfile macrodata into L->d_md (valid cells, one ghost
// layer filled) and vorticity (4) and dQCorr (3) into der (7 fab comps)
int fab_plot_fields(mrb_level *L, double *der, hipStream_t s);
// the plotfile macrodata of the last batched step (pull of the pre-step
// buffer, gathering coal_from's coalesced values on a coarse level, and at
// the listed ring-1 ghosts of a fine level) into L->d_md, valid cells and
// ring; the caller fills the covered ghost layer
int fab_plot_md_pull(mrb_level *L, const mrb_level *coal_from, hipStream_t s);
int compact_tag(const mrb_level *L, const double *field, int criterion, double value, uint8_t *tags,
                hipStream_t s);

}  // namespace mrb

struct mrb_level {
    // geometry
    int ncell[3];
    int periodic[3];
    double prob_lo[3], prob_hi[3], dx[3], inv_dx[3];
    int nbox = 0;
    int ng = 0;
    int layout = MRB_LAYOUT_FAB;
    int nlattice = 2;
    int storage = MRB_STORAGE_F64;
    std::vector<mrb::BoxDev> boxes;  // host copy
    int64_t total_cells = 0;         // fab cells over all boxes
    // device memory
    mrb::BoxDev *d_boxes = nullptr;
    int64_t *d_box_base = nullptr;  // nbox+1 prefix sums (for lookups)
    void *pop[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};  // [lattice][A/B]
    int cur = 0;                      // index of the current buffer
    int32_t *d_isf = nullptr;         // is_fluid, 2 comps
    uint32_t *d_mask = nullptr;       // compact: stream mask per cell
    double *d_md = nullptr;           // FAB: macrodata, 19 comps
    double *d_dq = nullptr;           // FAB: dQ, 3 comps
    double *d_q = nullptr;            // compact: Q_corr, 3 comps
    double *d_red = nullptr;          // reduction partials
    double *d_small = nullptr;        // [fshift | gshift | init values | inlet f | inlet g] (6*NQ)
    mrb::GhostPair *d_gmap = nullptr;   // FillBoundary map, ng layers
    int64_t n_gmap = 0;
    mrb::GhostPair *d_gmap1 = nullptr;  // FillBoundary map, 1 layer (macrodata)
    int64_t n_gmap1 = 0;
    // fine level of a linked pair: the FillBoundary entries the two-level
    // step reads (the first layer, and deeper ghosts next to uncovered or
    // out-of-domain ghosts); built by mrb_amr_link
    mrb::GhostPair *d_gmap_step = nullptr;
    int64_t n_gmap_step = 0;
    int64_t *d_bcells = nullptr;        // fab cells outside non-periodic faces
    int64_t n_bcells = 0;
    // d_bcells is ordered by the distance beyond the domain (sum over the
    // non-periodic directions of how many cells out); cells of distance
    // p are [bcell_off[p-1], bcell_off[p])
    std::vector<int64_t> bcell_off;
    // physics
    mrb::CollideParams cp{};
    mrb_phys phys{};
    mrb_bc bc{};
    mrb::BCDev bcdev{};
    bool phys_set = false;
    bool flags_ready = false;
    double shift_rho0 = 1.0, shift_T0 = 1.0 / 3.0;
    bool pops_set = false;  // populations uploaded or initialised
    // compact layout: the spare buffer holds the state before the last
    // step (set by a step, cleared when the populations are set)
    bool prev_valid = false;
    // fab layout: where the plotfile's macrodata of the current state come
    // from -- 0: the moments of the stored populations (initial state), 1:
    // d_md of the last stage collide, 2: the pull of the spare buffer,
    // which holds the state before the last batched step; -1: none
    int md_src = 0;
    bool in_hierarchy = false;  // linked by mrb_amr_link
    mrb_level *amr_fine = nullptr;  // on the coarse level of a hierarchy
    double fshift[mrb::NQ] = {}, gshift[mrb::NQ] = {};
    // FP16 storage (compact layout): power-of-two scale exponent and the
    // largest |population - shift| over fluid cells per (scale box, lattice),
    // one pair of arrays per population buffer; saturation counter
    int *d_e16[2] = {nullptr, nullptr};
    uint32_t *d_amax16[2] = {nullptr, nullptr};  // float bits
    uint32_t *d_ovf16 = nullptr;
    int nsb16 = 0;  // scale boxes (z-slabs of kScaleSlab planes)
    double *d_mac = nullptr;  // compact: macro fields scratch (5 N) for derived quantities
    double *d_bctab = nullptr;  // compact: velocity-profile ghost tables of the faces
    // timing events (created with the level)
    hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    // hot-path kernel variant of the compact layout (compact_variant_name)
    int hot = 0;
    bool hot_default = false;  // hot is the storage's default, not a caller's choice
    // FP32-arithmetic variant set aside while the level has a PRESSURE face
    // (those collides have none); restored when the faces change back
    int hot_h32 = -1;
    // single-pass z-sweep on 64-cell blocks (k_sweep_x64, compact FP64
    // levels): ring of kSweepRingPlanes planes of Q_y, per-block count of
    // published planes; count of waits that gave up (both z-sweeps)
    double *d_sweep_q = nullptr;
    double *d_sweep_qx = nullptr;  // Q_x of the end cells of 64-cell blocks (k_sweep_x64)
    uint32_t *d_sweep_flags = nullptr;
    uint32_t *d_sweep_fault = nullptr;
    // k_sweep's ring of tagged Q_y (two 64-bit words per cell and plane) and
    // the launch count its tags start from
    uint64_t *d_sweep_qt = nullptr;
    uint32_t *d_sweep_epoch = nullptr;
    // two-level link, held by the fine level (mrb_amr_link)
    mrb_level *amr_crse = nullptr;
    void *d_coal = nullptr;  // mrb::CoalesceEntry[n_coal]
    int64_t n_coal = 0;
    void *d_expl = nullptr;  // mrb::ExplodeEntry[n_expl]
    int64_t n_expl = 0;
    void *d_expl_par = nullptr;  // mrb::ExplodeParent[n_expl_par]
    int64_t n_expl_par = 0;
    double *d_expl_slopes = nullptr;  // [4 * NQ][parent]: value and limited slopes of one lattice
    int expl_crse_depth = 1;          // coarse FillBoundary depth the interpolation reads
    int32_t *d_coal_cell = nullptr;  // per coarse fab cell: coalescence entry or -1
    double *d_coal_vals = nullptr;   // [2q + lattice][entry]: masked means of the children
    uint64_t *d_coal_has = nullptr;  // per entry: bit 2q + lattice set if some child is not -1
    // explosion entries are ordered ring <= 2 first (n_expl_le2 of them);
    // d_qring lists the ring-1 entries again
    int64_t n_expl_le2 = 0;
    void *d_qring = nullptr;  // mrb::ExplodeEntry[n_qring]
    int64_t n_qring = 0;
    bool stream_rings_le2 = false;  // the listed ghost stream covers rings 1-2 only
    // batched fab-layout step: tile list of the valid cells and the number
    // of persistent blocks (0 = one block per tile)
    mrb::FabTile *d_tiles_c = nullptr;
    int n_tiles_c = 0;
    int fab_blocks = 0;
    // fab layout: ghost layers 1..fresh_depth that FillBoundary writes hold
    // their owners' current values (0: stale; set by FillBoundary, cleared
    // by anything that changes valid data otherwise)
    int fresh_depth = 0;
    // bumped whenever captured kernel arguments would go stale (physics,
    // boundary values, buffer pointers, launch shapes, links)
    uint64_t gen = 1;
    // hipGraph of one coarse step, held by the fine level of a linked pair:
    // [coarse buffer index][coarse ghosts fresh]
    struct AmrGraph {
        hipGraphExec_t exec = nullptr;
        uint64_t gen_c = 0, gen_f = 0;
        int fine_cur = 0;
        int post_cur_c = 0, post_cur_f = 0;
        int post_fresh_c = 0, post_fresh_f = 0;
    } amr_graph[2][2];
    hipStream_t cap_stream = nullptr;
};

namespace mrb {

int set_error(int code, const std::string &msg);
int hip_fail(hipError_t e, const char *what);

#define MRB_HIP(call)                                       \
    do {                                                    \
        hipError_t e_ = (call);                             \
        if (e_ != hipSuccess) return ::mrb::hip_fail(e_, #call); \
    } while (0)

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

// The block index with the grid renumbered so that each XCD takes one
// contiguous run of indices (the hardware hands block b to XCD b mod 8):
// list kernels whose neighbouring entries gather neighbouring data then
// share that XCD's L2 instead of spreading each neighbourhood over eight.
// A bijection on [0, gridDim.x).
__device__ __forceinline__ uint32_t xcd_index(uint32_t b, uint32_t nb)
{
    const uint32_t q = nb / 8u, r = nb % 8u, x = b % 8u, k = b / 8u;
    return x * q + (x < r ? x : r) + k;
}
__device__ __forceinline__ int64_t xcd_block() { return int64_t(xcd_index(blockIdx.x, gridDim.x)); }

// compact-layout kernels (mrb_compact.hip)
int compact_stream_collide(mrb_level *L, hipStream_t s);
// one step with an event recorded before kernel slot k (ev[k]) and after the
// last slot; returns the number of slots
int compact_stream_collide_events(mrb_level *L, hipStream_t s, int &nslots);
int compact_stream_only(mrb_level *L, hipStream_t s);
int compact_build_mask(mrb_level *L, hipStream_t s);
const char *compact_kernel_name(const mrb_level *L);
// single-pass kernels (mrb_fused.hip): launch reads the current buffer and
// writes the spare one; step also flips the buffers
int fused_launch(mrb_level *L, int variant, hipStream_t s);
int fused_step(mrb_level *L, int variant, hipStream_t s);
// compact-layout step variants (mrb_compact.hip): index < kNumTwoPassVariants
// is a two-pass step (register budgets differ) or, at kOnePassVariant, the
// single pass that recomputes the neighbours' Q_corr; the rest the fused
// variants; launch reads the current buffers, writes the spare ones, does
// not flip
constexpr int kNumTwoPassVariants = 14;
constexpr int kOnePassVariant = 2;
constexpr int kRows8Variant = 3;  // two passes on 64x8 blocks
constexpr int kRows2Variant = 4;  // two passes on 64x2 blocks
constexpr int kHalf32Variant = 5; // scaled FP16 storage, FP32 arithmetic in pass 2
constexpr int kX128Variant = 6;   // two passes on 128x2 blocks
constexpr int kX256Variant = 7;   // two passes on 256x1 blocks
constexpr int kHalf32X256Variant = 8; // variant 5 on 256x1 blocks
constexpr int kF32Variant = 9;         // shifted FP32 storage, FP32 arithmetic in pass 2 (64x2 blocks)
constexpr int kF32X256Variant = 10;    // variant 9 on 256x1 blocks
constexpr int kSweepVariant = 11;      // single-pass z-sweep, one block per x-row (FP64 / FP32, rows <= 256 cells)
constexpr int kSweepRingPlanes = 8;    // planes of Q_y the z-sweep keeps for the rows beside
constexpr int kSweepX64Variant = 12;   // the z-sweep on single-wave blocks of 64 cells of a row (FP64)
constexpr int kSweepF32Variant = 13;   // the z-sweep with the FP32-arithmetic collide (shifted FP32)
// the variants whose pass 2 computes in FP32 (their results differ from the
// FP64-arithmetic variants'), and the FP64-arithmetic variant with the same
// block shape that replaces one on a level with PRESSURE faces
inline bool fp32_collide_variant(int v)
{
    return v == kHalf32Variant || v == kHalf32X256Variant || v == kF32Variant || v == kF32X256Variant ||
           v == kSweepF32Variant;
}
inline int fp64_counterpart(int v)
{
    return v == kHalf32Variant     ? 0
           : v == kF32Variant      ? kRows2Variant
           : v == kSweepF32Variant ? kSweepVariant
                                   : kX256Variant;
}
// conservation sums (mrb_level.hip): block partial sums (5 per block, at
// most kConservationBlocks blocks, grid-strided over any level size) into
// `part`; the caller adds them in block order
constexpr unsigned kConservationBlocks = 4096;
unsigned conservation_blocks(const mrb_level *L);
int conservation_partials(const mrb_level *L, double *part, hipStream_t s);
int compact_two_pass_launch(mrb_level *L, int variant, hipStream_t s);
int compact_variant_launch(mrb_level *L, int index, hipStream_t s);
const char *compact_variant_name(int index);
// FillBoundary of one lattice's current fab buffer (mrb_level.hip)
int fab_fill_boundary(mrb_level *L, int lattice, hipStream_t s);
// FillBoundary of one lattice's current fab buffer over the given map
int fab_fill_boundary_map(mrb_level *L, int lattice, const GhostPair *map, int64_t n, hipStream_t s);
// every ghost cell (within `depth` layers) that FillBoundary writes, with its owner
void build_ghost_map(const mrb_level *L, int depth, std::vector<GhostPair> &map);
// FillBoundary of both lattices, `depth` ghost layers (1 or ng); sets fresh_depth
int fab_fill_boundary_depth(mrb_level *L, int depth, hipStream_t s);
// FillBoundary of the batched step's Q array (3 comps in d_dq), one layer
int fab_fill_boundary_q(mrb_level *L, hipStream_t s);
// frees the two-level maps held by a fine level (mrb_amr.hip)
void amr_release(mrb_level *L);
// batched fab-layout step (mrb_batched.hip): pass 1 + pass 2 into the spare
// buffer (coalescing from `coal_from` when given), ghost streaming over the
// explosion map of a fine level, flip, FillBoundary of out_depth layers
// (0: none; 1: what the next step's pulls read; ng: all)
int fab_batched_step(mrb_level *L, const mrb_level *coal_from, int out_depth, hipStream_t s);
int fab_build_mask(mrb_level *L, hipStream_t s);
int fab_batched_passes(mrb_level *L, const mrb_level *coal_from, hipStream_t s);
// persistent-block counts the fab-layout autotuner tries (0 = one block per tile)
extern const int kFabBlockChoices[];
extern const int kNumFabBlockChoices;
void build_fab_tiles(const mrb_level *L, int grow, std::vector<FabTile> &tiles);

}  // namespace mrb
