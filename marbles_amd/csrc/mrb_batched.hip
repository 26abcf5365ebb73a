// Batched step on the reference (fab) layout: every box of a level in one
// pass, driven by a flat tile work-list built when the level is created.
//
//   pass 1 (k_fab_qpass):   valid cells of every box: pull 54 populations,
//                           moments -> Q_corr; FillBoundary of Q (one layer)
//   pass 2 (k_fab_collide): valid cells: pull 54 populations, moments, dQ
//                           from the Q of the face neighbours, collide,
//                           write the spare buffer
//
// followed by FillBoundary of the spare buffer and a buffer flip.  This is
// LBM::stream + LBM::collide (LBM.cpp:558-807) without materialising the
// streamed state.  The Q of a ring-1 ghost cell comes from its owning box
// by FillBoundary, as the reference's macrodata does (LBM.cpp:905); on a
// fine level the ring-1 ghost cells over uncovered domain have no owner
// and compute Q from their own pulled populations (k_fab_qpass_listed).
// Those uncovered ghost cells also keep the streamed values that the
// reference keeps (they feed the second substep and the coalescence), so
// they are streamed from the explosion map; all other ghost values are
// rewritten by FillBoundary or the boundary functor before anything reads
// them.
//
// On a coarse level linked to a fine one, coalescence (average_down_with_
// ghosts, Utilities.cpp:5-28) is fused into both passes: a cell under the
// fine footprint (or its one-cell ring) takes the masked mean of its fine
// children instead of its pulled value, population by population.
//
// A tile is kFabTile consecutive cells (x fastest) of one box region.
// Blocks are grouped by XCD (block b runs on XCD b % 8): each XCD walks a
// contiguous chunk of the tile list so that neighbouring tiles, which
// share the cache lines of their halo pulls, meet in the same L2.
#include <algorithm>
#include <vector>

#include "mrb_internal.h"

namespace mrb {

namespace {

struct FabArgs {
    const BoxDev *boxes;
    const uint32_t *mask;  // per fab cell: bit q = is_fluid(x - c_q) != 0, bit 0 own fluid
    int dom[3];
    int grow;
    int ntiles;
};

// coalescence source for the fused coarse passes: the masked means of each
// entry's children, computed once per coarse step by k_coalesce_vals
struct CoalArgs {
    const double *vals;    // [2q + lattice][entry]
    const uint64_t *has;   // per entry: bit 2q + lattice = some child not -1
    const int32_t *cell;   // per coarse fab cell: entry index or -1
    int64_t n;             // entries
};

__device__ __forceinline__ bool tile_cell(
    const FabArgs &A, const FabTile &t, int r, const BoxDev &B, int &i, int &j, int &k, int64_t &loc)
{
    const int g = A.grow;
    const int rx = B.vhi[0] - B.vlo[0] + 1 + 2 * g;
    const int ry = B.vhi[1] - B.vlo[1] + 1 + 2 * g;
    const int rz = B.vhi[2] - B.vlo[2] + 1 + 2 * g;
    const int c = t.first + r;
    if (c >= rx * ry * rz) return false;
    const int kk = c / (rx * ry);
    const int rem = c - kk * rx * ry;
    const int jj = rem / rx;
    const int ii = rem - jj * rx;
    i = B.vlo[0] - g + ii;
    j = B.vlo[1] - g + jj;
    k = B.vlo[2] - g + kk;
    loc = (int64_t(k - B.lo[2]) * B.n[1] + (j - B.lo[1])) * B.n[0] + (i - B.lo[0]);
    return true;
}

// masked_avgdown (Utilities.H:315-350) of one population; false if all 8
// children are the -1 sentinel
__device__ __forceinline__ bool masked_mean(const double *__restrict__ c0, int64_t sy, int64_t sz, double &out)
{
    double sum = 0;
    double vol = 0;
    for (int kref = 0; kref < 2; ++kref)
        for (int jref = 0; jref < 2; ++jref)
            for (int iref = 0; iref < 2; ++iref) {
                const double fval = c0[kref * sz + jref * sy + iref];
                if (fabs(fval - (-1.0)) > kSmallNum) {
                    sum += fval;
                    vol += 1.0;
                }
            }
    if (vol > 0.0) {
        out = sum / vol;
        return true;
    }
    return false;
}

template <bool COAL>
__device__ __forceinline__ void gather_cell(
    const double *__restrict__ f, const double *__restrict__ g, const BoxDev &B, int64_t loc, uint32_t m,
    const CoalArgs &C, double fv[NQ], double gv[NQ])
{
    if constexpr (COAL) {
        const int32_t ci = C.cell[B.base + loc];
        if (ci >= 0) {
            // coalesced cell: the children's masked mean, or the pulled
            // value where all 8 children are -1 (masked_avgdown leaves the
            // coarse value unchanged)
            const uint64_t has = C.has[ci];
            const double *fb = f + B.base * NQ;
            const double *gb = g + B.base * NQ;
            const int64_t csy = B.n[0], csz = int64_t(B.n[0]) * B.n[1];
            static_for<NQ>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                constexpr int cx = ev(q, 0), cy = ev(q, 1), cz = ev(q, 2);
                const int64_t src = loc - (cx + cy * csy + cz * csz);
                const int64_t pe = ((m >> q) & 1u) ? int64_t(q) * B.ncell + src
                                                   : int64_t(kBounce[q]) * B.ncell + loc;
                fv[q] = ((has >> (2 * q)) & 1u) ? C.vals[(2 * q) * C.n + ci] : fb[pe];
                gv[q] = ((has >> (2 * q + 1)) & 1u) ? C.vals[(2 * q + 1) * C.n + ci] : gb[pe];
            });
            return;
        }
    }
    pull_fab(f, g, B, loc, m, fv, gv);
}

// masked means of every coalescence entry's children (both lattices)
__global__ void __launch_bounds__(256) k_coalesce_vals(
    const double *__restrict__ ff, const double *__restrict__ fg, const BoxDev *__restrict__ fboxes,
    const CoalesceEntry *__restrict__ map, int64_t n, double *__restrict__ vals, uint64_t *__restrict__ has)
{
    const int64_t t = xcd_block() * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const CoalesceEntry e = map[t];
    const BoxDev &FB = fboxes[e.fine_box];
    const int64_t sy = FB.n[0], sz = int64_t(FB.n[0]) * FB.n[1];
    const double *src[2] = {ff + FB.base * NQ + e.fine_local, fg + FB.base * NQ + e.fine_local};
    uint64_t bits = 0;
    for (int q = 0; q < NQ; ++q)
        for (int l = 0; l < 2; ++l) {
            double v;
            if (masked_mean(src[l] + q * FB.ncell, sy, sz, v)) {
                vals[(2 * q + l) * n + t] = v;
                bits |= uint64_t(1) << (2 * q + l);
            }
        }
    has[t] = bits;
}

__device__ __forceinline__ int xcd_tile_begin(int ntiles, int &step, int &end)
{
    // gridDim.x is a multiple of 8; XCD x = blockIdx.x % 8 walks tiles
    // [x*ntiles/8, (x+1)*ntiles/8)
    const int x = int(blockIdx.x & 7u);
    const int per = int(gridDim.x >> 3);
    const int lo = int(int64_t(ntiles) * x / 8);
    end = int(int64_t(ntiles) * (x + 1) / 8);
    step = per;
    return lo + int(blockIdx.x >> 3);
}

template <bool COAL>
__global__ void __launch_bounds__(kFabTile) k_fab_qpass(
    const double *__restrict__ f, const double *__restrict__ g, double *__restrict__ Q,
    const FabTile *__restrict__ tiles, FabArgs A, CoalArgs C, CollideParams P)
{
    int step, end;
    for (int t = xcd_tile_begin(A.ntiles, step, end); t < end; t += step) {
        const FabTile tile = tiles[t];
        const BoxDev &B = A.boxes[tile.box];
        int i, j, k;
        int64_t loc;
        if (!tile_cell(A, tile, int(threadIdx.x), B, i, j, k, loc)) continue;
        const uint32_t m = A.mask[B.base + loc];
        if (!(m & 1u)) continue;
        double fv[NQ], gv[NQ];
        gather_cell<COAL>(f, g, B, loc, m, C, fv, gv);
        double rho, jx, jy, jz, E;
        accumulate_q_moments(fv, gv, rho, jx, jy, jz, E);
        const Macro M = macro_from(rho, jx, jy, jz, E, P);
        double *q = Q + B.base * 3 + loc;
        q[0] = M.Qx;
        q[B.ncell] = M.Qy;
        q[2 * B.ncell] = M.Qz;
    }
}

// pass 1 at the listed ghost cells (the first n entries of the explosion
// map, which are the ring-1 cells)
__global__ void __launch_bounds__(256) k_fab_qpass_listed(
    const double *__restrict__ f, const double *__restrict__ g, double *__restrict__ Q,
    const BoxDev *__restrict__ boxes, const uint32_t *__restrict__ mask, const ExplodeEntry *__restrict__ list,
    int64_t n, CollideParams P)
{
    const int64_t t = xcd_block() * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const ExplodeEntry e = list[t];
    const BoxDev &B = boxes[e.fine_box];
    const int64_t loc = e.fine_local;
    const uint32_t m = mask[B.base + loc];
    if (!(m & 1u)) return;
    double fv[NQ], gv[NQ];
    pull_fab(f, g, B, loc, m, fv, gv);
    double rho, jx, jy, jz, E;
    accumulate_q_moments(fv, gv, rho, jx, jy, jz, E);
    const Macro M = macro_from(rho, jx, jy, jz, E, P);
    double *q = Q + B.base * 3 + loc;
    q[0] = M.Qx;
    q[B.ncell] = M.Qy;
    q[2 * B.ncell] = M.Qz;
}

template <bool COAL>
__global__ void __launch_bounds__(kFabTile) k_fab_collide(
    const double *__restrict__ f, const double *__restrict__ g, double *__restrict__ fo,
    double *__restrict__ go, const double *__restrict__ Q, const FabTile *__restrict__ tiles, FabArgs A,
    CoalArgs C, CollideParams P)
{
    int step, end;
    for (int t = xcd_tile_begin(A.ntiles, step, end); t < end; t += step) {
        const FabTile tile = tiles[t];
        const BoxDev &B = A.boxes[tile.box];
        int i, j, k;
        int64_t loc;
        if (!tile_cell(A, tile, int(threadIdx.x), B, i, j, k, loc)) continue;
        const uint32_t m = A.mask[B.base + loc];
        double *fob = fo + B.base * NQ + loc;
        double *gob = go + B.base * NQ + loc;
        const int64_t s = B.ncell;
        if (!(m & 1u)) {
            // solid cell: the stream leaves the -1 sentinel (LBM.cpp:565)
            for (int q = 0; q < NQ; ++q) {
                fob[q * s] = -1.0;
                gob[q * s] = -1.0;
            }
            continue;
        }
        double fv[NQ], gv[NQ];
        gather_cell<COAL>(f, g, B, loc, m, C, fv, gv);
        // the six neighbours' Q_corr, loaded with the pull (a valid cell's
        // neighbours along each axis lie in the fab, at most in its first
        // ghost layer; a value the gradient does not use is loaded anyway),
        // so that the gradient does not wait on loads issued after the
        // moments
        const double *q = Q + B.base * 3 + loc;
        const int64_t sy = B.n[0], sz = int64_t(B.n[0]) * B.n[1];
        const double qxp = q[1], qxm = q[-1], qyp = q[s + sy], qym = q[s - sy], qzp = q[2 * s + sz],
                     qzm = q[2 * s - sz];
        Moments mo;
        accumulate_moments(fv, gv, mo);
        const Macro M = macro_from(mo.rho, mo.jx, mo.jy, mo.jz, mo.E, P);
        // gradient(): neighbour inside the domain box and fluid (bit 2 of
        // the mask is the source x + e_x, bit 1 the source x - e_x, ...)
        const bool xp = i + 1 < A.dom[0] && ((m >> 2) & 1u), xm = i - 1 >= 0 && ((m >> 1) & 1u);
        const bool yp = j + 1 < A.dom[1] && ((m >> 4) & 1u), ym = j - 1 >= 0 && ((m >> 3) & 1u);
        const bool zp = k + 1 < A.dom[2] && ((m >> 6) & 1u), zm = k - 1 >= 0 && ((m >> 5) & 1u);
        const double dqx = one_sided_gradient(xp, xm, xp ? qxp : 0.0, M.Qx, xm ? qxm : 0.0, P.inv_dx[0]);
        const double dqy = one_sided_gradient(yp, ym, yp ? qyp : 0.0, M.Qy, ym ? qym : 0.0, P.inv_dx[1]);
        const double dqz = one_sided_gradient(zp, zm, zp ? qzp : 0.0, M.Qz, zm ? qzm : 0.0, P.inv_dx[2]);
        collide_cell(fv, gv, mo, M, dqx, dqy, dqz, P, [&](auto qc, double fnew, double gnew) {
            constexpr int qq = decltype(qc)::value;
            __builtin_nontemporal_store(fnew, fob + qq * s);
            __builtin_nontemporal_store(gnew, gob + qq * s);
        });
    }
}

// LBM::stream (LBM.cpp:558-604) at the listed ghost cells (fine cells over
// uncovered domain, from the explosion map): pull with the in-fab test
__global__ void __launch_bounds__(256) k_stream_listed(
    const double *__restrict__ f, const double *__restrict__ g, double *__restrict__ fo,
    double *__restrict__ go, const uint32_t *__restrict__ mask, const BoxDev *__restrict__ boxes,
    const ExplodeEntry *__restrict__ list, int64_t n)
{
    const int64_t t = xcd_block() * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const ExplodeEntry e = list[t];
    const BoxDev &B = boxes[e.fine_box];
    const int64_t loc = e.fine_local;
    const int64_t nxy = int64_t(B.n[0]) * B.n[1];
    const int kk = int(loc / nxy);
    const int jj = int((loc - kk * nxy) / B.n[0]);
    const int ii = int(loc - kk * nxy - int64_t(jj) * B.n[0]);
    // bit q of the mask: the source x - c_q is inside the fab and fluid
    const uint32_t m = mask[B.base + loc];
    const bool fluid = m & 1u;
    const double *fb = f + B.base * NQ;
    const double *gb = g + B.base * NQ;
    double *fob = fo + B.base * NQ;
    double *gob = go + B.base * NQ;
    const int64_t sy = B.n[0], sz = nxy;
    // every load issued unconditionally (a source outside the fab, or any
    // source of a solid cell, reads the cell itself), the -1 sentinel
    // selected afterwards
    double vf[NQ], vg[NQ];
    uint32_t keep = 0;
    static_for<NQ>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        constexpr int cx = ev(q, 0), cy = ev(q, 1), cz = ev(q, 2);
        const int a = ii - cx, b = jj - cy, c = kk - cz;
        const bool in_fab = a >= 0 && a < B.n[0] && b >= 0 && b < B.n[1] && c >= 0 && c < B.n[2];
        const bool use = fluid && in_fab;
        const bool direct = use && ((m >> q) & 1u);
        const int64_t idx = direct ? int64_t(q) * B.ncell + loc - (cx + cy * sy + cz * sz)
                                   : int64_t(kBounce[q]) * B.ncell + loc;
        vf[q] = fb[idx];
        vg[q] = gb[idx];
        keep |= uint32_t(use) << q;
    });
    static_for<NQ>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        const bool use = (keep >> q) & 1u;
        fob[q * B.ncell + loc] = use ? vf[q] : -1.0;
        gob[q * B.ncell + loc] = use ? vg[q] : -1.0;
    });
}

// per fab cell: bit q = is_fluid(x - c_q) != 0 (sources inside the fab),
// bit 0 = own is_fluid == 1
__global__ void k_fab_mask(
    const int32_t *__restrict__ isf, uint32_t *__restrict__ mask, const BoxDev *__restrict__ boxes, int nbox,
    const int64_t *__restrict__ base, int64_t total)
{
    const int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= total) return;
    int lo = 0, hi = nbox - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (base[mid] <= c)
            lo = mid;
        else
            hi = mid - 1;
    }
    const BoxDev B = boxes[lo];
    const int64_t loc = c - B.base;
    const int64_t nxy = int64_t(B.n[0]) * B.n[1];
    const int kk = int(loc / nxy);
    const int jj = int((loc - kk * nxy) / B.n[0]);
    const int ii = int(loc - kk * nxy - int64_t(jj) * B.n[0]);
    const int32_t *fl = isf + B.base * 2;
    uint32_t m = fl[loc] == 1 ? 1u : 0u;
    for (int q = 1; q < NQ; ++q) {
        const int a = ii - kEV[q][0], b = jj - kEV[q][1], e = kk - kEV[q][2];
        if (a >= 0 && a < B.n[0] && b >= 0 && b < B.n[1] && e >= 0 && e < B.n[2] &&
            fl[(int64_t(e) * B.n[1] + b) * B.n[0] + a] != 0)
            m |= 1u << q;
    }
    mask[c] = m;
}

unsigned grid_for(int ntiles, int blocks)
{
    // a multiple of 8 (one group per XCD), at most one block per tile per XCD
    const int full = (ntiles + 7) / 8 * 8;
    int g = blocks > 0 ? std::min(blocks, full) : full;
    g = std::max(8, (g + 7) / 8 * 8);
    return unsigned(g);
}

}  // namespace

void build_fab_tiles(const mrb_level *L, int grow, std::vector<FabTile> &tiles)
{
    tiles.clear();
    for (int b = 0; b < L->nbox; ++b) {
        const BoxDev &B = L->boxes[b];
        int64_t n = 1;
        for (int d = 0; d < 3; ++d) n *= B.vhi[d] - B.vlo[d] + 1 + 2 * grow;
        for (int64_t c = 0; c < n; c += kFabTile) tiles.push_back(FabTile{b, int32_t(c)});
    }
}

int fab_build_mask(mrb_level *L, hipStream_t s)
{
    const int64_t T = L->total_cells;
    k_fab_mask<<<unsigned((T + 255) / 256), 256, 0, s>>>(L->d_isf, L->d_mask, L->d_boxes, L->nbox,
                                                          L->d_box_base, T);
    MRB_HIP(hipGetLastError());
    return MRB_OK;
}

// LBM::f_to_macrodata (LBM.cpp:810-906) for the plotfile: the 19
// macrodata components of a cell's post-stream state (as pass 1 gathers
// it) into md; 0 for a solid cell
__device__ __forceinline__ void plot_md_store(const double fv[NQ], const double gv[NQ], bool fluid, double *out,
                                              int64_t s, const CollideParams &P)
{
    double o[19];
    for (int n = 0; n < 19; ++n) o[n] = 0.0;
    if (fluid) {
        Moments mo;
        accumulate_moments(fv, gv, mo);
        const Macro M = macro_from(mo.rho, mo.jx, mo.jy, mo.jz, mo.E, P);
        o[0] = mo.rho;
        o[1] = M.u;
        o[2] = M.v;
        o[3] = M.w;
        o[4] = sqrt(M.u * M.u + M.v * M.v + M.w * M.w);
        o[5] = mo.E;
        o[6] = M.Qx;
        o[7] = M.Qy;
        o[8] = M.Qz;
        o[9] = mo.pxx;
        o[10] = mo.pyy;
        o[11] = mo.pzz;
        o[12] = mo.pxy;
        o[13] = mo.pxz;
        o[14] = mo.pyz;
        o[15] = mo.qx;
        o[16] = mo.qy;
        o[17] = mo.qz;
        o[18] = M.T;
    }
    for (int n = 0; n < 19; ++n) out[n * s] = o[n];
}

// the valid cells: the gather of pass 1 from the buffer before the step
template <bool COAL>
__global__ void __launch_bounds__(kFabTile) k_fab_plot_md(
    const double *__restrict__ f, const double *__restrict__ g, double *__restrict__ md,
    const FabTile *__restrict__ tiles, FabArgs A, CoalArgs C, CollideParams P)
{
    int step, end;
    for (int t = xcd_tile_begin(A.ntiles, step, end); t < end; t += step) {
        const FabTile tile = tiles[t];
        const BoxDev &B = A.boxes[tile.box];
        int i, j, k;
        int64_t loc;
        if (!tile_cell(A, tile, int(threadIdx.x), B, i, j, k, loc)) continue;
        const uint32_t m = A.mask[B.base + loc];
        double fv[NQ], gv[NQ];
        if (m & 1u) gather_cell<COAL>(f, g, B, loc, m, C, fv, gv);
        plot_md_store(fv, gv, (m & 1u) != 0, md + B.base * 19 + loc, B.ncell, P);
    }
}

// the listed ring-1 ghost cells of a fine level (their pass 1)
__global__ void __launch_bounds__(256) k_fab_plot_md_listed(
    const double *__restrict__ f, const double *__restrict__ g, double *__restrict__ md,
    const BoxDev *__restrict__ boxes, const uint32_t *__restrict__ mask, const ExplodeEntry *__restrict__ list,
    int64_t n, CollideParams P)
{
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const ExplodeEntry e = list[t];
    const BoxDev &B = boxes[e.fine_box];
    const int64_t loc = e.fine_local;
    const uint32_t m = mask[B.base + loc];
    double fv[NQ], gv[NQ];
    if (m & 1u) pull_fab(f, g, B, loc, m, fv, gv);
    plot_md_store(fv, gv, (m & 1u) != 0, md + B.base * 19 + loc, B.ncell, P);
}

int fab_plot_md_pull(mrb_level *L, const mrb_level *coal_from, hipStream_t s)
{
    const int prev = 1 - L->cur;
    const double *f = (const double *)L->pop[0][prev];
    const double *g = (const double *)L->pop[1][prev];
    FabArgs A{};
    A.boxes = L->d_boxes;
    A.mask = L->d_mask;
    for (int d = 0; d < 3; ++d) A.dom[d] = L->ncell[d];
    A.grow = 0;
    A.ntiles = L->n_tiles_c;
    CoalArgs C{};
    if (coal_from) {
        // the values the last coarse pass gathered (the fine level has not
        // stepped since)
        C.vals = coal_from->d_coal_vals;
        C.has = coal_from->d_coal_has;
        C.cell = coal_from->d_coal_cell;
        C.n = coal_from->n_coal;
    }
    const unsigned gc = grid_for(A.ntiles, L->fab_blocks);
    if (coal_from && C.n > 0)
        k_fab_plot_md<true><<<gc, kFabTile, 0, s>>>(f, g, L->d_md, L->d_tiles_c, A, C, L->cp);
    else
        k_fab_plot_md<false><<<gc, kFabTile, 0, s>>>(f, g, L->d_md, L->d_tiles_c, A, C, L->cp);
    MRB_HIP(hipGetLastError());
    if (L->n_qring > 0) {
        k_fab_plot_md_listed<<<unsigned((L->n_qring + 255) / 256), 256, 0, s>>>(
            f, g, L->d_md, L->d_boxes, L->d_mask, (const ExplodeEntry *)L->d_qring, L->n_qring, L->cp);
        MRB_HIP(hipGetLastError());
    }
    return MRB_OK;
}

const int kFabBlockChoices[] = {0, 1024, 2048, 4096, 8192};
const int kNumFabBlockChoices = int(sizeof(kFabBlockChoices) / sizeof(kFabBlockChoices[0]));

int fab_batched_passes(mrb_level *L, const mrb_level *coal_from, hipStream_t s);

int fab_batched_step(mrb_level *L, const mrb_level *coal_from, int out_depth, hipStream_t s)
{
    int rc = fab_batched_passes(L, coal_from, s);
    if (rc) return rc;
    L->cur = 1 - L->cur;
    L->fresh_depth = 0;
    L->md_src = 2;
    if (out_depth > 0) return fab_fill_boundary_depth(L, out_depth, s);
    return MRB_OK;
}

// passes 1 and 2 (and the listed ghost stream) into the spare buffer; the
// current buffer is not modified
int fab_batched_passes(mrb_level *L, const mrb_level *coal_from, hipStream_t s)
{
    const int cur = L->cur, nxt = 1 - L->cur;
    const double *f = (const double *)L->pop[0][cur];
    const double *g = (const double *)L->pop[1][cur];
    double *fo = (double *)L->pop[0][nxt];
    double *go = (double *)L->pop[1][nxt];
    FabArgs A{};
    A.boxes = L->d_boxes;
    A.mask = L->d_mask;
    for (int d = 0; d < 3; ++d) A.dom[d] = L->ncell[d];
    CoalArgs C{};
    if (coal_from) {
        C.vals = coal_from->d_coal_vals;
        C.has = coal_from->d_coal_has;
        C.cell = coal_from->d_coal_cell;
        C.n = coal_from->n_coal;
        if (C.n > 0) {
            k_coalesce_vals<<<unsigned((C.n + 255) / 256), 256, 0, s>>>(
                (const double *)coal_from->pop[0][coal_from->cur], (const double *)coal_from->pop[1][coal_from->cur],
                coal_from->d_boxes, (const CoalesceEntry *)coal_from->d_coal, C.n, coal_from->d_coal_vals,
                coal_from->d_coal_has);
            MRB_HIP(hipGetLastError());
        }
    }
    A.grow = 0;
    A.ntiles = L->n_tiles_c;
    const unsigned gc = grid_for(A.ntiles, L->fab_blocks);
    if (coal_from)
        k_fab_qpass<true><<<gc, kFabTile, 0, s>>>(f, g, L->d_dq, L->d_tiles_c, A, C, L->cp);
    else
        k_fab_qpass<false><<<gc, kFabTile, 0, s>>>(f, g, L->d_dq, L->d_tiles_c, A, C, L->cp);
    MRB_HIP(hipGetLastError());
    if (L->n_qring > 0) {
        k_fab_qpass_listed<<<unsigned((L->n_qring + 255) / 256), 256, 0, s>>>(
            f, g, L->d_dq, L->d_boxes, L->d_mask, (const ExplodeEntry *)L->d_qring, L->n_qring, L->cp);
        MRB_HIP(hipGetLastError());
    }
    int rc = fab_fill_boundary_q(L, s);
    if (rc) return rc;
    if (coal_from)
        k_fab_collide<true><<<gc, kFabTile, 0, s>>>(f, g, fo, go, L->d_dq, L->d_tiles_c, A, C, L->cp);
    else
        k_fab_collide<false><<<gc, kFabTile, 0, s>>>(f, g, fo, go, L->d_dq, L->d_tiles_c, A, C, L->cp);
    MRB_HIP(hipGetLastError());
    const int64_t nst = L->stream_rings_le2 ? L->n_expl_le2 : L->n_expl;
    if (nst > 0) {
        k_stream_listed<<<unsigned((nst + 255) / 256), 256, 0, s>>>(
            f, g, fo, go, L->d_mask, L->d_boxes, (const ExplodeEntry *)L->d_expl, nst);
        MRB_HIP(hipGetLastError());
    }
    return MRB_OK;
}

}  // namespace mrb
