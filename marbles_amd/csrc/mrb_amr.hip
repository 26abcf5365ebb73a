// Two levels with refinement ratio 2: coalescence (coarse <- masked mean of
// the fine children), explosion (fine ghost cells <- fine neighbours or the
// coarse level) and the coarse-step schedule.
//
// Both transfers run from gather maps built once on the host when the
// levels are linked: one entry per written cell, so every launch is a
// flat, race-free pass over its entries with no box search.
#include <algorithm>
#include <cmath>
#include <vector>

#include <array>
#include <tuple>

#include "mrb_internal.h"

namespace mrb {

namespace {

constexpr int kAmrBlock = 256;

inline unsigned blocks_for(int64_t n) { return unsigned((n + kAmrBlock - 1) / kAmrBlock); }

// masked_avgdown (Utilities.H:315-350), ratio 2, (kref, jref, iref) order
__global__ void __launch_bounds__(kAmrBlock) k_coalesce(
    const double *__restrict__ fine, double *__restrict__ crse, const BoxDev *__restrict__ fboxes,
    const BoxDev *__restrict__ cboxes, const CoalesceEntry *__restrict__ map, int64_t n)
{
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const CoalesceEntry e = map[t];
    const BoxDev &FB = fboxes[e.fine_box];
    const BoxDev &CB = cboxes[e.crse_box];
    const double *f = fine + FB.base * NQ + e.fine_local;
    double *c = crse + CB.base * NQ + e.crse_local;
    const int64_t sy = FB.n[0], sz = int64_t(FB.n[0]) * FB.n[1];
    for (int q = 0; q < NQ; ++q) {
        const double *fq = f + q * FB.ncell;
        double sum = 0;
        double vol = 0;
        for (int kref = 0; kref < 2; ++kref)
            for (int jref = 0; jref < 2; ++jref)
                for (int iref = 0; iref < 2; ++iref) {
                    const double fval = fq[kref * sz + jref * sy + iref];
                    if (fabs(fval - (-1.0)) > kSmallNum) {
                        sum += fval;
                        vol += 1.0;
                    }
                }
        if (vol > 0.0) c[q * CB.ncell] = sum / vol;
    }
}

// Coarse-to-fine interpolation of FillPatchTwoLevels / InterpFromCoarseLevel
// with amrex::cell_cons_interp (FillPatchOps.H:124-128,174-177,190): AMReX's
// conservative linear scheme without linear limiting, ratio 2.  Per coarse
// parent and component (k_explode_slopes): monotonised-central slopes, the
// centred difference replaced at an ext_dir face (every non-periodic face,
// LBM.cpp:108-150) by the one-sided second-order form through the value
// beyond the face; then one factor alpha <= 1 so that no child leaves the
// min/max of the 3x3x3 neighbourhood.  The coarse patch AMReX interpolates
// from holds -1 beyond a non-periodic face: BCFill only writes the boundary
// value where iv + c_q lies inside the patch shrunk by the level's 3 ghost
// cells (BC.H:375-380), which a ghost's neighbour never does.  Per child
// (k_explode): ((c + o_x s_x) + o_y s_y) + o_z s_z with o = -+1/4.
// Every load is issued unconditionally: the fab's ghost layers (ng >= 2,
// checked when the map is built) hold P +- 2 in every direction.
__global__ void __launch_bounds__(kAmrBlock) k_explode_slopes(
    const double *__restrict__ crse, const BoxDev *__restrict__ cboxes, const ExplodeParent *__restrict__ par,
    int64_t npar, double *__restrict__ out)
{
    const int64_t t = xcd_block() * blockDim.x + threadIdx.x;
    if (t >= npar * NQ) return;
    const int q = int(t / npar);
    const int64_t p = t - int64_t(q) * npar;
    const ExplodeParent P = par[p];
    const BoxDev &CB = cboxes[P.crse_box];
    const double *c = crse + CB.base * NQ + int64_t(q) * CB.ncell + P.crse_local;
    const int64_t st[3] = {1, CB.n[0], int64_t(CB.n[0]) * CB.n[1]};
    const int fl = P.flags;
    double u[27];
    for (int oz = -1; oz <= 1; ++oz)
        for (int oy = -1; oy <= 1; ++oy)
            for (int ox = -1; ox <= 1; ++ox) {
                const double v = c[oz * st[2] + oy * st[1] + ox * st[0]];
                const bool beyond = (ox < 0 && (fl & 1)) || (ox > 0 && (fl & 8)) || (oy < 0 && (fl & 2)) ||
                                    (oy > 0 && (fl & 16)) || (oz < 0 && (fl & 4)) || (oz > 0 && (fl & 32));
                u[(oz + 1) * 9 + (oy + 1) * 3 + (ox + 1)] = beyond ? -1.0 : v;
            }
    double u2p[3], u2m[3];
    for (int d = 0; d < 3; ++d) {
        u2p[d] = c[2 * st[d]];
        u2m[d] = c[-2 * st[d]];
    }
    const double uc = u[13];
    double s[3];
    for (int d = 0; d < 3; ++d) {
        const int sn = d == 0 ? 1 : (d == 1 ? 3 : 9);
        const double um = u[13 - sn], up = u[13 + sn];
        double dc = 0.5 * (up - um);
        if (fl & (1 << d)) {
            if (fl & (1 << (6 + d)))
                dc = -16. / 15. * um + 0.5 * uc + 2. / 3. * up - 0.1 * u2p[d];
            else
                dc = 0.25 * (up + 5. * uc - 6. * um);
        }
        if (fl & (1 << (3 + d))) {
            if (fl & (1 << (9 + d)))
                dc = 16. / 15. * up - 0.5 * uc - 2. / 3. * um + 0.1 * u2m[d];
            else
                dc = -0.25 * (um + 5. * uc - 6. * up);
        }
        const double df = 2.0 * (up - uc);
        const double db = 2.0 * (uc - um);
        const double sl = (df * db >= 0.0) ? fmin(fabs(df), fabs(db)) : 0.0;
        s[d] = copysign(1.0, dc) * fmin(sl, fabs(dc));
    }
    double alpha = 1.0;
    if (s[0] != 0.0 || s[1] != 0.0 || s[2] != 0.0) {
        const double dumax = fabs(s[0]) * 1.0 / 4.0 + fabs(s[1]) * 1.0 / 4.0 + fabs(s[2]) * 1.0 / 4.0;
        double umax = uc, umin = uc;
        for (int n = 0; n < 27; ++n) {
            umin = fmin(umin, u[n]);
            umax = fmax(umax, u[n]);
        }
        if (dumax * alpha > (umax - uc)) alpha = (umax - uc) / dumax;
        if (dumax * alpha > (uc - umin)) alpha = (uc - umin) / dumax;
    }
    out[(int64_t(0) * NQ + q) * npar + p] = uc;
    for (int d = 0; d < 3; ++d) out[(int64_t(1 + d) * NQ + q) * npar + p] = s[d] * alpha;
}

__global__ void __launch_bounds__(kAmrBlock) k_explode(
    const double *__restrict__ slopes, int64_t npar, double *__restrict__ fine, const BoxDev *__restrict__ fboxes,
    const ExplodeEntry *__restrict__ map, int64_t n)
{
    const int64_t t = xcd_block() * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const ExplodeEntry e = map[t];
    const BoxDev &FB = fboxes[e.fine_box];
    double *fo = fine + FB.base * NQ + e.fine_local;
    const double ox = (e.bits & 1) ? 0.25 : -0.25;
    const double oy = (e.bits & 2) ? 0.25 : -0.25;
    const double oz = (e.bits & 4) ? 0.25 : -0.25;
    const int64_t p = e.parent;
    double v[NQ][4];
    static_for<NQ>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        for (int k = 0; k < 4; ++k) v[q][k] = slopes[(int64_t(k) * NQ + q) * npar + p];
    });
    static_for<NQ>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        fo[int64_t(q) * FB.ncell] = v[q][0] + ox * v[q][1] + oy * v[q][2] + oz * v[q][3];
    });
}

// regrid: new fine cell <- the old fine level's cell (NQ components)
struct LevelCopy {
    int64_t dst_local, src_local;
    int32_t dst_box, src_box;
};

__global__ void __launch_bounds__(kAmrBlock) k_copy_levels(
    const double *__restrict__ src, double *__restrict__ dst, const BoxDev *__restrict__ sboxes,
    const BoxDev *__restrict__ dboxes, const LevelCopy *__restrict__ map, int64_t n)
{
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const LevelCopy m = map[t];
    const BoxDev &S = sboxes[m.src_box];
    const BoxDev &D = dboxes[m.dst_box];
    const double *sp = src + S.base * NQ + m.src_local;
    double *dp = dst + D.base * NQ + m.dst_local;
    double v[NQ];
    static_for<NQ>([&](auto qc) { v[decltype(qc)::value] = sp[int64_t(decltype(qc)::value) * S.ncell]; });
    static_for<NQ>([&](auto qc) { dp[int64_t(decltype(qc)::value) * D.ncell] = v[decltype(qc)::value]; });
}

// fill_f_inside_eb (LBM.cpp:1278-1298): 0 in every solid fab cell
__global__ void __launch_bounds__(kAmrBlock) k_zero_solid(
    double *__restrict__ f, const int32_t *__restrict__ isf, const BoxDev *__restrict__ boxes,
    const int64_t *__restrict__ base, int nbox, int64_t total)
{
    const int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= total) return;
    int lo = 0, hi = nbox - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (base[mid] <= c) lo = mid; else hi = mid - 1;
    }
    const BoxDev &B = boxes[lo];
    const int64_t loc = c - B.base;
    if (isf[B.base * 2 + loc] != 0) return;
    for (int q = 0; q < NQ; ++q) f[B.base * NQ + int64_t(q) * B.ncell + loc] = 0.0;
}

int64_t local_of(const BoxDev &B, int i, int j, int k)
{
    return (int64_t(k - B.lo[2]) * B.n[1] + (j - B.lo[1])) * B.n[0] + (i - B.lo[0]);
}

// image one period away (AMReX's periodic shifts span one period); a cell
// more than one period outside a short periodic direction has none
bool periodic_image(const mrb_level *L, const int iv[3], int s[3])
{
    for (int d = 0; d < 3; ++d) {
        s[d] = iv[d];
        if (s[d] < 0) {
            if (!L->periodic[d]) return false;
            s[d] += L->ncell[d];
        } else if (s[d] >= L->ncell[d]) {
            if (!L->periodic[d]) return false;
            s[d] -= L->ncell[d];
        }
        if (s[d] < 0 || s[d] >= L->ncell[d]) return false;
    }
    return true;
}

// domain cell -> owning box (-1 if uncovered)
std::vector<int32_t> owner_map(const mrb_level *L)
{
    const int64_t n0 = L->ncell[0], n1 = L->ncell[1];
    std::vector<int32_t> own(size_t(n0 * n1 * L->ncell[2]), -1);
    for (int b = 0; b < L->nbox; ++b) {
        const BoxDev &B = L->boxes[b];
        for (int k = B.vlo[2]; k <= B.vhi[2]; ++k)
            for (int j = B.vlo[1]; j <= B.vhi[1]; ++j)
                for (int i = B.vlo[0]; i <= B.vhi[0]; ++i) own[size_t((k * n1 + j) * n0 + i)] = b;
    }
    return own;
}

int64_t dom_index(const mrb_level *L, const int s[3])
{
    return (int64_t(s[2]) * L->ncell[1] + s[1]) * L->ncell[0] + s[0];
}

int check_pair(const mrb_level *crse, const mrb_level *fine)
{
    if (!crse || !fine) return set_error(MRB_E_ARG, "null level");
    if (crse->layout != MRB_LAYOUT_FAB || fine->layout != MRB_LAYOUT_FAB)
        return set_error(MRB_E_UNSUPPORTED, "two-level transfers need the fab layout");
    if (crse->storage != MRB_STORAGE_F64 || fine->storage != MRB_STORAGE_F64)
        return set_error(MRB_E_UNSUPPORTED, "two-level transfers need FP64 storage");
    for (int d = 0; d < 3; ++d) {
        if (fine->ncell[d] != 2 * crse->ncell[d] || fine->periodic[d] != crse->periodic[d])
            return set_error(MRB_E_ARG, "fine level is not the coarse domain refined by 2");
    }
    if (fine->ng < 2) return set_error(MRB_E_ARG, "fine level needs at least 2 ghost layers");
    if (crse->ng < 1) return set_error(MRB_E_ARG, "coarse level needs at least 1 ghost layer");
    for (const BoxDev &B : fine->boxes)
        for (int d = 0; d < 3; ++d)
            if ((B.vlo[d] & 1) || !(B.vhi[d] & 1))
                return set_error(MRB_E_ARG, "fine boxes must be coarse-aligned (even lo, odd hi)");
    return MRB_OK;
}

// parent list of an interpolation: coarse cell P (its owner box and local
// index) -> index; flags as ExplodeParent
struct ParentSet {
    std::vector<ExplodeParent> list;
    std::vector<int32_t> index;  // per coarse domain cell, -1 if not a parent
    int depth = 1;               // coarse FillBoundary depth the slopes read
};

int32_t parent_of(const mrb_level *crse, const std::vector<int32_t> &cown, const int P[3], ParentSet &ps)
{
    if (ps.index.empty()) ps.index.assign(size_t(int64_t(crse->ncell[0]) * crse->ncell[1] * crse->ncell[2]), -1);
    const int64_t di = dom_index(crse, P);
    if (ps.index[size_t(di)] >= 0) return ps.index[size_t(di)];
    ExplodeParent e;
    e.crse_box = cown[size_t(di)];
    const BoxDev &CB = crse->boxes[size_t(e.crse_box)];
    e.crse_local = local_of(CB, P[0], P[1], P[2]);
    int fl = 0;
    for (int d = 0; d < 3; ++d) {
        if (crse->periodic[d]) continue;
        if (P[d] == 0) fl |= 1 << d;
        if (P[d] == crse->ncell[d] - 1) fl |= 1 << (3 + d);
        if (P[d] + 2 <= crse->ncell[d] - 1) fl |= 1 << (6 + d);
        if (P[d] - 2 >= 0) fl |= 1 << (9 + d);
        // the 4-point form reads P -+ 2 inward: a box one cell wide there
        // puts it in the second ghost layer
        if (((fl >> d) & 1) && ((fl >> (6 + d)) & 1) && P[d] + 2 > CB.vhi[d] + 1) ps.depth = crse->ng;
        if (((fl >> (3 + d)) & 1) && ((fl >> (9 + d)) & 1) && P[d] - 2 < CB.vlo[d] - 1) ps.depth = crse->ng;
    }
    e.flags = fl;
    ps.index[size_t(di)] = int32_t(ps.list.size());
    ps.list.push_back(e);
    return ps.index[size_t(di)];
}

// the slopes kernel reads P +- 2 along every direction and the 3x3x3 block
// unconditionally: those cells must lie inside the parent's fab
int check_parents(const mrb_level *crse, const ParentSet &ps)
{
    if (crse->ng < 2) return set_error(MRB_E_UNSUPPORTED, "the interpolation needs 2 coarse ghost layers");
    for (const ExplodeParent &e : ps.list) {
        const BoxDev &CB = crse->boxes[size_t(e.crse_box)];
        const int64_t nxy = int64_t(CB.n[0]) * CB.n[1];
        const int c[3] = {int(e.crse_local % CB.n[0]), int((e.crse_local % nxy) / CB.n[0]), int(e.crse_local / nxy)};
        for (int d = 0; d < 3; ++d)
            if (c[d] - 2 < 0 || c[d] + 2 >= CB.n[d])
                return set_error(MRB_E_ARG, "interpolation parent too close to its fab edge");
    }
    return MRB_OK;
}

}  // namespace

static void drop_graphs(mrb_level *L)
{
    for (auto &row : L->amr_graph)
        for (auto &G : row) {
            if (G.exec) (void)hipGraphExecDestroy(G.exec);
            G = mrb_level::AmrGraph{};
        }
}

void amr_release(mrb_level *L)
{
    drop_graphs(L);
    if (L->cap_stream) (void)hipStreamDestroy(L->cap_stream);
    L->cap_stream = nullptr;
    if (L->d_coal) (void)hipFree(L->d_coal);
    if (L->d_expl) (void)hipFree(L->d_expl);
    if (L->d_expl_par) (void)hipFree(L->d_expl_par);
    if (L->d_expl_slopes) (void)hipFree(L->d_expl_slopes);
    L->d_expl_par = nullptr;
    L->d_expl_slopes = nullptr;
    L->n_expl_par = 0;
    L->expl_crse_depth = 1;
    if (L->d_coal_cell) (void)hipFree(L->d_coal_cell);
    if (L->d_coal_vals) (void)hipFree(L->d_coal_vals);
    if (L->d_qring) (void)hipFree(L->d_qring);
    L->d_qring = nullptr;
    if (L->d_gmap_step) (void)hipFree(L->d_gmap_step);
    L->d_gmap_step = nullptr;
    L->n_gmap_step = 0;
    L->n_expl_le2 = 0;
    if (L->d_coal_has) (void)hipFree(L->d_coal_has);
    L->d_coal = L->d_expl = nullptr;
    L->d_coal_cell = nullptr;
    L->d_coal_vals = nullptr;
    L->d_coal_has = nullptr;
    L->n_coal = L->n_expl = L->n_qring = 0;
    L->amr_crse = nullptr;
}

}  // namespace mrb

using namespace mrb;

extern "C" {

int mrb_amr_link(mrb_level *crse, mrb_level *fine)
{
    int rc = check_pair(crse, fine);
    if (rc) return rc;
    crse->in_hierarchy = fine->in_hierarchy = true;
    crse->amr_fine = fine;
    const std::vector<int32_t> cown = owner_map(crse);
    for (int32_t o : cown)
        if (o < 0) return set_error(MRB_E_ARG, "the coarse level must cover its domain");
    const std::vector<int32_t> fown = owner_map(fine);

    // coalescence: coarsened fine boxes grown by one, box order, the last
    // box wins where rings overlap (average_down_with_ghosts,
    // Utilities.cpp:5-28, ParallelCopy back with periodicity)
    // A coarse cell under a fine box takes its children from that box's
    // valid cells: any other box's copies of them (FillBoundary ghosts) are
    // equal, so the last-box rule only decides the ring cells over
    // uncovered domain, whose children differ between boxes.
    std::vector<int32_t> under(cown.size(), -1);
    for (int b = 0; b < fine->nbox; ++b) {
        const BoxDev &FB = fine->boxes[b];
        for (int k = FB.vlo[2] / 2; k <= FB.vhi[2] / 2; ++k)
            for (int j = FB.vlo[1] / 2; j <= FB.vhi[1] / 2; ++j)
                for (int i = FB.vlo[0] / 2; i <= FB.vhi[0] / 2; ++i) {
                    const int s[3] = {i, j, k};
                    under[size_t(dom_index(crse, s))] = b;
                }
    }
    std::vector<CoalesceEntry> coal;
    std::vector<int64_t> slot(cown.size(), -1);
    for (int b = 0; b < fine->nbox; ++b) {
        const BoxDev &FB = fine->boxes[b];
        for (int k = FB.vlo[2] / 2 - 1; k <= FB.vhi[2] / 2 + 1; ++k)
            for (int j = FB.vlo[1] / 2 - 1; j <= FB.vhi[1] / 2 + 1; ++j)
                for (int i = FB.vlo[0] / 2 - 1; i <= FB.vhi[0] / 2 + 1; ++i) {
                    const int iv[3] = {i, j, k};
                    int s[3];
                    if (!periodic_image(crse, iv, s)) continue;
                    const int64_t di = dom_index(crse, s);
                    // under a footprint: only the owner's own valid children
                    // (not a periodic image, whose children are ghosts)
                    if (under[size_t(di)] >= 0 &&
                        (under[size_t(di)] != b || s[0] != i || s[1] != j || s[2] != k))
                        continue;
                    const int ob = cown[size_t(di)];
                    CoalesceEntry e;
                    e.crse_box = ob;
                    e.crse_local = local_of(crse->boxes[ob], s[0], s[1], s[2]);
                    e.fine_box = b;
                    e.fine_local = local_of(FB, 2 * i, 2 * j, 2 * k);
                    if (slot[size_t(di)] >= 0) {
                        coal[size_t(slot[size_t(di)])] = e;
                    } else {
                        slot[size_t(di)] = int64_t(coal.size());
                        coal.push_back(e);
                    }
                }
    }
    // explosion: fine ghost cells whose periodic image is inside the domain
    // and not covered by a fine box
    std::vector<ExplodeEntry> expl;
    ParentSet parents;
    for (int b = 0; b < fine->nbox; ++b) {
        const BoxDev &FB = fine->boxes[b];
        for (int k = FB.lo[2]; k < FB.lo[2] + FB.n[2]; ++k)
            for (int j = FB.lo[1]; j < FB.lo[1] + FB.n[1]; ++j)
                for (int i = FB.lo[0]; i < FB.lo[0] + FB.n[0]; ++i) {
                    if (i >= FB.vlo[0] && i <= FB.vhi[0] && j >= FB.vlo[1] && j <= FB.vhi[1] &&
                        k >= FB.vlo[2] && k <= FB.vhi[2])
                        continue;
                    const int iv[3] = {i, j, k};
                    int s[3];
                    if (!periodic_image(fine, iv, s)) continue;
                    if (fown[size_t(dom_index(fine, s))] >= 0) continue;
                    const int P[3] = {s[0] / 2, s[1] / 2, s[2] / 2};
                    const int ob = cown[size_t(dom_index(crse, P))];
                    ExplodeEntry e{};
                    e.fine_box = b;
                    e.fine_local = local_of(FB, i, j, k);
                    e.crse_box = ob;
                    e.crse_local = local_of(crse->boxes[ob], P[0], P[1], P[2]);
                    e.parent = parent_of(crse, cown, P, parents);
                    for (int d = 0; d < 3; ++d) e.bits |= (s[d] & 1) << d;
                    expl.push_back(e);
                }
    }
    // Order the uncovered ghost cells as (ring <= 2 first, then ring 3),
    // each group by (box, fab index): consecutive threads of the listed
    // kernels then touch consecutive cells of a row.  The second substep
    // streams rings 1-2 only (ring-3 values streamed then are overwritten
    // by the next explosion before anything reads them).  Ring-1 cells are
    // also listed apart: the batched step computes their Q from their own
    // pulls (they have no owning box to copy it from).
    auto ring_of = [&](const ExplodeEntry &e) {
        const BoxDev &FB = fine->boxes[e.fine_box];
        const int64_t nxy = int64_t(FB.n[0]) * FB.n[1];
        const int c[3] = {int(e.fine_local % FB.n[0]) + FB.lo[0], int((e.fine_local % nxy) / FB.n[0]) + FB.lo[1],
                          int(e.fine_local / nxy) + FB.lo[2]};
        int r = 0;
        for (int d = 0; d < 3; ++d) r = std::max(r, std::max(FB.vlo[d] - c[d], c[d] - FB.vhi[d]));
        return r;
    };
    std::vector<uint8_t> ring(expl.size());
    std::vector<std::pair<uint64_t, uint32_t>> key(expl.size());
    for (size_t n = 0; n < expl.size(); ++n) {
        ring[n] = uint8_t(ring_of(expl[n]));
        key[n] = {(uint64_t(ring[n] >= 3) << 63) | (uint64_t(expl[n].fine_box) << 40) | uint64_t(expl[n].fine_local),
                  uint32_t(n)};
    }
    std::sort(key.begin(), key.end());
    std::vector<ExplodeEntry> sorted(expl.size());
    int64_t n_le2 = 0;
    std::vector<ExplodeEntry> qring;
    for (size_t n = 0; n < key.size(); ++n) {
        const uint32_t o = key[n].second;
        sorted[n] = expl[o];
        if (ring[o] <= 2) ++n_le2;
        if (ring[o] == 1) qring.push_back(expl[o]);
    }
    expl.swap(sorted);
    // The fine level's FillBoundary inside the two-level step: every first-
    // layer ghost (the pulls of valid cells), and deeper ghosts only where
    // a kernel of the step reads them: next to an uncovered ghost (the
    // listed ghost stream and Q pass pull its 26 neighbours) or within the
    // ghost depth of a non-periodic domain face (the boundary functor).
    // The other deeper ghosts keep the values of the last full fill.
    std::vector<GhostPair> step_map;
    {
        std::vector<uint8_t> need(size_t(fine->total_cells), 0);
        auto fab_ijk = [](const BoxDev &B, int64_t loc, int c[3]) {
            const int64_t nxy = int64_t(B.n[0]) * B.n[1];
            c[0] = int(loc % B.n[0]);
            c[1] = int((loc % nxy) / B.n[0]);
            c[2] = int(loc / nxy);
        };
        for (const ExplodeEntry &e : expl) {
            const BoxDev &FB = fine->boxes[e.fine_box];
            int c[3];
            fab_ijk(FB, e.fine_local, c);
            for (int dz = -1; dz <= 1; ++dz)
                for (int dy = -1; dy <= 1; ++dy)
                    for (int dx = -1; dx <= 1; ++dx) {
                        const int a = c[0] + dx, b = c[1] + dy, k = c[2] + dz;
                        if (a < 0 || b < 0 || k < 0 || a >= FB.n[0] || b >= FB.n[1] || k >= FB.n[2]) continue;
                        need[size_t(FB.base + (int64_t(k) * FB.n[1] + b) * FB.n[0] + a)] = 1;
                    }
        }
        std::vector<GhostPair> full;
        build_ghost_map(fine, fine->ng, full);
        for (const GhostPair &p : full) {
            const BoxDev &FB = fine->boxes[p.dst_box];
            int c[3];
            fab_ijk(FB, p.dst_local, c);
            int layer = 0;
            bool near_face = false;
            for (int d = 0; d < 3; ++d) {
                const int g = FB.lo[d] + c[d];  // domain index
                layer = std::max(layer, std::max(FB.vlo[d] - g, g - FB.vhi[d]));
                if (!fine->periodic[d] && (g < fine->ng || g >= fine->ncell[d] - fine->ng)) near_face = true;
            }
            if (layer <= 1 || near_face || need[size_t(FB.base + p.dst_local)]) step_map.push_back(p);
        }
    }
    // per coarse fab cell: the coalescence entry of valid cells under the
    // fine footprint or its ring
    std::vector<int32_t> coal_cell(size_t(crse->total_cells), -1);
    for (int b = 0; b < crse->nbox; ++b) {
        const BoxDev &CB = crse->boxes[b];
        for (int k = CB.vlo[2]; k <= CB.vhi[2]; ++k)
            for (int j = CB.vlo[1]; j <= CB.vhi[1]; ++j)
                for (int i = CB.vlo[0]; i <= CB.vhi[0]; ++i) {
                    const int s[3] = {i, j, k};
                    const int64_t sl = slot[size_t(dom_index(crse, s))];
                    if (sl >= 0) coal_cell[size_t(CB.base + local_of(CB, i, j, k))] = int32_t(sl);
                }
    }
    // every address the kernels form must lie inside its fab
    for (const CoalesceEntry &e : coal) {
        const BoxDev &FB = fine->boxes[e.fine_box];
        const BoxDev &CB = crse->boxes[e.crse_box];
        const int64_t last = e.fine_local + int64_t(FB.n[0]) * FB.n[1] + FB.n[0] + 1;
        if (e.fine_local < 0 || last >= FB.ncell || e.crse_local < 0 || e.crse_local >= CB.ncell)
            return set_error(MRB_E_ARG, "coalescence map leaves a fab (fine ghost layers too few)");
    }
    for (const ExplodeEntry &e : expl) {
        const BoxDev &FB = fine->boxes[e.fine_box];
        if (e.fine_local < 0 || e.fine_local >= FB.ncell || e.parent < 0 ||
            e.parent >= int32_t(parents.list.size()))
            return set_error(MRB_E_ARG, "explosion map leaves a fab");
    }
    if ((rc = check_parents(crse, parents))) return rc;
    amr_release(fine);
    if (!parents.list.empty()) {
        MRB_HIP(hipMalloc(&fine->d_expl_par, sizeof(ExplodeParent) * parents.list.size()));
        MRB_HIP(hipMemcpy(fine->d_expl_par, parents.list.data(), sizeof(ExplodeParent) * parents.list.size(),
                          hipMemcpyHostToDevice));
        MRB_HIP(hipMalloc(&fine->d_expl_slopes, sizeof(double) * 4 * NQ * parents.list.size()));
    }
    fine->n_expl_par = int64_t(parents.list.size());
    fine->expl_crse_depth = parents.depth;
    if (!coal.empty()) {
        MRB_HIP(hipMalloc(&fine->d_coal, sizeof(CoalesceEntry) * coal.size()));
        MRB_HIP(hipMemcpy(fine->d_coal, coal.data(), sizeof(CoalesceEntry) * coal.size(), hipMemcpyHostToDevice));
    }
    if (!expl.empty()) {
        MRB_HIP(hipMalloc(&fine->d_expl, sizeof(ExplodeEntry) * expl.size()));
        MRB_HIP(hipMemcpy(fine->d_expl, expl.data(), sizeof(ExplodeEntry) * expl.size(), hipMemcpyHostToDevice));
    }
    if (!step_map.empty()) {
        MRB_HIP(hipMalloc(&fine->d_gmap_step, sizeof(GhostPair) * step_map.size()));
        MRB_HIP(hipMemcpy(fine->d_gmap_step, step_map.data(), sizeof(GhostPair) * step_map.size(),
                          hipMemcpyHostToDevice));
    }
    fine->n_gmap_step = int64_t(step_map.size());
    if (!qring.empty()) {
        MRB_HIP(hipMalloc(&fine->d_qring, sizeof(ExplodeEntry) * qring.size()));
        MRB_HIP(hipMemcpy(fine->d_qring, qring.data(), sizeof(ExplodeEntry) * qring.size(), hipMemcpyHostToDevice));
    }
    if (!coal.empty()) {
        MRB_HIP(hipMalloc(&fine->d_coal_vals, sizeof(double) * 2 * NQ * coal.size()));
        MRB_HIP(hipMalloc(&fine->d_coal_has, sizeof(uint64_t) * coal.size()));
    }
    MRB_HIP(hipMalloc(&fine->d_coal_cell, sizeof(int32_t) * coal_cell.size()));
    MRB_HIP(hipMemcpy(fine->d_coal_cell, coal_cell.data(), sizeof(int32_t) * coal_cell.size(),
                      hipMemcpyHostToDevice));
    fine->n_coal = int64_t(coal.size());
    fine->n_expl = int64_t(expl.size());
    fine->n_qring = int64_t(qring.size());
    fine->n_expl_le2 = n_le2;
    fine->amr_crse = crse;
    ++fine->gen;
    return MRB_OK;
}

int mrb_coalesce(const mrb_level *fine, mrb_level *crse, void *stream)
{
    if (!fine || !crse) return set_error(MRB_E_ARG, "null level");
    if (fine->amr_crse != crse) return set_error(MRB_E_STATE, "levels are not linked (mrb_amr_link)");
    if (fine->n_coal == 0) return MRB_OK;
    hipStream_t s = as_stream(stream);
    crse->fresh_depth = 0;
    for (int l = 0; l < 2; ++l) {
        k_coalesce<<<blocks_for(fine->n_coal), kAmrBlock, 0, s>>>(
            (const double *)fine->pop[l][fine->cur], (double *)crse->pop[l][crse->cur], fine->d_boxes,
            crse->d_boxes, (const CoalesceEntry *)fine->d_coal, fine->n_coal);
        MRB_HIP(hipGetLastError());
    }
    return MRB_OK;
}

namespace mrb {
// the two-level fillpatch of one lattice; `step_map`: the fine FillBoundary
// covers what the two-level step reads (the linked pair's step map) rather
// than every ghost layer
static int explode_impl(const mrb_level *crse, mrb_level *fine, int lattice, hipStream_t s, bool step_map);
}  // namespace mrb

int mrb_explode(const mrb_level *crse, mrb_level *fine, int lattice, void *stream)
{
    if (!fine || !crse) return set_error(MRB_E_ARG, "null level");
    if (fine->amr_crse != crse) return set_error(MRB_E_STATE, "levels are not linked (mrb_amr_link)");
    if (lattice < 0 || lattice > 1) return set_error(MRB_E_ARG, "lattice must be 0 or 1");
    return explode_impl(crse, fine, lattice, as_stream(stream), false);
}

namespace mrb {
static int explode_impl(const mrb_level *crse, mrb_level *fine, int lattice, hipStream_t s, bool step_map)
{
    // the slopes read the coarse parents' neighbourhoods: FillBoundary of
    // the coarse level (ghost layers only, so its valid data is unchanged)
    const int depth = std::max(1, fine->expl_crse_depth);
    int rc;
    if (crse->fresh_depth < depth &&
        (rc = fab_fill_boundary_depth(const_cast<mrb_level *>(crse), depth, s)))
        return rc;
    // fine-fine ghosts first, then the uncovered ones, then the domain
    // boundary (the three sets of cells are disjoint; physbc reads valid
    // and in-domain ghost cells only)
    const bool part = step_map && fine->d_gmap_step;
    rc = part ? fab_fill_boundary_map(fine, lattice, fine->d_gmap_step, fine->n_gmap_step, s)
              : fab_fill_boundary(fine, lattice, s);
    if (rc) return rc;
    if (lattice == 1) fine->fresh_depth = part ? 1 : fine->ng;
    if (fine->n_expl > 0) {
        k_explode_slopes<<<blocks_for(fine->n_expl_par * NQ), kAmrBlock, 0, s>>>(
            (const double *)crse->pop[lattice][crse->cur], crse->d_boxes, (const ExplodeParent *)fine->d_expl_par,
            fine->n_expl_par, fine->d_expl_slopes);
        MRB_HIP(hipGetLastError());
        k_explode<<<blocks_for(fine->n_expl), kAmrBlock, 0, s>>>(
            fine->d_expl_slopes, fine->n_expl_par, (double *)fine->pop[lattice][fine->cur], fine->d_boxes,
            (const ExplodeEntry *)fine->d_expl, fine->n_expl);
        MRB_HIP(hipGetLastError());
    }
    return mrb_physbc(fine, lattice, s);
}
}  // namespace mrb

static int check_advance(mrb_level *crse, mrb_level *fine)
{
    if (!fine || !crse) return set_error(MRB_E_ARG, "null level");
    if (fine->amr_crse != crse) return set_error(MRB_E_STATE, "levels are not linked (mrb_amr_link)");
    if (!crse->flags_ready || !fine->flags_ready) return set_error(MRB_E_STATE, "flags not set");
    if (fine->phys.dt != 0.5 * crse->phys.dt)
        return set_error(MRB_E_STATE, "fine level dt must be half the coarse dt");
    return MRB_OK;
}

// one coarse step: LBM::evolve (LBM.cpp:416-418) and time_step(0) (498-512).
// FillBoundary depths follow the readers: the coarse level's pulls and the
// explosion read one ghost layer of the coarse level; the fine level's
// uncovered ghosts are streamed from up to ng layers during the first
// substep; after the second substep nothing reads fine ghosts covered by
// fine boxes (the coalescence takes those cells' children from the owning
// box) before the next explosion refills them.
static int amr_step(mrb_level *crse, mrb_level *fine, void *stream)
{
    int rc;
    hipStream_t s = as_stream(stream);
    const int depth = std::max(1, fine->expl_crse_depth);
    if (crse->fresh_depth < depth && (rc = fab_fill_boundary_depth(crse, depth, s))) return rc;
    for (int l = 0; l < 2; ++l)
        if ((rc = mrb_physbc(crse, l, stream))) return rc;
    for (int l = 0; l < 2; ++l)
        if ((rc = explode_impl(crse, fine, l, s, true))) return rc;
    for (int sub = 0; sub < 2; ++sub) {
        for (int l = 0; l < 2; ++l)
            if ((rc = mrb_physbc(fine, l, stream))) return rc;
        // time_step(1) -> advance(1) (LBM.cpp:523-544): stream f, g and
        // collide as one batched step
        fine->stream_rings_le2 = sub == 1;
        rc = fab_batched_step(fine, nullptr, 0, s);
        fine->stream_rings_le2 = false;
        if (rc) return rc;
        if (sub == 0) {
            // the ghosts the second substep reads
            for (int l = 0; l < 2; ++l) {
                rc = fine->d_gmap_step ? fab_fill_boundary_map(fine, l, fine->d_gmap_step, fine->n_gmap_step, s)
                                       : fab_fill_boundary(fine, l, s);
                if (rc) return rc;
            }
            fine->fresh_depth = fine->d_gmap_step ? 1 : fine->ng;
        }
    }
    // advance(0): stream, average_down_to, collide, with the coalescence
    // fused into the coarse passes
    return fab_batched_step(crse, fine, 1, s);
}

int mrb_amr_regrid_fill(mrb_level *crse, const mrb_level *old_fine, mrb_level *fine)
{
    int rc = check_pair(crse, fine);
    if (rc) return rc;
    if (!old_fine) return set_error(MRB_E_ARG, "null level");
    for (int d = 0; d < 3; ++d)
        if (old_fine->ncell[d] != fine->ncell[d]) return set_error(MRB_E_ARG, "old and new fine domains differ");
    if (old_fine->layout != MRB_LAYOUT_FAB || old_fine->storage != MRB_STORAGE_F64)
        return set_error(MRB_E_UNSUPPORTED, "regrid needs fab levels with FP64 storage");
    if (!fine->flags_ready) return set_error(MRB_E_STATE, "the new fine level's flags are not set");
    const std::vector<int32_t> cown = owner_map(crse);
    const std::vector<int32_t> oown = owner_map(old_fine);
    // FillPatchTwoLevels (FillPatchOps.H:118-129) over the new valid cells:
    // the old fine level where it has the cell, else the interpolation from
    // the coarse parent
    std::vector<LevelCopy> copy;
    std::vector<ExplodeEntry> interp;
    ParentSet parents;
    for (int b = 0; b < fine->nbox; ++b) {
        const BoxDev &FB = fine->boxes[size_t(b)];
        for (int k = FB.vlo[2]; k <= FB.vhi[2]; ++k)
            for (int j = FB.vlo[1]; j <= FB.vhi[1]; ++j)
                for (int i = FB.vlo[0]; i <= FB.vhi[0]; ++i) {
                    const int iv[3] = {i, j, k};
                    const int ob = oown[size_t(dom_index(fine, iv))];
                    if (ob >= 0) {
                        LevelCopy m;
                        m.dst_box = b;
                        m.dst_local = local_of(FB, i, j, k);
                        m.src_box = ob;
                        m.src_local = local_of(old_fine->boxes[size_t(ob)], i, j, k);
                        copy.push_back(m);
                        continue;
                    }
                    const int P[3] = {i / 2, j / 2, k / 2};
                    ExplodeEntry e{};
                    e.fine_box = b;
                    e.fine_local = local_of(FB, i, j, k);
                    e.crse_box = cown[size_t(dom_index(crse, P))];
                    e.crse_local = local_of(crse->boxes[size_t(e.crse_box)], P[0], P[1], P[2]);
                    e.parent = parent_of(crse, cown, P, parents);
                    for (int d = 0; d < 3; ++d) e.bits |= (iv[d] & 1) << d;
                    interp.push_back(e);
                }
    }
    if ((rc = check_parents(crse, parents))) return rc;
    // the interpolation reads the coarse parents' neighbourhoods
    const int depth = std::max(1, parents.depth);
    if (crse->fresh_depth < depth && (rc = fab_fill_boundary_depth(crse, depth, nullptr))) return rc;
    const int64_t npar = int64_t(parents.list.size());
    void *d_copy = nullptr, *d_interp = nullptr, *d_par = nullptr;
    double *d_slopes = nullptr;
    if (npar > 0) {
        MRB_HIP(hipMalloc(&d_par, sizeof(ExplodeParent) * size_t(npar)));
        MRB_HIP(hipMemcpy(d_par, parents.list.data(), sizeof(ExplodeParent) * size_t(npar), hipMemcpyHostToDevice));
        MRB_HIP(hipMalloc(&d_slopes, sizeof(double) * 4 * NQ * size_t(npar)));
    }
    if (!copy.empty()) {
        MRB_HIP(hipMalloc(&d_copy, sizeof(LevelCopy) * copy.size()));
        MRB_HIP(hipMemcpy(d_copy, copy.data(), sizeof(LevelCopy) * copy.size(), hipMemcpyHostToDevice));
    }
    if (!interp.empty()) {
        MRB_HIP(hipMalloc(&d_interp, sizeof(ExplodeEntry) * interp.size()));
        MRB_HIP(hipMemcpy(d_interp, interp.data(), sizeof(ExplodeEntry) * interp.size(), hipMemcpyHostToDevice));
    }
    for (int l = 0; l < 2; ++l) {
        double *dst = (double *)fine->pop[l][fine->cur];
        if (!copy.empty())
            k_copy_levels<<<blocks_for(int64_t(copy.size())), kAmrBlock>>>(
                (const double *)old_fine->pop[l][old_fine->cur], dst, old_fine->d_boxes, fine->d_boxes,
                (const LevelCopy *)d_copy, int64_t(copy.size()));
        if (!interp.empty()) {
            k_explode_slopes<<<blocks_for(npar * NQ), kAmrBlock>>>((const double *)crse->pop[l][crse->cur],
                                                                  crse->d_boxes, (const ExplodeParent *)d_par, npar,
                                                                  d_slopes);
            k_explode<<<blocks_for(int64_t(interp.size())), kAmrBlock>>>(
                d_slopes, npar, dst, fine->d_boxes, (const ExplodeEntry *)d_interp, int64_t(interp.size()));
        }
        k_zero_solid<<<blocks_for(fine->total_cells), kAmrBlock>>>(dst, fine->d_isf, fine->d_boxes,
                                                                     fine->d_box_base, fine->nbox, fine->total_cells);
    }
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipDeviceSynchronize();
    (void)hipFree(d_copy);
    (void)hipFree(d_interp);
    (void)hipFree(d_par);
    (void)hipFree(d_slopes);
    if (e != hipSuccess) return hip_fail(e, "mrb_amr_regrid_fill");
    fine->fresh_depth = 0;
    fine->pops_set = true;
    fine->prev_valid = false;
    fine->md_src = 0;
    ++fine->gen;
    return MRB_OK;
}

// Berger-Rigoutsos clustering of tagged points (integer coordinates in a
// block grid): a cluster whose bounding box is filled to at least grid_eff
// is kept; otherwise it is cut at a plane without points (the one nearest
// the middle of the longest such direction), else at the strongest sign
// change of the second difference of the point counts per plane, else in
// the middle of its longest side.
static void br_cluster(std::vector<std::array<int, 3>> pts, double grid_eff, std::vector<std::array<int, 6>> &out)
{
    std::vector<std::vector<std::array<int, 3>>> stack;
    stack.push_back(std::move(pts));
    while (!stack.empty()) {
        std::vector<std::array<int, 3>> P = std::move(stack.back());
        stack.pop_back();
        if (P.empty()) continue;
        int lo[3] = {P[0][0], P[0][1], P[0][2]}, hi[3] = {lo[0], lo[1], lo[2]};
        for (const auto &p : P)
            for (int d = 0; d < 3; ++d) {
                lo[d] = std::min(lo[d], p[d]);
                hi[d] = std::max(hi[d], p[d]);
            }
        const double vol = double(hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1) * (hi[2] - lo[2] + 1);
        if (double(P.size()) >= grid_eff * vol || (lo[0] == hi[0] && lo[1] == hi[1] && lo[2] == hi[2])) {
            out.push_back({lo[0], lo[1], lo[2], hi[0], hi[1], hi[2]});
            continue;
        }
        std::vector<int> sig[3];
        for (int d = 0; d < 3; ++d) sig[d].assign(size_t(hi[d] - lo[d] + 1), 0);
        for (const auto &p : P)
            for (int d = 0; d < 3; ++d) ++sig[d][size_t(p[d] - lo[d])];
        int dims[3] = {0, 1, 2};
        std::stable_sort(dims, dims + 3, [&](int a, int b) { return hi[a] - lo[a] > hi[b] - lo[b]; });
        int cut_d = -1, cut_at = 0;  // points with coordinate < cut_at go left
        for (int di = 0; di < 3 && cut_d < 0; ++di) {
            const int d = dims[di], n = int(sig[d].size());
            int best = -1;
            for (int i = 1; i + 1 < n; ++i)
                if (sig[d][size_t(i)] == 0 && (best < 0 || std::abs(2 * i - (n - 1)) < std::abs(2 * best - (n - 1))))
                    best = i;
            if (best >= 0) {
                cut_d = d;
                cut_at = lo[d] + best;
            }
        }
        if (cut_d < 0) {
            int strength = 0;
            for (int di = 0; di < 3; ++di) {
                const int d = dims[di], n = int(sig[d].size());
                for (int i = 1; i + 2 < n; ++i) {
                    const int l0 = sig[d][size_t(i - 1)] - 2 * sig[d][size_t(i)] + sig[d][size_t(i + 1)];
                    const int l1 = sig[d][size_t(i)] - 2 * sig[d][size_t(i + 1)] + sig[d][size_t(i + 2)];
                    if ((l0 < 0 && l1 > 0) || (l0 > 0 && l1 < 0)) {
                        const int st = std::abs(l1 - l0);
                        if (st > strength) {
                            strength = st;
                            cut_d = d;
                            cut_at = lo[d] + i + 1;
                        }
                    }
                }
            }
        }
        if (cut_d < 0) {
            cut_d = dims[0];
            cut_at = lo[cut_d] + (hi[cut_d] - lo[cut_d] + 1) / 2;
        }
        std::vector<std::array<int, 3>> L, R;
        for (const auto &p : P) (p[cut_d] < cut_at ? L : R).push_back(p);
        stack.push_back(std::move(R));
        stack.push_back(std::move(L));
    }
}

int mrb_cluster_tags(const uint8_t *tags, const int ncell[3], int n_error_buf, int blocking_factor,
                     int max_grid_size, double grid_eff, int *nbox, int *box_lo, int *box_hi, int max_boxes)
{
    if (!tags || !ncell || !nbox) return set_error(MRB_E_ARG, "null argument");
    if (blocking_factor < 2 || (blocking_factor & 1) || max_grid_size < blocking_factor ||
        max_grid_size % blocking_factor || n_error_buf < 0 || !(grid_eff > 0.0 && grid_eff <= 1.0))
        return set_error(MRB_E_ARG, "need an even blocking factor, max_grid_size a multiple of it, 0 < grid_eff <= 1");
    const int bc = blocking_factor / 2;  // block edge in coarse cells
    for (int d = 0; d < 3; ++d)
        if (ncell[d] < 1 || ncell[d] % bc) return set_error(MRB_E_ARG, "blocking factor must divide the fine domain");
    const int nb[3] = {ncell[0] / bc, ncell[1] / bc, ncell[2] / bc};
    // tags grown by n_error_buf cells (TagBoxArray::buffer), then coarsened
    // to blocks of blocking_factor fine cells
    std::vector<uint8_t> blk(size_t(nb[0]) * nb[1] * nb[2], 0);
    const int64_t n0 = ncell[0], n1 = ncell[1];
    for (int k = 0; k < ncell[2]; ++k)
        for (int j = 0; j < ncell[1]; ++j)
            for (int i = 0; i < ncell[0]; ++i) {
                if (!tags[size_t((k * n1 + j) * n0 + i)]) continue;
                const int lo[3] = {std::max(0, i - n_error_buf) / bc, std::max(0, j - n_error_buf) / bc,
                                   std::max(0, k - n_error_buf) / bc};
                const int hi[3] = {std::min(ncell[0] - 1, i + n_error_buf) / bc,
                                   std::min(ncell[1] - 1, j + n_error_buf) / bc,
                                   std::min(ncell[2] - 1, k + n_error_buf) / bc};
                for (int c = lo[2]; c <= hi[2]; ++c)
                    for (int b = lo[1]; b <= hi[1]; ++b)
                        for (int a = lo[0]; a <= hi[0]; ++a) blk[size_t((c * nb[1] + b) * nb[0] + a)] = 1;
            }
    std::vector<std::array<int, 3>> pts;
    for (int c = 0; c < nb[2]; ++c)
        for (int b = 0; b < nb[1]; ++b)
            for (int a = 0; a < nb[0]; ++a)
                if (blk[size_t((c * nb[1] + b) * nb[0] + a)]) pts.push_back({a, b, c});
    std::vector<std::array<int, 6>> clusters;
    br_cluster(std::move(pts), grid_eff, clusters);
    // blocks -> fine cells (ratio 2), chopped to max_grid_size along each axis
    std::vector<std::array<int, 6>> boxes;
    const int step = max_grid_size / blocking_factor;  // blocks per piece
    for (const auto &cb : clusters) {
        for (int z = cb[2]; z <= cb[5]; z += step)
            for (int y = cb[1]; y <= cb[4]; y += step)
                for (int x = cb[0]; x <= cb[3]; x += step) {
                    const int l[3] = {x, y, z};
                    const int h[3] = {std::min(cb[3], x + step - 1), std::min(cb[4], y + step - 1),
                                      std::min(cb[5], z + step - 1)};
                    std::array<int, 6> fb;
                    for (int d = 0; d < 3; ++d) {
                        fb[size_t(d)] = l[d] * blocking_factor;
                        fb[size_t(3 + d)] = (h[d] + 1) * blocking_factor - 1;
                    }
                    boxes.push_back(fb);
                }
    }
    std::sort(boxes.begin(), boxes.end(), [](const std::array<int, 6> &a, const std::array<int, 6> &b) {
        return std::make_tuple(a[2], a[1], a[0]) < std::make_tuple(b[2], b[1], b[0]);
    });
    *nbox = int(boxes.size());
    if (!box_lo && !box_hi) return MRB_OK;
    if (!box_lo || !box_hi || max_boxes < int(boxes.size())) return set_error(MRB_E_ARG, "box arrays too short");
    for (size_t b = 0; b < boxes.size(); ++b)
        for (int d = 0; d < 3; ++d) {
            box_lo[3 * b + size_t(d)] = boxes[b][size_t(d)];
            box_hi[3 * b + size_t(d)] = boxes[b][size_t(3 + d)];
        }
    return MRB_OK;
}

int mrb_amr_advance(mrb_level *crse, mrb_level *fine, int nsteps, void *stream)
{
    int rc = check_advance(crse, fine);
    if (rc) return rc;
    for (int n = 0; n < nsteps; ++n)
        if ((rc = amr_step(crse, fine, stream))) return rc;
    return MRB_OK;
}

int mrb_amr_advance_graph(mrb_level *crse, mrb_level *fine, int nsteps, void *stream)
{
    int rc = check_advance(crse, fine);
    if (rc) return rc;
    hipStream_t s = as_stream(stream);
    for (int n = 0; n < nsteps; ++n) {
        mrb_level::AmrGraph &G = fine->amr_graph[crse->cur][crse->fresh_depth >= 1 ? 1 : 0];
        if (!G.exec || G.gen_c != crse->gen || G.gen_f != fine->gen || G.fine_cur != fine->cur) {
            if (G.exec) (void)hipGraphExecDestroy(G.exec);
            G = mrb_level::AmrGraph{};
            if (!fine->cap_stream) MRB_HIP(hipStreamCreateWithFlags(&fine->cap_stream, hipStreamNonBlocking));
            // capture one step: the host-side state advances during the
            // capture as if the step had run; it is put back afterwards and
            // set to the recorded post-step values after each launch
            const int cur_c = crse->cur, cur_f = fine->cur;
            const int fresh_c = crse->fresh_depth, fresh_f = fine->fresh_depth;
            G.gen_c = crse->gen;
            G.gen_f = fine->gen;
            G.fine_cur = cur_f;
            MRB_HIP(hipStreamBeginCapture(fine->cap_stream, hipStreamCaptureModeThreadLocal));
            rc = amr_step(crse, fine, fine->cap_stream);
            hipGraph_t graph = nullptr;
            const hipError_t ec = hipStreamEndCapture(fine->cap_stream, &graph);
            G.post_cur_c = crse->cur;
            G.post_cur_f = fine->cur;
            crse->cur = cur_c;
            fine->cur = cur_f;
            G.post_fresh_c = crse->fresh_depth;
            G.post_fresh_f = fine->fresh_depth;
            crse->fresh_depth = fresh_c;
            fine->fresh_depth = fresh_f;
            if (rc) {
                if (graph) (void)hipGraphDestroy(graph);
                return rc;
            }
            if (ec != hipSuccess) return hip_fail(ec, "hipStreamEndCapture");
            const hipError_t ei = hipGraphInstantiate(&G.exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (ei != hipSuccess) {
                G = mrb_level::AmrGraph{};
                return hip_fail(ei, "hipGraphInstantiate");
            }
        }
        MRB_HIP(hipGraphLaunch(G.exec, s));
        crse->cur = G.post_cur_c;
        fine->cur = G.post_cur_f;
        crse->fresh_depth = G.post_fresh_c;
        fine->fresh_depth = G.post_fresh_f;
    }
    return MRB_OK;
}

}  // extern "C"
