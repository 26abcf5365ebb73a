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

// Conservative linear interpolation with monotonised-central limited
// slopes: child = ((c + s_x o_x) + s_y o_y) + s_z o_z, o = -+1/4.
__global__ void __launch_bounds__(kAmrBlock) k_explode(
    const double *__restrict__ crse, double *__restrict__ fine, const BoxDev *__restrict__ fboxes,
    const BoxDev *__restrict__ cboxes, const ExplodeEntry *__restrict__ map, int64_t n)
{
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const ExplodeEntry e = map[t];
    const BoxDev &FB = fboxes[e.fine_box];
    const BoxDev &CB = cboxes[e.crse_box];
    const double *c0 = crse + CB.base * NQ + e.crse_local;
    double *fo = fine + FB.base * NQ + e.fine_local;
    // a direction without both neighbours reads the cell itself (offset 0):
    // then dl = dr = 0 and the limiter gives the zero slope, with every load
    // in bounds and issued unconditionally
    int64_t stride[3] = {1, CB.n[0], int64_t(CB.n[0]) * CB.n[1]};
    double o[3];
    for (int d = 0; d < 3; ++d) {
        o[d] = ((e.bits >> d) & 1) ? 0.25 : -0.25;
        if (!((e.bits >> (3 + d)) & 1)) stride[d] = 0;
    }
    double v[NQ][7];
    static_for<NQ>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        const double *cq = c0 + int64_t(q) * CB.ncell;
        v[q][0] = cq[0];
        for (int d = 0; d < 3; ++d) {
            v[q][1 + 2 * d] = cq[-stride[d]];
            v[q][2 + 2 * d] = cq[stride[d]];
        }
    });
    static_for<NQ>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        const double c = v[q][0];
        double s[3];
        for (int d = 0; d < 3; ++d) {
            const double cm = v[q][1 + 2 * d], cp = v[q][2 + 2 * d];
            const double dc = 0.5 * (cp - cm);
            const double dl = c - cm;
            const double dr = cp - c;
            s[d] = dl * dr > 0.0 ? copysign(fmin(fabs(dc), fmin(2.0 * fabs(dl), 2.0 * fabs(dr))), dc) : 0.0;
        }
        fo[int64_t(q) * FB.ncell] = ((c + s[0] * o[0]) + s[1] * o[1]) + s[2] * o[2];
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
    if (L->d_coal_cell) (void)hipFree(L->d_coal_cell);
    if (L->d_coal_vals) (void)hipFree(L->d_coal_vals);
    if (L->d_qring) (void)hipFree(L->d_qring);
    L->d_qring = nullptr;
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
                    for (int d = 0; d < 3; ++d) {
                        const bool hm = crse->periodic[d] || P[d] - 1 >= 0;
                        const bool hp = crse->periodic[d] || P[d] + 1 < crse->ncell[d];
                        e.bits |= (s[d] & 1) << d;
                        e.bits |= int(hm && hp) << (3 + d);
                    }
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
        const BoxDev &CB = crse->boxes[e.crse_box];
        const int64_t sz = int64_t(CB.n[0]) * CB.n[1];
        const int64_t lo = (e.bits & (7 << 3)) ? e.crse_local - sz : e.crse_local;
        const int64_t hi = (e.bits & (7 << 3)) ? e.crse_local + sz : e.crse_local;
        if (e.fine_local < 0 || e.fine_local >= FB.ncell || lo < 0 || hi >= CB.ncell)
            return set_error(MRB_E_ARG, "explosion map leaves a fab");
    }
    amr_release(fine);
    if (!coal.empty()) {
        MRB_HIP(hipMalloc(&fine->d_coal, sizeof(CoalesceEntry) * coal.size()));
        MRB_HIP(hipMemcpy(fine->d_coal, coal.data(), sizeof(CoalesceEntry) * coal.size(), hipMemcpyHostToDevice));
    }
    if (!expl.empty()) {
        MRB_HIP(hipMalloc(&fine->d_expl, sizeof(ExplodeEntry) * expl.size()));
        MRB_HIP(hipMemcpy(fine->d_expl, expl.data(), sizeof(ExplodeEntry) * expl.size(), hipMemcpyHostToDevice));
    }
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

int mrb_explode(const mrb_level *crse, mrb_level *fine, int lattice, void *stream)
{
    if (!fine || !crse) return set_error(MRB_E_ARG, "null level");
    if (fine->amr_crse != crse) return set_error(MRB_E_STATE, "levels are not linked (mrb_amr_link)");
    if (lattice < 0 || lattice > 1) return set_error(MRB_E_ARG, "lattice must be 0 or 1");
    hipStream_t s = as_stream(stream);
    // fine-fine ghosts first, then the uncovered ones, then the domain
    // boundary (the three sets of cells are disjoint; physbc reads valid
    // and in-domain ghost cells only)
    int rc = fab_fill_boundary(fine, lattice, s);
    if (rc) return rc;
    if (lattice == 1) fine->fresh_depth = fine->ng;
    if (fine->n_expl > 0) {
        k_explode<<<blocks_for(fine->n_expl), kAmrBlock, 0, s>>>(
            (const double *)crse->pop[lattice][crse->cur], (double *)fine->pop[lattice][fine->cur],
            fine->d_boxes, crse->d_boxes, (const ExplodeEntry *)fine->d_expl, fine->n_expl);
        MRB_HIP(hipGetLastError());
    }
    return mrb_physbc(fine, lattice, stream);
}

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
    if (crse->fresh_depth < 1 && (rc = fab_fill_boundary_depth(crse, 1, s))) return rc;
    for (int l = 0; l < 2; ++l)
        if ((rc = mrb_physbc(crse, l, stream))) return rc;
    for (int l = 0; l < 2; ++l)
        if ((rc = mrb_explode(crse, fine, l, stream))) return rc;
    for (int sub = 0; sub < 2; ++sub) {
        for (int l = 0; l < 2; ++l)
            if ((rc = mrb_physbc(fine, l, stream))) return rc;
        // time_step(1) -> advance(1) (LBM.cpp:523-544): stream f, g and
        // collide as one batched step
        fine->stream_rings_le2 = sub == 1;
        rc = fab_batched_step(fine, nullptr, sub == 0 ? fine->ng : 0, s);
        fine->stream_rings_le2 = false;
        if (rc) return rc;
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
    // the interpolation reads the coarse parents' neighbours
    if (crse->fresh_depth < 1 && (rc = fab_fill_boundary_depth(crse, 1, nullptr))) return rc;
    const std::vector<int32_t> cown = owner_map(crse);
    const std::vector<int32_t> oown = owner_map(old_fine);
    // FillPatchTwoLevels (FillPatchOps.H:118-129) over the new valid cells:
    // the old fine level where it has the cell, else the interpolation from
    // the coarse parent
    std::vector<LevelCopy> copy;
    std::vector<ExplodeEntry> interp;
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
                    for (int d = 0; d < 3; ++d) {
                        const bool hm = crse->periodic[d] || P[d] - 1 >= 0;
                        const bool hp = crse->periodic[d] || P[d] + 1 < crse->ncell[d];
                        e.bits |= (iv[d] & 1) << d;
                        e.bits |= int(hm && hp) << (3 + d);
                    }
                    interp.push_back(e);
                }
    }
    void *d_copy = nullptr, *d_interp = nullptr;
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
        if (!interp.empty())
            k_explode<<<blocks_for(int64_t(interp.size())), kAmrBlock>>>(
                (const double *)crse->pop[l][crse->cur], dst, fine->d_boxes, crse->d_boxes,
                (const ExplodeEntry *)d_interp, int64_t(interp.size()));
        k_zero_solid<<<blocks_for(fine->total_cells), kAmrBlock>>>(dst, fine->d_isf, fine->d_boxes,
                                                                     fine->d_box_base, fine->nbox, fine->total_cells);
    }
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipDeviceSynchronize();
    (void)hipFree(d_copy);
    (void)hipFree(d_interp);
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
