// Hot path on the compact layout (one box == the domain, no ghost cells):
// pull-stream with halfway bounce-back, periodic wrap and physical-boundary
// ghost values resolved inside the kernel, followed by the thermal collide.
//
// HBM layout: populations SoA, cell = (z*ny + y)*nx + x.  FP64: f and g
// interleaved by component, pop[(2q + lattice)*N + cell], so f_q =
// f[q*cs + cell], g_q = g[q*cs + cell] with g = f + N and cs = 2N.  Shifted
// FP32: pop[(q*N + cell)*2 + lattice], f_q = f[q*cs + 2*cell] with g = f + 1.
// Two buffers (A/B); a 32-bit stream mask per cell (bit q = the pull
// source x - c_q is fluid).
//
// The collide needs the Q-correction gradient of post-stream moments at the
// six face neighbours (LBM.cpp:959-991), so one step is two passes:
//   k_qpass:   pull 54 populations, moments -> Q_corr (3 doubles per cell)
//   k_collide: pull 54 populations, moments, dQ from the Q array of the
//              neighbours, equilibria, relax, write 54 populations to B.
#include <type_traits>

#include "mrb_internal.h"

namespace mrb {
namespace {

struct CompactArgs {
    int nx, ny, nz;
    int per[3];
    int64_t N;   // cells
    int64_t cs;  // component stride of the population arrays (2N)
    int bc[6];
    uint32_t bcpack;  // bc[face] in bits 4*face .. 4*face+3
};

// Source coordinates of the pull for one direction component c in
// {-1,0,1} along a dimension of extent n.  Returns false when the source
// lies beyond a non-periodic face (a physical-boundary ghost).
__device__ __forceinline__ void wrap3(int x, int n, int periodic, int s[3], bool ghost[3])
{
    // index c+1: source x - c
    s[0] = x + 1;  // c = -1
    s[1] = x;      // c = 0
    s[2] = x - 1;  // c = +1
    ghost[0] = ghost[1] = ghost[2] = false;
    if (s[0] >= n) {
        if (periodic)
            s[0] -= n;
        else
            ghost[0] = true;
    }
    if (s[2] < 0) {
        if (periodic)
            s[2] += n;
        else
            ghost[2] = true;
    }
}

// Rest-state populations subtracted by the shifted storages (stored value =
// population - shift); one value per speed class |c_q|^2 = 0..3, so the
// eight values sit in scalar registers; unused for FP64 storage.
struct ShiftTab {
    double fk[4], gk[4];
    __device__ __forceinline__ double f(int q) const { return fk[csq(q)]; }
    __device__ __forceinline__ double g(int q) const { return gk[csq(q)]; }
};

// Element stride between neighbouring cells of one component: FP64
// storage keeps the components of a lattice contiguous
// (pop[(2q + lattice)*N + cell]); shifted FP32 storage keeps f_q and g_q
// of a cell side by side (pop[(q*N + cell)*2 + lattice]) so that one
// 8-byte load or store moves both.
template <class T>
constexpr int kCellStride = std::is_same<T, float>::value ? 2 : 1;

// store of a cell's f_q and g_q
template <class T>
__device__ __forceinline__ void store_pair(T *__restrict__ fo, T *__restrict__ go, int64_t e, T fv, T gv)
{
    if constexpr (kCellStride<T> == 2) {
        *reinterpret_cast<float2 *>(fo + e) = make_float2(fv, gv);
    } else {
        fo[e] = fv;
        go[e] = gv;
    }
}

// storage <-> FP64 working value
template <class T>
__device__ __forceinline__ double widen(T v, double shift)
{
    if constexpr (std::is_same<T, double>::value) {
        (void)shift;
        return v;
    } else {
        return double(v) + shift;
    }
}
template <class T>
__device__ __forceinline__ T narrow(double v, double shift)
{
    if constexpr (std::is_same<T, double>::value) {
        (void)shift;
        return v;
    } else {
        return T(v - shift);
    }
}

// Per dimension, for the three lattice-velocity components c = -1, 0, +1
// (index c + 1): t = the coordinate the pull reads, x - c wrapped on a
// periodic axis and, beyond a non-periodic face, the boundary cell next to
// the ghost (the cell an outflow ghost copies); g = the source is a ghost.
__device__ __forceinline__ void targets3(int x, int n, int periodic, int t[3], bool g[3])
{
    const int p = x + 1, m = x - 1;
    const bool hi = p >= n, lo = m < 0;
    g[0] = hi && !periodic;
    g[1] = false;
    g[2] = lo && !periodic;
    t[0] = hi ? (periodic ? p - n : n - 1) : p;
    t[1] = x;
    t[2] = lo ? (periodic ? m + n : 0) : m;
}

// Pull of one cell, without branches so that all 54 loads are in flight
// together.  For a pull source beyond a physical face the reference's
// BCFill value is resolved: the face applied is the last non-periodic face
// in BCFill's loop order (BC.H:365-372, z over y over x); NOSLIP and SLIP
// read the cell's own pre-stream populations, OUTFLOW the boundary cell
// next to the ghost (at a corner of two non-periodic faces, clamped in
// both), VELOCITY a constant.
template <class T>
__device__ __forceinline__ void pull_cell(
    const T *__restrict__ f, const T *__restrict__ g, const CompactArgs &A, const BCDev &bc,
    const ShiftTab &sh, int x, int y, int z, int64_t cell, uint32_t mask, double fv[NQ], double gv[NQ])
{
    int tx[3], ty[3], tz[3];
    bool gx[3], gy[3], gz[3];
    targets3(x, A.nx, A.per[0], tx, gx);
    targets3(y, A.ny, A.per[1], ty, gy);
    targets3(z, A.nz, A.per[2], tz, gz);
    const uint32_t c32 = uint32_t(cell);
    T rf[NQ], rg[NQ];
    uint32_t cmask = 0;
    static_for<NQ>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        constexpr int cx = ev(q, 0), cy = ev(q, 1), cz = ev(q, 2);
        constexpr int qb = kBounce[q];
        const bool hx = gx[cx + 1], hy = gy[cy + 1], hz = gz[cz + 1];
        const bool ghost = hx || hy || hz;
        const uint32_t src = (uint32_t(tz[cz + 1]) * uint32_t(A.ny) + uint32_t(ty[cy + 1])) * uint32_t(A.nx) +
                             uint32_t(tx[cx + 1]);
        int face = cx > 0 ? 0 : 3;
        face = hy ? (cy > 0 ? 1 : 4) : face;
        face = hz ? (cz > 0 ? 2 : 5) : face;
        const uint32_t t = (A.bcpack >> (4 * face)) & 15u;
        const bool direct = ghost ? t == MRB_BC_OUTFLOW : ((mask >> q) & 1u) != 0;
        int comp = direct ? q : qb;
        comp = (ghost && t == MRB_BC_SLIPX) ? reflect(q, 0) : comp;
        comp = (ghost && t == MRB_BC_SLIPY) ? reflect(q, 1) : comp;
        comp = (ghost && t == MRB_BC_SLIPZ) ? reflect(q, 2) : comp;
        const int64_t e = int64_t(comp) * A.cs + int64_t(direct ? src : c32) * kCellStride<T>;
        if constexpr (kCellStride<T> == 2) {
            // f_q and g_q side by side: one 8-byte load (g == f + 1)
            const float2 v = *reinterpret_cast<const float2 *>(f + e);
            rf[q] = v.x;
            rg[q] = v.y;
        } else {
            rf[q] = f[e];
            rg[q] = g[e];
        }
        cmask |= uint32_t(ghost && t == MRB_BC_VELOCITY) << q;
    });
    static_for<NQ>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        const bool use_const = (cmask >> q) & 1u;
        fv[q] = use_const ? bc.fval[q] : widen(rf[q], sh.f(q));
        gv[q] = use_const ? bc.gval[q] : widen(rg[q], sh.g(q));
    });
}

__device__ __forceinline__ bool cell_coords(const CompactArgs &A, int &x, int &y, int &z, int64_t &cell)
{
    x = int(blockIdx.x * blockDim.x + threadIdx.x);
    y = int(blockIdx.y * blockDim.y + threadIdx.y);
    z = int(blockIdx.z);
    if (x >= A.nx || y >= A.ny) return false;
    cell = (int64_t(z) * A.ny + y) * A.nx + x;
    return true;
}

// pass 1: Q_corr of the post-stream state
template <class T>
__global__ void __launch_bounds__(256) k_qpass(
    const T *__restrict__ f, const T *__restrict__ g, const uint32_t *__restrict__ mask,
    double *__restrict__ Q, CompactArgs A, BCDev bc, ShiftTab sh, CollideParams P)
{
    int x, y, z;
    int64_t cell;
    if (!cell_coords(A, x, y, z, cell)) return;
    const uint32_t m = mask[cell];
    if (!(m & 1u)) return;
    double fv[NQ], gv[NQ];
    pull_cell(f, g, A, bc, sh, x, y, z, cell, m, fv, gv);
    double rho, jx, jy, jz, E;
    accumulate_q_moments(fv, gv, rho, jx, jy, jz, E);
    const Macro M = macro_from(rho, jx, jy, jz, E, P);
    Q[cell] = M.Qx;
    Q[A.N + cell] = M.Qy;
    Q[2 * A.N + cell] = M.Qz;
}

// pass 2: stream + collide, writes the B buffers
template <class T>
__device__ __forceinline__ void collide_body(
    const T *__restrict__ f, const T *__restrict__ g, T *__restrict__ fo, T *__restrict__ go,
    const uint32_t *__restrict__ mask, const double *__restrict__ Q, const CompactArgs &A, const BCDev &bc,
    const ShiftTab &sh, const CollideParams &P)
{
    int x, y, z;
    int64_t cell;
    if (!cell_coords(A, x, y, z, cell)) return;
    const uint32_t m = mask[cell];
    if (!(m & 1u)) {
        // solid cell: the stream leaves the -1 sentinel (LBM.cpp:565)
        static_for<NQ>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            store_pair(fo, go, int64_t(q) * A.cs + cell * kCellStride<T>, narrow<T>(-1.0, sh.f(q)),
                       narrow<T>(-1.0, sh.g(q)));
        });
        return;
    }
    double fv[NQ], gv[NQ];
    pull_cell(f, g, A, bc, sh, x, y, z, cell, m, fv, gv);
    Moments mo;
    accumulate_moments(fv, gv, mo);
    const Macro M = macro_from(mo.rho, mo.jx, mo.jy, mo.jz, mo.E, P);
    // gradient(): neighbour inside the domain box and fluid
    const bool xp = x < A.nx - 1 && ((m >> 2) & 1u), xm = x > 0 && ((m >> 1) & 1u);
    const bool yp = y < A.ny - 1 && ((m >> 4) & 1u), ym = y > 0 && ((m >> 3) & 1u);
    const bool zp = z < A.nz - 1 && ((m >> 6) & 1u), zm = z > 0 && ((m >> 5) & 1u);
    const int64_t sy = A.nx, sz = int64_t(A.nx) * A.ny;
    const double dqx = one_sided_gradient(xp, xm, xp ? Q[cell + 1] : 0.0, M.Qx, xm ? Q[cell - 1] : 0.0,
                                          P.inv_dx[0]);
    const double dqy = one_sided_gradient(yp, ym, yp ? Q[A.N + cell + sy] : 0.0, M.Qy,
                                          ym ? Q[A.N + cell - sy] : 0.0, P.inv_dx[1]);
    const double dqz = one_sided_gradient(zp, zm, zp ? Q[2 * A.N + cell + sz] : 0.0, M.Qz,
                                          zm ? Q[2 * A.N + cell - sz] : 0.0, P.inv_dx[2]);
    collide_cell(fv, gv, mo, M, dqx, dqy, dqz, P, [&](auto qc, double fnew, double gnew) {
        constexpr int q = decltype(qc)::value;
        store_pair(fo, go, int64_t(q) * A.cs + cell * kCellStride<T>, narrow<T>(fnew, sh.f(q)),
                   narrow<T>(gnew, sh.g(q)));
    });
}

// register budget as the compiler chooses (2 waves per SIMD)
template <class T>
__global__ void __launch_bounds__(256) k_collide(
    const T *__restrict__ f, const T *__restrict__ g, T *__restrict__ fo, T *__restrict__ go,
    const uint32_t *__restrict__ mask, const double *__restrict__ Q, CompactArgs A, BCDev bc, ShiftTab sh,
    CollideParams P)
{
    collide_body(f, g, fo, go, mask, Q, A, bc, sh, P);
}

// at most 168 registers (3 waves per SIMD), some values spilled to scratch
template <class T>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 3))) k_collide_w3(
    const T *__restrict__ f, const T *__restrict__ g, T *__restrict__ fo, T *__restrict__ go,
    const uint32_t *__restrict__ mask, const double *__restrict__ Q, CompactArgs A, BCDev bc, ShiftTab sh,
    CollideParams P)
{
    collide_body(f, g, fo, go, mask, Q, A, bc, sh, P);
}

// stream only (parity of the pull with the reference push)
template <class T>
__global__ void __launch_bounds__(256) k_stream_only(
    const T *__restrict__ f, const T *__restrict__ g, T *__restrict__ fo, T *__restrict__ go,
    const uint32_t *__restrict__ mask, CompactArgs A, BCDev bc, ShiftTab sh)
{
    int x, y, z;
    int64_t cell;
    if (!cell_coords(A, x, y, z, cell)) return;
    const uint32_t m = mask[cell];
    double fv[NQ], gv[NQ];
    if (m & 1u) {
        pull_cell(f, g, A, bc, sh, x, y, z, cell, m, fv, gv);
    } else {
        for (int q = 0; q < NQ; ++q) fv[q] = gv[q] = -1.0;
    }
    static_for<NQ>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        store_pair(fo, go, int64_t(q) * A.cs + cell * kCellStride<T>, narrow<T>(fv[q], sh.f(q)),
                   narrow<T>(gv[q], sh.g(q)));
    });
}

// stream mask: bit q = fluid(x - c_q); a physical-boundary ghost is fluid
__global__ void k_build_mask(const int32_t *__restrict__ isf, uint32_t *__restrict__ mask, CompactArgs A)
{
    int x, y, z;
    int64_t cell;
    if (!cell_coords(A, x, y, z, cell)) return;
    int sx[3], sy[3], sz[3];
    bool gx[3], gy[3], gz[3];
    wrap3(x, A.nx, A.per[0], sx, gx);
    wrap3(y, A.ny, A.per[1], sy, gy);
    wrap3(z, A.nz, A.per[2], sz, gz);
    uint32_t m = 0;
    static_for<NQ>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        constexpr int cx = ev(q, 0), cy = ev(q, 1), cz = ev(q, 2);
        bool fl = true;
        if (!(gx[cx + 1] || gy[cy + 1] || gz[cz + 1])) {
            const int64_t src = (int64_t(sz[cz + 1]) * A.ny + sy[cy + 1]) * A.nx + sx[cx + 1];
            fl = isf[src] != 0;
        }
        if (fl) m |= 1u << q;
    });
    // bit 0 is the cell's own is_fluid (c_0 = 0)
    mask[cell] = m;
}

CompactArgs make_args(const mrb_level *L)
{
    CompactArgs A{};
    A.nx = L->ncell[0];
    A.ny = L->ncell[1];
    A.nz = L->ncell[2];
    for (int d = 0; d < 3; ++d) A.per[d] = L->periodic[d];
    A.N = L->total_cells;
    A.cs = 2 * A.N;
    for (int n = 0; n < 6; ++n) A.bc[n] = L->bc.bc_type[n];
    A.bcpack = 0;
    for (int n = 0; n < 6; ++n) A.bcpack |= uint32_t(A.bc[n] & 15) << (4 * n);
    return A;
}

dim3 grid_of(const CompactArgs &A) { return dim3((A.nx + 63) / 64, (A.ny + 3) / 4, A.nz); }

}  // namespace

int compact_build_mask(mrb_level *L, hipStream_t s)
{
    const CompactArgs A = make_args(L);
    k_build_mask<<<grid_of(A), dim3(64, 4, 1), 0, s>>>(L->d_isf, L->d_mask, A);
    MRB_HIP(hipGetLastError());
    return MRB_OK;
}

// Two-pass step; reads the current buffers, writes the spare ones, does not
// flip.  ev (if not null) receives events before, between and after.
static ShiftTab shift_tab(const mrb_level *L)
{
    ShiftTab t{};
    if (L->storage != MRB_STORAGE_F64) {
        const int rep[4] = {0, 1, 7, 19};
        for (int k = 0; k < 4; ++k) {
            t.fk[k] = L->fshift[rep[k]];
            t.gk[k] = L->gshift[rep[k]];
        }
    }
    return t;
}

template <class T>
static int two_pass_t(mrb_level *L, int variant, hipStream_t s, hipEvent_t *ev)
{
    const CompactArgs A = make_args(L);
    const ShiftTab sh = shift_tab(L);
    const T *f = (const T *)L->pop[0][L->cur];
    const T *g = (const T *)L->pop[1][L->cur];
    T *fo = (T *)L->pop[0][1 - L->cur];
    T *go = (T *)L->pop[1][1 - L->cur];
    if (ev) MRB_HIP(hipEventRecord(ev[0], s));
    k_qpass<T><<<grid_of(A), dim3(64, 4, 1), 0, s>>>(f, g, L->d_mask, L->d_q, A, L->bcdev, sh, L->cp);
    MRB_HIP(hipGetLastError());
    if (ev) MRB_HIP(hipEventRecord(ev[1], s));
    if (variant == 1)
        k_collide_w3<T><<<grid_of(A), dim3(64, 4, 1), 0, s>>>(f, g, fo, go, L->d_mask, L->d_q, A, L->bcdev, sh,
                                                              L->cp);
    else
        k_collide<T><<<grid_of(A), dim3(64, 4, 1), 0, s>>>(f, g, fo, go, L->d_mask, L->d_q, A, L->bcdev, sh, L->cp);
    MRB_HIP(hipGetLastError());
    if (ev) MRB_HIP(hipEventRecord(ev[2], s));
    return MRB_OK;
}

static int two_pass(mrb_level *L, int variant, hipStream_t s, hipEvent_t *ev)
{
    if (L->storage == MRB_STORAGE_F64) return two_pass_t<double>(L, variant, s, ev);
    if (L->storage == MRB_STORAGE_F32_SHIFTED) return two_pass_t<float>(L, variant, s, ev);
    return set_error(MRB_E_UNSUPPORTED, "storage not supported by this kernel");
}

const char *const kTwoPassNames[kNumTwoPassVariants] = {"k_qpass+k_collide", "k_qpass+k_collide_w3"};

int compact_two_pass_launch(mrb_level *L, int variant, hipStream_t s) { return two_pass(L, variant, s, nullptr); }

int compact_variant_launch(mrb_level *L, int index, hipStream_t s)
{
    return index < kNumTwoPassVariants ? compact_two_pass_launch(L, index, s)
                                       : fused_launch(L, index - kNumTwoPassVariants, s);
}

const char *compact_variant_name(int index)
{
    if (index < 0 || index >= kNumTwoPassVariants + kNumFusedVariants) return "";
    return index < kNumTwoPassVariants ? kTwoPassNames[index] : kFusedVariants[index - kNumTwoPassVariants].name;
}

static int step_impl(mrb_level *L, hipStream_t s, bool events, int &nslots)
{
    int rc;
    if (L->hot < kNumTwoPassVariants) {
        rc = two_pass(L, L->hot, s, events ? L->ev : nullptr);
        nslots = 2;
    } else {
        if (events) MRB_HIP(hipEventRecord(L->ev[0], s));
        rc = fused_launch(L, L->hot - kNumTwoPassVariants, s);
        if (events && !rc) MRB_HIP(hipEventRecord(L->ev[1], s));
        nslots = 1;
    }
    if (rc) return rc;
    L->cur = 1 - L->cur;
    return MRB_OK;
}

int compact_stream_collide(mrb_level *L, hipStream_t s)
{
    int n;
    return step_impl(L, s, false, n);
}

int compact_stream_collide_events(mrb_level *L, hipStream_t s, int &nslots)
{
    return step_impl(L, s, true, nslots);
}

int compact_stream_only(mrb_level *L, hipStream_t s)
{
    const CompactArgs A = make_args(L);
    const ShiftTab sh = shift_tab(L);
    if (L->storage == MRB_STORAGE_F64) {
        k_stream_only<double><<<grid_of(A), dim3(64, 4, 1), 0, s>>>(
            (const double *)L->pop[0][L->cur], (const double *)L->pop[1][L->cur],
            (double *)L->pop[0][1 - L->cur], (double *)L->pop[1][1 - L->cur], L->d_mask, A, L->bcdev, sh);
    } else if (L->storage == MRB_STORAGE_F32_SHIFTED) {
        k_stream_only<float><<<grid_of(A), dim3(64, 4, 1), 0, s>>>(
            (const float *)L->pop[0][L->cur], (const float *)L->pop[1][L->cur], (float *)L->pop[0][1 - L->cur],
            (float *)L->pop[1][1 - L->cur], L->d_mask, A, L->bcdev, sh);
    } else {
        return set_error(MRB_E_UNSUPPORTED, "storage not supported by this kernel");
    }
    MRB_HIP(hipGetLastError());
    L->cur = 1 - L->cur;
    return MRB_OK;
}

const char *compact_kernel_name(const mrb_level *L)
{
    return compact_variant_name(L->hot);
}

}  // namespace mrb
