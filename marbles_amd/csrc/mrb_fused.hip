// Single-pass stream + collide on the compact layout: a z-sweep that keeps
// the Q-correction halo on chip.
//
// Mapping.  A workgroup owns an x-y tile and sweeps a chunk of z planes.  Each
// wave is one x-row of 32 cells; the two lanes of a pair hold the same cell,
// the even lane its f populations and the odd lane its g populations (f and g
// are interleaved by component in HBM, so both lanes of a load instruction
// share one base address).  Cells 1..30 of a row and rows 1..TY-2 of the
// tile are updated; the outer ring of the tile only supplies Q_corr for the
// gradient (LBM.cpp:959-991).
//
// Per plane z:
//   1. pull plane z+1, moment sums, exchange with the partner lane, Q(z+1)
//   2. Q_x, Q_y of plane z to LDS; barrier
//   3. interior cells: moments of plane z, gradient (x/y neighbours from LDS,
//      z neighbours from registers), equilibria, relax, write the B buffer.
// HOLD = true keeps the pulled populations of plane z+1 in registers for
// step 3 of the next plane (each population leaves HBM once); HOLD = false
// keeps only Q(z+1) and pulls plane z again in step 3 (one plane of
// registers, the second pull served by L2 for the planes the look-ahead just
// touched).
//
// Addressing: every access is (uniform base of direction q) + (32-bit byte
// offset of the lane), so a load or store needs one VGPR for its address.
//
// Workgroups are dealt round-robin to the 8 XCDs; the tile index is remapped
// so each XCD works on a contiguous range of tiles and neighbouring tiles
// (whose rings overlap) share that XCD's L2.
#include "mrb_internal.h"

namespace mrb {
namespace {

constexpr int kRow = 32;  // cells per wave row (64 lanes = 32 f/g pairs)
constexpr int kInner = kRow - 2;

struct FusedArgs {
    int nx, ny, nz;
    int per[3];
    int64_t N;   // cells
    int64_t cs;  // component stride (2N)
    int bc[6];
    int ntx, nty, ntz, zc;
    int ntiles;
};

template <class T>
__device__ __forceinline__ T load_at(const T *base, uint32_t byte_off)
{
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + byte_off);
}

// Source coordinate of the pull along one dimension for the direction
// component c (index c+1: source x - c) and the boundary kind of that source:
// 0 = inside the domain (or periodic image), else the bc type of the face it
// lies beyond.  An outflow ghost reads the boundary cell next to it
// (BC.H:323-342), so its coordinate is clamped onto the face.
struct DimSrc {
    int s[3];
    int kind[3];
};

__device__ __forceinline__ DimSrc dim_src(int x, int n, int periodic, int bc_lo, int bc_hi)
{
    DimSrc d;
    d.s[1] = x;
    d.kind[1] = 0;
    d.s[0] = x + 1;  // c = -1
    d.kind[0] = 0;
    if (x + 1 >= n) {
        d.kind[0] = periodic ? 0 : bc_hi;
        d.s[0] = periodic ? 0 : (bc_hi == MRB_BC_OUTFLOW ? n - 1 : x);
    }
    d.s[2] = x - 1;  // c = +1
    d.kind[2] = 0;
    if (x - 1 < 0) {
        d.kind[2] = periodic ? 0 : bc_lo;
        d.s[2] = periodic ? n - 1 : (bc_lo == MRB_BC_OUTFLOW ? 0 : x);
    }
    return d;
}

// Pull the 27 populations of the lane's lattice for cell (x,y,z).
// lat_off = 0 (f) or 1 (g): f_q and g_q of a cell are neighbours,
// element (q*N + cell)*2 + lattice.  Cells away from the domain faces take a
// branch-free path with uniform offsets; face cells resolve periodic wraps
// and physical-boundary ghosts with selects.  The face applied to a ghost is
// the last non-periodic face in BCFill's loop order (BC.H:365-372); at a
// corner of two non-periodic faces the other out-of-domain coordinate is
// the cell's own.  bcv = [f inlet values | g inlet values].
__device__ __forceinline__ void pull_lattice(
    const double *__restrict__ f, const double *__restrict__ bcv, int lat, uint32_t lat_off,
    const FusedArgs &A, int x, int y, int z, uint32_t cell, uint32_t mask, double v[NQ])
{
    const uint32_t cs = uint32_t(A.cs);
    const uint32_t dy = uint32_t(A.nx), dz = uint32_t(A.nx) * uint32_t(A.ny);
    const bool face = x == 0 || x == A.nx - 1 || y == 0 || y == A.ny - 1 || z == 0 || z == A.nz - 1;
    if (!face) {
        static_for<NQ>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            constexpr int cx = ev(q, 0), cy = ev(q, 1), cz = ev(q, 2);
            constexpr int qb = kBounce[q];
            constexpr int qlo = q < qb ? q : qb;
            // base: component qlo; direct source at +(q-qlo)*cs, bounce-back
            // (own cell, opposite direction) at +(qb-qlo)*cs
            const uint32_t back = uint32_t(cx) + uint32_t(cy) * dy + uint32_t(cz) * dz;
            const uint32_t e_direct = uint32_t(q - qlo) * cs + 2u * (cell - back);
            const uint32_t e_bounce = uint32_t(qb - qlo) * cs + 2u * cell;
            const uint32_t e = ((mask >> q) & 1u) ? e_direct : e_bounce;
            v[q] = load_at(f + int64_t(qlo) * A.cs, (e + lat_off) * 8u);
        });
        return;
    }
    const DimSrc X = dim_src(x, A.nx, A.per[0], A.bc[0], A.bc[3]);
    const DimSrc Y = dim_src(y, A.ny, A.per[1], A.bc[1], A.bc[4]);
    const DimSrc Z = dim_src(z, A.nz, A.per[2], A.bc[2], A.bc[5]);
    static_for<NQ>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        constexpr int cx = ev(q, 0), cy = ev(q, 1), cz = ev(q, 2);
        constexpr int qb = kBounce[q];
        constexpr int qlo = q < qb ? q : qb;
        const int kz = Z.kind[cz + 1], ky = Y.kind[cy + 1], kx = X.kind[cx + 1];
        const int kind = kz ? kz : (ky ? ky : kx);
        const uint32_t src = uint32_t(Z.s[cz + 1]) * dz + uint32_t(Y.s[cy + 1]) * dy + uint32_t(X.s[cx + 1]);
        const bool bounce = kind == MRB_BC_NOSLIP || (kind == 0 && !((mask >> q) & 1u));
        const uint32_t e_direct = uint32_t(q - qlo) * cs + 2u * src;
        const uint32_t e_bounce = uint32_t(qb - qlo) * cs + 2u * cell;
        const uint32_t e = bounce ? e_bounce : e_direct;
        const double val = load_at(f + int64_t(qlo) * A.cs, (e + lat_off) * 8u);
        const double inlet = lat ? bcv[NQ + q] : bcv[q];
        v[q] = kind == MRB_BC_VELOCITY ? inlet : val;
    });
}

// Generic direction sums of one lattice: S0 = sum v, S1..3 = sum c v,
// S4..6 = sum c_a^2 v, S7..9 = sum c_x c_y v, c_x c_z v, c_y c_z v; each in
// direction order (the order of LBM.cpp:841-862).
template <int NS>
__device__ __forceinline__ void lattice_sums(const double v[NQ], double S[NS])
{
    for (int k = 0; k < NS; ++k) S[k] = 0.0;
    static_for<NQ>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        constexpr int cx = ev(q, 0), cy = ev(q, 1), cz = ev(q, 2);
        const double fq = v[q];
        S[0] += fq;
        if constexpr (cx == 1) S[1] += fq;
        if constexpr (cx == -1) S[1] += -fq;
        if constexpr (cy == 1) S[2] += fq;
        if constexpr (cy == -1) S[2] += -fq;
        if constexpr (cz == 1) S[3] += fq;
        if constexpr (cz == -1) S[3] += -fq;
        if constexpr (NS > 4) {
            if constexpr (cx != 0) S[4] += fq;
            if constexpr (cy != 0) S[5] += fq;
            if constexpr (cz != 0) S[6] += fq;
            if constexpr (cx * cy == 1) S[7] += fq;
            if constexpr (cx * cy == -1) S[7] += -fq;
            if constexpr (cx * cz == 1) S[8] += fq;
            if constexpr (cx * cz == -1) S[8] += -fq;
            if constexpr (cy * cz == 1) S[9] += fq;
            if constexpr (cy * cz == -1) S[9] += -fq;
        }
    });
}

__device__ __forceinline__ double partner(double v) { return __shfl_xor(v, 1); }

// Q_corr of a pulled cell from the pair's sums (both lanes get the result).
__device__ __forceinline__ Macro pair_macro(const double v[NQ], int lat, const CollideParams &P)
{
    double S[4];
    lattice_sums<4>(v, S);
    double T[4];
    for (int k = 0; k < 4; ++k) T[k] = partner(S[k]);
    const double rho = lat ? T[0] : S[0];
    const double jx = lat ? T[1] : S[1];
    const double jy = lat ? T[2] : S[2];
    const double jz = lat ? T[3] : S[3];
    const double E = lat ? S[0] : T[0];
    return macro_from(rho, jx, jy, jz, E, P);
}

template <int TY, bool HOLD>
__global__ void __launch_bounds__(64 * TY) k_fused(
    const double *__restrict__ f, double *__restrict__ fo, const uint32_t *__restrict__ mask,
    const double *__restrict__ bcv, FusedArgs A, CollideParams P)
{
    __shared__ double sQx[TY][kRow];
    __shared__ double sQy[TY][kRow];

    // XCD-aware tile index (grid padded to a multiple of 8)
    const unsigned nb = gridDim.x;
    const unsigned b = blockIdx.x;
    const int t = int((b & 7u) * (nb >> 3) + (b >> 3));
    if (t >= A.ntiles) return;  // uniform per workgroup, before any barrier
    const int tx = t % A.ntx;
    const int ty = (t / A.ntx) % A.nty;
    const int tz = t / (A.ntx * A.nty);

    const int lane = int(threadIdx.x);
    const int row = int(threadIdx.y);
    const int lat = lane & 1;
    const int c = lane >> 1;
    const int x = tx * kInner - 1 + c;
    const int y = ty * (TY - 2) - 1 + row;
    const bool in_dom = x >= 0 && x < A.nx && y >= 0 && y < A.ny;
    const bool interior = in_dom && c >= 1 && c <= kInner && row >= 1 && row <= TY - 2;
    const int z0 = tz * A.zc;
    const int z1 = min(z0 + A.zc, A.nz);
    const uint32_t lat_off = lat ? 1u : 0u;
    const int64_t plane = int64_t(A.nx) * A.ny;
    const int64_t col = int64_t(y) * A.nx + x;

    double Pc[NQ], Pn[NQ];
    uint32_t mc = 0, mn = 0;
    double Qc[3] = {0.0, 0.0, 0.0}, Qn[3] = {0.0, 0.0, 0.0};
    double Qzp = 0.0;

    if (in_dom && z0 > 0) {
        const int64_t cell = int64_t(z0 - 1) * plane + col;
        const uint32_t m = mask[cell];
        if (m & 1u) {
            pull_lattice(f, bcv, lat, lat_off, A, x, y, z0 - 1, uint32_t(cell), m, Pn);
            Qzp = pair_macro(Pn, lat, P).Qz;
        }
    }
    if (in_dom) {
        const int64_t cell = int64_t(z0) * plane + col;
        mc = mask[cell];
        if (mc & 1u) {
            pull_lattice(f, bcv, lat, lat_off, A, x, y, z0, uint32_t(cell), mc, Pc);
            const Macro M = pair_macro(Pc, lat, P);
            Qc[0] = M.Qx;
            Qc[1] = M.Qy;
            Qc[2] = M.Qz;
        }
    }

    for (int z = z0; z < z1; ++z) {
        if (in_dom && z + 1 < A.nz) {
            const int64_t cell = int64_t(z + 1) * plane + col;
            mn = mask[cell];
            if (mn & 1u) {
                pull_lattice(f, bcv, lat, lat_off, A, x, y, z + 1, uint32_t(cell), mn, Pn);
                const Macro M = pair_macro(Pn, lat, P);
                Qn[0] = M.Qx;
                Qn[1] = M.Qy;
                Qn[2] = M.Qz;
            }
        }
        if (lat == 0) {
            sQx[row][c] = Qc[0];
            sQy[row][c] = Qc[1];
        }
        __syncthreads();
        if (interior) {
            const int64_t cell = int64_t(z) * plane + col;
            const uint32_t soff = (2u * uint32_t(cell) + lat_off) * 8u;
            if (mc & 1u) {
                if constexpr (!HOLD) pull_lattice(f, bcv, lat, lat_off, A, x, y, z, uint32_t(cell), mc, Pc);
                double S[10];
                lattice_sums<10>(Pc, S);
                double T[10];
                for (int k = 0; k < 10; ++k) T[k] = partner(S[k]);
                // f-lattice sums live in the even lane, g-lattice sums in the odd one
                Moments mo;
                mo.rho = lat ? T[0] : S[0];
                mo.jx = lat ? T[1] : S[1];
                mo.jy = lat ? T[2] : S[2];
                mo.jz = lat ? T[3] : S[3];
                mo.pxx = lat ? T[4] : S[4];
                mo.pyy = lat ? T[5] : S[5];
                mo.pzz = lat ? T[6] : S[6];
                mo.pxy = lat ? T[7] : S[7];
                mo.pxz = lat ? T[8] : S[8];
                mo.pyz = lat ? T[9] : S[9];
                mo.E = lat ? S[0] : T[0];
                mo.qx = lat ? S[1] : T[1];
                mo.qy = lat ? S[2] : T[2];
                mo.qz = lat ? S[3] : T[3];
                const Macro M = macro_from(mo.rho, mo.jx, mo.jy, mo.jz, mo.E, P);
                const bool xp = x < A.nx - 1 && ((mc >> 2) & 1u), xm = x > 0 && ((mc >> 1) & 1u);
                const bool yp = y < A.ny - 1 && ((mc >> 4) & 1u), ym = y > 0 && ((mc >> 3) & 1u);
                const bool zp = z < A.nz - 1 && ((mc >> 6) & 1u), zm = z > 0 && ((mc >> 5) & 1u);
                const double dqx = one_sided_gradient(xp, xm, xp ? sQx[row][c + 1] : 0.0, M.Qx,
                                                      xm ? sQx[row][c - 1] : 0.0, P.inv_dx[0]);
                const double dqy = one_sided_gradient(yp, ym, yp ? sQy[row + 1][c] : 0.0, M.Qy,
                                                      ym ? sQy[row - 1][c] : 0.0, P.inv_dx[1]);
                const double dqz = one_sided_gradient(zp, zm, zp ? Qn[2] : 0.0, M.Qz, zm ? Qzp : 0.0,
                                                      P.inv_dx[2]);
                const CollideCoeffs C = collide_coeffs(mo, M, dqx, dqy, dqz, P);
                static_for<NQ>([&](auto qc) {
                    constexpr int q = decltype(qc)::value;
                    const double eq = lat ? geq_of<q>(C) : feq_of<q>(C);
                    const double v = Pc[q];
                    double *p = reinterpret_cast<double *>(reinterpret_cast<char *>(fo + int64_t(q) * A.cs) + soff);
                    __builtin_nontemporal_store(v + C.omega * (eq - v), p);
                });
            } else {
                static_for<NQ>([&](auto qc) {
                    constexpr int q = decltype(qc)::value;
                    double *p = reinterpret_cast<double *>(reinterpret_cast<char *>(fo + int64_t(q) * A.cs) + soff);
                    __builtin_nontemporal_store(-1.0, p);
                });
            }
        }
        __syncthreads();
        Qzp = Qc[2];
        Qc[0] = Qn[0];
        Qc[1] = Qn[1];
        Qc[2] = Qn[2];
        mc = mn;
        if constexpr (HOLD) {
            static_for<NQ>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                Pc[q] = Pn[q];
            });
        }
    }
}

FusedArgs make_fused_args(const mrb_level *L, int ty, int zc)
{
    FusedArgs A{};
    A.nx = L->ncell[0];
    A.ny = L->ncell[1];
    A.nz = L->ncell[2];
    for (int d = 0; d < 3; ++d) A.per[d] = L->periodic[d];
    A.N = L->total_cells;
    A.cs = 2 * A.N;
    for (int n = 0; n < 6; ++n) A.bc[n] = L->bc.bc_type[n];
    A.ntx = (A.nx + kInner - 1) / kInner;
    A.nty = (A.ny + (ty - 2) - 1) / (ty - 2);
    A.zc = zc;
    A.ntz = (A.nz + zc - 1) / zc;
    A.ntiles = A.ntx * A.nty * A.ntz;
    return A;
}

template <int TY, bool HOLD>
int launch_fused(mrb_level *L, int zc, hipStream_t s)
{
    const FusedArgs A = make_fused_args(L, TY, zc);
    // byte offsets from a component pair's base are 32-bit: (2N + 2N) * 8 < 2^32
    if (4 * A.N * 8 >= (int64_t(1) << 32))
        return set_error(MRB_E_UNSUPPORTED, "fused kernel: domain too large for 32-bit offsets");
    const unsigned nb = unsigned((A.ntiles + 7) / 8 * 8);
    const double *f = (const double *)L->pop[0][L->cur];
    double *fo = (double *)L->pop[0][1 - L->cur];
    k_fused<TY, HOLD><<<dim3(nb), dim3(64, TY, 1), 0, s>>>(f, fo, L->d_mask, L->d_small + 4 * NQ, A, L->cp);
    MRB_HIP(hipGetLastError());
    return MRB_OK;
}

}  // namespace

// Variants of the fused kernel: register mode, rows per tile, z planes per chunk.
const FusedVariant kFusedVariants[] = {
    {"fused_hold_ty4_z32", 4, 32, 1},   {"fused_hold_ty6_z32", 6, 32, 1},
    {"fused_hold_ty8_z32", 8, 32, 1},   {"fused_repull_ty4_z32", 4, 32, 0},
    {"fused_repull_ty6_z32", 6, 32, 0}, {"fused_repull_ty8_z32", 8, 32, 0},
    {"fused_repull_ty8_z64", 8, 64, 0}, {"fused_repull_ty16_z32", 16, 32, 0},
};
const int kNumFusedVariants = int(sizeof(kFusedVariants) / sizeof(kFusedVariants[0]));

int fused_launch(mrb_level *L, int variant, hipStream_t s)
{
    if (L->storage != MRB_STORAGE_F64)
        return set_error(MRB_E_UNSUPPORTED, "fused kernel: FP64 storage only");
    for (int n = 0; n < 6; ++n)
        if (L->bc.bc_type[n] == MRB_BC_SLIPX || L->bc.bc_type[n] == MRB_BC_SLIPY ||
            L->bc.bc_type[n] == MRB_BC_SLIPZ)
            return set_error(MRB_E_UNSUPPORTED, "fused kernel: slip walls use the two-pass step");
    if (L->bcdev.prof.kind)
        return set_error(MRB_E_UNSUPPORTED, "fused kernel: velocity profiles use the two-pass step");
    if (L->bcdev.pdir >= 0)
        return set_error(MRB_E_UNSUPPORTED, "fused kernel: pressure boundaries use the two-pass step");
    if (variant < 0 || variant >= kNumFusedVariants) return set_error(MRB_E_ARG, "bad fused variant");
    const FusedVariant &v = kFusedVariants[variant];
    if (v.hold) {
        switch (v.ty) {
        case 4: return launch_fused<4, true>(L, v.zc, s);
        case 6: return launch_fused<6, true>(L, v.zc, s);
        case 8: return launch_fused<8, true>(L, v.zc, s);
        default: break;
        }
    } else {
        switch (v.ty) {
        case 4: return launch_fused<4, false>(L, v.zc, s);
        case 6: return launch_fused<6, false>(L, v.zc, s);
        case 8: return launch_fused<8, false>(L, v.zc, s);
        case 16: return launch_fused<16, false>(L, v.zc, s);
        default: break;
        }
    }
    return set_error(MRB_E_ARG, "bad fused tile");
}

int fused_step(mrb_level *L, int variant, hipStream_t s)
{
    const int rc = fused_launch(L, variant, s);
    if (rc) return rc;
    L->cur = 1 - L->cur;
    L->prev_valid = true;
    return MRB_OK;
}

}  // namespace mrb
