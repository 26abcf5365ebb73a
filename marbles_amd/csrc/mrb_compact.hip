// Hot path on the compact layout (one box == the domain, no ghost cells):
// pull-stream with halfway bounce-back, periodic wrap and physical-boundary
// ghost values resolved inside the kernel, followed by the thermal collide.
//
// HBM layout: populations SoA by direction, cell = (z*ny + y)*nx + x, with
// f_q and g_q of a cell side by side: pop[(q*N + cell)*2 + lattice], so
// f_q = f[q*cs + 2*cell] with g = f + 1 and cs = 2N; one 16-byte (FP64),
// 8-byte (shifted FP32) or 4-byte (scaled FP16) access moves both.
// Two buffers (A/B); a 32-bit stream mask per cell (bit q = the pull
// source x - c_q is fluid).
//
// The collide needs the Q-correction gradient of post-stream moments at the
// six face neighbours (LBM.cpp:959-991), so one step is two passes:
//   k_qpass:   pull 54 populations, moments -> Q_corr (3 doubles per cell)
//   k_collide: pull 54 populations, moments, dQ from the Q array of the
//              neighbours, equilibria, relax, write 54 populations to B.
#include <type_traits>

#include <algorithm>
#include <cmath>

#include <map>
#include <mutex>

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
    int fast_ok;      // byte offsets of a lattice's cells fit 32 bits
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

// Rest state of the shifted FP32 storage, for its FP32-arithmetic collide
// (k_collide_f32): the product-form factors and the Grad coefficients of
// the subtracted equilibrium f_eq(rho0, 0, T0), g_eq(rho0, 0, T0), and the
// raw moments of those shifts; FP64.
struct DevRest {
    double A0[3];   // rho0 phi0(c), c = -1, 0, 1 (the shift's x factor)
    double B0[3];   // phi0(c) (its y and z factors)
    double E0;      // 2 rho0 E of the rest state
    double t0[2];   // (c^2 - theta0) a2_aa of the rest state, c^2 = 0, 1
    double mo[14];  // moments of the shifts, in MomentsR order
};

// Element stride between neighbouring cells of one component: every
// storage keeps f_q and g_q of a cell side by side
// (pop[(q*N + cell)*2 + lattice]) so that one load or store moves both.
template <class T>
constexpr int kCellStride = 2;
template <class T>
constexpr bool kIsHalf = std::is_same<T, _Float16>::value;

// store of a cell's f_q and g_q
template <class T>
__device__ __forceinline__ void store_pair(T *__restrict__ fo, T *__restrict__ go, int64_t e, T fv, T gv)
{
    if constexpr (kIsHalf<T>) {
        const uint32_t lo = __builtin_bit_cast(uint16_t, fv), hi = __builtin_bit_cast(uint16_t, gv);
        *reinterpret_cast<uint32_t *>(fo + e) = lo | (hi << 16);
    } else if constexpr (std::is_same<T, double>::value) {
        // non-temporal: the next step reads the buffer back from HBM
        // (measured at 256^3: z-sweep 3.039 -> 3.024 ms FP64, 2.34 -> 2.30 ms
        // shifted FP32)
        (void)go;
        typedef double d2 __attribute__((ext_vector_type(2)));
        d2 v;
        v.x = fv;
        v.y = gv;
        __builtin_nontemporal_store(v, reinterpret_cast<d2 *>(fo + e));
    } else {
        (void)go;
        typedef float f2 __attribute__((ext_vector_type(2)));
        f2 v;
        v.x = fv;
        v.y = gv;
        __builtin_nontemporal_store(v, reinterpret_cast<f2 *>(fo + e));
    }
}

// FP16 storage: scale exponents and box maxima of the read (cur) and the
// written (next) buffers, per (scale box, lattice); unused by the other
// storages
struct Half16 {
    const int *__restrict__ e_cur;
    int *__restrict__ e_next;
    const uint32_t *__restrict__ amax_cur;
    uint32_t *__restrict__ amax_next;
    uint32_t *__restrict__ ovf;
    // the largest |population - shift| the boundary faces bring in (inlet
    // and pressure-face equilibria), per lattice: a lower bound for the box
    // maxima that set the next scales, so that a box at rest next to an
    // inlet keeps room for the inflow
    float floor_f, floor_g;
};

// Largest |population - shift| that scale box b's written cells can be
// pulled from: the maxima of boxes b-1, b and b+1 in the read buffer (a pull
// reaches one plane into each neighbouring box; the neighbours wrap on a
// periodic z axis), no smaller than the boundary floor.  A box exactly at
// the shift state (maximum 0) next to a disturbed one thereby takes the
// disturbed box's scale instead of the largest exponent.
__device__ __forceinline__ float f16_pull_max(const uint32_t *__restrict__ amax, int b, int lattice, int nz,
                                              int periodic_z, float floor)
{
    const int nb = (nz + kScaleSlab - 1) / kScaleSlab;
    float m = fmaxf(__uint_as_float(amax[2 * b + lattice]), floor);
    int bm = b - 1, bp = b + 1;
    if (periodic_z) {
        bm = bm < 0 ? nb - 1 : bm;
        bp = bp >= nb ? 0 : bp;
    }
    if (bm >= 0) m = fmaxf(m, __uint_as_float(amax[2 * bm + lattice]));
    if (bp < nb) m = fmaxf(m, __uint_as_float(amax[2 * bp + lattice]));
    return m;
}

// FP16 value of (v - shift) * 2^e: the difference is taken in FP64, then
// rounded to FP32, scaled there (exact for a power of two) and clamped to
// the FP16 range; amax tracks |v - shift|
__device__ __forceinline__ _Float16 to_half(double v, double shift, float scale, float &amax)
{
    const float d = float(v - shift);
    amax = fmaxf(amax, fabsf(d));
    return _Float16(__builtin_amdgcn_fmed3f(d * scale, -65504.0f, 65504.0f));
}

// the same with FP32 arithmetic: v and the shift are FP32 values
__device__ __forceinline__ _Float16 to_half(float v, float shift, float scale, float &amax)
{
    const float d = v - shift;
    amax = fmaxf(amax, fabsf(d));
    return _Float16(__builtin_amdgcn_fmed3f(d * scale, -65504.0f, 65504.0f));
}

// storage <-> FP64 working value (FP64 and shifted FP32 storage)
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
// the directions q with c_q[d] == c, as a bit set
constexpr uint32_t qset(int d, int c)
{
    uint32_t m = 0;
    for (int q = 0; q < NQ; ++q)
        if (ev(q, d) == c) m |= 1u << q;
    return m;
}
constexpr uint32_t kAllQ = (1u << NQ) - 1u;

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

// VELOCITY ghost value of direction q for the cell (x, y, z) from the
// face tables of a Channel or Parabolic profile: the face is the first of
// z, y, x (BCFill's loop order) in which the source x - c_q is a ghost, the
// ghost cell the unwrapped x - c_q
template <int q>
__device__ __forceinline__ void profile_ghost(const BCDev &bc, const bool gx[3], const bool gy[3], const bool gz[3],
                                              int x, int y, int z, double &fv, double &gv)
{
    constexpr int cx = ev(q, 0), cy = ev(q, 1), cz = ev(q, 2);
    const int iv[3] = {x - cx, y - cy, z - cz};
    const bool hz = gz[cz + 1], hy = gy[cy + 1];
    const int dd = hz ? 2 : (hy ? 1 : 0);
    const int c = dd == 2 ? cz : (dd == 1 ? cy : cx);
    const int face = c > 0 ? dd : dd + 3;
    const int d1 = dd == 0 ? 1 : 0, d2 = dd == 2 ? 1 : 2;
    const double *t = bc.tab[face] + (int64_t(iv[d1] + 1) * bc.tabw[face] + (iv[d2] + 1)) * (2 * NQ);
    fv = t[q];
    gv = t[NQ + q];
}

// Compile-time boundary features of a level's pull (template argument BCX):
// a Channel or Parabolic inlet profile, PRESSURE faces
constexpr int kBcxProfile = 1;
constexpr int kBcxPressure = 2;

// pressure_bc_f (BC.H:222-271) for the fluid cell (x, y, z) next to a
// PRESSURE face of axis pd: the ghosts of the directions in Pm take
// f_eq(rho_bc, u, R T) with u along pd from the populations the cell
// receives.  The reference reads them as data(x + c_qq, qqbar): for the
// tangential and outgoing directions those are the pulled values (a
// periodic image, or a NOSLIP ghost whose value is the cell's own
// population), except from a solid cell, which holds its stored value (0
// in the initial state, the -1 of the stream afterwards; FP16 storage keeps
// 0 there, so -1 is used)
template <class T>
__device__ __forceinline__ void pressure_ghosts(
    const T *__restrict__ f, const CompactArgs &A, const BCDev &bc, const ShiftTab &sh, const int tx[3],
    const int ty[3], const int tz[3], int x, int y, int z, uint32_t mask, uint32_t Pm, double fv[NQ],
    double gv[NQ])
{
    const int pd = bc.pdir;
    const int xc = pd == 0 ? x : (pd == 1 ? y : z);
    const int ndir = xc == 0 ? 1 : -1;
    double rho_out = 0.0, rho_tan = 0.0;
    static_for<NQ>([&](auto qc) {
        constexpr int qq = decltype(qc)::value;
        constexpr int p = kBounce[qq];
        constexpr int cx = ev(p, 0), cy = ev(p, 1), cz = ev(p, 2);
        const int cin = pd == 0 ? cx : (pd == 1 ? cy : cz);
        if (cin == -ndir || cin == 0) {
            double v = fv[p];
            if (!((mask >> p) & 1u)) {
                if constexpr (kIsHalf<T>) {
                    v = -1.0;
                } else {
                    const int64_t src =
                        (int64_t(tz[cz + 1]) * A.ny + ty[cy + 1]) * A.nx + tx[cx + 1];
                    v = widen(f[int64_t(p) * A.cs + src * kCellStride<T>], sh.f(p));
                }
            }
            if (cin == -ndir)
                rho_out += 2.0 * v;
            else
                rho_tan += 1.0 * v;
        }
    });
    const double un = ndir * (1.0 - (rho_out + rho_tan) / bc.prho);
    const double vel[3] = {pd == 0 ? un : 0.0, pd == 1 ? un : 0.0, pd == 2 ? un : 0.0};
    static_for<NQ>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        if ((Pm >> q) & 1u) {
            fv[q] = set_equilibrium_value(bc.prho, vel, bc.prt, q);
            gv[q] = bc.pgval[q];
        }
    });
}

// Pull of one cell, without branches so that all 54 loads are in flight
// together.  For a pull source beyond a physical face the reference's
// BCFill value is resolved: the face applied is the last non-periodic face
// in BCFill's loop order (BC.H:365-372, z over y over x); NOSLIP and SLIP
// read the cell's own pre-stream populations, OUTFLOW the boundary cell
// next to the ghost (at a corner of two non-periodic faces, clamped in
// both), VELOCITY a constant.
// W = the working type of the pulled values (double); BCX: the level's
// boundary features (kBcxProfile: Channel or Parabolic inlet profile, ghost
// values from the face tables; kBcxPressure: PRESSURE faces)
// DEV (shifted FP32 storage only): return the stored deviations f - shift
// themselves (boundary constants less the shift) instead of populations.
// The pull in two phases: pull_raw issues the loads of the stored values
// (rf, rg) and returns which sources take a boundary constant (cmask), are
// read directly (dmask), were moved inward across a z OUTFLOW face (zmove)
// or lie beyond a PRESSURE face (Pm); pull_finish turns them into
// populations.  A caller with other work to do (the z-sweep) places that
// work between the two, so that the loads stay in flight meanwhile.
struct PullBits {
    uint32_t cmask, dmask, zmove, Pm;
};

template <int BCX, class T, bool UNIFORM_LOADS = false>
__device__ __forceinline__ PullBits pull_raw(
    const T *__restrict__ f, const CompactArgs &A, int x, int y, int z, int64_t cell, uint32_t mask, T rf[NQ],
    T rg[NQ])
{
    int tx[3], ty[3], tz[3];
    bool gx[3], gy[3], gz[3];
    targets3(x, A.nx, A.per[0], tx, gx);
    targets3(y, A.ny, A.per[1], ty, gy);
    targets3(z, A.nz, A.per[2], tz, gz);
    const uint32_t c32 = uint32_t(cell);
    uint32_t cmask = 0, dmask = 0, zmove = 0;
    // The boundary handling of the 27 sources as bit sets over q, per cell.
    // BCFill applies the faces a ghost lies beyond in the order x, y, z, the
    // last one winning (BC.H:365-372); an OUTFLOW face copies the ghost one
    // step inward, whose own value the remaining faces set (the ghosts are
    // filled in order of their distance beyond the domain).  So, scanning
    // z, y, x: G = sources beyond a non-periodic face; Tm = those whose value
    // a non-OUTFLOW face fixes (the first such face in the scan), with V, Pm
    // the VELOCITY and PRESSURE ones; O = the rest of G, beyond OUTFLOW
    // faces only, which read the source clamped into the domain.  A source
    // is read directly unless it is a solid cell or Tm holds it with a wall.
    uint32_t G = 0, Tm = 0, V = 0, Pm = 0;
    static_for<3>([&](auto dc) {
        constexpr int d = 2 - decltype(dc)::value;
        const bool *gd = d == 0 ? gx : (d == 1 ? gy : gz);
        const uint32_t Gd = (gd[2] ? qset(d, 1) : 0u) | (gd[0] ? qset(d, -1) : 0u);
        const uint32_t Ud = Gd & ~Tm;
        G |= Gd;
        const uint32_t tl = (A.bcpack >> (4 * d)) & 15u, th = (A.bcpack >> (4 * (d + 3))) & 15u;
        const uint32_t lo = Ud & qset(d, 1), hi = Ud & qset(d, -1);
        Tm |= (tl != MRB_BC_OUTFLOW ? lo : 0u) | (th != MRB_BC_OUTFLOW ? hi : 0u);
        V |= (tl == MRB_BC_VELOCITY ? lo : 0u) | (th == MRB_BC_VELOCITY ? hi : 0u);
        if constexpr ((BCX & kBcxPressure) != 0)
            Pm |= (tl == MRB_BC_PRESSURE ? lo : 0u) | (th == MRB_BC_PRESSURE ? hi : 0u);
    });
    const uint32_t O = G & ~Tm;
    const uint32_t nondirect = (G & ~(V | O)) | (~G & ~mask & kAllQ);
    // one (f_q, g_q) pair from element e of the f buffer (g == f + 1)
    auto load_pair = [&](int q, int64_t e) {
        if constexpr (kIsHalf<T>) {
            const uint32_t v = *reinterpret_cast<const uint32_t *>(f + e);
            rf[q] = __builtin_bit_cast(_Float16, uint16_t(v & 0xffffu));
            rg[q] = __builtin_bit_cast(_Float16, uint16_t(v >> 16));
        } else if constexpr (std::is_same<T, double>::value) {
            const double2 v = *reinterpret_cast<const double2 *>(f + e);  // one 16-byte load
            rf[q] = v.x;
            rg[q] = v.y;
        } else {
            const float2 v = *reinterpret_cast<const float2 *>(f + e);  // one 8-byte load
            rf[q] = v.x;
            rg[q] = v.y;
        }
    };
    // the element of direction q's source on the general path: the pull's
    // source, its bounce-back, a reflection or a cell moved inward
    auto general_e = [&](auto qc) -> int64_t {
        constexpr int q = decltype(qc)::value;
        constexpr int cx = ev(q, 0), cy = ev(q, 1), cz = ev(q, 2);
        constexpr int qb = kBounce[q];
        const bool hx = gx[cx + 1], hy = gy[cy + 1], hz = gz[cz + 1];
        const bool ghost = hx || hy || hz;
        const uint32_t src = (uint32_t(tz[cz + 1]) * uint32_t(A.ny) + uint32_t(ty[cy + 1])) * uint32_t(A.nx) +
                             uint32_t(tx[cx + 1]);
        // the face that fixes a ghost source's value (scan z, y, x past
        // OUTFLOW faces, as above) and the cell its wall fill reads: the cell
        // itself moved one step inward across each OUTFLOW face passed
        const uint32_t tfz = (A.bcpack >> (4 * (cz > 0 ? 2 : 5))) & 15u;
        const uint32_t tfy = (A.bcpack >> (4 * (cy > 0 ? 1 : 4))) & 15u;
        const uint32_t tfx = (A.bcpack >> (4 * (cx > 0 ? 0 : 3))) & 15u;
        const bool pz = hz && tfz == MRB_BC_OUTFLOW;
        const bool py = hy && tfy == MRB_BC_OUTFLOW && (pz || !hz);
        const bool px = hx && tfx == MRB_BC_OUTFLOW && (py || !hy) && (pz || !hz);
        const uint32_t t = (hz && !pz) ? tfz : (hy && !py) ? tfy : (hx && !px) ? tfx : uint32_t(MRB_BC_OUTFLOW);
        const int32_t roff = (pz ? cz * A.nx * A.ny : 0) + (py ? cy * A.nx : 0) + (px ? cx : 0);
        zmove |= uint32_t(pz) << q;
        const bool direct = ghost ? t == MRB_BC_OUTFLOW : ((mask >> q) & 1u) != 0;
        int comp = direct ? q : qb;
        comp = (ghost && t == MRB_BC_SLIPX) ? reflect(q, 0) : comp;
        comp = (ghost && t == MRB_BC_SLIPY) ? reflect(q, 1) : comp;
        comp = (ghost && t == MRB_BC_SLIPZ) ? reflect(q, 2) : comp;
        dmask |= uint32_t(direct) << q;
        cmask |= uint32_t(ghost && t == MRB_BC_VELOCITY) << q;
        return int64_t(comp) * A.cs + int64_t(direct ? src : uint32_t(int32_t(c32) + roff)) * kCellStride<T>;
    };
    const bool all_direct = __all(nondirect == 0u);
    if constexpr (UNIFORM_LOADS) {
        // the 27 source elements on the direct path or the general one (no
        // memory access on either), then the loads in one fixed sequence:
        // the number of loads in flight is then the same on every path
        int64_t e[NQ];
        if (all_direct) {
            const uint32_t nxy = uint32_t(A.nx) * uint32_t(A.ny);
            uint32_t xi[3], yi[3], zi[3];
            for (int k = 0; k < 3; ++k) {
                xi[k] = uint32_t(tx[k]);
                yi[k] = uint32_t(ty[k]) * uint32_t(A.nx);
                zi[k] = uint32_t(tz[k]) * nxy;
            }
            static_for<NQ>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                constexpr int cx = ev(q, 0), cy = ev(q, 1), cz = ev(q, 2);
                e[q] = int64_t(q) * A.cs + int64_t(zi[cz + 1] + yi[cy + 1] + xi[cx + 1]) * kCellStride<T>;
            });
            cmask = V;
            dmask = kAllQ;
        } else {
            static_for<NQ>([&](auto qc) { e[decltype(qc)::value] = general_e(qc); });
        }
        static_for<NQ>([&](auto qc) { load_pair(decltype(qc)::value, e[decltype(qc)::value]); });
    } else if (A.fast_ok && all_direct) {
        // every source of every cell of the wave is read directly: one
        // wave-uniform base per direction and a 32-bit byte offset per lane
        constexpr uint32_t bytes = kCellStride<T> * sizeof(T);
        const uint32_t nxb = uint32_t(A.nx) * bytes, nxyb = nxb * uint32_t(A.ny);
        uint32_t xo[3], yo[3], zo[3];
        for (int k = 0; k < 3; ++k) {
            xo[k] = uint32_t(tx[k]) * bytes;
            yo[k] = uint32_t(ty[k]) * nxb;
            zo[k] = uint32_t(tz[k]) * nxyb;
        }
        static_for<NQ>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            constexpr int cx = ev(q, 0), cy = ev(q, 1), cz = ev(q, 2);
            const uint32_t off = zo[cz + 1] + yo[cy + 1] + xo[cx + 1];
            const char *bf = reinterpret_cast<const char *>(f) + int64_t(q) * A.cs * int64_t(sizeof(T));
            if constexpr (kIsHalf<T>) {
                const uint32_t v = *reinterpret_cast<const uint32_t *>(bf + off);
                rf[q] = __builtin_bit_cast(_Float16, uint16_t(v & 0xffffu));
                rg[q] = __builtin_bit_cast(_Float16, uint16_t(v >> 16));
            } else if constexpr (std::is_same<T, double>::value) {
                const double2 v = *reinterpret_cast<const double2 *>(bf + off);
                rf[q] = v.x;
                rg[q] = v.y;
            } else {
                const float2 v = *reinterpret_cast<const float2 *>(bf + off);
                rf[q] = v.x;
                rg[q] = v.y;
            }
        });
        cmask = V;
        dmask = kAllQ;
    } else {
        static_for<NQ>([&](auto qc) { load_pair(decltype(qc)::value, general_e(qc)); });
    }
    return PullBits{cmask, dmask, zmove, Pm};
}

template <int BCX, class T, class W, bool DEV = false>
__device__ __forceinline__ void pull_finish(
    const T *__restrict__ f, const CompactArgs &A, const BCDev &bc, const ShiftTab &sh, const Half16 &h16,
    int x, int y, int z, uint32_t mask, const PullBits &pb, const T rf[NQ], const T rg[NQ], W fv[NQ], W gv[NQ])
{
    static_assert(!DEV || (std::is_same<T, float>::value && (BCX & kBcxPressure) == 0),
                  "deviation pull: shifted FP32 storage without PRESSURE faces");
    int tx[3], ty[3], tz[3];
    bool gx[3], gy[3], gz[3];
    targets3(x, A.nx, A.per[0], tx, gx);
    targets3(y, A.ny, A.per[1], ty, gy);
    targets3(z, A.nz, A.per[2], tz, gz);
    const uint32_t cmask = pb.cmask, dmask = pb.dmask, zmove = pb.zmove, Pm = pb.Pm;
    (void)zmove;
    (void)tx;
    (void)ty;
    if constexpr (kIsHalf<T>) {
        // inverse scales of the source planes z - c_z (index c_z + 1); a
        // source that is not the pull's (bounce-back, wall ghosts) is the
        // cell itself, in plane z (index 1).  Wave-uniform values.
        float invf[3], invg[3];
        for (int k = 0; k < 3; ++k) {
            const int b = tz[k] / kScaleSlab;
            invf[k] = ldexpf(1.0f, -h16.e_cur[2 * b]);
            invg[k] = ldexpf(1.0f, -h16.e_cur[2 * b + 1]);
        }
        static_for<NQ>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            constexpr int k = ev(q, 2) + 1;
            const bool use_const = (cmask >> q) & 1u;
            const bool direct = (dmask >> q) & 1u;
            // a wall fill read moved inward across a z OUTFLOW face lies in
            // plane z + c_z (index 1 - c_z)
            const int kn = ((zmove >> q) & 1u) ? 1 - ev(q, 2) : 1;
            const float sf = direct ? invf[k] : invf[kn], sg = direct ? invg[k] : invg[kn];
            // unscaling in FP32 is exact (a power of two on an FP16 value)
            const W vf = W(float(rf[q]) * sf) + W(sh.f(q));
            const W vg = W(float(rg[q]) * sg) + W(sh.g(q));
            double cf = bc.fval[q], cg = bc.gval[q];
            if constexpr ((BCX & kBcxProfile) != 0)
                if (use_const) profile_ghost<q>(bc, gx, gy, gz, x, y, z, cf, cg);
            fv[q] = use_const ? W(cf) : vf;
            gv[q] = use_const ? W(cg) : vg;
        });
    } else {
        (void)dmask;
        static_for<NQ>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            const bool use_const = (cmask >> q) & 1u;
            double cf = bc.fval[q], cg = bc.gval[q];
            if constexpr ((BCX & kBcxProfile) != 0)
                if (use_const) profile_ghost<q>(bc, gx, gy, gz, x, y, z, cf, cg);
            if constexpr (DEV) {
                fv[q] = use_const ? W(cf - sh.f(q)) : W(rf[q]);
                gv[q] = use_const ? W(cg - sh.g(q)) : W(rg[q]);
            } else {
                fv[q] = use_const ? cf : widen(rf[q], sh.f(q));
                gv[q] = use_const ? cg : widen(rg[q], sh.g(q));
            }
        });
    }
    if constexpr ((BCX & kBcxPressure) != 0) {
        // (a cell with a PRESSURE ghost source never takes the fast path)
        if (Pm) pressure_ghosts(f, A, bc, sh, tx, ty, tz, x, y, z, mask, Pm, fv, gv);
    } else {
        (void)Pm;
    }
}

template <int BCX, class T, class W, bool DEV = false>
__device__ __forceinline__ void pull_cell(
    const T *__restrict__ f, const T *__restrict__ g, const CompactArgs &A, const BCDev &bc,
    const ShiftTab &sh, const Half16 &h16, int x, int y, int z, int64_t cell, uint32_t mask, W fv[NQ],
    W gv[NQ])
{
    (void)g;  // g == f + 1 (f_q and g_q of a cell side by side)
    T rf[NQ], rg[NQ];
    const PullBits pb = pull_raw<BCX>(f, A, x, y, z, cell, mask, rf, rg);
    pull_finish<BCX, T, W, DEV>(f, A, bc, sh, h16, x, y, z, mask, pb, rf, rg, fv, gv);
}

// The six neighbours' Q_corr of a cell, loaded unconditionally with the
// pull (indices clamped into the domain: a value past a face is loaded but
// not used), so that the gradient does not wait on loads issued after the
// moments.
struct QNbr {
    double xp, xm, yp, ym, zp, zm;
};
__device__ __forceinline__ QNbr load_q_nbr(const double *__restrict__ Q, const CompactArgs &A, int x, int y, int z,
                                           int64_t cell)
{
    const int64_t sy = A.nx, sz = int64_t(A.nx) * A.ny;
    QNbr r;
    r.xp = Q[cell + (x + 1 < A.nx ? 1 : 0)];
    r.xm = Q[cell - (x > 0 ? 1 : 0)];
    r.yp = Q[A.N + cell + (y + 1 < A.ny ? sy : 0)];
    r.ym = Q[A.N + cell - (y > 0 ? sy : 0)];
    r.zp = Q[2 * A.N + cell + (z + 1 < A.nz ? sz : 0)];
    r.zm = Q[2 * A.N + cell - (z > 0 ? sz : 0)];
    return r;
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

// pass 1: Q_corr of the post-stream state.  With FP16 storage it also
// clears the box maxima that pass 2 accumulates for the buffer it writes.
template <class T, int BCX>
__device__ __forceinline__ void qpass_body(
    const T *__restrict__ f, const T *__restrict__ g, const uint32_t *__restrict__ mask,
    double *__restrict__ Q, const CompactArgs &A, const BCDev &bc, const ShiftTab &sh, const Half16 &h16,
    const CollideParams &P)
{
    if constexpr (kIsHalf<T>) {
        if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0 && threadIdx.y == 0 &&
            blockIdx.z % kScaleSlab == 0) {
            const int b = int(blockIdx.z) / kScaleSlab;
            h16.amax_next[2 * b] = 0u;
            h16.amax_next[2 * b + 1] = 0u;
        }
    }
    int x, y, z;
    int64_t cell;
    if (!cell_coords(A, x, y, z, cell)) return;
    const uint32_t m = mask[cell];
    if (!(m & 1u)) return;
    double fv[NQ], gv[NQ];
    pull_cell<BCX>(f, g, A, bc, sh, h16, x, y, z, cell, m, fv, gv);
    double rho, jx, jy, jz, E;
    accumulate_q_moments(fv, gv, rho, jx, jy, jz, E);
    const Macro M = macro_from(rho, jx, jy, jz, E, P);
    Q[cell] = M.Qx;
    Q[A.N + cell] = M.Qy;
    Q[2 * A.N + cell] = M.Qz;
}

// pass 1 on blocks of BX x BY cells (one name per block shape)
#define MRB_QPASS_KERNEL(NAME, BX, BY)                                                                        \
    template <class T, int BCX>                                                                               \
    __global__ void __launch_bounds__(BX * BY) NAME(                                                          \
        const T *__restrict__ f, const T *__restrict__ g, const uint32_t *__restrict__ mask,                  \
        double *__restrict__ Q, CompactArgs A, BCDev bc, ShiftTab sh, Half16 h16, CollideParams P)           \
    {                                                                                                         \
        qpass_body<T, BCX>(f, g, mask, Q, A, bc, sh, h16, P);                                                 \
    }
MRB_QPASS_KERNEL(k_qpass, 64, 4)
MRB_QPASS_KERNEL(k_qpass_y8, 64, 8)
MRB_QPASS_KERNEL(k_qpass_y2, 64, 2)
MRB_QPASS_KERNEL(k_qpass_x128, 128, 2)
MRB_QPASS_KERNEL(k_qpass_x256, 256, 1)
#undef MRB_QPASS_KERNEL

// FP16 storage: the block's largest |population - shift| per lattice goes
// to the written buffer's box entries (every block lies in one z-plane, so
// in one scale box); a block whose largest value exceeded the FP16 range
// at its scale (so that some value was clamped) is counted; the box's new
// exponent is stored by each of its blocks (the same value)
__device__ __forceinline__ void half_block_stats(const Half16 &h16, int ef, int eg, float mf, float mg)
{
    __shared__ float smf[16], smg[16];  // one per wave, blocks up to 1024 threads
    for (int o = 32; o > 0; o >>= 1) {
        mf = fmaxf(mf, __shfl_xor(mf, o));
        mg = fmaxf(mg, __shfl_xor(mg, o));
    }
    const int t = threadIdx.y * blockDim.x + threadIdx.x;
    const int wv = t / 64;
    if ((t & 63) == 0) {
        smf[wv] = mf;
        smg[wv] = mg;
    }
    __syncthreads();
    if (t == 0) {
        const int nw = int(blockDim.x * blockDim.y) / 64;
        for (int i = 1; i < nw; ++i) {
            mf = fmaxf(mf, smf[i]);
            mg = fmaxf(mg, smg[i]);
        }
        const int b = int(blockIdx.z) / kScaleSlab;
        atomicMax(&h16.amax_next[2 * b], __float_as_uint(mf));
        atomicMax(&h16.amax_next[2 * b + 1], __float_as_uint(mg));
        const uint32_t nsat = uint32_t(ldexpf(mf, ef) > 65504.0f) + uint32_t(ldexpf(mg, eg) > 65504.0f);
        if (nsat) atomicAdd(h16.ovf, nsat);
        h16.e_next[2 * b] = ef;
        h16.e_next[2 * b + 1] = eg;
    }
}

// pass 2: stream + collide, writes the B buffers
template <class T, int BCX, class R = double>
__device__ __forceinline__ void collide_body(
    const T *__restrict__ f, const T *__restrict__ g, T *__restrict__ fo, T *__restrict__ go,
    const uint32_t *__restrict__ mask, const double *__restrict__ Q, const CompactArgs &A, const BCDev &bc,
    const ShiftTab &sh, const Half16 &h16, const CollideParams &P)
{
    int x, y, z;
    int64_t cell;
    const bool in = cell_coords(A, x, y, z, cell);
    // FP16: scales of the written buffer from the read buffer's box maxima
    int ef = 0, eg = 0;
    float scf = 1.0f, scg = 1.0f;
    float mf = 0.0f, mg = 0.0f;
    if constexpr (kIsHalf<T>) {
        const int b = int(blockIdx.z) / kScaleSlab;
        ef = f16_exponent(f16_pull_max(h16.amax_cur, b, 0, A.nz, A.per[2], h16.floor_f));
        eg = f16_exponent(f16_pull_max(h16.amax_cur, b, 1, A.nz, A.per[2], h16.floor_g));
        scf = ldexpf(1.0f, ef);
        scg = ldexpf(1.0f, eg);
    }
    auto put = [&](auto qc, R fnew, R gnew) {
        constexpr int q = decltype(qc)::value;
        const int64_t e = int64_t(q) * A.cs + cell * kCellStride<T>;
        if constexpr (kIsHalf<T>) {
            store_pair(fo, go, e, to_half(fnew, R(sh.f(q)), scf, mf), to_half(gnew, R(sh.g(q)), scg, mg));
        } else {
            store_pair(fo, go, e, narrow<T>(fnew, sh.f(q)), narrow<T>(gnew, sh.g(q)));
        }
    };
    const uint32_t m = in ? mask[cell] : 0u;
    if (in && !(m & 1u)) {
        // solid cell: the stream leaves the -1 sentinel (LBM.cpp:565); FP16
        // storage holds 0 there (downloads return the sentinel)
        static_for<NQ>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            const int64_t e = int64_t(q) * A.cs + cell * kCellStride<T>;
            if constexpr (kIsHalf<T>)
                store_pair(fo, go, e, T(0.0f), T(0.0f));
            else
                store_pair(fo, go, e, narrow<T>(-1.0, sh.f(q)), narrow<T>(-1.0, sh.g(q)));
        });
    } else if (in) {
        R fv[NQ], gv[NQ];
        pull_cell<BCX>(f, g, A, bc, sh, h16, x, y, z, cell, m, fv, gv);
        const QNbr qn = load_q_nbr(Q, A, x, y, z, cell);
        MomentsR<R> mo;
        accumulate_moments(fv, gv, mo);
        const MacroR<R> M = macro_from_t<R>(mo.rho, mo.jx, mo.jy, mo.jz, mo.E, P);
        // gradient(): neighbour inside the domain box and fluid
        const bool xp = x < A.nx - 1 && ((m >> 2) & 1u), xm = x > 0 && ((m >> 1) & 1u);
        const bool yp = y < A.ny - 1 && ((m >> 4) & 1u), ym = y > 0 && ((m >> 3) & 1u);
        const bool zp = z < A.nz - 1 && ((m >> 6) & 1u), zm = z > 0 && ((m >> 5) & 1u);
        const double dqx = one_sided_gradient(xp, xm, xp ? qn.xp : 0.0, M.Qx, xm ? qn.xm : 0.0, P.inv_dx[0]);
        const double dqy = one_sided_gradient(yp, ym, yp ? qn.yp : 0.0, M.Qy, ym ? qn.ym : 0.0, P.inv_dx[1]);
        const double dqz = one_sided_gradient(zp, zm, zp ? qn.zp : 0.0, M.Qz, zm ? qn.zm : 0.0, P.inv_dx[2]);
        if constexpr (std::is_same<R, double>::value) {
            collide_cell_t<R>(fv, gv, mo, M, dqx, dqy, dqz, P, [&](auto qc, R fnew, R gnew) { put(qc, fnew, gnew); });
        } else {
            // FP32 equilibria; their sums would miss rho and E by the FP32
            // rounding of the coefficients, the same for similar cells
            // every step.  The rest direction's equilibria are rho and E
            // less the FP64 sums of the other 26, so that each cell's
            // equilibria carry its mass and energy; it is relaxed last,
            // which keeps only the two running sums live.
            const CollideCoeffsR<R> C = collide_coeffs_t<R>(mo, M, R(dqx), R(dqy), R(dqz), P);
            const R omega = C.omega;
            // rho and E as FP64 sums of the pulled values
            double rho_d = 0.0, e_d = 0.0;
            static_for<NQ>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                rho_d += double(fv[q]);
                e_d += double(gv[q]);
            });
            double sf = 0.0, sg = 0.0;
            static_for<NQ - 1>([&](auto pc) {
                constexpr int q = decltype(pc)::value + 1;
                const R fe = feq_of<q>(C), ge = geq_of<q>(C);
                sf += double(fe);
                sg += double(ge);
                put(std::integral_constant<int, q>{}, fv[q] + omega * (fe - fv[q]), gv[q] + omega * (ge - gv[q]));
            });
            const R fe0 = R(rho_d - sf), ge0 = R(e_d - sg);
            put(std::integral_constant<int, 0>{}, fv[0] + omega * (fe0 - fv[0]), gv[0] + omega * (ge0 - gv[0]));
        }
    }
    if constexpr (kIsHalf<T>) half_block_stats(h16, ef, eg, mf, mg);
}

// Pass 2 of the shifted FP32 storage with FP32 arithmetic, in deviation
// form.  The populations are handled as their stored differences d = f - s
// from the rest equilibrium s (O(1e-2) of f), so that FP32 rounding applies
// to the differences, not to the populations:
//  - moments: FP64 sums of the pulled differences plus the shift's moments;
//    macroscopic fields, the gradient and the equilibrium coefficients in
//    FP64, as the FP64 collide computes them;
//  - f equilibria: with A, B, C the x, y, z factors of the product form and
//    A0, B0, C0 the shift's, A B C - A0 B0 C0 = (dA B + A0 dB) C + A0 B0 dC,
//    the factor differences taken in FP64 and the products in FP32;
//  - g equilibria: w_q (dE + c.a1 + (dS)/2), the differences of the Grad
//    coefficients from the shift's taken in FP64;
//  - relaxation d' = d + omega (d_eq - d) in FP32;
//  - the rest direction's equilibrium difference is the FP64 sum of the
//    pulled differences less the FP64 sum of the other 26 (for f and g),
//    so that each cell's equilibria carry its mass and energy exactly;
//  - the FP64 momentum residual of the 26 FP32 f equilibria goes, half and
//    half with opposite signs, to the pair c = +-e_a of each axis, so that
//    they carry the pulled momentum too (without it the rounding, alike in
//    neighbouring cells, drifts the velocity: 2e-6 after 100 steps against
//    4e-7 with it).
// the FP64 moments of the pulled differences of a cell (fv, gv: arrays or
// LDS rows of FP32 values) plus the shift's; their mass, energy and
// momentum sums kept apart for the residuals of the FP32 equilibria
struct F32Sums {
    double rho_d, e_d, jd[3];
};
template <class SF, class SG>
__device__ __forceinline__ void f32_moments(const SF &fv, const SG &gv, const DevRest &D, MomentsR<double> &mo,
                                            F32Sums &ds)
{
    accumulate_moments(fv, gv, mo);
    ds.rho_d = mo.rho;  // sums of the differences
    ds.e_d = mo.E;
    ds.jd[0] = mo.jx;
    ds.jd[1] = mo.jy;
    ds.jd[2] = mo.jz;
    mo.rho += D.mo[0];
    mo.jx += D.mo[1];
    mo.jy += D.mo[2];
    mo.jz += D.mo[3];
    mo.pxx += D.mo[4];
    mo.pyy += D.mo[5];
    mo.pzz += D.mo[6];
    mo.pxy += D.mo[7];
    mo.pxz += D.mo[8];
    mo.pyz += D.mo[9];
    mo.E += D.mo[10];
    mo.qx += D.mo[11];
    mo.qy += D.mo[12];
    mo.qz += D.mo[13];
}

// the equilibria in FP32 from the FP64 coefficients and the relaxation of
// the differences; put(q, f', g') receives the post-collision differences
template <class SF, class SG, class Put>
__device__ __forceinline__ void f32_relax(const SF &fv, const SG &gv, const MomentsR<double> &mo, const Macro &M,
                                          const F32Sums &ds, double dqx, double dqy, double dqz, const DevRest &D,
                                          const CollideParams &P, Put &&put)
{
    const double rho_d = ds.rho_d, e_d = ds.e_d;
    const double *jd = ds.jd;
    const CollideCoeffs C = collide_coeffs(mo, M, dqx, dqy, dqz, P);
    // factor and coefficient differences from the rest state (FP64, then FP32)
    float dA[3], Bf[3], dB[3], Cf[3], dC[3], A0f[3], B0f[3];
    for (int k = 0; k < 3; ++k) {
        dA[k] = float(C.rphx[k] - D.A0[k]);
        Bf[k] = float(C.phy[k]);
        dB[k] = float(C.phy[k] - D.B0[k]);
        Cf[k] = float(C.phz[k]);
        dC[k] = float(C.phz[k] - D.B0[k]);
        A0f[k] = float(D.A0[k]);
        B0f[k] = float(D.B0[k]);
    }
    const float dE = float(C.E - D.E0);
    const float a1x = float(C.a1x), a1y = float(C.a1y), a1z = float(C.a1z);
    float dtx[2], dty[2], dtz[2];
    for (int k = 0; k < 2; ++k) {
        dtx[k] = float(C.tx[k] - D.t0[k]);
        dty[k] = float(C.ty[k] - D.t0[k]);
        dtz[k] = float(C.tz[k] - D.t0[k]);
    }
    const float dxy = float(C.dxy), dxz = float(C.dxz), dyz = float(C.dyz);
    const float omega = float(C.omega);
    // the 26 moving directions' equilibrium differences, with FP64 sums of
    // their momentum
    float fe[NQ], ge[NQ];
    double sj[3] = {0.0, 0.0, 0.0};
    static_for<NQ - 1>([&](auto pc) {
        constexpr int q = decltype(pc)::value + 1;
        constexpr int cx = ev(q, 0), cy = ev(q, 1), cz = ev(q, 2);
        fe[q] = (dA[cx + 1] * Bf[cy + 1] + A0f[cx + 1] * dB[cy + 1]) * Cf[cz + 1] +
                (A0f[cx + 1] * B0f[cy + 1]) * dC[cz + 1];
        float gq = dE;
        if constexpr (cx == 1) gq = gq + a1x;
        if constexpr (cx == -1) gq = gq - a1x;
        if constexpr (cy == 1) gq = gq + a1y;
        if constexpr (cy == -1) gq = gq - a1y;
        if constexpr (cz == 1) gq = gq + a1z;
        if constexpr (cz == -1) gq = gq - a1z;
        float S = dtx[cx * cx] + dty[cy * cy];
        if constexpr (cx * cy == 1) S = S + dxy;
        if constexpr (cx * cy == -1) S = S - dxy;
        S = S + dtz[cz * cz];
        if constexpr (cx * cz == 1) S = S + dxz;
        if constexpr (cx * cz == -1) S = S - dxz;
        if constexpr (cy * cz == 1) S = S + dyz;
        if constexpr (cy * cz == -1) S = S - dyz;
        ge[q] = (gq + 0.5f * S) * float(weight(q));
        if constexpr (cx != 0) sj[0] += cx * double(fe[q]);
        if constexpr (cy != 0) sj[1] += cy * double(fe[q]);
        if constexpr (cz != 0) sj[2] += cz * double(fe[q]);
    });
    // Their momentum is the pulled momentum less the rest state's (zero)
    // in exact arithmetic; the FP32 rounding of the 26 products leaves a
    // residual that the same rounding repeats in neighbouring cells, so the
    // velocity would drift coherently.  The residual of each axis goes to
    // that axis's pair of directions c = +-e_a (half each, opposite signs:
    // mass and the second moments are unchanged).
    static_for<3>([&](auto ac) {
        constexpr int a = decltype(ac)::value;
        constexpr int qp = a == 0 ? qof(1, 0, 0) : (a == 1 ? qof(0, 1, 0) : qof(0, 0, 1));
        constexpr int qm = kBounce[qp];
        const float h = float(0.5 * (jd[a] - sj[a]));
        fe[qp] = fe[qp] + h;
        fe[qm] = fe[qm] - h;
    });
    double sf = 0.0, sg = 0.0;
    static_for<NQ - 1>([&](auto pc) {
        constexpr int q = decltype(pc)::value + 1;
        sf += double(fe[q]);
        sg += double(ge[q]);
        put(std::integral_constant<int, q>{}, fv[q] + omega * (fe[q] - fv[q]), gv[q] + omega * (ge[q] - gv[q]));
    });
    const float fe0 = float(rho_d - sf), ge0 = float(e_d - sg);
    put(std::integral_constant<int, 0>{}, fv[0] + omega * (fe0 - fv[0]), gv[0] + omega * (ge0 - gv[0]));
}

template <int BCX>
__device__ __forceinline__ void collide_f32_body(
    const float *__restrict__ f, const float *__restrict__ g, float *__restrict__ fo, float *__restrict__ go,
    const uint32_t *__restrict__ mask, const double *__restrict__ Q, const CompactArgs &A, const BCDev &bc,
    const ShiftTab &sh, const DevRest &D, const CollideParams &P)
{
    int x, y, z;
    int64_t cell;
    if (!cell_coords(A, x, y, z, cell)) return;
    const uint32_t m = mask[cell];
    if (!(m & 1u)) {
        // solid cell: the stream's -1 sentinel (LBM.cpp:565)
        static_for<NQ>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            store_pair(fo, go, int64_t(q) * A.cs + cell * 2, narrow<float>(-1.0, sh.f(q)),
                       narrow<float>(-1.0, sh.g(q)));
        });
        return;
    }
    float fv[NQ], gv[NQ];
    pull_cell<BCX, float, float, true>(f, g, A, bc, sh, Half16{}, x, y, z, cell, m, fv, gv);
    const QNbr qn = load_q_nbr(Q, A, x, y, z, cell);
    MomentsR<double> mo;
    F32Sums ds;
    f32_moments(fv, gv, D, mo, ds);
    const Macro M = macro_from(mo.rho, mo.jx, mo.jy, mo.jz, mo.E, P);
    const bool xp = x < A.nx - 1 && ((m >> 2) & 1u), xm = x > 0 && ((m >> 1) & 1u);
    const bool yp = y < A.ny - 1 && ((m >> 4) & 1u), ym = y > 0 && ((m >> 3) & 1u);
    const bool zp = z < A.nz - 1 && ((m >> 6) & 1u), zm = z > 0 && ((m >> 5) & 1u);
    const double dqx = one_sided_gradient(xp, xm, xp ? qn.xp : 0.0, M.Qx, xm ? qn.xm : 0.0, P.inv_dx[0]);
    const double dqy = one_sided_gradient(yp, ym, yp ? qn.yp : 0.0, M.Qy, ym ? qn.ym : 0.0, P.inv_dx[1]);
    const double dqz = one_sided_gradient(zp, zm, zp ? qn.zp : 0.0, M.Qz, zm ? qn.zm : 0.0, P.inv_dx[2]);
    f32_relax(fv, gv, mo, M, ds, dqx, dqy, dqz, D, P, [&](auto qc, float fn, float gn) {
        constexpr int q = decltype(qc)::value;
        store_pair(fo, go, int64_t(q) * A.cs + cell * 2, fn, gn);
    });
}

// the FP32-arithmetic pass 2 of the shifted FP32 storage on 64x2 and 256x1
// blocks
template <int BCX>
__global__ void __launch_bounds__(128) k_collide_f32(
    const float *__restrict__ f, const float *__restrict__ g, float *__restrict__ fo, float *__restrict__ go,
    const uint32_t *__restrict__ mask, const double *__restrict__ Q, CompactArgs A, BCDev bc, ShiftTab sh,
    DevRest D, CollideParams P)
{
    collide_f32_body<BCX>(f, g, fo, go, mask, Q, A, bc, sh, D, P);
}
template <int BCX>
__global__ void __launch_bounds__(256) k_collide_f32_x256(
    const float *__restrict__ f, const float *__restrict__ g, float *__restrict__ fo, float *__restrict__ go,
    const uint32_t *__restrict__ mask, const double *__restrict__ Q, CompactArgs A, BCDev bc, ShiftTab sh,
    DevRest D, CollideParams P)
{
    collide_f32_body<BCX>(f, g, fo, go, mask, Q, A, bc, sh, D, P);
}

// ---------------------------------------------------------------------------
// Single pass as a z-sweep (FP64 and shifted FP32 storage, rows of at most
// 256 cells): one block per x-row y, all rows resident at once (checked on
// the host from the kernel's occupancy).  The block walks z; at step z it
// issues the loads of plane z+1, collides plane z-1 from the pull held in
// LDS once the rows y -+ 1 have published plane z-1's Q_y (the loads are in
// flight meanwhile), moves plane z from registers into LDS, then forms
// plane z+1's Q_corr and publishes its Q_y for the rows beside.  Each
// population is read from HBM once and written once; the neighbours'
// Q_corr travel through LDS (x), registers (z) and a ring of 8 planes of
// Q_y in global memory (y), read and written past the non-coherent L2
// (agent-scope accesses).
//
// A ring entry is two 64-bit words, each holding the plane's tag in its
// high half and one half of Q_y's bits in its low half; each word is
// written and read atomically, so a reader that sees the expected tag in
// both words holds that plane's Q_y, and the writer needs no fence (its
// population stores keep flowing while the next plane's loads issue).
// Tags: epoch * (nz + 1) + plane + 1, where the epoch (a counter a one-thread
// kernel advances before every launch) makes the tags of one launch differ
// from every earlier launch's: a stale entry of the previous launch carries
// a tag (nz + 1) + plane - plane' away from the expected one, never equal.
// Every lane of a row waits for both rows beside, fluid or not, so a row
// publishes plane k+1 only after it saw its neighbours' plane k-1, which
// they published after they had read plane k-3 of this row: the ring's 8
// planes (4 would do) are never overwritten while read.
// The cell arithmetic is the two-pass step's (same functions, same order).
// ---------------------------------------------------------------------------
constexpr int kSweepRing = kSweepRingPlanes;
constexpr int kSweepRow = 256;
constexpr uint32_t kSweepSpinLimit = 1u << 22;

// f_q (or g_q) of the block's row, cell x, held in LDS at [2q + lattice][x]
// (stride: 2 x the block's row width)
template <class H>
struct LdsRowT {
    const H *p;
    int stride = 2 * kSweepRow;
    __device__ __forceinline__ H operator[](int q) const { return p[q * stride]; }
};
using LdsRow = LdsRowT<double>;

__device__ __forceinline__ void store_agent(double *p, double v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double load_agent(double *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// every memory access of the wave issued so far has completed
__device__ __forceinline__ void drain_wave()
{
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_waitcnt(0);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// tagged ring entry: word 0 = tag : high 32 bits of v, word 1 = tag : low
__device__ __forceinline__ void put_tagged(uint64_t *p, uint32_t tag, double v)
{
    const uint64_t b = __builtin_bit_cast(uint64_t, v), t = uint64_t(tag) << 32;
    __hip_atomic_store(p, t | (b >> 32), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(p + 1, t | (b & 0xffffffffull), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
struct Tagged {
    uint64_t w0 = 0, w1 = 0;
    __device__ __forceinline__ void load(uint64_t *p)
    {
        w0 = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        w1 = __hip_atomic_load(p + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __device__ __forceinline__ bool has(uint32_t tag) const
    {
        return uint32_t(w0 >> 32) == tag && uint32_t(w1 >> 32) == tag;
    }
    __device__ __forceinline__ double value() const
    {
        return __builtin_bit_cast(double, (w0 << 32) | (w1 & 0xffffffffull));
    }
};

__global__ void k_bump_u32(uint32_t *a) { a[0] += 1u; }

// F32A (shifted FP32 storage): the collide of the two-pass step's
// FP32-arithmetic pass 2 (k_collide_f32, deviation form), the held plane
// kept as the stored FP32 differences; the neighbours' Q_corr as pass 1
// forms it (FP64, from the populations).
template <class T, int BCX, bool F32A = false>
__global__ void __launch_bounds__(kSweepRow, 1) k_sweep(
    const T *__restrict__ f, const T *__restrict__ g, T *__restrict__ fo, T *__restrict__ go,
    const uint32_t *__restrict__ mask, uint64_t *qring, const uint32_t *epoch, uint32_t *fault, CompactArgs A,
    BCDev bc, ShiftTab sh, CollideParams P, DevRest D)
{
    static_assert(!F32A || (std::is_same<T, float>::value && (BCX & kBcxPressure) == 0),
                  "FP32-arithmetic sweep: shifted FP32 storage without PRESSURE faces");
    using H = std::conditional_t<F32A, float, double>;  // held values
    __shared__ H hold[2 * NQ * kSweepRow];
    __shared__ double qxrow[kSweepRow];
    __shared__ int gave_up;  // a wait for a neighbour row hit the spin limit
    if (threadIdx.x == 0) gave_up = 0;
    // No branch of the sweep depends on the lane, so that the compiler
    // counts the memory operations in flight exactly and the collide runs
    // while the next plane's loads are outstanding: lanes past the row's
    // end repeat the last cell's work (the same values to the same
    // addresses), and solid cells pull and collide like fluid ones, their
    // results replaced by the solid marker when stored.
    const int x = min(int(threadIdx.x), A.nx - 1), y = int(blockIdx.x);
    const int64_t nxy = int64_t(A.nx) * A.ny;
    const int64_t row = int64_t(y) * A.nx + x;
    // the rows beside (this row itself at the y faces, read but not used)
    const bool wm = y > 0, wp = y + 1 < A.ny;
    const int64_t rowm = wm ? row - A.nx : row, rowp = wp ? row + A.nx : row;
    const uint32_t tag0 =
        __hip_atomic_load(epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) * uint32_t(A.nz + 1) + 1u;
    // two planes in registers: the pull of plane p (rf, rg: its Q_corr is
    // known) and the stored values of plane p+1 in flight (nf, ng) while
    // plane p-1, held in LDS, is collided; they become populations (or FP32
    // differences) after the collide (pull_finish), so the collide does not
    // wait for them
    H rf[NQ], rg[NQ];
    T nf[NQ], ng[NQ];
    PullBits pb{0u, 0u, 0u, 0u};
    // planes past the last one are replaced by the last (their loads and
    // results go unused), and a plane's stream mask is loaded one plane
    // ahead of its pull, so that the pull's loads do not wait for it
    auto mask_of = [&](int zp) { return mask[int64_t(min(zp, A.nz - 1)) * nxy + row]; };
    // (a solid cell pulls as if all its sources were fluid, so that it does
    // not keep its wave off the direct-load path)
    auto pull_issue = [&](int zp, uint32_t m) {
        zp = min(zp, A.nz - 1);
        pb = pull_raw<BCX, T, true>(f, A, x, y, zp, int64_t(zp) * nxy + row, (m & 1u) ? m : (m | kAllQ), nf, ng);
    };
    auto pull_done = [&](int zp, uint32_t m) {
        zp = min(zp, A.nz - 1);
        pull_finish<BCX, T, H, F32A>(f, A, bc, sh, Half16{}, x, y, zp, m, pb, nf, ng, rf, rg);
    };
    // Q_corr of plane zp from its populations (with F32A formed from the
    // stored values still in nf, ng)
    auto q_of = [&](int zp, uint32_t m, double &qx, double &qy, double &qz) {
        double rho, jx, jy, jz, E;
        if constexpr (F32A) {
            zp = min(zp, A.nz - 1);
            double pf[NQ], pg[NQ];
            pull_finish<BCX, T, double>(f, A, bc, sh, Half16{}, x, y, zp, m, pb, nf, ng, pf, pg);
            accumulate_q_moments(pf, pg, rho, jx, jy, jz, E);
        } else {
            (void)zp;
            accumulate_q_moments(rf, rg, rho, jx, jy, jz, E);
        }
        const Macro M = macro_from(rho, jx, jy, jz, E, P);
        const bool fluid = (m & 1u) != 0;
        qx = fluid ? M.Qx : 0.0;
        qy = fluid ? M.Qy : 0.0;
        qz = fluid ? M.Qz : 0.0;
    };
    auto entry = [&](int zp, int64_t r) { return qring + 2 * (int64_t(zp % kSweepRing) * nxy + r); };
    auto publish = [&](int zp, double qy) { put_tagged(entry(zp, row), tag0 + uint32_t(zp), qy); };
    uint32_t m0 = 0u, m1 = mask_of(0), mnext = mask_of(1);
    double qx1, qy1, qz1;      // Q_corr of plane p (pulled, in rf/rg)
    double qz0 = 0.0, qzm = 0.0;  // Q_z of planes p-1 and p-2
    // plane p becomes the held one, plane p+1 (in flight) is finished,
    // its Q_corr formed and published
    auto advance = [&](int p, uint32_t m2) {
        static_for<NQ>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            hold[(2 * q) * kSweepRow + x] = rf[q];
            hold[(2 * q + 1) * kSweepRow + x] = rg[q];
        });
        qxrow[x] = qx1;
        qzm = qz0;
        qz0 = qz1;
        m0 = m1;
        pull_done(p + 1, m2);
        q_of(p + 1, m2, qx1, qy1, qz1);
        publish(p + 1, qy1);
        m1 = m2;
    };
    // collide plane zc (held in LDS) once the rows beside have published
    // its Q_y (em, ep: their entries, loaded before); bounded wait: past the
    // spin limit the block counts a fault and writes NaN from then on
    auto collide = [&](int zc, Tagged &em, Tagged &ep) {
        const uint32_t tag = tag0 + uint32_t(zc);
        bool need_m = wm && !em.has(tag), need_p = wp && !ep.has(tag);
        if (need_m || need_p) {
            uint32_t it = 0;
            while (!gave_up) {
                if (++it > kSweepSpinLimit) {
                    atomicAdd(fault, 1u);
                    gave_up = 1;
                    break;
                }
                __builtin_amdgcn_s_sleep(2);
                if (need_m) {
                    em.load(entry(zc, rowm));
                    need_m = !em.has(tag);
                }
                if (need_p) {
                    ep.load(entry(zc, rowp));
                    need_p = !ep.has(tag);
                }
                if (!need_m && !need_p) break;
            }
        }
        __syncthreads();
        const bool poisoned = gave_up != 0;
        const int64_t cell = int64_t(zc) * nxy + row;
        const uint32_t m = m0;
        const bool fluid = (m & 1u) != 0;
        const bool xp = x < A.nx - 1 && ((m >> 2) & 1u), xm = x > 0 && ((m >> 1) & 1u);
        const bool yp = wp && ((m >> 4) & 1u), ym = wm && ((m >> 3) & 1u);
        const bool zp = zc < A.nz - 1 && ((m >> 6) & 1u), zm = zc > 0 && ((m >> 5) & 1u);
        const double qyp = yp ? ep.value() : 0.0;
        const double qym = ym ? em.value() : 0.0;
        const LdsRowT<H> F{&hold[x]}, G{&hold[kSweepRow + x]};
        MomentsR<double> mo;
        F32Sums ds;
        if constexpr (F32A)
            f32_moments(F, G, D, mo, ds);
        else
            accumulate_moments(F, G, mo);
        const Macro M = macro_from(mo.rho, mo.jx, mo.jy, mo.jz, mo.E, P);
        const double dqx =
            one_sided_gradient(xp, xm, xp ? qxrow[x + 1] : 0.0, M.Qx, xm ? qxrow[x - 1] : 0.0, P.inv_dx[0]);
        const double dqy = one_sided_gradient(yp, ym, qyp, M.Qy, qym, P.inv_dx[1]);
        const double dqz = one_sided_gradient(zp, zm, zp ? qz1 : 0.0, M.Qz, zm ? qzm : 0.0, P.inv_dx[2]);
        const double nan = __builtin_nan("");
        if constexpr (F32A) {
            (void)ds;
            f32_relax(F, G, mo, M, ds, dqx, dqy, dqz, D, P, [&](auto qc, float fn, float gn) {
                constexpr int q = decltype(qc)::value;
                const float vf = fluid ? (poisoned ? float(nan) : fn) : narrow<float>(-1.0, sh.f(q));
                const float vg = fluid ? (poisoned ? float(nan) : gn) : narrow<float>(-1.0, sh.g(q));
                store_pair(fo, go, int64_t(q) * A.cs + cell * kCellStride<T>, vf, vg);
            });
        } else {
            (void)D;
            collide_cell_t<double>(F, G, mo, M, dqx, dqy, dqz, P, [&](auto qc, double fnew, double gnew) {
                constexpr int q = decltype(qc)::value;
                const double vf = fluid ? (poisoned ? nan : fnew) : -1.0;
                const double vg = fluid ? (poisoned ? nan : gnew) : -1.0;
                store_pair(fo, go, int64_t(q) * A.cs + cell * kCellStride<T>, narrow<T>(vf, sh.f(q)),
                           narrow<T>(vg, sh.g(q)));
            });
        }
        __syncthreads();
    };
    // plane 0 pulled, its Q_corr published; plane 1 in flight
    pull_issue(0, m1);
    pull_done(0, m1);
    q_of(0, m1, qx1, qy1, qz1);
    publish(0, qy1);
    uint32_t m2 = mnext;
    mnext = mask_of(2);
    pull_issue(1, m2);
    __syncthreads();  // gave_up
    advance(0, m2);
    // step z: plane z+1's loads issued, plane z-1 collided, plane z held
    for (int z = 1; z < A.nz; ++z) {
        Tagged em, ep;
        em.load(entry(z - 1, rowm));
        ep.load(entry(z - 1, rowp));
        m2 = mnext;
        mnext = mask_of(z + 2);
        pull_issue(z + 1, m2);
        collide(z - 1, em, ep);
        advance(z, m2);
    }
    Tagged em, ep;
    em.load(entry(A.nz - 1, rowm));
    ep.load(entry(A.nz - 1, rowp));
    collide(A.nz - 1, em, ep);
}

// The z-sweep on single-wave blocks of 64 cells of an x-row: four blocks,
// independent of each other, share a CU (the 256-wide block runs its four
// waves in lockstep).  A block exchanges Q_y with the blocks of rows y -+ 1
// and the Q_x of its two end cells with the blocks beside in x, through
// the same ring (Q_x of the end cells in a second ring), and waits for all
// four; inside the block Q_x moves between lanes.
constexpr int kSweepW = 64;

template <int BCX>
__global__ void __launch_bounds__(kSweepW, 1) k_sweep_x64(
    const double *__restrict__ f, const double *__restrict__ g, double *__restrict__ fo, double *__restrict__ go,
    const uint32_t *__restrict__ mask, double *qring, double *qxring, uint32_t *flags, uint32_t *fault,
    CompactArgs A, BCDev bc, ShiftTab sh, CollideParams P)
{
    __shared__ double hold[2 * NQ * kSweepW];
    const int lane = int(threadIdx.x), xb = int(blockIdx.x), y = int(blockIdx.y), nxb = int(gridDim.x);
    const int x = xb * kSweepW + lane;
    const bool in = x < A.nx;
    const int last = min(kSweepW, A.nx - xb * kSweepW) - 1;  // the block's last lane inside the domain
    const int64_t nxy = int64_t(A.nx) * A.ny;
    const int64_t row = int64_t(y) * A.nx + x;
    const int64_t edge_plane = int64_t(A.ny) * nxb * 2;  // Q_x ring entries per plane
    uint32_t *const own_flag = &flags[y * nxb + xb];
    double rf[NQ], rg[NQ], nf[NQ], ng[NQ];
    auto pull_into = [&](int zp, uint32_t &m, double (&pf)[NQ], double (&pg)[NQ]) {
        m = 0u;
        if (!in) return;
        const int64_t cell = int64_t(zp) * nxy + row;
        m = mask[cell];
        if (!(m & 1u)) return;
        pull_cell<BCX>(f, g, A, bc, sh, Half16{}, x, y, zp, cell, m, pf, pg);
    };
    auto q_of = [&](uint32_t m, const double (&pf)[NQ], const double (&pg)[NQ], double &qx, double &qy,
                    double &qz) {
        qx = qy = qz = 0.0;
        if (!in || !(m & 1u)) return;
        double rho, jx, jy, jz, E;
        accumulate_q_moments(pf, pg, rho, jx, jy, jz, E);
        const Macro M = macro_from(rho, jx, jy, jz, E, P);
        qx = M.Qx;
        qy = M.Qy;
        qz = M.Qz;
    };
    auto publish = [&](int zp, double qx, double qy) {
        const int slot = zp % kSweepRing;
        if (in) store_agent(&qring[int64_t(slot) * nxy + row], qy);
        if (lane == 0 || lane == last)
            store_agent(&qxring[int64_t(slot) * edge_plane + (int64_t(y) * nxb + xb) * 2 + (lane == 0 ? 0 : 1)], qx);
        drain_wave();
        if (lane == 0) __hip_atomic_store(own_flag, uint32_t(zp + 1), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    };
    bool gave_up = false;  // wave-uniform
    auto wait_block = [&](int bx, int by, uint32_t target) {
        uint32_t it = 0;
        while (__hip_atomic_load(&flags[by * nxb + bx], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            if (++it > kSweepSpinLimit) {
                if (lane == 0) atomicAdd(fault, 1u);
                gave_up = true;
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
    };
    uint32_t m1 = 0u, m0 = 0u;
    double qx1, qy1, qz1;
    pull_into(0, m1, rf, rg);
    q_of(m1, rf, rg, qx1, qy1, qz1);
    publish(0, qx1, qy1);
    double qx0 = 0.0, qz0 = 0.0, qzm = 0.0;  // Q_x of the held plane, Q_z of planes p-1, p-2
    for (int z = 0; z <= A.nz; ++z) {
        uint32_t m2 = 0u;
        if (z + 1 < A.nz) pull_into(z + 1, m2, nf, ng);
        if (z >= 1) {
            const int zc = z - 1;
            if (!gave_up) {
                if (y > 0) wait_block(xb, y - 1, uint32_t(zc + 1));
                if (y + 1 < A.ny) wait_block(xb, y + 1, uint32_t(zc + 1));
                if (xb > 0) wait_block(xb - 1, y, uint32_t(zc + 1));
                if (xb + 1 < nxb) wait_block(xb + 1, y, uint32_t(zc + 1));
            }
            const int slot = zc % kSweepRing;
            // Q_x of x -+ 1: the lanes beside, or the end cells of the blocks beside
            double qxp = __shfl_down(qx0, 1, kSweepW), qxm = __shfl_up(qx0, 1, kSweepW);
            const double *er = qxring + int64_t(slot) * edge_plane + int64_t(y) * nxb * 2;
            if (lane == last && xb + 1 < nxb) qxp = load_agent(const_cast<double *>(&er[(xb + 1) * 2 + 0]));
            if (lane == 0 && xb > 0) qxm = load_agent(const_cast<double *>(&er[(xb - 1) * 2 + 1]));
            if (in) {
                const int64_t cell = int64_t(zc) * nxy + row;
                const uint32_t m = m0;
                if (!(m & 1u)) {
                    static_for<NQ>([&](auto qc) {
                        constexpr int q = decltype(qc)::value;
                        store_pair(fo, go, int64_t(q) * A.cs + cell * kCellStride<double>, -1.0, -1.0);
                    });
                } else {
                    const LdsRow F{&hold[lane], 2 * kSweepW}, G{&hold[kSweepW + lane], 2 * kSweepW};
                    MomentsR<double> mo;
                    accumulate_moments(F, G, mo);
                    const Macro M = macro_from(mo.rho, mo.jx, mo.jy, mo.jz, mo.E, P);
                    const bool xp = x < A.nx - 1 && ((m >> 2) & 1u), xm = x > 0 && ((m >> 1) & 1u);
                    const bool yp = y < A.ny - 1 && ((m >> 4) & 1u), ym = y > 0 && ((m >> 3) & 1u);
                    const bool zp = zc < A.nz - 1 && ((m >> 6) & 1u), zm = zc > 0 && ((m >> 5) & 1u);
                    double *ring = qring + int64_t(slot) * nxy + row;
                    const double dqx = one_sided_gradient(xp, xm, xp ? qxp : 0.0, M.Qx, xm ? qxm : 0.0, P.inv_dx[0]);
                    const double dqy = one_sided_gradient(yp, ym, yp ? load_agent(ring + A.nx) : 0.0, M.Qy,
                                                          ym ? load_agent(ring - A.nx) : 0.0, P.inv_dx[1]);
                    const double dqz = one_sided_gradient(zp, zm, zp ? qz1 : 0.0, M.Qz, zm ? qzm : 0.0,
                                                          P.inv_dx[2]);
                    const double nan = __builtin_nan("");
                    const bool poisoned = gave_up;
                    collide_cell_t<double>(F, G, mo, M, dqx, dqy, dqz, P,
                                           [&](auto qc, double fnew, double gnew) {
                                               constexpr int q = decltype(qc)::value;
                                               store_pair(fo, go, int64_t(q) * A.cs + cell * kCellStride<double>,
                                                          poisoned ? nan : fnew,
                                                          poisoned ? nan : gnew);
                                           });
                }
            }
        }
        if (z == A.nz) break;
        // plane z becomes the held one (each lane its own column: no barrier)
        static_for<NQ>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            hold[(2 * q) * kSweepW + lane] = rf[q];
            hold[(2 * q + 1) * kSweepW + lane] = rg[q];
        });
        qx0 = qx1;
        qzm = qz0;
        qz0 = qz1;
        m0 = m1;
        if (z + 1 < A.nz) {
            q_of(m2, nf, ng, qx1, qy1, qz1);
            publish(z + 1, qx1, qy1);
            static_for<NQ>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                rf[q] = nf[q];
                rg[q] = ng[q];
            });
            m1 = m2;
        } else {
            qz1 = 0.0;
        }
    }
}

__global__ void k_zero_u32(uint32_t *__restrict__ a, int n)
{
    const int i = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) a[i] = 0u;
}

// pass 2 on blocks of BX x BY cells, register budget as the compiler
// chooses (2 waves per SIMD); one name per block shape
#define MRB_COLLIDE_KERNEL(NAME, BX, BY)                                                                      \
    template <class T, int BCX>                                                                               \
    __global__ void __launch_bounds__(BX * BY) NAME(                                                          \
        const T *__restrict__ f, const T *__restrict__ g, T *__restrict__ fo, T *__restrict__ go,             \
        const uint32_t *__restrict__ mask, const double *__restrict__ Q, CompactArgs A, BCDev bc, ShiftTab sh, \
        Half16 h16, CollideParams P)                                                                          \
    {                                                                                                         \
        collide_body<T, BCX>(f, g, fo, go, mask, Q, A, bc, sh, h16, P);                                       \
    }
MRB_COLLIDE_KERNEL(k_collide, 64, 4)
MRB_COLLIDE_KERNEL(k_collide_y8, 64, 8)
MRB_COLLIDE_KERNEL(k_collide_y2, 64, 2)
MRB_COLLIDE_KERNEL(k_collide_x128, 128, 2)
MRB_COLLIDE_KERNEL(k_collide_x256, 256, 1)
#undef MRB_COLLIDE_KERNEL

// scaled FP16 storage with FP32 arithmetic in pass 2 (pass 1 and the
// gradient stay FP64): the FP16 rounding of the stored deviations, about
// 2.4e-4 of each, dominates the FP32 rounding of the populations
template <class T, int BCX>
__global__ void __launch_bounds__(256) k_collide_h32(
    const T *__restrict__ f, const T *__restrict__ g, T *__restrict__ fo, T *__restrict__ go,
    const uint32_t *__restrict__ mask, const double *__restrict__ Q, CompactArgs A, BCDev bc, ShiftTab sh,
    Half16 h16, CollideParams P)
{
    collide_body<T, BCX, float>(f, g, fo, go, mask, Q, A, bc, sh, h16, P);
}
// the same on 256x1 blocks
template <class T, int BCX>
__global__ void __launch_bounds__(256) k_collide_h32_x256(
    const T *__restrict__ f, const T *__restrict__ g, T *__restrict__ fo, T *__restrict__ go,
    const uint32_t *__restrict__ mask, const double *__restrict__ Q, CompactArgs A, BCDev bc, ShiftTab sh,
    Half16 h16, CollideParams P)
{
    collide_body<T, BCX, float>(f, g, fo, go, mask, Q, A, bc, sh, h16, P);
}

// The two on levels without inlet profiles, held to 128 registers (four
// waves per SIMD; they need 118 then, and 129 without the bound, which
// leaves three).  With a profile inlet the bound would spill, so those
// levels use the kernels above (193 registers).
template <class T>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) k_collide_h32_w4(
    const T *__restrict__ f, const T *__restrict__ g, T *__restrict__ fo, T *__restrict__ go,
    const uint32_t *__restrict__ mask, const double *__restrict__ Q, CompactArgs A, BCDev bc, ShiftTab sh,
    Half16 h16, CollideParams P)
{
    collide_body<T, 0, float>(f, g, fo, go, mask, Q, A, bc, sh, h16, P);
}
template <class T>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) k_collide_h32_x256_w4(
    const T *__restrict__ f, const T *__restrict__ g, T *__restrict__ fo, T *__restrict__ go,
    const uint32_t *__restrict__ mask, const double *__restrict__ Q, CompactArgs A, BCDev bc, ShiftTab sh,
    Half16 h16, CollideParams P)
{
    collide_body<T, 0, float>(f, g, fo, go, mask, Q, A, bc, sh, h16, P);
}

// at most 168 registers (3 waves per SIMD), some values spilled to scratch
template <class T, int BCX>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 3))) k_collide_w3(
    const T *__restrict__ f, const T *__restrict__ g, T *__restrict__ fo, T *__restrict__ go,
    const uint32_t *__restrict__ mask, const double *__restrict__ Q, CompactArgs A, BCDev bc, ShiftTab sh,
    Half16 h16, CollideParams P)
{
    collide_body<T, BCX>(f, g, fo, go, mask, Q, A, bc, sh, h16, P);
}

// Q_corr of an inside fluid cell from its own pull, as k_qpass computes it
template <class T, int BCX>
__device__ __forceinline__ Macro pulled_macro(const T *__restrict__ f, const T *__restrict__ g,
                                              const uint32_t *__restrict__ mask, const CompactArgs &A,
                                              const BCDev &bc, const ShiftTab &sh, const CollideParams &P, int x,
                                              int y, int z)
{
    const int64_t cell = (int64_t(z) * A.ny + y) * A.nx + x;
    double fv[NQ], gv[NQ];
    pull_cell<BCX>(f, g, A, bc, sh, Half16{}, x, y, z, cell, mask[cell], fv, gv);
    double rho, jx, jy, jz, E;
    accumulate_q_moments(fv, gv, rho, jx, jy, jz, E);
    return macro_from(rho, jx, jy, jz, E, P);
}

// single pass: stream + collide with the neighbours' Q_corr recomputed in
// place of pass 1.  A wave is one row segment of 64 cells along x
// (blockDim.x == 64): the x neighbours' Q comes from the lanes beside (the
// end lanes pull the cells beyond the segment), the y and z neighbours' Q
// from their own pulls, which the blocks of the neighbouring rows and
// planes also read, so that they come mostly from L2.  The values equal
// pass 1's bit for bit (same pull, same sums, same operation order).
// FP64 and shifted FP32 storage.
template <class T, int BCX>
__global__ void __launch_bounds__(256) k_onepass(
    const T *__restrict__ f, const T *__restrict__ g, T *__restrict__ fo, T *__restrict__ go,
    const uint32_t *__restrict__ mask, CompactArgs A, BCDev bc, ShiftTab sh, CollideParams P)
{
    int x, y, z;
    int64_t cell;
    const bool in = cell_coords(A, x, y, z, cell);
    const uint32_t m = in ? mask[cell] : 0u;
    const bool fluid = in && (m & 1u);
    // gradient(): neighbour inside the domain box and fluid
    const bool xp = fluid && x < A.nx - 1 && ((m >> 2) & 1u), xm = fluid && x > 0 && ((m >> 1) & 1u);
    const bool yp = fluid && y < A.ny - 1 && ((m >> 4) & 1u), ym = fluid && y > 0 && ((m >> 3) & 1u);
    const bool zp = fluid && z < A.nz - 1 && ((m >> 6) & 1u), zm = fluid && z > 0 && ((m >> 5) & 1u);
    double qyp = 0.0, qym = 0.0, qzp = 0.0, qzm = 0.0, qxe = 0.0;
    if (yp) qyp = pulled_macro<T, BCX>(f, g, mask, A, bc, sh, P, x, y + 1, z).Qy;
    if (ym) qym = pulled_macro<T, BCX>(f, g, mask, A, bc, sh, P, x, y - 1, z).Qy;
    if (zp) qzp = pulled_macro<T, BCX>(f, g, mask, A, bc, sh, P, x, y, z + 1).Qz;
    if (zm) qzm = pulled_macro<T, BCX>(f, g, mask, A, bc, sh, P, x, y, z - 1).Qz;
    const int lane = int(threadIdx.x);
    if ((lane == 0 && xm) || (lane == 63 && xp))
        qxe = pulled_macro<T, BCX>(f, g, mask, A, bc, sh, P, lane == 0 ? x - 1 : x + 1, y, z).Qx;
    double fv[NQ], gv[NQ];
    Moments mo{};
    Macro M{};
    if (fluid) {
        pull_cell<BCX>(f, g, A, bc, sh, Half16{}, x, y, z, cell, m, fv, gv);
        accumulate_moments(fv, gv, mo);
        M = macro_from(mo.rho, mo.jx, mo.jy, mo.jz, mo.E, P);
    }
    // every lane of the wave takes part in the exchange
    const double qx_lo = __shfl_up(M.Qx, 1), qx_hi = __shfl_down(M.Qx, 1);
    const double qxm = lane == 0 ? qxe : qx_lo, qxp = lane == 63 ? qxe : qx_hi;
    if (in && !fluid) {
        // solid cell: the stream leaves the -1 sentinel (LBM.cpp:565)
        static_for<NQ>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            store_pair(fo, go, int64_t(q) * A.cs + cell * kCellStride<T>, narrow<T>(-1.0, sh.f(q)),
                       narrow<T>(-1.0, sh.g(q)));
        });
    } else if (fluid) {
        const double dqx = one_sided_gradient(xp, xm, xp ? qxp : 0.0, M.Qx, xm ? qxm : 0.0, P.inv_dx[0]);
        const double dqy = one_sided_gradient(yp, ym, qyp, M.Qy, qym, P.inv_dx[1]);
        const double dqz = one_sided_gradient(zp, zm, qzp, M.Qz, qzm, P.inv_dx[2]);
        collide_cell(fv, gv, mo, M, dqx, dqy, dqz, P, [&](auto qc, double fnew, double gnew) {
            constexpr int q = decltype(qc)::value;
            store_pair(fo, go, int64_t(q) * A.cs + cell * kCellStride<T>, narrow<T>(fnew, sh.f(q)),
                       narrow<T>(gnew, sh.g(q)));
        });
    }
}

// stream only (parity of the pull with the reference push); FP64 and
// shifted FP32 storage
template <class T, int BCX>
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
        pull_cell<BCX>(f, g, A, bc, sh, Half16{}, x, y, z, cell, m, fv, gv);
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
    // bytes of one cell's (f_q, g_q) pair
    const int64_t cell_bytes = L->storage == MRB_STORAGE_F64 ? 16 : (L->storage == MRB_STORAGE_F32_SHIFTED ? 8 : 4);
    A.fast_ok = A.N * cell_bytes < (int64_t(1) << 32) ? 1 : 0;
    return A;
}

dim3 grid_of(const CompactArgs &A, int by = 4, int bx = 64)
{
    return dim3((A.nx + bx - 1) / bx, (A.ny + by - 1) / by, A.nz);
}

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

// the boundary floor of the FP16 box maxima: |f_eq - shift| and
// |g_eq - shift| of the inlet state (for a Channel or Parabolic profile the
// reference speed bounds the profile) and of the pressure faces' state at
// rest and at their g velocity
static void f16_floors(const mrb_level *L, float &ff, float &fg)
{
    double mf = 0.0, mg = 0.0;
    const BCDev &d = L->bcdev;
    auto take = [&](const double *fv, const double *gv) {
        for (int q = 0; q < NQ; ++q) {
            mf = std::max(mf, std::fabs(fv[q] - L->fshift[q]));
            mg = std::max(mg, std::fabs(gv[q] - L->gshift[q]));
        }
    };
    bool velocity = false;
    for (int n = 0; n < 6; ++n) velocity = velocity || L->bc.bc_type[n] == MRB_BC_VELOCITY;
    if (velocity) {
        if (d.prof.kind) {
            const VelProfile &P = d.prof;
            double fv[NQ], gv[NQ];
            for (int dir = 0; dir < 3; ++dir)
                for (int sgn = -1; sgn <= 1; sgn += 2) {
                    double v[3] = {0.0, 0.0, 0.0};
                    v[dir] = sgn * P.vref;
                    for (int q = 0; q < NQ; ++q) {
                        fv[q] = set_equilibrium_value(P.rho, v, P.R * P.T, q);
                        gv[q] = thermal_g_equilibrium(P.rho, v, P.T, P.R, P.gamma, q);
                    }
                    take(fv, gv);
                }
        } else {
            take(d.fval, d.gval);
        }
    }
    if (d.pdir >= 0) {
        const double zero[3] = {0.0, 0.0, 0.0};
        for (int q = 0; q < NQ; ++q) {
            mf = std::max(mf, std::fabs(set_equilibrium_value(d.prho, zero, d.prt, q) - L->fshift[q]));
            if (kEV[q][d.pdir] != 0) mg = std::max(mg, std::fabs(d.pgval[q] - L->gshift[q]));
        }
    }
    ff = float(mf);
    fg = float(mg);
}

static DevRest dev_rest(const mrb_level *L)
{
    DevRest d{};
    if (L->storage != MRB_STORAGE_F32_SHIFTED) return d;
    const double R = kAirMolarMass / (L->phys_set ? L->phys.m_bar : kAirMolarMass);
    const double gamma = L->phys_set ? L->phys.adiabatic_exponent : 5.0 / 3.0;
    const double rho0 = L->shift_rho0, T0 = L->shift_T0, rt = R * T0;
    // the factors of set_equilibrium_value at u = 0 (the f shift is
    // ((rho0 phi_x) phi_y) phi_z of these)
    for (int c = -1; c <= 1; ++c) {
        const double pxx = 0.0 * 0.0 + rt;
        const double phi = c * 0.5 * 0.0 + iabs(c) * (1.50 * pxx - 1.0) - pxx + 1.0;
        d.B0[c + 1] = phi;
        d.A0[c + 1] = rho0 * phi;
    }
    // the Grad coefficients of thermal_g_equilibrium at u = 0
    const double cv = R / (gamma - 1.0);
    const double E = rho0 * (2.0 * cv * T0 + 0.0 * 0.0 + 0.0 * 0.0 + 0.0 * 0.0);
    const double energy = E / (2.0 * rho0);
    const double Tv = (0.50 / cv) * ((E / rho0) - 0.0 * 0.0 + 0.0 * 0.0 + 0.0 * 0.0);
    const double p = rho0 * R * Tv;
    const double h = energy + (p / rho0);
    const double rxx = 2.0 * rho0 * 0.0 * 0.0 * (h + p / rho0) + 2.0 * p * h;
    const double inv = 1.0 / kTheta0;
    const double a2xx = (rxx - E * 1.0 * kTheta0) * inv * inv;
    d.E0 = E;
    d.t0[0] = (0 - kTheta0) * a2xx;
    d.t0[1] = (1 - kTheta0) * a2xx;
    MomentsR<double> m;
    accumulate_moments(L->fshift, L->gshift, m);
    const double mo[14] = {m.rho, m.jx, m.jy, m.jz, m.pxx, m.pyy, m.pzz, m.pxy, m.pxz, m.pyz, m.E, m.qx, m.qy, m.qz};
    for (int k = 0; k < 14; ++k) d.mo[k] = mo[k];
    return d;
}

static Half16 half16_of(mrb_level *L)
{
    Half16 h{};
    if (L->storage == MRB_STORAGE_F16_SCALED) {
        h.e_cur = L->d_e16[L->cur];
        h.e_next = L->d_e16[1 - L->cur];
        h.amax_cur = L->d_amax16[L->cur];
        h.amax_next = L->d_amax16[1 - L->cur];
        h.ovf = L->d_ovf16;
        f16_floors(L, h.floor_f, h.floor_g);
    }
    return h;
}

// blocks of a kernel that the GPU holds at once (occupancy x CUs): the
// z-sweeps wait on each other and need every block resident; they are
// launched only when the occupancy calculation says so (their waits are
// bounded, so a block that is not resident after all turns into NaN
// results and a fault count, not a hang)
static int resident_blocks(const void *kernel, int threads, int &fit)
{
    static std::mutex mu;
    static std::map<std::pair<int, const void *>, int> cache;  // (device, kernel) -> blocks
    int dev = 0;
    MRB_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(mu);
    const auto key = std::make_pair(dev, kernel);
    const auto it = cache.find(key);
    if (it != cache.end()) {
        fit = it->second;
        return MRB_OK;
    }
    int per_cu = 0, cus = 0;
    MRB_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    MRB_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0));
    fit = per_cu * cus;
    cache[key] = fit;
    return MRB_OK;
}

template <class T, int BCX>
static int two_pass_tp(mrb_level *L, int variant, hipStream_t s, hipEvent_t *ev)
{
    int rc = MRB_OK;
    const CompactArgs A = make_args(L);
    const ShiftTab sh = shift_tab(L);
    const Half16 h16 = half16_of(L);
    const T *f = (const T *)L->pop[0][L->cur];
    const T *g = (const T *)L->pop[1][L->cur];
    T *fo = (T *)L->pop[0][1 - L->cur];
    T *go = (T *)L->pop[1][1 - L->cur];
    if (variant == kSweepX64Variant) {
        if constexpr (std::is_same<T, double>::value) {
            if (!L->d_sweep_q) return set_error(MRB_E_UNSUPPORTED, "k_sweep_x64: FP64 storage only");
            const int nxb = (A.nx + kSweepW - 1) / kSweepW;
            k_zero_u32<<<(nxb * A.ny + 255) / 256, 256, 0, s>>>(L->d_sweep_flags, nxb * A.ny);
            MRB_HIP(hipGetLastError());
            if (ev) MRB_HIP(hipEventRecord(ev[0], s));
            double *qr = L->d_sweep_q, *qxr = L->d_sweep_qx;
            uint32_t *fl = L->d_sweep_flags, *fa = L->d_sweep_fault;
            const uint32_t *mk = L->d_mask;
            CompactArgs Ac = A;
            BCDev bcc = L->bcdev;
            ShiftTab shc = sh;
            CollideParams Pc = L->cp;
            int fit = 0;
            if ((rc = resident_blocks(reinterpret_cast<const void *>(&k_sweep_x64<BCX>), kSweepW, fit))) return rc;
            if (int64_t(nxb) * A.ny > fit)
                return set_error(MRB_E_UNSUPPORTED, "k_sweep_x64: more blocks than the GPU holds at once");
            k_sweep_x64<BCX><<<dim3(nxb, A.ny), dim3(kSweepW), 0, s>>>(f, g, fo, go, mk, qr, qxr, fl, fa, Ac, bcc, shc,
                                                                      Pc);
            MRB_HIP(hipGetLastError());
            if (ev) MRB_HIP(hipEventRecord(ev[1], s));
            return MRB_OK;
        } else {
            return set_error(MRB_E_UNSUPPORTED, "k_sweep_x64: FP64 storage only");
        }
    }
    if (variant == kSweepVariant || variant == kSweepF32Variant) {
        const bool f32a = variant == kSweepF32Variant;
        if (f32a && !std::is_same<T, float>::value)
            return set_error(MRB_E_UNSUPPORTED, "k_sweep_f32: shifted FP32 storage only");
        if constexpr (!kIsHalf<T>) {
            if (!L->d_sweep_qt || A.nx > kSweepRow)
                return set_error(MRB_E_UNSUPPORTED, "k_sweep: rows of at most 256 cells only");
            auto go_sweep = [&](auto kernel) -> int {
                int fit = 0;
                int rc2 = resident_blocks(reinterpret_cast<const void *>(kernel), kSweepRow, fit);
                if (rc2) return rc2;
                if (A.ny > fit) return set_error(MRB_E_UNSUPPORTED, "k_sweep: more rows than the GPU holds at once");
                k_bump_u32<<<1, 1, 0, s>>>(L->d_sweep_epoch);
                MRB_HIP(hipGetLastError());
                if (ev) MRB_HIP(hipEventRecord(ev[0], s));
                uint64_t *qr = L->d_sweep_qt;
                uint32_t *fl = L->d_sweep_epoch, *fa = L->d_sweep_fault;
                const uint32_t *mk = L->d_mask;
                CompactArgs Ac = A;
                BCDev bcc = L->bcdev;
                ShiftTab shc = sh;
                CollideParams Pc = L->cp;
                DevRest Dc = dev_rest(L);
                kernel<<<dim3(A.ny), dim3(kSweepRow), 0, s>>>(f, g, fo, go, mk, qr, fl, fa, Ac, bcc, shc, Pc, Dc);
                MRB_HIP(hipGetLastError());
                if (ev) MRB_HIP(hipEventRecord(ev[1], s));
                return MRB_OK;
            };
            if (!f32a) return go_sweep(&k_sweep<T, BCX, false>);
            if constexpr (std::is_same<T, float>::value && (BCX & kBcxPressure) == 0)
                return go_sweep(&k_sweep<T, BCX, true>);
            return set_error(MRB_E_UNSUPPORTED, "k_sweep_f32: no PRESSURE faces");
        } else {
            return set_error(MRB_E_UNSUPPORTED, "k_sweep: FP64 and shifted FP32 storage only");
        }
    }
    if (variant == kOnePassVariant) {
        if constexpr (kIsHalf<T>) {
            return set_error(MRB_E_UNSUPPORTED, "k_onepass: FP64 and shifted FP32 storage only");
        } else {
            if (ev) MRB_HIP(hipEventRecord(ev[0], s));
            k_onepass<T, BCX><<<grid_of(A), dim3(64, 4, 1), 0, s>>>(f, g, fo, go, L->d_mask, A, L->bcdev, sh,
                                                                     L->cp);
            MRB_HIP(hipGetLastError());
            if (ev) MRB_HIP(hipEventRecord(ev[1], s));
            return MRB_OK;
        }
    }
    if (variant == kX128Variant || variant == kX256Variant) {
        // the same two passes on 128x2 or 256x1 blocks (two or four waves
        // per row): fewer cache lines shared with the blocks beside in x,
        // which run on other XCDs
        const int bx = variant == kX128Variant ? 128 : 256, by = variant == kX128Variant ? 2 : 1;
        const dim3 grid = grid_of(A, by, bx), block(bx, by, 1);
        if (ev) MRB_HIP(hipEventRecord(ev[0], s));
        if (bx == 128)
            k_qpass_x128<T, BCX><<<grid, block, 0, s>>>(f, g, L->d_mask, L->d_q, A, L->bcdev, sh, h16, L->cp);
        else
            k_qpass_x256<T, BCX><<<grid, block, 0, s>>>(f, g, L->d_mask, L->d_q, A, L->bcdev, sh, h16, L->cp);
        MRB_HIP(hipGetLastError());
        if (ev) MRB_HIP(hipEventRecord(ev[1], s));
        if (bx == 128)
            k_collide_x128<T, BCX><<<grid, block, 0, s>>>(f, g, fo, go, L->d_mask, L->d_q, A, L->bcdev, sh, h16,
                                                           L->cp);
        else
            k_collide_x256<T, BCX><<<grid, block, 0, s>>>(f, g, fo, go, L->d_mask, L->d_q, A, L->bcdev, sh, h16,
                                                           L->cp);
        MRB_HIP(hipGetLastError());
        if (ev) MRB_HIP(hipEventRecord(ev[2], s));
        return MRB_OK;
    }
    if (variant == kRows8Variant || variant == kRows2Variant) {
        // the same two passes on 64x8 or 64x2 blocks
        const int by = variant == kRows8Variant ? 8 : 2;
        const dim3 grid = grid_of(A, by), block(64, by, 1);
        if (ev) MRB_HIP(hipEventRecord(ev[0], s));
        if (by == 8)
            k_qpass_y8<T, BCX><<<grid, block, 0, s>>>(f, g, L->d_mask, L->d_q, A, L->bcdev, sh, h16, L->cp);
        else
            k_qpass_y2<T, BCX><<<grid, block, 0, s>>>(f, g, L->d_mask, L->d_q, A, L->bcdev, sh, h16, L->cp);
        MRB_HIP(hipGetLastError());
        if (ev) MRB_HIP(hipEventRecord(ev[1], s));
        if (by == 8)
            k_collide_y8<T, BCX><<<grid, block, 0, s>>>(f, g, fo, go, L->d_mask, L->d_q, A, L->bcdev, sh, h16,
                                                         L->cp);
        else
            k_collide_y2<T, BCX><<<grid, block, 0, s>>>(f, g, fo, go, L->d_mask, L->d_q, A, L->bcdev, sh, h16,
                                                         L->cp);
        MRB_HIP(hipGetLastError());
        if (ev) MRB_HIP(hipEventRecord(ev[2], s));
        return MRB_OK;
    }
    if (variant == kF32Variant || variant == kF32X256Variant) {
        if constexpr (std::is_same<T, float>::value && (BCX & kBcxPressure) == 0) {
            const DevRest D = dev_rest(L);
            const bool wide = variant == kF32X256Variant;
            const dim3 grid = wide ? grid_of(A, 1, 256) : grid_of(A, 2);
            if (ev) MRB_HIP(hipEventRecord(ev[0], s));
            if (wide)
                k_qpass_x256<T, BCX><<<grid, dim3(256, 1, 1), 0, s>>>(f, g, L->d_mask, L->d_q, A, L->bcdev, sh,
                                                                     h16, L->cp);
            else
                k_qpass_y2<T, BCX><<<grid, dim3(64, 2, 1), 0, s>>>(f, g, L->d_mask, L->d_q, A, L->bcdev, sh, h16,
                                                                  L->cp);
            MRB_HIP(hipGetLastError());
            if (ev) MRB_HIP(hipEventRecord(ev[1], s));
            if (wide)
                k_collide_f32_x256<BCX><<<grid, dim3(256, 1, 1), 0, s>>>(f, g, fo, go, L->d_mask, L->d_q, A,
                                                                        L->bcdev, sh, D, L->cp);
            else
                k_collide_f32<BCX><<<grid, dim3(64, 2, 1), 0, s>>>(f, g, fo, go, L->d_mask, L->d_q, A, L->bcdev,
                                                                  sh, D, L->cp);
            MRB_HIP(hipGetLastError());
            if (ev) MRB_HIP(hipEventRecord(ev[2], s));
            return MRB_OK;
        } else {
            return set_error(MRB_E_UNSUPPORTED,
                             "k_collide_f32: shifted FP32 storage without PRESSURE faces only");
        }
    }
    if (variant == kHalf32Variant || variant == kHalf32X256Variant) {
        if constexpr (kIsHalf<T> && (BCX & kBcxPressure) == 0) {
            if (ev) MRB_HIP(hipEventRecord(ev[0], s));
            if (variant == kHalf32Variant)
                k_qpass<T, BCX><<<grid_of(A), dim3(64, 4, 1), 0, s>>>(f, g, L->d_mask, L->d_q, A, L->bcdev, sh,
                                                                       h16, L->cp);
            else
                k_qpass_x256<T, BCX><<<grid_of(A, 1, 256), dim3(256, 1, 1), 0, s>>>(f, g, L->d_mask, L->d_q, A,
                                                                                   L->bcdev, sh, h16, L->cp);
            MRB_HIP(hipGetLastError());
            if (ev) MRB_HIP(hipEventRecord(ev[1], s));
            if (variant == kHalf32Variant) {
                if constexpr (BCX == 0)
                    k_collide_h32_w4<T><<<grid_of(A), dim3(64, 4, 1), 0, s>>>(f, g, fo, go, L->d_mask, L->d_q, A,
                                                                               L->bcdev, sh, h16, L->cp);
                else
                    k_collide_h32<T, BCX><<<grid_of(A), dim3(64, 4, 1), 0, s>>>(f, g, fo, go, L->d_mask, L->d_q,
                                                                                 A, L->bcdev, sh, h16, L->cp);
            } else {
                if constexpr (BCX == 0)
                    k_collide_h32_x256_w4<T><<<grid_of(A, 1, 256), dim3(256, 1, 1), 0, s>>>(
                        f, g, fo, go, L->d_mask, L->d_q, A, L->bcdev, sh, h16, L->cp);
                else
                    k_collide_h32_x256<T, BCX><<<grid_of(A, 1, 256), dim3(256, 1, 1), 0, s>>>(
                        f, g, fo, go, L->d_mask, L->d_q, A, L->bcdev, sh, h16, L->cp);
            }
            MRB_HIP(hipGetLastError());
            if (ev) MRB_HIP(hipEventRecord(ev[2], s));
            return MRB_OK;
        } else {
            return set_error(MRB_E_UNSUPPORTED, "k_collide_h32: scaled FP16 storage without PRESSURE faces only");
        }
    }
    if (ev) MRB_HIP(hipEventRecord(ev[0], s));
    k_qpass<T, BCX><<<grid_of(A), dim3(64, 4, 1), 0, s>>>(f, g, L->d_mask, L->d_q, A, L->bcdev, sh, h16, L->cp);
    MRB_HIP(hipGetLastError());
    if (ev) MRB_HIP(hipEventRecord(ev[1], s));
    if (variant == 1)
        k_collide_w3<T, BCX><<<grid_of(A), dim3(64, 4, 1), 0, s>>>(f, g, fo, go, L->d_mask, L->d_q, A, L->bcdev,
                                                                    sh, h16, L->cp);
    else
        k_collide<T, BCX><<<grid_of(A), dim3(64, 4, 1), 0, s>>>(f, g, fo, go, L->d_mask, L->d_q, A, L->bcdev, sh,
                                                                 h16, L->cp);
    MRB_HIP(hipGetLastError());
    if (ev) MRB_HIP(hipEventRecord(ev[2], s));
    return MRB_OK;
}

// the inlet-profile lookups and the pressure ghosts are compiled only into
// the kernels of levels that have them
static int bcx_of(const mrb_level *L)
{
    return (L->bcdev.prof.kind ? kBcxProfile : 0) | (L->bcdev.pdir >= 0 ? kBcxPressure : 0);
}

template <class T>
static int two_pass_t(mrb_level *L, int variant, hipStream_t s, hipEvent_t *ev)
{
    switch (bcx_of(L)) {
    case kBcxProfile: return two_pass_tp<T, kBcxProfile>(L, variant, s, ev);
    case kBcxPressure: return two_pass_tp<T, kBcxPressure>(L, variant, s, ev);
    case kBcxProfile | kBcxPressure: return two_pass_tp<T, kBcxProfile | kBcxPressure>(L, variant, s, ev);
    default: return two_pass_tp<T, 0>(L, variant, s, ev);
    }
}

static int two_pass(mrb_level *L, int variant, hipStream_t s, hipEvent_t *ev)
{
    if (L->storage == MRB_STORAGE_F64) return two_pass_t<double>(L, variant, s, ev);
    if (L->storage == MRB_STORAGE_F32_SHIFTED) return two_pass_t<float>(L, variant, s, ev);
    if (L->storage == MRB_STORAGE_F16_SCALED) return two_pass_t<_Float16>(L, variant, s, ev);
    return set_error(MRB_E_UNSUPPORTED, "storage not supported by this kernel");
}

const char *const kTwoPassNames[kNumTwoPassVariants] = {"k_qpass+k_collide", "k_qpass+k_collide_w3", "k_onepass",
                                                           "k_qpass_y8+k_collide_y8", "k_qpass_y2+k_collide_y2",
                                                           "k_qpass+k_collide_h32",
                                                           "k_qpass_x128+k_collide_x128",
                                                           "k_qpass_x256+k_collide_x256",
                                                           "k_qpass_x256+k_collide_h32_x256",
                                                           "k_qpass_y2+k_collide_f32",
                                                           "k_qpass_x256+k_collide_f32_x256",
                                                           "k_sweep", "k_sweep_x64", "k_sweep_f32"};

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
        nslots = (L->hot == kOnePassVariant || L->hot == kSweepVariant || L->hot == kSweepX64Variant ||
                  L->hot == kSweepF32Variant)
                     ? 1
                     : 2;
    } else {
        if (events) MRB_HIP(hipEventRecord(L->ev[0], s));
        rc = fused_launch(L, L->hot - kNumTwoPassVariants, s);
        if (events && !rc) MRB_HIP(hipEventRecord(L->ev[1], s));
        nslots = 1;
    }
    if (rc) return rc;
    L->cur = 1 - L->cur;
    L->prev_valid = true;
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

template <class T, int BCX>
static void launch_stream_only(mrb_level *L, const CompactArgs &A, const ShiftTab &sh, hipStream_t s)
{
    const T *f = (const T *)L->pop[0][L->cur], *g = (const T *)L->pop[1][L->cur];
    T *fo = (T *)L->pop[0][1 - L->cur], *go = (T *)L->pop[1][1 - L->cur];
    k_stream_only<T, BCX><<<grid_of(A), dim3(64, 4, 1), 0, s>>>(f, g, fo, go, L->d_mask, A, L->bcdev, sh);
}

template <class T>
static void stream_only_t(mrb_level *L, const CompactArgs &A, const ShiftTab &sh, hipStream_t s)
{
    switch (bcx_of(L)) {
    case kBcxProfile: launch_stream_only<T, kBcxProfile>(L, A, sh, s); break;
    case kBcxPressure: launch_stream_only<T, kBcxPressure>(L, A, sh, s); break;
    case kBcxProfile | kBcxPressure: launch_stream_only<T, kBcxProfile | kBcxPressure>(L, A, sh, s); break;
    default: launch_stream_only<T, 0>(L, A, sh, s); break;
    }
}

int compact_stream_only(mrb_level *L, hipStream_t s)
{
    const CompactArgs A = make_args(L);
    const ShiftTab sh = shift_tab(L);
    if (L->storage == MRB_STORAGE_F64)
        stream_only_t<double>(L, A, sh, s);
    else if (L->storage == MRB_STORAGE_F32_SHIFTED)
        stream_only_t<float>(L, A, sh, s);
    else
        return set_error(MRB_E_UNSUPPORTED, "storage not supported by this kernel");
    MRB_HIP(hipGetLastError());
    L->cur = 1 - L->cur;
    L->prev_valid = false;  // a stream without a collide
    return MRB_OK;
}

const char *compact_kernel_name(const mrb_level *L)
{
    return compact_variant_name(L->hot);
}

// ---------------------------------------------------------------------------
// FP16 storage: FP64 populations (component stride cs) <-> the scaled FP16
// pair layout of one lattice of the current buffer
// ---------------------------------------------------------------------------
namespace {

__device__ __forceinline__ float block_max(float m)
{
    __shared__ float sm[4];
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
    const int t = threadIdx.y * blockDim.x + threadIdx.x;
    if ((t & 63) == 0) sm[t / 64] = m;
    __syncthreads();
    if (t == 0)
        for (int i = 1; i < int(blockDim.x * blockDim.y) / 64; ++i) m = fmaxf(m, sm[i]);
    return m;
}

__global__ void k_clear_lattice(uint32_t *__restrict__ a, int n, int lattice)
{
    const int i = int(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) a[2 * i + lattice] = 0u;
}

// largest |population - shift| over the fluid cells of each scale box
__global__ void __launch_bounds__(256) k_amax16(
    const double *__restrict__ src, int64_t cs, const double *__restrict__ shift, const int32_t *__restrict__ isf,
    uint32_t *__restrict__ amax, int lattice, CompactArgs A)
{
    int x, y, z;
    int64_t cell;
    float m = 0.0f;
    if (cell_coords(A, x, y, z, cell) && isf[cell] == 1)
        for (int q = 0; q < NQ; ++q) m = fmaxf(m, float(fabs(src[q * cs + cell] - shift[q])));
    m = block_max(m);
    if (threadIdx.x == 0 && threadIdx.y == 0)
        atomicMax(&amax[2 * (int(blockIdx.z) / kScaleSlab) + lattice], __float_as_uint(m));
}

__global__ void __launch_bounds__(256) k_pack16(
    const double *__restrict__ src, int64_t cs, const double *__restrict__ shift, const int32_t *__restrict__ isf,
    const uint32_t *__restrict__ amax, int *__restrict__ e, _Float16 *__restrict__ dst, uint32_t *__restrict__ ovf,
    int lattice, CompactArgs A)
{
    const int b = int(blockIdx.z) / kScaleSlab;
    const int ee = f16_exponent(__uint_as_float(amax[2 * b + lattice]));
    if (threadIdx.x == 0 && threadIdx.y == 0) e[2 * b + lattice] = ee;
    int x, y, z;
    int64_t cell;
    if (!cell_coords(A, x, y, z, cell)) return;
    const bool solid = isf[cell] != 1;
    const float scale = ldexpf(1.0f, ee);
    float mx = 0.0f;
    for (int q = 0; q < NQ; ++q)
        dst[(int64_t(q) * A.N + cell) * 2] = solid ? _Float16(0.0f) : to_half(src[q * cs + cell], shift[q], scale, mx);
    if (ldexpf(mx, ee) > 65504.0f) atomicAdd(ovf, 1u);
}

__global__ void __launch_bounds__(256) k_unpack16(
    const _Float16 *__restrict__ src, const double *__restrict__ shift, const int *__restrict__ e,
    const int32_t *__restrict__ isf, double *__restrict__ dst, int lattice, CompactArgs A)
{
    int x, y, z;
    int64_t cell;
    if (!cell_coords(A, x, y, z, cell)) return;
    const bool solid = isf[cell] != 1;
    const float inv = ldexpf(1.0f, -e[2 * (z / kScaleSlab) + lattice]);
    for (int q = 0; q < NQ; ++q) {
        const float h = float(src[(int64_t(q) * A.N + cell) * 2]);
        dst[int64_t(q) * A.N + cell] = solid ? -1.0 : double(h * inv) + shift[q];
    }
}

// density, velocity and temperature of the stored state per cell (the
// macrodata of LBM::f_to_macrodata, LBM.cpp:810-906, that the precision
// comparisons use); solid cells get zeros
template <class T>
__global__ void __launch_bounds__(256) k_macro(
    const T *__restrict__ f, const T *__restrict__ g, const uint32_t *__restrict__ mask, CompactArgs A,
    ShiftTab sh, const int *__restrict__ e16, CollideParams P, double *__restrict__ out)
{
    int x, y, z;
    int64_t cell;
    if (!cell_coords(A, x, y, z, cell)) return;
    double o[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (mask[cell] & 1u) {
        float sf = 1.0f, sg = 1.0f;
        if constexpr (kIsHalf<T>) {
            sf = ldexpf(1.0f, -e16[2 * (z / kScaleSlab)]);
            sg = ldexpf(1.0f, -e16[2 * (z / kScaleSlab) + 1]);
        }
        double fv[NQ], gv[NQ];
        static_for<NQ>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            const int64_t e = int64_t(q) * A.cs + cell * kCellStride<T>;
            if constexpr (kIsHalf<T>) {
                fv[q] = double(float(f[e]) * sf) + sh.f(q);
                gv[q] = double(float(g[e]) * sg) + sh.g(q);
            } else {
                fv[q] = widen(f[e], sh.f(q));
                gv[q] = widen(g[e], sh.g(q));
            }
        });
        double rho, jx, jy, jz, E;
        accumulate_q_moments(fv, gv, rho, jx, jy, jz, E);
        const Macro M = macro_from(rho, jx, jy, jz, E, P);
        o[0] = rho;
        o[1] = M.u;
        o[2] = M.v;
        o[3] = M.w;
        o[4] = M.T;
    }
    for (int k = 0; k < 5; ++k) out[k * A.N + cell] = o[k];
}

// vorticity (x, y, z, magnitude) of the velocity in mac (fields 1..3) with
// the reference's gradient(): central inside, one-sided where the neighbour
// lies outside the domain box or is solid (LBM::compute_derived,
// LBM.cpp:909-955); 0 in solid cells
__global__ void __launch_bounds__(256) k_vorticity(
    const double *__restrict__ mac, const uint32_t *__restrict__ mask, CompactArgs A, CollideParams P,
    double *__restrict__ out)
{
    int x, y, z;
    int64_t cell;
    if (!cell_coords(A, x, y, z, cell)) return;
    const uint32_t m = mask[cell];
    double w[4] = {0.0, 0.0, 0.0, 0.0};
    if (m & 1u) {
        // bits of the stream mask: 2/1 = x+1/x-1 fluid, 4/3 = y, 6/5 = z
        const bool ok[3][2] = {{x < A.nx - 1 && ((m >> 2) & 1u), x > 0 && ((m >> 1) & 1u)},
                               {y < A.ny - 1 && ((m >> 4) & 1u), y > 0 && ((m >> 3) & 1u)},
                               {z < A.nz - 1 && ((m >> 6) & 1u), z > 0 && ((m >> 5) & 1u)}};
        const int64_t st[3] = {1, A.nx, int64_t(A.nx) * A.ny};
        auto grad = [&](int d, int comp) {
            const double *u = mac + comp * A.N;
            const bool p = ok[d][0], mm = ok[d][1];
            return one_sided_gradient(p, mm, p ? u[cell + st[d]] : 0.0, u[cell], mm ? u[cell - st[d]] : 0.0,
                                      P.inv_dx[d]);
        };
        const double vx = grad(0, 2), wx = grad(0, 3);
        const double uy = grad(1, 1), wy = grad(1, 3);
        const double uz = grad(2, 1), vz = grad(2, 2);
        w[0] = wy - vz;
        w[1] = uz - wx;
        w[2] = vx - uy;
        w[3] = sqrt((wy - vz) * (wy - vz) + (uz - wx) * (uz - wx) + (vx - uy) * (vx - uy));
    }
    for (int k = 0; k < 4; ++k) out[k * A.N + cell] = w[k];
}

// LBM::f_to_macrodata (LBM.cpp:810-906) for the plotfile: the 19 macrodata
// components (rho, u, v, w, |u|, 2 rho e, QCorr, the pressure tensor, the
// heat flux, T; 0 in solid cells) of the post-stream state of the last step,
// which the reference's m_macrodata holds after the collide: the pull of
// the buffer that held the state before the step (PULL), or, before any
// step, of the stored populations themselves
template <class T, int BCX, bool PULL>
__global__ void __launch_bounds__(256) k_plot_md(
    const T *__restrict__ f, const T *__restrict__ g, const uint32_t *__restrict__ mask, CompactArgs A, BCDev bc,
    ShiftTab sh, Half16 h16, CollideParams P, double *__restrict__ out)
{
    int x, y, z;
    int64_t cell;
    if (!cell_coords(A, x, y, z, cell)) return;
    const uint32_t m = mask[cell];
    double o[19];
    for (int k = 0; k < 19; ++k) o[k] = 0.0;
    if (m & 1u) {
        double fv[NQ], gv[NQ];
        if constexpr (PULL) {
            pull_cell<BCX>(f, g, A, bc, sh, h16, x, y, z, cell, m, fv, gv);
        } else {
            float sf = 1.0f, sg = 1.0f;
            if constexpr (kIsHalf<T>) {
                sf = ldexpf(1.0f, -h16.e_cur[2 * (z / kScaleSlab)]);
                sg = ldexpf(1.0f, -h16.e_cur[2 * (z / kScaleSlab) + 1]);
            }
            static_for<NQ>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                const int64_t e = int64_t(q) * A.cs + cell * kCellStride<T>;
                if constexpr (kIsHalf<T>) {
                    fv[q] = double(float(f[e]) * sf) + sh.f(q);
                    gv[q] = double(float(g[e]) * sg) + sh.g(q);
                } else {
                    fv[q] = widen(f[e], sh.f(q));
                    gv[q] = widen(g[e], sh.g(q));
                }
            });
        }
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
    for (int k = 0; k < 19; ++k) out[k * A.N + cell] = o[k];
}

// LBM::compute_q_corrections (LBM.cpp:959-991): dQCorr x, y, z, the
// gradient() of the macrodata's QCorr components (6, 7, 8) along x, y, z;
// 0 in solid cells
__global__ void __launch_bounds__(256) k_plot_dq(
    const double *__restrict__ md, const uint32_t *__restrict__ mask, CompactArgs A, CollideParams P,
    double *__restrict__ out)
{
    int x, y, z;
    int64_t cell;
    if (!cell_coords(A, x, y, z, cell)) return;
    const uint32_t m = mask[cell];
    double d[3] = {0.0, 0.0, 0.0};
    if (m & 1u) {
        const bool ok[3][2] = {{x < A.nx - 1 && ((m >> 2) & 1u), x > 0 && ((m >> 1) & 1u)},
                               {y < A.ny - 1 && ((m >> 4) & 1u), y > 0 && ((m >> 3) & 1u)},
                               {z < A.nz - 1 && ((m >> 6) & 1u), z > 0 && ((m >> 5) & 1u)}};
        const int64_t st[3] = {1, A.nx, int64_t(A.nx) * A.ny};
        for (int k = 0; k < 3; ++k) {
            const double *u = md + (6 + k) * A.N;
            const bool p = ok[k][0], mm = ok[k][1];
            d[k] = one_sided_gradient(p, mm, p ? u[cell + st[k]] : 0.0, u[cell], mm ? u[cell - st[k]] : 0.0,
                                      P.inv_dx[k]);
        }
    }
    for (int k = 0; k < 3; ++k) out[k * A.N + cell] = d[k];
}

// refinement tags (AMReX AMRErrorTag criteria): 1 where the field is
// greater / less than the value, or where the largest absolute difference
// to a face neighbour (wrapped on periodic axes, skipped beyond a
// non-periodic face) is greater than it
__global__ void __launch_bounds__(256) k_tag(
    const double *__restrict__ fld, CompactArgs A, int criterion, double value, uint8_t *__restrict__ tags)
{
    int x, y, z;
    int64_t cell;
    if (!cell_coords(A, x, y, z, cell)) return;
    const double v = fld[cell];
    bool t;
    if (criterion == MRB_TAG_GREATER) {
        t = v > value;
    } else if (criterion == MRB_TAG_LESS) {
        t = v < value;
    } else {
        const int pos[3] = {x, y, z}, n[3] = {A.nx, A.ny, A.nz};
        const int64_t st[3] = {1, A.nx, int64_t(A.nx) * A.ny};
        double dmax = 0.0;
        for (int d = 0; d < 3; ++d) {
            for (int sgn = -1; sgn <= 1; sgn += 2) {
                int p = pos[d] + sgn;
                bool in = p >= 0 && p < n[d];
                if (!in && A.per[d]) {
                    p = p < 0 ? p + n[d] : p - n[d];
                    in = true;
                }
                if (in) dmax = fmax(dmax, fabs(fld[cell + int64_t(p - pos[d]) * st[d]] - v));
            }
        }
        t = dmax > value;
    }
    tags[cell] = t ? 1 : 0;
}

}  // namespace

int compact_vorticity(mrb_level *L, double *out, hipStream_t s)
{
    const int64_t N = L->total_cells;
    if (!L->d_mac) MRB_HIP(hipMalloc(&L->d_mac, sizeof(double) * 5 * size_t(N)));
    const int rc = compact_macro(L, L->d_mac, s);
    if (rc) return rc;
    const CompactArgs A = make_args(L);
    k_vorticity<<<grid_of(A), dim3(64, 4, 1), 0, s>>>(L->d_mac, L->d_mask, A, L->cp, out);
    MRB_HIP(hipGetLastError());
    return MRB_OK;
}

template <class T, int BCX>
static void launch_plot_md(mrb_level *L, const CompactArgs &A, const ShiftTab &sh, double *md, hipStream_t s)
{
    // the state the macrodata are taken from: the buffer before the last
    // step (pulled), or the current one before any step
    const int src = L->prev_valid ? 1 - L->cur : L->cur;
    Half16 h{};
    if (L->storage == MRB_STORAGE_F16_SCALED) h.e_cur = L->d_e16[src];
    const T *f = (const T *)L->pop[0][src], *g = (const T *)L->pop[1][src];
    if (L->prev_valid)
        k_plot_md<T, BCX, true><<<grid_of(A), dim3(64, 4, 1), 0, s>>>(f, g, L->d_mask, A, L->bcdev, sh, h, L->cp,
                                                                      md);
    else
        k_plot_md<T, 0, false><<<grid_of(A), dim3(64, 4, 1), 0, s>>>(f, g, L->d_mask, A, L->bcdev, sh, h, L->cp,
                                                                     md);
}

template <class T>
static void plot_md_t(mrb_level *L, const CompactArgs &A, const ShiftTab &sh, double *md, hipStream_t s)
{
    switch (bcx_of(L)) {
    case kBcxProfile: launch_plot_md<T, kBcxProfile>(L, A, sh, md, s); break;
    case kBcxPressure: launch_plot_md<T, kBcxPressure>(L, A, sh, md, s); break;
    case kBcxProfile | kBcxPressure: launch_plot_md<T, kBcxProfile | kBcxPressure>(L, A, sh, md, s); break;
    default: launch_plot_md<T, 0>(L, A, sh, md, s); break;
    }
}

int compact_plot_fields(mrb_level *L, double *md, double *der, hipStream_t s)
{
    const CompactArgs A = make_args(L);
    const ShiftTab sh = shift_tab(L);
    if (L->storage == MRB_STORAGE_F64)
        plot_md_t<double>(L, A, sh, md, s);
    else if (L->storage == MRB_STORAGE_F32_SHIFTED)
        plot_md_t<float>(L, A, sh, md, s);
    else
        plot_md_t<_Float16>(L, A, sh, md, s);
    MRB_HIP(hipGetLastError());
    // LBM::compute_derived (LBM.cpp:909-955) on the macrodata velocities
    k_vorticity<<<grid_of(A), dim3(64, 4, 1), 0, s>>>(md, L->d_mask, A, L->cp, der);
    MRB_HIP(hipGetLastError());
    k_plot_dq<<<grid_of(A), dim3(64, 4, 1), 0, s>>>(md, L->d_mask, A, L->cp, der + 4 * A.N);
    MRB_HIP(hipGetLastError());
    return MRB_OK;
}

int compact_tag(const mrb_level *L, const double *field, int criterion, double value, uint8_t *tags, hipStream_t s)
{
    const CompactArgs A = make_args(L);
    k_tag<<<grid_of(A), dim3(64, 4, 1), 0, s>>>(field, A, criterion, value, tags);
    MRB_HIP(hipGetLastError());
    return MRB_OK;
}

int compact_macro(const mrb_level *L, double *out, hipStream_t s)
{
    const CompactArgs A = make_args(L);
    const ShiftTab sh = shift_tab(L);
    const dim3 g = grid_of(A), b(64, 4, 1);
    if (L->storage == MRB_STORAGE_F64)
        k_macro<double><<<g, b, 0, s>>>((const double *)L->pop[0][L->cur], (const double *)L->pop[1][L->cur],
                                        L->d_mask, A, sh, nullptr, L->cp, out);
    else if (L->storage == MRB_STORAGE_F32_SHIFTED)
        k_macro<float><<<g, b, 0, s>>>((const float *)L->pop[0][L->cur], (const float *)L->pop[1][L->cur],
                                       L->d_mask, A, sh, nullptr, L->cp, out);
    else
        k_macro<_Float16><<<g, b, 0, s>>>((const _Float16 *)L->pop[0][L->cur],
                                          (const _Float16 *)L->pop[1][L->cur], L->d_mask, A, sh,
                                          L->d_e16[L->cur], L->cp, out);
    MRB_HIP(hipGetLastError());
    return MRB_OK;
}

int f16_pack(mrb_level *L, int lattice, const double *src, int64_t src_cs, hipStream_t s)
{
    if (L->storage != MRB_STORAGE_F16_SCALED) return set_error(MRB_E_STATE, "not an FP16 level");
    const CompactArgs A = make_args(L);
    uint32_t *amax = L->d_amax16[L->cur];
    const double *shift = L->d_small + lattice * NQ;
    k_clear_lattice<<<(L->nsb16 + 255) / 256, 256, 0, s>>>(amax, L->nsb16, lattice);
    MRB_HIP(hipGetLastError());
    k_amax16<<<grid_of(A), dim3(64, 4, 1), 0, s>>>(src, src_cs, shift, L->d_isf, amax, lattice, A);
    MRB_HIP(hipGetLastError());
    k_pack16<<<grid_of(A), dim3(64, 4, 1), 0, s>>>(src, src_cs, shift, L->d_isf, amax, L->d_e16[L->cur],
                                                    static_cast<_Float16 *>(L->pop[lattice][L->cur]), L->d_ovf16,
                                                    lattice, A);
    MRB_HIP(hipGetLastError());
    return MRB_OK;
}

int f16_unpack(const mrb_level *L, int lattice, double *dst, hipStream_t s)
{
    if (L->storage != MRB_STORAGE_F16_SCALED) return set_error(MRB_E_STATE, "not an FP16 level");
    const CompactArgs A = make_args(L);
    k_unpack16<<<grid_of(A), dim3(64, 4, 1), 0, s>>>(static_cast<const _Float16 *>(L->pop[lattice][L->cur]),
                                                      L->d_small + lattice * NQ, L->d_e16[L->cur], L->d_isf, dst,
                                                      lattice, A);
    MRB_HIP(hipGetLastError());
    return MRB_OK;
}

}  // namespace mrb
