// Level object, data movement, flags, initial state, and the reference-layout
// (MRB_LAYOUT_FAB) stages of the hot path: ghost fill, stream, collide.
//
// The FAB-layout kernels restate the reference stages one to one so that
// every stage can be checked bit for bit against the reference CPU path;
// the fused hot path lives in mrb_compact.hip.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <memory>
#include <thread>

#include <string>

#include "mrb_internal.h"

namespace mrb {

static thread_local std::string g_err;

int set_error(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

int hip_fail(hipError_t e, const char *what)
{
    return set_error(MRB_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

namespace {

constexpr int kBlock = 256;

inline unsigned nblocks(int64_t n) { return unsigned((n + kBlock - 1) / kBlock); }

// ---------------------------------------------------------------------------
// small device helpers
// ---------------------------------------------------------------------------

// global fab cell -> (box, local index) by binary search on the prefix sums
__device__ __forceinline__ int find_box(const int64_t *__restrict__ base, int nbox, int64_t c)
{
    int lo = 0, hi = nbox - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (base[mid] <= c)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ void local_to_ijk(const BoxDev &B, int64_t loc, int &i, int &j, int &k)
{
    const int64_t nxy = int64_t(B.n[0]) * B.n[1];
    k = int(loc / nxy);
    const int64_t r = loc - int64_t(k) * nxy;
    j = int(r / B.n[0]);
    i = int(r - int64_t(j) * B.n[0]);
    i += B.lo[0];
    j += B.lo[1];
    k += B.lo[2];
}

__device__ __forceinline__ bool in_fab(const BoxDev &B, int i, int j, int k)
{
    return i >= B.lo[0] && i < B.lo[0] + B.n[0] && j >= B.lo[1] && j < B.lo[1] + B.n[1] &&
           k >= B.lo[2] && k < B.lo[2] + B.n[2];
}

__device__ __forceinline__ bool in_valid(const BoxDev &B, int i, int j, int k)
{
    return i >= B.vlo[0] && i <= B.vhi[0] && j >= B.vlo[1] && j <= B.vhi[1] && k >= B.vlo[2] &&
           k <= B.vhi[2];
}

__device__ __forceinline__ int64_t loc_of(const BoxDev &B, int i, int j, int k)
{
    return (int64_t(k - B.lo[2]) * B.n[1] + (j - B.lo[1])) * B.n[0] + (i - B.lo[0]);
}

// ---------------------------------------------------------------------------
// FillBoundary: ghost <- owner valid cell, all components
// ---------------------------------------------------------------------------
template <class T, int NC>
__global__ void k_fill_map(
    T *__restrict__ arr, const BoxDev *__restrict__ boxes, const GhostPair *__restrict__ map, int64_t nmap)
{
    const int64_t t = xcd_block() * blockDim.x + threadIdx.x;
    if (t >= nmap) return;
    const GhostPair p = map[t];
    const int64_t dn = boxes[p.dst_box].ncell, sn = boxes[p.src_box].ncell;
    T *dst = arr + boxes[p.dst_box].base * NC + p.dst_local;
    const T *src = arr + boxes[p.src_box].base * NC + p.src_local;
    // all NC loads in flight before the stores
    T v[NC];
    static_for<NC>([&](auto cc) { v[decltype(cc)::value] = src[decltype(cc)::value * sn]; });
    static_for<NC>([&](auto cc) { dst[decltype(cc)::value * dn] = v[decltype(cc)::value]; });
}

// ---------------------------------------------------------------------------
// LBM::stream (LBM.cpp:558-604) as a pull over every fab cell y:
//   y fluid and y-c_p inside the fab:  fluid(y-c_p) ? f_p(y-c_p) : f_pbar(y)
//   otherwise:                         -1
// which is exactly the value the reference's push leaves in its -1
// pre-filled temporary (each (y,p) has at most one writer).
// ---------------------------------------------------------------------------
__global__ void k_stream_fab(
    const double *__restrict__ fin, double *__restrict__ fout, const int32_t *__restrict__ isf,
    const BoxDev *__restrict__ boxes, const int64_t *__restrict__ base, int nbox, int64_t total)
{
    const int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= total) return;
    const int b = find_box(base, nbox, c);
    const BoxDev B = boxes[b];
    const int64_t loc = c - B.base;
    int i, j, k;
    local_to_ijk(B, loc, i, j, k);
    const double *f = fin + B.base * NQ;
    double *fo = fout + B.base * NQ;
    const int32_t *fl = isf + B.base * 2;
    const bool fluid = fl[loc] == 1;
    static_for<NQ>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        const int a = i - ev(q, 0), bb = j - ev(q, 1), e = k - ev(q, 2);
        double v = -1.0;
        if (fluid && in_fab(B, a, bb, e)) {
            const int64_t sl = loc_of(B, a, bb, e);
            v = fl[sl] != 0 ? f[q * B.ncell + sl] : f[kBounce[q] * B.ncell + loc];
        }
        fo[q * B.ncell + loc] = v;
    });
}

// pressure_bc_f (BC.H:222-271) for the ghost whose direction q feeds the
// fluid cell ivn = (n0, n1, n2): twice the populations that cell receives
// travelling out through the face plus those it receives tangentially set
// the normal velocity of the ghost's equilibrium at the boundary density.
// A read beyond a (NOSLIP) tangential face, data(ivn + c_qq, qqbar), finds
// the bounce prefill of FillPatchOps.H:92-110, which the wall's own fill
// repeats: the population qq of ivn itself.  That prefill is not run here,
// so the value is read from ivn.
__device__ __noinline__ double pressure_f(const double *__restrict__ data, const BoxDev &B, int n0, int n1, int n2,
                                          int idir, int ndir, const int domlo[3], const int domhi[3], const int *type,
                                          double rho_bc, double rt, int q)
{
    const int64_t own = loc_of(B, n0, n1, n2);
    double rho_out = 0.0, rho_tan = 0.0;
    for (int qq = 0; qq < NQ; ++qq) {
        const int bqq = kBounce[qq];
        const int cin = kEV[bqq][idir];
        const int a[3] = {n0 + kEV[qq][0], n1 + kEV[qq][1], n2 + kEV[qq][2]};
        bool wall = false;
        for (int d = 0; d < 3; ++d)
            wall = wall || ((a[d] < domlo[d] || a[d] > domhi[d]) && type[d] != MRB_BC_PERIODIC);
        const int64_t r = wall ? int64_t(qq) * B.ncell + own : int64_t(bqq) * B.ncell + loc_of(B, a[0], a[1], a[2]);
        if (cin == -ndir) {
            rho_out += 2.0 * data[r];
        } else if (cin == 0) {
            rho_tan += 1.0 * data[r];
        }
    }
    double vel[3] = {0.0, 0.0, 0.0};
    vel[idir] = ndir * (1.0 - (rho_out + rho_tan) / rho_bc);
    return set_equilibrium_value(rho_bc, vel, rt, q);
}

// ---------------------------------------------------------------------------
// Physical boundary functor BCFill::operator() (BC.H:344-471) for the listed
// fab cells (outside the domain in a non-periodic direction).
// ---------------------------------------------------------------------------
__global__ void k_physbc_fab(
    double *__restrict__ arr, const BoxDev *__restrict__ boxes, const int64_t *__restrict__ base,
    int nbox, const int64_t *__restrict__ cells, int64_t ncells, int dlo0, int dlo1, int dlo2,
    int dhi0, int dhi1, int dhi2, BCDev bc, int is_energy)
{
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= ncells) return;
    const int64_t c = cells[t];
    const int b = find_box(base, nbox, c);
    const BoxDev B = boxes[b];
    const int64_t loc = c - B.base;
    int iv[3];
    local_to_ijk(B, loc, iv[0], iv[1], iv[2]);
    double *data = arr + B.base * NQ;
    const int domlo[3] = {dlo0, dlo1, dlo2};
    const int domhi[3] = {dhi0, dhi1, dhi2};
    const int normal_dirs[2] = {1, -1};
    for (int idir = 0; idir < 3; ++idir) {
        for (int lohi = 0; lohi < 2; ++lohi) {
            const int normal_dir = normal_dirs[lohi];
            const int bc_idx = idir + lohi * 3;
            if (!((lohi == 0 && iv[idir] < domlo[idir]) || (lohi == 1 && iv[idir] > domhi[idir])))
                continue;
            const int t_bc = bc.type[bc_idx];
            static_for<NQ>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                const int n0 = iv[0] + ev(q, 0), n1 = iv[1] + ev(q, 1), n2 = iv[2] + ev(q, 2);
                double *dq = data + q * B.ncell + loc;
                if (in_valid(B, n0, n1, n2)) {
                    if (t_bc == MRB_BC_NOSLIP) {
                        *dq = data[kBounce[q] * B.ncell + loc_of(B, n0, n1, n2)];
                    } else if (t_bc == MRB_BC_VELOCITY) {
                        if (bc.prof.kind) {
                            // Channel / Parabolic: the velocity of this ghost cell
                            double v[3];
                            profile_velocity(bc.prof, iv, v);
                            const VelProfile &P = bc.prof;
                            *dq = is_energy ? thermal_g_equilibrium(P.rho, v, P.T, P.R, P.gamma, q)
                                            : set_equilibrium_value(P.rho, v, P.R * P.T, q);
                        } else {
                            *dq = is_energy ? bc.gval[q] : bc.fval[q];
                        }
                    } else if (t_bc == MRB_BC_PRESSURE) {
                        *dq = is_energy ? bc.pgval[q]
                                        : pressure_f(data, B, n0, n1, n2, idir, normal_dir, domlo, domhi,
                                                     bc.type, bc.prho, bc.prt, q);
                    } else if (t_bc == MRB_BC_OUTFLOW) {
                        int s[3] = {iv[0], iv[1], iv[2]};
                        s[idir] += normal_dir;
                        *dq = data[q * B.ncell + loc_of(B, s[0], s[1], s[2])];
                    } else if (t_bc == MRB_BC_SLIPX) {
                        *dq = data[reflect(q, 0) * B.ncell + loc_of(B, n0, n1, n2)];
                    } else if (t_bc == MRB_BC_SLIPY) {
                        *dq = data[reflect(q, 1) * B.ncell + loc_of(B, n0, n1, n2)];
                    } else if (t_bc == MRB_BC_SLIPZ) {
                        *dq = data[reflect(q, 2) * B.ncell + loc_of(B, n0, n1, n2)];
                    }
                } else {
                    *dq = -1.0;
                }
            });
        }
    }
}

// ---------------------------------------------------------------------------
// LBM::f_to_macrodata (LBM.cpp:810-906) over valid + 1 ghost
// ---------------------------------------------------------------------------
__global__ void k_moments_fab(
    const double *__restrict__ fin, const double *__restrict__ gin,
    const int32_t *__restrict__ isf, double *__restrict__ md, const BoxDev *__restrict__ boxes,
    const int64_t *__restrict__ base, int nbox, int64_t total, CollideParams P)
{
    const int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= total) return;
    const int b = find_box(base, nbox, c);
    const BoxDev B = boxes[b];
    const int64_t loc = c - B.base;
    int i, j, k;
    local_to_ijk(B, loc, i, j, k);
    if (i < B.vlo[0] - 1 || i > B.vhi[0] + 1 || j < B.vlo[1] - 1 || j > B.vhi[1] + 1 ||
        k < B.vlo[2] - 1 || k > B.vhi[2] + 1)
        return;
    if (isf[B.base * 2 + loc] != 1) return;
    double fv[NQ], gv[NQ];
    const double *f = fin + B.base * NQ + loc;
    const double *g = gin + B.base * NQ + loc;
    for (int q = 0; q < NQ; ++q) {
        fv[q] = f[q * B.ncell];
        gv[q] = g[q * B.ncell];
    }
    Moments m;
    accumulate_moments(fv, gv, m);
    const Macro M = macro_from(m.rho, m.jx, m.jy, m.jz, m.E, P);
    double *o = md + B.base * 19 + loc;
    const int64_t s = B.ncell;
    o[0 * s] = m.rho;
    o[1 * s] = M.u;
    o[2 * s] = M.v;
    o[3 * s] = M.w;
    o[4 * s] = sqrt(M.u * M.u + M.v * M.v + M.w * M.w);
    o[5 * s] = m.E;
    o[6 * s] = M.Qx;
    o[7 * s] = M.Qy;
    o[8 * s] = M.Qz;
    o[9 * s] = m.pxx;
    o[10 * s] = m.pyy;
    o[11 * s] = m.pzz;
    o[12 * s] = m.pxy;
    o[13 * s] = m.pxz;
    o[14 * s] = m.pyz;
    o[15 * s] = m.qx;
    o[16 * s] = m.qy;
    o[17 * s] = m.qz;
    o[18 * s] = M.T;
}

// ---------------------------------------------------------------------------
// LBM::compute_q_corrections (LBM.cpp:959-991), valid cells
// ---------------------------------------------------------------------------
__global__ void k_dq_fab(
    const double *__restrict__ md, const int32_t *__restrict__ isf, double *__restrict__ dq,
    const BoxDev *__restrict__ boxes, const int64_t *__restrict__ base, int nbox, int64_t total,
    int n0, int n1, int n2, CollideParams P)
{
    const int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= total) return;
    const int b = find_box(base, nbox, c);
    const BoxDev B = boxes[b];
    const int64_t loc = c - B.base;
    int iv[3];
    local_to_ijk(B, loc, iv[0], iv[1], iv[2]);
    if (!in_valid(B, iv[0], iv[1], iv[2])) return;
    const int32_t *fl = isf + B.base * 2;
    if (fl[loc] != 1) return;
    const double *m = md + B.base * 19;
    const int nd[3] = {n0, n1, n2};
    for (int d = 0; d < 3; ++d) {
        int p[3] = {iv[0], iv[1], iv[2]}, q[3] = {iv[0], iv[1], iv[2]};
        p[d] += 1;
        q[d] -= 1;
        const bool p_ok = p[d] >= 0 && p[d] < nd[d] && p[(d + 1) % 3] >= 0 &&
                          p[(d + 1) % 3] < nd[(d + 1) % 3] && p[(d + 2) % 3] >= 0 &&
                          p[(d + 2) % 3] < nd[(d + 2) % 3] &&
                          fl[loc_of(B, p[0], p[1], p[2])] == 1;
        const bool m_ok = q[d] >= 0 && q[d] < nd[d] && q[(d + 1) % 3] >= 0 &&
                          q[(d + 1) % 3] < nd[(d + 1) % 3] && q[(d + 2) % 3] >= 0 &&
                          q[(d + 2) % 3] < nd[(d + 2) % 3] &&
                          fl[loc_of(B, q[0], q[1], q[2])] == 1;
        const int comp = 6 + d;
        const double qp = p_ok ? m[comp * B.ncell + loc_of(B, p[0], p[1], p[2])] : 0.0;
        const double qm = m_ok ? m[comp * B.ncell + loc_of(B, q[0], q[1], q[2])] : 0.0;
        const double qcv = m[comp * B.ncell + loc];
        dq[B.base * 3 + d * B.ncell + loc] = one_sided_gradient(p_ok, m_ok, qp, qcv, qm, P.inv_dx[d]);
    }
}

// gradient() (Utilities.H:279-312) of macrodata component comp along d at
// the valid fluid cell iv of box B, as k_dq_fab takes it: a neighbour counts
// where it lies inside the domain box and is fluid
__device__ double fab_gradient(const double *__restrict__ m, const int32_t *__restrict__ fl, const BoxDev &B,
                               const int iv[3], int d, int comp, const int nd[3], double inv_dx)
{
    int p[3] = {iv[0], iv[1], iv[2]}, q[3] = {iv[0], iv[1], iv[2]};
    p[d] += 1;
    q[d] -= 1;
    bool p_ok = true, m_ok = true;
    for (int k = 0; k < 3; ++k) {
        p_ok = p_ok && p[k] >= 0 && p[k] < nd[k];
        m_ok = m_ok && q[k] >= 0 && q[k] < nd[k];
    }
    p_ok = p_ok && fl[loc_of(B, p[0], p[1], p[2])] == 1;
    m_ok = m_ok && fl[loc_of(B, q[0], q[1], q[2])] == 1;
    const double vp = p_ok ? m[comp * B.ncell + loc_of(B, p[0], p[1], p[2])] : 0.0;
    const double vm = m_ok ? m[comp * B.ncell + loc_of(B, q[0], q[1], q[2])] : 0.0;
    const double vc = m[comp * B.ncell + loc_of(B, iv[0], iv[1], iv[2])];
    return one_sided_gradient(p_ok, m_ok, vp, vc, vm, inv_dx);
}

// LBM::compute_derived (LBM.cpp:909-955) and compute_q_corrections
// (959-991) on the macrodata: vort_x, vort_y, vort_z, vort_mag, dQCorrX..Z
// per valid cell (fab comps of der); 0 in solid cells
__global__ void k_plot_derived_fab(
    const double *__restrict__ md, const int32_t *__restrict__ isf, double *__restrict__ der,
    const BoxDev *__restrict__ boxes, const int64_t *__restrict__ base, int nbox, int64_t total, int n0, int n1,
    int n2, CollideParams P)
{
    const int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= total) return;
    const int b = find_box(base, nbox, c);
    const BoxDev B = boxes[b];
    const int64_t loc = c - B.base;
    int iv[3];
    local_to_ijk(B, loc, iv[0], iv[1], iv[2]);
    if (!in_valid(B, iv[0], iv[1], iv[2])) return;
    const int32_t *fl = isf + B.base * 2;
    const double *m = md + B.base * 19;
    double o[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (fl[loc] == 1) {
        const int nd[3] = {n0, n1, n2};
        const double vx = fab_gradient(m, fl, B, iv, 0, 2, nd, P.inv_dx[0]);
        const double wx = fab_gradient(m, fl, B, iv, 0, 3, nd, P.inv_dx[0]);
        const double uy = fab_gradient(m, fl, B, iv, 1, 1, nd, P.inv_dx[1]);
        const double wy = fab_gradient(m, fl, B, iv, 1, 3, nd, P.inv_dx[1]);
        const double uz = fab_gradient(m, fl, B, iv, 2, 1, nd, P.inv_dx[2]);
        const double vz = fab_gradient(m, fl, B, iv, 2, 2, nd, P.inv_dx[2]);
        o[0] = wy - vz;
        o[1] = uz - wx;
        o[2] = vx - uy;
        o[3] = sqrt((wy - vz) * (wy - vz) + (uz - wx) * (uz - wx) + (vx - uy) * (vx - uy));
        for (int d = 0; d < 3; ++d) o[4 + d] = fab_gradient(m, fl, B, iv, d, 6 + d, nd, P.inv_dx[d]);
    }
    double *out = der + B.base * 7 + loc;
    for (int k = 0; k < 7; ++k) out[k * B.ncell] = o[k];
}

// ---------------------------------------------------------------------------
// macrodata_to_equilibrium + relax_f_to_equilibrium (LBM.cpp:621-807),
// valid fluid cells, in place
// ---------------------------------------------------------------------------
__global__ void k_eqrelax_fab(
    double *__restrict__ f, double *__restrict__ g, const double *__restrict__ md,
    const double *__restrict__ dq, const int32_t *__restrict__ isf,
    const BoxDev *__restrict__ boxes, const int64_t *__restrict__ base, int nbox, int64_t total,
    CollideParams P)
{
    const int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= total) return;
    const int b = find_box(base, nbox, c);
    const BoxDev B = boxes[b];
    const int64_t loc = c - B.base;
    int i, j, k;
    local_to_ijk(B, loc, i, j, k);
    if (!in_valid(B, i, j, k)) return;
    if (isf[B.base * 2 + loc] != 1) return;
    const int64_t s = B.ncell;
    const double *m = md + B.base * 19 + loc;
    Moments mo;
    mo.rho = m[0];
    mo.E = m[5 * s];
    mo.pxx = m[9 * s];
    mo.pyy = m[10 * s];
    mo.pzz = m[11 * s];
    mo.pxy = m[12 * s];
    mo.pxz = m[13 * s];
    mo.pyz = m[14 * s];
    mo.qx = m[15 * s];
    mo.qy = m[16 * s];
    mo.qz = m[17 * s];
    Macro M;
    M.u = m[1 * s];
    M.v = m[2 * s];
    M.w = m[3 * s];
    M.T = m[18 * s];
    const double *d = dq + B.base * 3 + loc;
    double *fp = f + B.base * NQ + loc;
    double *gp = g + B.base * NQ + loc;
    double fv[NQ], gv[NQ];
    for (int q = 0; q < NQ; ++q) {
        fv[q] = fp[q * s];
        gv[q] = gp[q * s];
    }
    collide_cell(fv, gv, mo, M, d[0], d[s], d[2 * s], P, [&](auto qc, double fo, double go) {
        constexpr int q = decltype(qc)::value;
        fp[q * s] = fo;
        gp[q * s] = go;
    });
}

// ---------------------------------------------------------------------------
// flags: EB sphere classification, eb_boundary ring (LBM.cpp:1213-1262)
// ---------------------------------------------------------------------------
// A cell is covered when all 8 of its corner nodes lie in the body; a node
// on the surface counts as body.  Sphere: |x - c| <= r.  Cylinder along
// `axis` (EB2 "cylinder"): the distance from the axis <= r and
// |x_axis - c_axis| <= height / 2.  has_fluid_inside takes the complement.
struct BodyArgs {
    double plo[3], dx[3], c[3], r;
    double half_height;  // cylinder only
    int axis;            // -1: sphere, else the cylinder's direction
    int fluid_inside;
};

__device__ __forceinline__ bool body_covered(const BodyArgs &S, double x0, double y0, double z0)
{
    const double r2 = S.r * S.r;
    for (int a = 0; a < 2; ++a)
        for (int c = 0; c < 2; ++c)
            for (int e = 0; e < 2; ++e) {
                const double x = x0 + a * S.dx[0] - S.c[0];
                const double y = y0 + c * S.dx[1] - S.c[1];
                const double z = z0 + e * S.dx[2] - S.c[2];
                bool inside;
                if (S.axis < 0) {
                    const double d2 = x * x + y * y + z * z;
                    inside = S.fluid_inside ? (d2 >= r2) : (d2 <= r2);
                } else {
                    const double p[3] = {x, y, z};
                    const int u = S.axis == 0 ? 1 : 0, v = S.axis == 2 ? 1 : 2;
                    const double d2 = p[u] * p[u] + p[v] * p[v];
                    const double h = fabs(p[S.axis]);
                    inside = S.fluid_inside ? (d2 >= r2 || h >= S.half_height)
                                            : (d2 <= r2 && h <= S.half_height);
                }
                if (!inside) return false;
            }
    return true;
}

__global__ void k_flags_body(
    int32_t *__restrict__ isf, const BoxDev *__restrict__ boxes, const int64_t *__restrict__ base,
    int nbox, int64_t total, BodyArgs S)
{
    const int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= total) return;
    const int b = find_box(base, nbox, c);
    const BoxDev B = boxes[b];
    const int64_t loc = c - B.base;
    int i, j, k;
    local_to_ijk(B, loc, i, j, k);
    const bool cov = body_covered(S, S.plo[0] + i * S.dx[0], S.plo[1] + j * S.dx[1], S.plo[2] + k * S.dx[2]);
    isf[B.base * 2 + loc] = cov ? 0 : 1;
    isf[B.base * 2 + B.ncell + loc] = 0;
}

// comp1 over the fab shrunk by one cell (LBM.cpp:1237-1259)
__global__ void k_flags_ring(
    int32_t *__restrict__ isf, const BoxDev *__restrict__ boxes, const int64_t *__restrict__ base,
    int nbox, int64_t total)
{
    const int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= total) return;
    const int b = find_box(base, nbox, c);
    const BoxDev B = boxes[b];
    const int64_t loc = c - B.base;
    int iv[3];
    local_to_ijk(B, loc, iv[0], iv[1], iv[2]);
    for (int d = 0; d < 3; ++d)
        if (iv[d] <= B.lo[d] || iv[d] >= B.lo[d] + B.n[d] - 1) return;
    int32_t *fl = isf + B.base * 2;
    bool all_covered = true;
    for (int d = 0; d < 3; ++d) {
        int m[3] = {iv[0], iv[1], iv[2]}, p[3] = {iv[0], iv[1], iv[2]};
        m[d] -= 1;
        p[d] += 1;
        all_covered &= (fl[loc_of(B, m[0], m[1], m[2])] == 0) && (fl[loc_of(B, p[0], p[1], p[2])] == 0);
    }
    fl[B.ncell + loc] = (all_covered || fl[loc] == 1) ? 0 : 1;
}

// ---------------------------------------------------------------------------
// conservation sums over valid fluid cells (FP64), two passes
// ---------------------------------------------------------------------------
template <class PopT>
__global__ void k_conservation(
    const PopT *__restrict__ f, const PopT *__restrict__ g, const int32_t *__restrict__ isf,
    const BoxDev *__restrict__ boxes, const int64_t *__restrict__ base, int nbox, int64_t total,
    const double *__restrict__ fshift, const double *__restrict__ gshift, double *__restrict__ part,
    int64_t cs, int cst, const int *__restrict__ e16, int64_t nxy)
{
    __shared__ double sh[5][kBlock];
    double acc[5] = {0, 0, 0, 0, 0};
    constexpr bool half = std::is_same<PopT, _Float16>::value;
    for (int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; c < total;
         c += int64_t(gridDim.x) * blockDim.x) {
        const int b = find_box(base, nbox, c);
        const BoxDev B = boxes[b];
        const int64_t loc = c - B.base;
        int i, j, k;
        local_to_ijk(B, loc, i, j, k);
        if (!in_valid(B, i, j, k)) continue;
        if (isf[B.base * 2 + loc] != 1) continue;
        const PopT *fp = f + B.base * NQ + loc * cst;
        const PopT *gp = g + B.base * NQ + loc * cst;
        const int64_t stride = cs ? cs : B.ncell;
        // FP16 storage (compact layout): per-box scale of the cell's z-slab
        float sf = 1.0f, sg = 1.0f;
        if constexpr (half) {
            const int b = int(c / nxy) / kScaleSlab;
            sf = ldexpf(1.0f, -e16[2 * b]);
            sg = ldexpf(1.0f, -e16[2 * b + 1]);
        }
        static_for<NQ>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            double fq, gq;
            if constexpr (half) {
                fq = double(float(fp[q * stride]) * sf) + fshift[q];
                gq = double(float(gp[q * stride]) * sg) + gshift[q];
            } else {
                fq = double(fp[q * stride]) + fshift[q];
                gq = double(gp[q * stride]) + gshift[q];
            }
            acc[0] += fq;
            if constexpr (ev(q, 0) != 0) acc[1] += ev(q, 0) * fq;
            if constexpr (ev(q, 1) != 0) acc[2] += ev(q, 1) * fq;
            if constexpr (ev(q, 2) != 0) acc[3] += ev(q, 2) * fq;
            acc[4] += gq;
        });
    }
    for (int n = 0; n < 5; ++n) sh[n][threadIdx.x] = acc[n];
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w)
            for (int n = 0; n < 5; ++n) sh[n][threadIdx.x] += sh[n][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int n = 0; n < 5; ++n) part[blockIdx.x * 5 + n] = sh[n][0];
}

__global__ void k_fill_const(double *__restrict__ a, int64_t n, double v)
{
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t < n) a[t] = v;
}

// uniform initial state: comp q of every cell = val[q]; solid cells 0.
// cs = component stride (0: the box's own fab size, the reference layout)
__global__ void k_init_uniform(
    double *__restrict__ pop, const int32_t *__restrict__ isf, const BoxDev *__restrict__ boxes,
    const int64_t *__restrict__ base, int nbox, int64_t total, const double *__restrict__ val, int64_t cs,
    int cst)
{
    const int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= total) return;
    const int b = find_box(base, nbox, c);
    const BoxDev B = boxes[b];
    const int64_t loc = c - B.base;
    const bool solid = isf[B.base * 2 + loc] == 0;
    const int64_t stride = cs ? cs : B.ncell;
    for (int q = 0; q < NQ; ++q) pop[B.base * NQ + q * stride + loc * cst] = solid ? 0.0 : val[q];
}

// LBM::compute_eb_forces (LBM.cpp:994-1044) on one level: over valid cells
// with is_fluid comp1 == 1, F += 2 c_q f_q(x - c_q) is_fluid(x - c_q, 0);
// per-block partial sums.  On the compact layout x - c_q wraps on periodic
// axes; a source beyond a non-periodic face contributes nothing (the
// reference reads a physical-boundary ghost there).
template <class PopT>
__global__ void k_eb_force(
    const PopT *__restrict__ f, const int32_t *__restrict__ isf, const BoxDev *__restrict__ boxes,
    const int64_t *__restrict__ base, int nbox, int64_t total, const double *__restrict__ fshift,
    double *__restrict__ part, int64_t cs, int cst, int compact, int n0, int n1, int n2, int p0, int p1,
    int p2)
{
    __shared__ double sh[3][kBlock];
    double acc[3] = {0, 0, 0};
    for (int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; c < total;
         c += int64_t(gridDim.x) * blockDim.x) {
        const int b = find_box(base, nbox, c);
        const BoxDev B = boxes[b];
        const int64_t loc = c - B.base;
        int i, j, k;
        local_to_ijk(B, loc, i, j, k);
        if (!in_valid(B, i, j, k)) continue;
        const int32_t *fl = isf + B.base * 2;
        if (fl[B.ncell + loc] != 1) continue;
        const int64_t stride = cs ? cs : B.ncell;
        double fs[3] = {0.0, 0.0, 0.0};
        static_for<NQ>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            int a = i - ev(q, 0), bb = j - ev(q, 1), e = k - ev(q, 2);
            bool ok = true;
            if (compact) {
                const int n[3] = {n0, n1, n2}, per[3] = {p0, p1, p2};
                int s[3] = {a, bb, e};
                for (int d = 0; d < 3; ++d) {
                    if (s[d] < 0 || s[d] >= n[d]) {
                        if (!per[d]) ok = false;
                        s[d] = s[d] < 0 ? s[d] + n[d] : s[d] - n[d];
                    }
                }
                a = s[0];
                bb = s[1];
                e = s[2];
            }
            if (ok) {
                const int64_t sl = loc_of(B, a, bb, e);
                const double fq = double(f[B.base * NQ + q * stride + sl * cst]) + fshift[q];
                const int32_t w = fl[sl];
                fs[0] += 2.0 * ev(q, 0) * fq * w;
                fs[1] += 2.0 * ev(q, 1) * fq * w;
                fs[2] += 2.0 * ev(q, 2) * fq * w;
            }
        });
        acc[0] += fs[0];
        acc[1] += fs[1];
        acc[2] += fs[2];
    }
    for (int n = 0; n < 3; ++n) sh[n][threadIdx.x] = acc[n];
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (int(threadIdx.x) < w)
            for (int n = 0; n < 3; ++n) sh[n][threadIdx.x] += sh[n][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int n = 0; n < 3; ++n) part[blockIdx.x * 5 + n] = sh[n][0];
}

// dst[i * ds] = src[i * ss]
template <class T>
__global__ void k_strided_copy(const T *__restrict__ src, int ss, T *__restrict__ dst, int ds, int64_t n)
{
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) dst[i * ds] = src[i * ss];
}

// initial state from per-cell fields (device copies of the host arrays):
// f = f_eq(rho, u, R T), g = Grad equilibrium, 0 in solid cells, stored with
// the level's storage (shifted FP32: value - shift)
template <class PopT>
__global__ void k_init_fields(
    PopT *__restrict__ f, PopT *__restrict__ g, const int32_t *__restrict__ isf,
    const BoxDev *__restrict__ boxes, const int64_t *__restrict__ base, int nbox, int64_t total,
    const double *__restrict__ fields, double R, double gamma, const double *__restrict__ shift, int64_t cs,
    int cst)
{
    const int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= total) return;
    const int b = find_box(base, nbox, c);
    const BoxDev B = boxes[b];
    const int64_t loc = c - B.base;
    const bool solid = isf[B.base * 2 + loc] == 0;
    const double rho = fields[c];
    const double vel[3] = {fields[total + c], fields[2 * total + c], fields[3 * total + c]};
    const double T = fields[4 * total + c];
    const int64_t stride = cs ? cs : B.ncell;
    for (int q = 0; q < NQ; ++q) {
        const double fv = solid ? 0.0 : set_equilibrium_value(rho, vel, R * T, q);
        const double gv = solid ? 0.0 : thermal_g_equilibrium(rho, vel, T, R, gamma, q);
        f[B.base * NQ + q * stride + loc * cst] = PopT(fv - shift[q]);
        g[B.base * NQ + q * stride + loc * cst] = PopT(gv - shift[NQ + q]);
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// host-side helpers
// ---------------------------------------------------------------------------

// Component stride of the population arrays: the compact layout keeps f_q
// and g_q of a cell side by side (component stride 2N); the reference
// layout uses each box's fab size (0 = per box).
static int64_t comp_stride(const mrb_level *L)
{
    return L->layout == MRB_LAYOUT_COMPACT ? 2 * L->total_cells : 0;
}

// Element stride between neighbouring cells: the compact layout keeps f_q
// and g_q of a cell side by side (pop[(q*N + cell)*2 + lattice]) so that
// one load fetches both.
static int cell_stride(const mrb_level *L) { return L->layout == MRB_LAYOUT_COMPACT ? 2 : 1; }

// Host <-> device copy of one lattice's 27 components (host: contiguous
// components of n elements; device: component stride dst/src_stride).
static hipError_t copy_components(
    void *dst, size_t dpitch, const void *src, size_t spitch, size_t width, hipMemcpyKind kind)
{
    return hipMemcpy2D(dst, dpitch, src, spitch, width, NQ, kind);
}

// Fills `map` with every ghost cell of every box (within `depth` layers) that
// AMReX FillBoundary would overwrite: its periodic image lies inside the
// domain and in some box's valid region.
void build_ghost_map(const mrb_level *L, int depth, std::vector<GhostPair> &map)
{
    const int64_t nd0 = L->ncell[0], nd1 = L->ncell[1], nd2 = L->ncell[2];
    std::vector<int32_t> owner(size_t(nd0 * nd1 * nd2), -1);
    for (int b = 0; b < L->nbox; ++b) {
        const BoxDev &B = L->boxes[b];
        for (int k = B.vlo[2]; k <= B.vhi[2]; ++k)
            for (int j = B.vlo[1]; j <= B.vhi[1]; ++j)
                for (int i = B.vlo[0]; i <= B.vhi[0]; ++i) owner[size_t((k * nd1 + j) * nd0 + i)] = b;
    }
    map.clear();
    for (int b = 0; b < L->nbox; ++b) {
        const BoxDev &B = L->boxes[b];
        for (int k = B.vlo[2] - depth; k <= B.vhi[2] + depth; ++k)
            for (int j = B.vlo[1] - depth; j <= B.vhi[1] + depth; ++j)
                for (int i = B.vlo[0] - depth; i <= B.vhi[0] + depth; ++i) {
                    if (i >= B.vlo[0] && i <= B.vhi[0] && j >= B.vlo[1] && j <= B.vhi[1] &&
                        k >= B.vlo[2] && k <= B.vhi[2])
                        continue;
                    // image one period away (AMReX's periodic shifts span
                    // one period; with ng > n a ghost can have none)
                    int s[3] = {i, j, k};
                    bool ok = true;
                    for (int d = 0; d < 3; ++d) {
                        if (s[d] < 0) {
                            if (!L->periodic[d]) ok = false;
                            s[d] += L->ncell[d];
                        } else if (s[d] >= L->ncell[d]) {
                            if (!L->periodic[d]) ok = false;
                            s[d] -= L->ncell[d];
                        }
                        if (s[d] < 0 || s[d] >= L->ncell[d]) ok = false;
                    }
                    if (!ok) continue;
                    const int ob = owner[size_t((int64_t(s[2]) * nd1 + s[1]) * nd0 + s[0])];
                    if (ob < 0) continue;
                    const BoxDev &S = L->boxes[ob];
                    GhostPair p;
                    p.dst_box = b;
                    p.src_box = ob;
                    p.dst_local = (int64_t(k - B.lo[2]) * B.n[1] + (j - B.lo[1])) * B.n[0] + (i - B.lo[0]);
                    p.src_local =
                        (int64_t(s[2] - S.lo[2]) * S.n[1] + (s[1] - S.lo[1])) * S.n[0] + (s[0] - S.lo[0]);
                    map.push_back(p);
                }
    }
}

static int fill_boundary(mrb_level *L, void *arr, int ncomp, bool is_int, bool one_layer, hipStream_t s)
{
    const GhostPair *map = one_layer ? L->d_gmap1 : L->d_gmap;
    const int64_t n = one_layer ? L->n_gmap1 : L->n_gmap;
    if (n == 0) return MRB_OK;
    if (is_int && ncomp == 2)
        k_fill_map<int32_t, 2><<<nblocks(n), kBlock, 0, s>>>((int32_t *)arr, L->d_boxes, map, n);
    else if (!is_int && ncomp == NQ)
        k_fill_map<double, NQ><<<nblocks(n), kBlock, 0, s>>>((double *)arr, L->d_boxes, map, n);
    else if (!is_int && ncomp == 3)
        k_fill_map<double, 3><<<nblocks(n), kBlock, 0, s>>>((double *)arr, L->d_boxes, map, n);
    else if (!is_int && ncomp == 19)
        k_fill_map<double, 19><<<nblocks(n), kBlock, 0, s>>>((double *)arr, L->d_boxes, map, n);
    else
        return set_error(MRB_E_ARG, "fill_boundary: unsupported component count");
    MRB_HIP(hipGetLastError());
    return MRB_OK;
}

int fab_fill_boundary(mrb_level *L, int lattice, hipStream_t s)
{
    return fill_boundary(L, L->pop[lattice][L->cur], NQ, false, false, s);
}

int fab_fill_boundary_map(mrb_level *L, int lattice, const GhostPair *map, int64_t n, hipStream_t s)
{
    if (n == 0) return MRB_OK;
    k_fill_map<double, NQ><<<nblocks(n), kBlock, 0, s>>>((double *)L->pop[lattice][L->cur], L->d_boxes, map, n);
    MRB_HIP(hipGetLastError());
    return MRB_OK;
}

int fab_fill_boundary_q(mrb_level *L, hipStream_t s) { return fill_boundary(L, L->d_dq, 3, false, true, s); }

// one fab component over the valid cells into a domain-sized array
// (x fastest), as the dense field of the tagging criteria
__global__ void k_fab_to_dense(const double *__restrict__ src, int ncomp, int comp, double *__restrict__ dense,
                               const BoxDev *__restrict__ boxes, const int64_t *__restrict__ base, int nbox,
                               int64_t total, int nx, int ny)
{
    const int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= total) return;
    const int b = find_box(base, nbox, c);
    const BoxDev B = boxes[b];
    const int64_t loc = c - B.base;
    int i, j, k;
    local_to_ijk(B, loc, i, j, k);
    if (!in_valid(B, i, j, k)) return;
    dense[(int64_t(k) * ny + j) * nx + i] = src[B.base * ncomp + int64_t(comp) * B.ncell + loc];
}

int fab_plot_fields(mrb_level *L, double *der, hipStream_t s)
{
    const unsigned nb = nblocks(L->total_cells);
    if (L->md_src < 0) return set_error(MRB_E_STATE, "plot fields: the stored state has no macrodata (stream without a collide)");
    if (L->md_src == 0) {
        // MakeNewLevelFromScratch: f_to_macrodata of the stored state
        k_moments_fab<<<nb, kBlock, 0, s>>>((const double *)L->pop[0][L->cur], (const double *)L->pop[1][L->cur],
                                            L->d_isf, L->d_md, L->d_boxes, L->d_box_base, L->nbox, L->total_cells,
                                            L->cp);
        MRB_HIP(hipGetLastError());
        const int rc = fill_boundary(L, L->d_md, 19, false, true, s);
        if (rc) return rc;
    } else if (L->md_src == 2) {
        // the coarse level of a hierarchy gathered the coalesced fine values
        const int rc = fab_plot_md_pull(L, L->amr_fine, s);
        if (rc) return rc;
        const int rc2 = fill_boundary(L, L->d_md, 19, false, true, s);
        if (rc2) return rc2;
    }
    k_plot_derived_fab<<<nb, kBlock, 0, s>>>(L->d_md, L->d_isf, der, L->d_boxes, L->d_box_base, L->nbox,
                                             L->total_cells, L->ncell[0], L->ncell[1], L->ncell[2], L->cp);
    MRB_HIP(hipGetLastError());
    return MRB_OK;
}

int fab_fill_boundary_depth(mrb_level *L, int depth, hipStream_t s)
{
    const bool one = depth <= 1;
    for (int l = 0; l < 2; ++l) {
        const int rc = fill_boundary(L, L->pop[l][L->cur], NQ, false, one, s);
        if (rc) return rc;
    }
    L->fresh_depth = one ? 1 : L->ng;
    return MRB_OK;
}

static void compute_bc_values(mrb_level *L)
{
    BCDev &d = L->bcdev;
    for (int n = 0; n < 6; ++n) d.type[n] = L->bc.bc_type[n];
    // BCFill::operator() (BC.H:395-419): R, T, gamma preset, then the
    // velocity operator overrides them (VelocityBC.H:253-271).
    double vel[3] = {0.0, 0.0, 0.0};
    double R = 1.0, T = 1.0 / 3.0, gamma = 5.0 / 3.0, rho = 0.0;
    if (L->bc.velbc_kind == 1) {
        const double sos = std::sqrt(
            L->bc.adiabatic_exponent * (kAirMolarMass / L->bc.m_bar) * L->bc.temperature);
        const double u0 = L->bc.mach_ref * sos;
        rho = L->bc.density;
        vel[L->bc.vel_dir] = u0;
        R = (kAirMolarMass / L->bc.m_bar);
        T = L->bc.temperature;
        gamma = L->bc.adiabatic_exponent;
    }
    for (int q = 0; q < NQ; ++q) {
        d.fval[q] = set_equilibrium_value(rho, vel, R * T, q);
        d.gval[q] = thermal_g_equilibrium(rho, vel, T, R, gamma, q);
    }
    // BCFill's PRESSURE case (BC.H:421-452): rho 1, R 1, T 1/3, gamma 5/3
    // unless a velocity operator sets them; the g ghost is the Grad
    // equilibrium at the normal velocity ndir (1 - 1 / rho_bc)
    // (pressure_bc_g with rho_bc_out = 1), ndir = c_q along the axis
    d.pdir = -1;
    for (int f = 0; f < 6; ++f)
        if (L->bc.bc_type[f] == MRB_BC_PRESSURE) d.pdir = f % 3;
    {
        double prho = 1.0, pR = 1.0, pT = 1.0 / 3.0, pgamma = 5.0 / 3.0;
        if (L->bc.velbc_kind >= 1) {
            prho = L->bc.density;
            pR = kAirMolarMass / L->bc.m_bar;
            pT = L->bc.temperature;
            pgamma = L->bc.adiabatic_exponent;
        }
        d.prho = prho;
        d.prt = pR * pT;
        for (int q = 0; q < NQ; ++q) {
            d.pgval[q] = 0.0;
            if (d.pdir < 0 || kEV[q][d.pdir] == 0) continue;
            const int ndir = kEV[q][d.pdir];
            double v[3] = {0.0, 0.0, 0.0};
            v[d.pdir] = ndir * (1.0 - (1.0) / prho);
            d.pgval[q] = thermal_g_equilibrium(prho, v, pT, pR, pgamma, q);
        }
    }
    if (L->d_small) {
        (void)hipMemcpy(L->d_small + 4 * NQ, d.fval, sizeof(double) * NQ, hipMemcpyHostToDevice);
        (void)hipMemcpy(L->d_small + 5 * NQ, d.gval, sizeof(double) * NQ, hipMemcpyHostToDevice);
    }
    // Channel / Parabolic (VelocityBC.cpp:23-54): the velocity depends on
    // the ghost cell
    VelProfile &P = d.prof;
    P = VelProfile{};
    for (int f = 0; f < 6; ++f) {
        d.tab[f] = nullptr;
        d.tabw[f] = 0;
    }
    (void)hipFree(L->d_bctab);
    L->d_bctab = nullptr;
    if (L->bc.velbc_kind == 2 || L->bc.velbc_kind == 3) {
        P.kind = L->bc.velbc_kind;
        const double sos = std::sqrt(
            L->bc.adiabatic_exponent * (kAirMolarMass / L->bc.m_bar) * L->bc.temperature);
        P.vref = L->bc.mach_ref * sos;
        P.rho = L->bc.density;
        P.R = kAirMolarMass / L->bc.m_bar;
        P.T = L->bc.temperature;
        P.gamma = L->bc.adiabatic_exponent;
        P.nd = L->bc.normal_dir;
        P.td = L->bc.vel_dir;
        for (int k = 0; k < 3; ++k) P.domhi[k] = L->ncell[k] - 1;
        P.plo = L->prob_lo[P.nd];
        P.phi = L->prob_hi[P.nd];
        P.dxn = L->dx[P.nd];
        if (L->layout == MRB_LAYOUT_COMPACT) {
            // ghost tables of the VELOCITY faces for the compact pull
            std::vector<double> host;
            int64_t offs[6];
            for (int f = 0; f < 6; ++f) {
                offs[f] = -1;
                if (L->bc.bc_type[f] != MRB_BC_VELOCITY) continue;
                const int dd = f % 3, d1 = dd == 0 ? 1 : 0, d2 = dd == 2 ? 1 : 2;
                const int n1 = L->ncell[d1] + 2, n2 = L->ncell[d2] + 2;
                offs[f] = int64_t(host.size());
                d.tabw[f] = n2;
                for (int a = 0; a < n1; ++a)
                    for (int b = 0; b < n2; ++b) {
                        int iv[3];
                        iv[dd] = f < 3 ? -1 : L->ncell[dd];
                        iv[d1] = a - 1;
                        iv[d2] = b - 1;
                        double v[3];
                        profile_velocity(P, iv, v);
                        for (int q = 0; q < NQ; ++q) host.push_back(set_equilibrium_value(P.rho, v, P.R * P.T, q));
                        for (int q = 0; q < NQ; ++q)
                            host.push_back(thermal_g_equilibrium(P.rho, v, P.T, P.R, P.gamma, q));
                    }
            }
            if (!host.empty() && hipMalloc(&L->d_bctab, sizeof(double) * host.size()) == hipSuccess &&
                hipMemcpy(L->d_bctab, host.data(), sizeof(double) * host.size(), hipMemcpyHostToDevice) ==
                    hipSuccess) {
                for (int f = 0; f < 6; ++f)
                    if (offs[f] >= 0) d.tab[f] = L->d_bctab + offs[f];
            }
        }
    }
}

static void compute_shift(mrb_level *L)
{
    // f_eq and g_eq of the rest state (rho0, u = 0, T0) with R and gamma of
    // the level physics: the subtracted reference of the shifted storages.
    const double R = kAirMolarMass / (L->phys_set ? L->phys.m_bar : kAirMolarMass);
    const double gamma = L->phys_set ? L->phys.adiabatic_exponent : 5.0 / 3.0;
    const double zero[3] = {0.0, 0.0, 0.0};
    // The rest state depends on |c_q|^2 only; one value per speed class
    // (taken from its first direction) so that kernels can hold the four
    // values of each lattice in scalar registers.
    const int rep[4] = {0, 1, 7, 19};
    for (int q = 0; q < NQ; ++q) {
        const int r = rep[csq(q)];
        L->fshift[q] = set_equilibrium_value(L->shift_rho0, zero, R * L->shift_T0, r);
        L->gshift[q] = thermal_g_equilibrium(L->shift_rho0, zero, L->shift_T0, R, gamma, r);
    }
    if (L->d_small) {
        double tab[2 * NQ];
        for (int q = 0; q < NQ; ++q) {
            tab[q] = L->storage == MRB_STORAGE_F64 ? 0.0 : L->fshift[q];
            tab[NQ + q] = L->storage == MRB_STORAGE_F64 ? 0.0 : L->gshift[q];
        }
        (void)hipMemcpy(L->d_small, tab, sizeof(tab), hipMemcpyHostToDevice);
    }
}

// A reduced-precision level holds population - shift: once populations
// are stored, a different shift would silently reinterpret them.
static bool shift_locked(const mrb_level *L, double m_bar, double gamma, double rho0, double T0)
{
    if (L->storage == MRB_STORAGE_F64 || !L->pops_set) return false;
    const double R_new = kAirMolarMass / m_bar;
    const double R_old = kAirMolarMass / (L->phys_set ? L->phys.m_bar : kAirMolarMass);
    const double g_old = L->phys_set ? L->phys.adiabatic_exponent : 5.0 / 3.0;
    return R_new != R_old || gamma != g_old || rho0 != L->shift_rho0 || T0 != L->shift_T0;
}

static size_t pop_elem_size(int storage)
{
    return storage == MRB_STORAGE_F64 ? 8 : storage == MRB_STORAGE_F32_SHIFTED ? 4 : 2;
}

}  // namespace mrb

using namespace mrb;

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

const char *mrb_last_error(void) { return g_err.c_str(); }

int mrb_level_create(const mrb_level_desc *desc, mrb_level **out)
{
    if (!desc || !out) return set_error(MRB_E_ARG, "null argument");
    *out = nullptr;
    if (desc->nbox < 1 || !desc->box_lo || !desc->box_hi)
        return set_error(MRB_E_ARG, "level needs at least one box");
    if (desc->nlattice != 2) return set_error(MRB_E_UNSUPPORTED, "nlattice must be 2 (f and g)");
    if (desc->layout != MRB_LAYOUT_FAB && desc->layout != MRB_LAYOUT_COMPACT)
        return set_error(MRB_E_ARG, "unknown layout");
    if (desc->storage != MRB_STORAGE_F64 && desc->layout == MRB_LAYOUT_FAB)
        return set_error(MRB_E_UNSUPPORTED, "reduced-precision storage needs the compact layout");
    for (int d = 0; d < 3; ++d)
        if (desc->ncell[d] < 1) return set_error(MRB_E_ARG, "ncell must be positive");
    std::unique_ptr<mrb_level> L(new mrb_level());
    for (int d = 0; d < 3; ++d) {
        L->ncell[d] = desc->ncell[d];
        L->periodic[d] = desc->periodic[d] ? 1 : 0;
        L->prob_lo[d] = desc->prob_lo[d];
        L->prob_hi[d] = desc->prob_hi[d];
        L->dx[d] = (desc->prob_hi[d] - desc->prob_lo[d]) / desc->ncell[d];
        L->inv_dx[d] = desc->ncell[d] / (desc->prob_hi[d] - desc->prob_lo[d]);
    }
    L->nbox = desc->nbox;
    L->ng = desc->ng;
    L->layout = desc->layout;
    L->nlattice = desc->nlattice;
    L->storage = desc->storage;
    // the default step variant per storage: two passes on 256x1-cell blocks
    // for FP64; for the reduced storages the FP32-arithmetic collide (64x2
    // blocks for shifted FP32, 64x4 for scaled FP16), which levels with
    // PRESSURE faces replace by FP64 arithmetic (mrb_set_bc);
    // mrb_set_hot_kernel and mrb_autotune select others
    L->hot = L->storage == MRB_STORAGE_F64            ? kX256Variant
             : L->storage == MRB_STORAGE_F32_SHIFTED ? kF32Variant
                                                      : kHalf32Variant;
    L->hot_default = true;
    L->hot_h32 = -1;
    if (L->layout == MRB_LAYOUT_COMPACT) {
        if (desc->nbox != 1 || desc->ng != 0)
            return set_error(MRB_E_ARG, "compact layout is one box without ghost cells");
        for (int d = 0; d < 3; ++d)
            if (desc->box_lo[d] != 0 || desc->box_hi[d] != desc->ncell[d] - 1)
                return set_error(MRB_E_ARG, "compact layout box must equal the domain");
        if (int64_t(desc->ncell[0]) * desc->ncell[1] * desc->ncell[2] >= (int64_t(1) << 31))
            return set_error(MRB_E_ARG, "compact layout limited to 2^31 cells");
    } else if (desc->ng < 1) {
        return set_error(MRB_E_ARG, "fab layout needs ng >= 1");
    }
    // boxes
    int64_t base = 0;
    std::vector<int64_t> bases;
    for (int b = 0; b < desc->nbox; ++b) {
        BoxDev B{};
        B.ncell = 1;
        for (int d = 0; d < 3; ++d) {
            B.vlo[d] = desc->box_lo[3 * b + d];
            B.vhi[d] = desc->box_hi[3 * b + d];
            if (B.vhi[d] < B.vlo[d] || B.vlo[d] < 0 || B.vhi[d] >= desc->ncell[d])
                return set_error(MRB_E_ARG, "box outside the domain");
            B.lo[d] = B.vlo[d] - L->ng;
            B.n[d] = B.vhi[d] - B.vlo[d] + 1 + 2 * L->ng;
            B.ncell *= B.n[d];
        }
        B.base = base;
        bases.push_back(base);
        base += B.ncell;
        L->boxes.push_back(B);
    }
    bases.push_back(base);
    L->total_cells = base;
    const int64_t T = base;
    const size_t es = pop_elem_size(L->storage);
    MRB_HIP(hipMalloc(&L->d_boxes, sizeof(BoxDev) * L->nbox));
    MRB_HIP(hipMemcpy(L->d_boxes, L->boxes.data(), sizeof(BoxDev) * L->nbox, hipMemcpyHostToDevice));
    MRB_HIP(hipMalloc(&L->d_box_base, sizeof(int64_t) * (L->nbox + 1)));
    MRB_HIP(hipMemcpy(L->d_box_base, bases.data(), sizeof(int64_t) * (L->nbox + 1), hipMemcpyHostToDevice));
    if (L->layout == MRB_LAYOUT_COMPACT) {
        // f and g in one allocation per buffer, f_q and g_q of a cell side
        // by side: element (q*N + cell)*2 + lattice (g = f + 1)
        for (int ab = 0; ab < 2; ++ab) {
            MRB_HIP(hipMalloc(&L->pop[0][ab], es * 2 * NQ * T));
            MRB_HIP(hipMemset(L->pop[0][ab], 0, es * 2 * NQ * T));
            L->pop[1][ab] = static_cast<char *>(L->pop[0][ab]) + es * (cell_stride(L.get()) == 2 ? 1 : T);
        }
    } else {
        for (int l = 0; l < 2; ++l)
            for (int ab = 0; ab < 2; ++ab) {
                MRB_HIP(hipMalloc(&L->pop[l][ab], es * NQ * T));
                MRB_HIP(hipMemset(L->pop[l][ab], 0, es * NQ * T));
            }
    }
    MRB_HIP(hipMalloc(&L->d_isf, sizeof(int32_t) * 2 * T));
    MRB_HIP(hipMemset(L->d_isf, 0, sizeof(int32_t) * 2 * T));
    MRB_HIP(hipMalloc(&L->d_red, sizeof(double) * 5 * kConservationBlocks));
    MRB_HIP(hipMalloc(&L->d_small, sizeof(double) * 6 * NQ));
    for (auto &e : L->ev) MRB_HIP(hipEventCreate(&e));
    MRB_HIP(hipMemset(L->d_small, 0, sizeof(double) * 6 * NQ));
    if (L->storage == MRB_STORAGE_F16_SCALED) {
        L->nsb16 = (L->ncell[2] + kScaleSlab - 1) / kScaleSlab;
        for (int ab = 0; ab < 2; ++ab) {
            MRB_HIP(hipMalloc(&L->d_e16[ab], sizeof(int) * 2 * L->nsb16));
            MRB_HIP(hipMemset(L->d_e16[ab], 0, sizeof(int) * 2 * L->nsb16));
            MRB_HIP(hipMalloc(&L->d_amax16[ab], sizeof(uint32_t) * 2 * L->nsb16));
            MRB_HIP(hipMemset(L->d_amax16[ab], 0, sizeof(uint32_t) * 2 * L->nsb16));
        }
        MRB_HIP(hipMalloc(&L->d_ovf16, sizeof(uint32_t)));
        MRB_HIP(hipMemset(L->d_ovf16, 0, sizeof(uint32_t)));
    }
    if (L->layout == MRB_LAYOUT_FAB) {
        MRB_HIP(hipMalloc(&L->d_md, sizeof(double) * 19 * T));
        MRB_HIP(hipMemset(L->d_md, 0, sizeof(double) * 19 * T));
        MRB_HIP(hipMalloc(&L->d_dq, sizeof(double) * 3 * T));
        MRB_HIP(hipMemset(L->d_dq, 0, sizeof(double) * 3 * T));
        std::vector<GhostPair> map;
        build_ghost_map(L.get(), L->ng, map);
        L->n_gmap = int64_t(map.size());
        if (!map.empty()) {
            MRB_HIP(hipMalloc(&L->d_gmap, sizeof(GhostPair) * map.size()));
            MRB_HIP(hipMemcpy(L->d_gmap, map.data(), sizeof(GhostPair) * map.size(), hipMemcpyHostToDevice));
        }
        build_ghost_map(L.get(), 1, map);
        L->n_gmap1 = int64_t(map.size());
        if (!map.empty()) {
            MRB_HIP(hipMalloc(&L->d_gmap1, sizeof(GhostPair) * map.size()));
            MRB_HIP(hipMemcpy(L->d_gmap1, map.data(), sizeof(GhostPair) * map.size(), hipMemcpyHostToDevice));
        }
        MRB_HIP(hipMalloc(&L->d_mask, sizeof(uint32_t) * T));
        MRB_HIP(hipMemset(L->d_mask, 0, sizeof(uint32_t) * T));
        std::vector<FabTile> tiles;
        build_fab_tiles(L.get(), 0, tiles);
        L->n_tiles_c = int(tiles.size());
        MRB_HIP(hipMalloc(&L->d_tiles_c, sizeof(FabTile) * tiles.size()));
        MRB_HIP(hipMemcpy(L->d_tiles_c, tiles.data(), sizeof(FabTile) * tiles.size(), hipMemcpyHostToDevice));
        // fab cells outside a non-periodic domain face
        // by distance beyond the domain: an OUTFLOW ghost copies the cell one
        // step closer, which a launch per distance has filled before
        const int maxdist = 3 * L->ng;
        std::vector<std::vector<int64_t>> bydist(size_t(maxdist + 1));
        for (int b = 0; b < L->nbox; ++b) {
            const BoxDev &B = L->boxes[b];
            for (int k = 0; k < B.n[2]; ++k)
                for (int j = 0; j < B.n[1]; ++j)
                    for (int i = 0; i < B.n[0]; ++i) {
                        const int iv[3] = {B.lo[0] + i, B.lo[1] + j, B.lo[2] + k};
                        int out = 0;
                        for (int d = 0; d < 3; ++d) {
                            if (L->periodic[d]) continue;
                            if (iv[d] < 0) out += -iv[d];
                            if (iv[d] >= L->ncell[d]) out += iv[d] - (L->ncell[d] - 1);
                        }
                        if (out > 0) bydist[size_t(out)].push_back(B.base + (int64_t(k) * B.n[1] + j) * B.n[0] + i);
                    }
        }
        std::vector<int64_t> cells;
        L->bcell_off.assign(size_t(maxdist + 1), 0);
        for (int p = 1; p <= maxdist; ++p) {
            cells.insert(cells.end(), bydist[size_t(p)].begin(), bydist[size_t(p)].end());
            L->bcell_off[size_t(p)] = int64_t(cells.size());
        }
        L->n_bcells = int64_t(cells.size());
        if (!cells.empty()) {
            MRB_HIP(hipMalloc(&L->d_bcells, sizeof(int64_t) * cells.size()));
            MRB_HIP(hipMemcpy(L->d_bcells, cells.data(), sizeof(int64_t) * cells.size(), hipMemcpyHostToDevice));
        }
    } else {
        MRB_HIP(hipMalloc(&L->d_mask, sizeof(uint32_t) * T));
        MRB_HIP(hipMemset(L->d_mask, 0, sizeof(uint32_t) * T));
        MRB_HIP(hipMalloc(&L->d_q, sizeof(double) * 3 * T));
        MRB_HIP(hipMemset(L->d_q, 0, sizeof(double) * 3 * T));
        if (L->storage == MRB_STORAGE_F64 || L->storage == MRB_STORAGE_F32_SHIFTED) {
            const size_t plane = size_t(L->ncell[0]) * size_t(L->ncell[1]);
            const size_t nblk = size_t((L->ncell[0] + 63) / 64) * size_t(L->ncell[1]);
            MRB_HIP(hipMalloc(&L->d_sweep_q, sizeof(double) * kSweepRingPlanes * plane));
            MRB_HIP(hipMemset(L->d_sweep_q, 0, sizeof(double) * kSweepRingPlanes * plane));
            MRB_HIP(hipMalloc(&L->d_sweep_qx, sizeof(double) * kSweepRingPlanes * 2 * nblk));
            MRB_HIP(hipMemset(L->d_sweep_qx, 0, sizeof(double) * kSweepRingPlanes * 2 * nblk));
            MRB_HIP(hipMalloc(&L->d_sweep_flags, sizeof(uint32_t) * nblk));
            MRB_HIP(hipMemset(L->d_sweep_flags, 0, sizeof(uint32_t) * nblk));
            MRB_HIP(hipMalloc(&L->d_sweep_fault, sizeof(uint32_t)));
            MRB_HIP(hipMemset(L->d_sweep_fault, 0, sizeof(uint32_t)));
            MRB_HIP(hipMalloc(&L->d_sweep_qt, sizeof(uint64_t) * 2 * kSweepRingPlanes * plane));
            MRB_HIP(hipMemset(L->d_sweep_qt, 0, sizeof(uint64_t) * 2 * kSweepRingPlanes * plane));
            MRB_HIP(hipMalloc(&L->d_sweep_epoch, sizeof(uint32_t)));
            MRB_HIP(hipMemset(L->d_sweep_epoch, 0, sizeof(uint32_t)));
        }
    }
    // defaults (LBM.H:272-318)
    mrb_phys p{};
    p.nu = 1.0;
    p.alpha = 1.0;
    p.adiabatic_exponent = 5.0 / 3.0;
    p.m_bar = kAirMolarMass;
    p.dt = 1.0;
    p.mesh_speed = 1.0;
    L->phys = p;
    for (int n = 0; n < 6; ++n) L->bc.bc_type[n] = MRB_BC_PERIODIC;
    L->bc.velbc_kind = 0;
    L->bc.vel_dir = 1;
    L->bc.mach_ref = 0.005;
    L->bc.density = 1.0;
    L->bc.temperature = 1.0 / 3.0;
    L->bc.adiabatic_exponent = 5.0 / 3.0;
    L->bc.m_bar = kAirMolarMass;
    compute_bc_values(L.get());
    mrb_set_phys(L.get(), &p);
    L->phys_set = false;
    compute_shift(L.get());
    *out = L.release();
    return MRB_OK;
}

int mrb_level_destroy(mrb_level *L)
{
    if (!L) return MRB_OK;
    (void)hipDeviceSynchronize();
    for (int l = 0; l < (L->layout == MRB_LAYOUT_COMPACT ? 1 : 2); ++l)
        for (int ab = 0; ab < 2; ++ab) (void)hipFree(L->pop[l][ab]);
    (void)hipFree(L->d_boxes);
    (void)hipFree(L->d_box_base);
    (void)hipFree(L->d_isf);
    (void)hipFree(L->d_mask);
    for (int ab = 0; ab < 2; ++ab) {
        (void)hipFree(L->d_e16[ab]);
        (void)hipFree(L->d_amax16[ab]);
    }
    (void)hipFree(L->d_ovf16);
    (void)hipFree(L->d_mac);
    (void)hipFree(L->d_bctab);
    (void)hipFree(L->d_md);
    (void)hipFree(L->d_dq);
    (void)hipFree(L->d_q);
    (void)hipFree(L->d_red);
    (void)hipFree(L->d_small);
    for (auto &e : L->ev)
        if (e) (void)hipEventDestroy(e);
    (void)hipFree(L->d_gmap);
    (void)hipFree(L->d_gmap1);
    (void)hipFree(L->d_bcells);
    (void)hipFree(L->d_sweep_q);
    (void)hipFree(L->d_sweep_qx);
    (void)hipFree(L->d_sweep_flags);
    (void)hipFree(L->d_sweep_fault);
    (void)hipFree(L->d_sweep_qt);
    (void)hipFree(L->d_sweep_epoch);
    (void)hipFree(L->d_tiles_c);
    amr_release(L);
    delete L;
    return MRB_OK;
}

int mrb_set_phys(mrb_level *L, const mrb_phys *p)
{
    if (!L || !p) return set_error(MRB_E_ARG, "null argument");
    if (!(p->dt > 0.0) || !(p->mesh_speed > 0.0) || !(p->adiabatic_exponent > 1.0) || !(p->m_bar > 0.0))
        return set_error(MRB_E_ARG, "invalid physics parameters");
    if (shift_locked(L, p->m_bar, p->adiabatic_exponent, L->shift_rho0, L->shift_T0))
        return set_error(MRB_E_STATE, "reduced-precision level: R and gamma must be set before the populations");
    L->phys = *p;
    L->phys_set = true;
    ++L->gen;
    CollideParams &c = L->cp;
    c.nu = p->nu;
    c.alpha = p->alpha;
    c.R = kAirMolarMass / p->m_bar;          // LBM.cpp:639
    c.cv = c.R / (p->adiabatic_exponent - 1.0);  // LBM.cpp:640
    c.dt = p->dt;
    c.ms = p->mesh_speed;
    for (int d = 0; d < 3; ++d) c.inv_dx[d] = L->inv_dx[d];
    compute_shift(L);
    return MRB_OK;
}

int mrb_set_bc(mrb_level *L, const mrb_bc *bc)
{
    if (!L || !bc) return set_error(MRB_E_ARG, "null argument");
    ++L->gen;
    for (int d = 0; d < 3; ++d) {
        for (int lohi = 0; lohi < 2; ++lohi) {
            const int t = bc->bc_type[d + 3 * lohi];
            if (L->periodic[d] && t != MRB_BC_PERIODIC)
                return set_error(MRB_E_ARG, "periodic direction needs bc 0 on both sides");
            if (!L->periodic[d] && t == MRB_BC_PERIODIC)
                return set_error(MRB_E_ARG, "bc 0 (interior) in a non-periodic direction");
            if (t != MRB_BC_PERIODIC && t != MRB_BC_NOSLIP && t != MRB_BC_VELOCITY && t != MRB_BC_OUTFLOW &&
                t != MRB_BC_PRESSURE && t != MRB_BC_SLIPX && t != MRB_BC_SLIPY && t != MRB_BC_SLIPZ)
                return set_error(MRB_E_ARG, "unknown bc type");
        }
    }
    // PRESSURE reads the populations around the fluid cell its ghost feeds
    // (BC.H:247-260), tangential ghosts included, while the same fill writes
    // ghost cells.  Those reads see values that do not depend on the fill
    // order when every tangential axis is periodic (FillBoundary images) or
    // NOSLIP on both sides (the prefill and the wall give the same value),
    // and the face's own axis has two cells or more.
    for (int f = 0; f < 6; ++f) {
        if (bc->bc_type[f] != MRB_BC_PRESSURE) continue;
        const int d = f % 3;
        if (L->ncell[d] < 2) return set_error(MRB_E_UNSUPPORTED, "pressure bc needs two cells or more along its axis");
        for (int t = 0; t < 3; ++t) {
            if (t == d || L->periodic[t]) continue;
            if (bc->bc_type[t] != MRB_BC_NOSLIP || bc->bc_type[t + 3] != MRB_BC_NOSLIP)
                return set_error(MRB_E_UNSUPPORTED,
                                 "pressure bc needs its tangential axes periodic or no-slip on both sides");
        }
    }
    if (bc->velbc_kind < 0 || bc->velbc_kind > 3) return set_error(MRB_E_ARG, "unknown velocity bc kind");
    // compact layout: a Channel or Parabolic inlet face meeting an OUTFLOW
    // face that BCFill applies after it (a higher direction) would read the
    // profile at an edge ghost the OUTFLOW face copies; not supported
    if (L->layout == MRB_LAYOUT_COMPACT && bc->velbc_kind >= 2)
        for (int f = 0; f < 6; ++f)
            for (int g = 0; g < 6; ++g)
                if (bc->bc_type[f] == MRB_BC_VELOCITY && bc->bc_type[g] == MRB_BC_OUTFLOW && g % 3 > f % 3)
                    return set_error(MRB_E_UNSUPPORTED,
                                     "velocity profile face next to a later OUTFLOW face (compact layout)");
    if ((bc->velbc_kind == 1 || bc->velbc_kind == 3) && (bc->vel_dir < 0 || bc->vel_dir > 2))
        return set_error(MRB_E_ARG, "velocity bc dir must be 0, 1 or 2");
    if (bc->velbc_kind == 3 && (bc->normal_dir < 0 || bc->normal_dir > 2))
        return set_error(MRB_E_ARG, "parabolic velocity bc normal_dir must be 0, 1 or 2");
    L->bc = *bc;
    compute_bc_values(L);
    // the FP32-arithmetic collide has no PRESSURE faces: such a level takes
    // the FP64-arithmetic two-pass step (and returns to the FP32 one when
    // its faces change back)
    if (L->storage != MRB_STORAGE_F64) {
        if (L->bcdev.pdir >= 0 && fp32_collide_variant(L->hot)) {
            L->hot_h32 = L->hot;
            L->hot = fp64_counterpart(L->hot);
        } else if (L->bcdev.pdir < 0 && L->hot_h32 >= 0 && L->hot == fp64_counterpart(L->hot_h32)) {
            L->hot = L->hot_h32;
        }
        if (L->bcdev.pdir < 0) L->hot_h32 = -1;
    }
    if (L->bc.velbc_kind >= 2 && L->layout == MRB_LAYOUT_COMPACT) {
        for (int f = 0; f < 6; ++f)
            if (L->bc.bc_type[f] == MRB_BC_VELOCITY && !L->bcdev.tab[f])
                return set_error(MRB_E_HIP, "velocity profile tables could not be allocated");
    }
    return MRB_OK;
}

int mrb_set_shift_state(mrb_level *L, double rho0, double T0)
{
    if (!L) return set_error(MRB_E_ARG, "null argument");
    if (!(rho0 > 0.0) || !(T0 > 0.0)) return set_error(MRB_E_ARG, "rho0 and T0 must be positive");
    const double m_bar = L->phys_set ? L->phys.m_bar : kAirMolarMass;
    const double gamma = L->phys_set ? L->phys.adiabatic_exponent : 5.0 / 3.0;
    if (shift_locked(L, m_bar, gamma, rho0, T0))
        return set_error(MRB_E_STATE, "reduced-precision level: set the shift state before the populations");
    L->shift_rho0 = rho0;
    L->shift_T0 = T0;
    compute_shift(L);
    return MRB_OK;
}

int64_t mrb_fab_size(const mrb_level *L, int arr, int b)
{
    if (!L || b < 0 || b >= L->nbox) return -1;
    const int64_t n = L->boxes[b].ncell;
    return arr == MRB_ARR_ISFLUID ? 2 * n : NQ * n;
}

// Copies one lattice between the host layout (components contiguous) and
// the device buffer of the current state (component stride of the layout).
static int copy_lattice(const mrb_level *L, int arr, void *host, bool to_device)
{
    const size_t es = pop_elem_size(L->storage);
    const int64_t N = L->total_cells;
    void *dev = L->pop[arr][L->cur];
    if (L->layout == MRB_LAYOUT_COMPACT && cell_stride(L) == 2) {
        // pair layout: element (q, cell) of the lattice at dev[(q*N + cell)*2]
        void *tmp = nullptr;
        MRB_HIP(hipMalloc(&tmp, es * NQ * size_t(N)));
        hipError_t e = hipSuccess;
        const int64_t n = NQ * N;
        const bool f64 = es == sizeof(double);
        if (to_device) {
            e = hipMemcpy(tmp, host, es * size_t(n), hipMemcpyHostToDevice);
            if (e == hipSuccess) {
                if (f64)
                    k_strided_copy<double><<<nblocks(n), kBlock>>>((const double *)tmp, 1, (double *)dev, 2, n);
                else
                    k_strided_copy<float><<<nblocks(n), kBlock>>>((const float *)tmp, 1, (float *)dev, 2, n);
                e = hipGetLastError();
            }
        } else {
            if (f64)
                k_strided_copy<double><<<nblocks(n), kBlock>>>((const double *)dev, 2, (double *)tmp, 1, n);
            else
                k_strided_copy<float><<<nblocks(n), kBlock>>>((const float *)dev, 2, (float *)tmp, 1, n);
            e = hipGetLastError();
            if (e == hipSuccess) e = hipMemcpy(host, tmp, es * size_t(n), hipMemcpyDeviceToHost);
        }
        if (e == hipSuccess) e = hipDeviceSynchronize();
        (void)hipFree(tmp);
        if (e != hipSuccess) return hip_fail(e, "copy of the pair layout");
        return MRB_OK;
    }
    if (L->layout == MRB_LAYOUT_COMPACT) {
        const size_t w = es * size_t(N), dp = es * size_t(2 * N);
        MRB_HIP(to_device ? copy_components(dev, dp, host, w, w, hipMemcpyHostToDevice)
                          : copy_components(host, w, dev, dp, w, hipMemcpyDeviceToHost));
    } else {
        const size_t bytes = es * NQ * size_t(N);
        MRB_HIP(to_device ? hipMemcpy(dev, host, bytes, hipMemcpyHostToDevice)
                          : hipMemcpy(host, dev, bytes, hipMemcpyDeviceToHost));
    }
    return MRB_OK;
}

int mrb_upload(mrb_level *L, int arr, const void *host)
{
    if (!L || !host) return set_error(MRB_E_ARG, "null argument");
    if (arr == MRB_ARR_ISFLUID) {
        MRB_HIP(hipMemcpy(L->d_isf, host, sizeof(int32_t) * 2 * L->total_cells, hipMemcpyHostToDevice));
        L->flags_ready = false;
        return MRB_OK;
    }
    if (arr != MRB_ARR_F && arr != MRB_ARR_G) return set_error(MRB_E_ARG, "unknown array");
    L->fresh_depth = 0;
    L->pops_set = true;
    L->prev_valid = false;
    L->md_src = 0;
    const int64_t n = NQ * L->total_cells;
    const double *h = static_cast<const double *>(host);
    if (L->storage == MRB_STORAGE_F64) return copy_lattice(L, arr, const_cast<void *>(host), true);
    // reduced precision: subtract the rest-state shift and round
    const double *sh = arr == MRB_ARR_F ? L->fshift : L->gshift;
    const int64_t N = L->total_cells;
    if (L->storage == MRB_STORAGE_F32_SHIFTED) {
        std::vector<float> buf(static_cast<size_t>(n));
        for (int q = 0; q < NQ; ++q)
            for (int64_t c = 0; c < N; ++c) buf[size_t(q * N + c)] = float(h[q * N + c] - sh[q]);
        return copy_lattice(L, arr, buf.data(), true);
    }
    // FP16 with per-box scale: scaled on the device (the box maxima skip
    // solid cells, so the flags come first)
    if (!L->flags_ready) return set_error(MRB_E_STATE, "fp16 storage: set the flags before the populations");
    double *d = nullptr;
    MRB_HIP(hipMalloc(&d, sizeof(double) * size_t(n)));
    hipError_t e = hipMemcpy(d, h, sizeof(double) * size_t(n), hipMemcpyHostToDevice);
    int rc = MRB_OK;
    if (e == hipSuccess) rc = f16_pack(L, arr, d, N, nullptr);
    if (e == hipSuccess && rc == MRB_OK) e = hipDeviceSynchronize();
    (void)hipFree(d);
    if (e != hipSuccess) return hip_fail(e, "fp16 upload");
    return rc;
}

int mrb_download(const mrb_level *L, int arr, void *host)
{
    if (!L || !host) return set_error(MRB_E_ARG, "null argument");
    MRB_HIP(hipDeviceSynchronize());
    if (arr == MRB_ARR_ISFLUID) {
        MRB_HIP(hipMemcpy(host, L->d_isf, sizeof(int32_t) * 2 * L->total_cells, hipMemcpyDeviceToHost));
        return MRB_OK;
    }
    if (arr != MRB_ARR_F && arr != MRB_ARR_G) return set_error(MRB_E_ARG, "unknown array");
    const int64_t n = NQ * L->total_cells;
    if (L->storage == MRB_STORAGE_F64) return copy_lattice(L, arr, host, false);
    double *h = static_cast<double *>(host);
    const double *sh = arr == MRB_ARR_F ? L->fshift : L->gshift;
    const int64_t N = L->total_cells;
    if (L->storage == MRB_STORAGE_F32_SHIFTED) {
        std::vector<float> buf(static_cast<size_t>(n));
        const int rc = copy_lattice(L, arr, buf.data(), false);
        if (rc) return rc;
        for (int q = 0; q < NQ; ++q)
            for (int64_t c = 0; c < N; ++c) h[q * N + c] = double(buf[size_t(q * N + c)]) + sh[q];
        return MRB_OK;
    }
    double *d = nullptr;
    MRB_HIP(hipMalloc(&d, sizeof(double) * size_t(n)));
    int rc = f16_unpack(L, arr, d, nullptr);
    hipError_t e = rc == MRB_OK ? hipDeviceSynchronize() : hipSuccess;
    if (rc == MRB_OK && e == hipSuccess) e = hipMemcpy(h, d, sizeof(double) * size_t(n), hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess) return hip_fail(e, "fp16 download");
    return rc;
}

// comp0 from the body, then the eb_boundary ring (LBM.cpp:1213-1262)
static int flags_of_body(mrb_level *L, const BodyArgs &S)
{
    k_flags_body<<<nblocks(L->total_cells), kBlock>>>(L->d_isf, L->d_boxes, L->d_box_base, L->nbox,
                                                      L->total_cells, S);
    MRB_HIP(hipGetLastError());
    if (L->layout == MRB_LAYOUT_FAB) {
        int rc = fill_boundary(L, L->d_isf, 2, true, false, nullptr);
        if (rc) return rc;
        k_flags_ring<<<nblocks(L->total_cells), kBlock>>>(L->d_isf, L->d_boxes, L->d_box_base, L->nbox,
                                                          L->total_cells);
        MRB_HIP(hipGetLastError());
        rc = fill_boundary(L, L->d_isf, 2, true, false, nullptr);
        if (rc) return rc;
    } else if (!S.fluid_inside) {
        // ghost cells beyond non-periodic faces must be fluid for the
        // compact layout's inline boundary handling: the body may not reach
        // within two cells of such a face (a sufficient check on the
        // centre's distance, along the axis on the cylinder's half height)
        for (int d = 0; d < 3; ++d) {
            if (L->periodic[d]) continue;
            const double extent = d == S.axis ? S.half_height : S.r;
            for (int side = 0; side < 2; ++side) {
                const double dist = side == 0 ? S.c[d] - L->prob_lo[d] : L->prob_hi[d] - S.c[d];
                if (dist - extent <= 2.0 * L->dx[d])
                    return set_error(MRB_E_UNSUPPORTED, "compact layout: EB body reaches a non-periodic face");
            }
        }
    }
    MRB_HIP(hipDeviceSynchronize());
    return mrb_finalize_flags(L);
}

static BodyArgs body_args(const mrb_level *L, const double center[3], double radius, int has_fluid_inside)
{
    BodyArgs S{};
    for (int d = 0; d < 3; ++d) {
        S.plo[d] = L->prob_lo[d];
        S.dx[d] = L->dx[d];
        S.c[d] = center[d];
    }
    S.r = radius;
    S.axis = -1;
    S.fluid_inside = has_fluid_inside;
    return S;
}

int mrb_compute_flags_sphere(mrb_level *L, const double center[3], double radius, int has_fluid_inside)
{
    if (!L || !center) return set_error(MRB_E_ARG, "null argument");
    return flags_of_body(L, body_args(L, center, radius, has_fluid_inside));
}

int mrb_compute_flags_cylinder(
    mrb_level *L, const double center[3], double radius, double height, int direction, int has_fluid_inside)
{
    if (!L || !center) return set_error(MRB_E_ARG, "null argument");
    if (direction < 0 || direction > 2) return set_error(MRB_E_ARG, "cylinder direction must be 0, 1 or 2");
    if (!(radius > 0.0) || !(height > 0.0)) return set_error(MRB_E_ARG, "cylinder radius and height must be > 0");
    BodyArgs S = body_args(L, center, radius, has_fluid_inside);
    S.axis = direction;
    S.half_height = 0.5 * height;
    return flags_of_body(L, S);
}

int mrb_stencil(int evs[NQ * 3], double weights[NQ], int bounce_dirs[NQ])
{
    if (!evs || !weights || !bounce_dirs) return set_error(MRB_E_ARG, "null argument");
    for (int q = 0; q < NQ; ++q) {
        for (int d = 0; d < 3; ++d) evs[3 * q + d] = kEV[q][d];
        weights[q] = weight(q);
        bounce_dirs[q] = kBounce[q];
    }
    return MRB_OK;
}

int mrb_finalize_flags(mrb_level *L)
{
    if (!L) return set_error(MRB_E_ARG, "null argument");
    if (L->layout == MRB_LAYOUT_COMPACT) {
        // eb_boundary ring with periodic wrap (ghosts beyond physical faces fluid)
        std::vector<int32_t> h(size_t(2 * L->total_cells));
        MRB_HIP(hipMemcpy(h.data(), L->d_isf, sizeof(int32_t) * h.size(), hipMemcpyDeviceToHost));
        const int nx = L->ncell[0], ny = L->ncell[1], nz = L->ncell[2];
        const int64_t N = L->total_cells;
        auto fluid = [&](int i, int j, int k) -> int {
            int s[3] = {i, j, k};
            const int n[3] = {nx, ny, nz};
            for (int d = 0; d < 3; ++d) {
                if (s[d] < 0 || s[d] >= n[d]) {
                    if (!L->periodic[d]) return 1;
                    s[d] = (s[d] + n[d]) % n[d];
                }
            }
            return h[size_t((int64_t(s[2]) * ny + s[1]) * nx + s[0])];
        };
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) {
                    const int64_t c = (int64_t(k) * ny + j) * nx + i;
                    bool all_covered = true;
                    all_covered &= fluid(i - 1, j, k) == 0 && fluid(i + 1, j, k) == 0;
                    all_covered &= fluid(i, j - 1, k) == 0 && fluid(i, j + 1, k) == 0;
                    all_covered &= fluid(i, j, k - 1) == 0 && fluid(i, j, k + 1) == 0;
                    h[size_t(N + c)] = (all_covered || h[size_t(c)] == 1) ? 0 : 1;
                }
        MRB_HIP(hipMemcpy(L->d_isf, h.data(), sizeof(int32_t) * h.size(), hipMemcpyHostToDevice));
        const int rc = compact_build_mask(L, nullptr);
        if (rc) return rc;
        MRB_HIP(hipDeviceSynchronize());
    } else {
        const int rc = fab_build_mask(L, nullptr);
        if (rc) return rc;
        MRB_HIP(hipDeviceSynchronize());
    }
    L->flags_ready = true;
    return MRB_OK;
}

// ---- initial state -------------------------------------------------------

static void parallel_for(int64_t n, const std::function<void(int64_t, int64_t)> &fn)
{
    const int nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    const int64_t chunk = (n + nt - 1) / nt;
    for (int t = 0; t < nt; ++t) {
        const int64_t a = t * chunk, b = std::min(n, a + chunk);
        if (a >= b) break;
        th.emplace_back(fn, a, b);
    }
    for (auto &x : th) x.join();
}

// Evaluates f and g for every fab cell on the host from a per-cell state
// function and uploads them; solid cells get 0 (LBM.cpp:1287-1295).
static int init_from_state(
    mrb_level *L,
    const std::function<void(int64_t, int, int, int, double &, double *, double &, double &, double &)> &state)
{
    if (!L->flags_ready) return set_error(MRB_E_STATE, "flags must be set before the initial state");
    const int64_t T = L->total_cells;
    std::vector<int32_t> isf(size_t(2 * T));
    MRB_HIP(hipMemcpy(isf.data(), L->d_isf, sizeof(int32_t) * isf.size(), hipMemcpyDeviceToHost));
    std::vector<double> f(size_t(NQ * T)), g(size_t(NQ * T));
    for (int b = 0; b < L->nbox; ++b) {
        const BoxDev &B = L->boxes[b];
        parallel_for(B.ncell, [&](int64_t a, int64_t e) {
            for (int64_t loc = a; loc < e; ++loc) {
                const int64_t nxy = int64_t(B.n[0]) * B.n[1];
                const int k = int(loc / nxy) + B.lo[2];
                const int j = int((loc % nxy) / B.n[0]) + B.lo[1];
                const int i = int(loc % B.n[0]) + B.lo[0];
                double rho = 1.0, vel[3] = {0.0, 0.0, 0.0}, T0 = 1.0 / 3.0, R = 1.0, gamma = 1.667;
                state(B.base + loc, i, j, k, rho, vel, T0, R, gamma);
                const bool solid = isf[size_t(B.base * 2 + loc)] == 0;
                for (int q = 0; q < NQ; ++q) {
                    const size_t o = size_t(B.base * NQ + q * B.ncell + loc);
                    f[o] = solid ? 0.0 : set_equilibrium_value(rho, vel, R * T0, q);
                    g[o] = solid ? 0.0 : thermal_g_equilibrium(rho, vel, T0, R, gamma, q);
                }
            }
        });
    }
    int rc = mrb_upload(L, MRB_ARR_F, f.data());
    if (rc) return rc;
    rc = mrb_upload(L, MRB_ARR_G, g.data());
    if (rc) return rc;
    if (L->layout == MRB_LAYOUT_FAB) {
        rc = fill_boundary(L, L->pop[0][L->cur], NQ, false, false, nullptr);
        if (rc) return rc;
        rc = fill_boundary(L, L->pop[1][L->cur], NQ, false, false, nullptr);
        if (rc) return rc;
        MRB_HIP(hipDeviceSynchronize());
    }
    return MRB_OK;
}

int mrb_init_taylorgreen(mrb_level *L, double rho0, double v0, const double omega[3])
{
    if (L) {
        L->pops_set = true;
        L->prev_valid = false;
        L->md_src = 0;
    }
    if (!L || !omega) return set_error(MRB_E_ARG, "null argument");
    const double Lw = 1.0 / M_PI;
    const double o0 = omega[0], o1 = omega[1], o2 = omega[2];
    return init_from_state(L, [&](int64_t, int i, int j, int k, double &rho, double *vel, double &T, double &R,
                                  double &gamma) {
        // ic::TaylorGreen::DeviceOp::thermal_initialize (IC.H:119-153)
        const double x = L->prob_lo[0] + (i + 0.5) * L->dx[0];
        const double y = L->prob_lo[1] + (j + 0.5) * L->dx[1];
        const double z = L->prob_lo[2] + (k + 0.5) * L->dx[2];
        rho = rho0 + rho0 * v0 * v0 / 16.0 * (std::cos(2.0 * o0 * x / Lw) + std::cos(2.0 * o1 * y / Lw)) *
                         (std::cos(2.0 * o2 * z / Lw) + 2.0);
        vel[0] = v0 * std::sin(o0 * x / Lw) * std::cos(o1 * y / Lw) * std::cos(o2 * z / Lw);
        vel[1] = -v0 * std::cos(o0 * x / Lw) * std::sin(o1 * y / Lw) * std::cos(o2 * z / Lw);
        vel[2] = 0.0;
        T = 1.0 / 3.0;
        R = kAirMolarMass / kAirMolarMass;
        gamma = 5.0 / 3.0;
    });
}

int mrb_init_constant(
    mrb_level *L, double density, const double mach[3], double temperature, double adiabatic_exponent,
    double m_bar)
{
    if (L) {
        L->pops_set = true;
        L->prev_valid = false;
        L->md_src = 0;
    }
    if (!L || !mach) return set_error(MRB_E_ARG, "null argument");
    if (!L->flags_ready) return set_error(MRB_E_STATE, "flags must be set before the initial state");
    // ic::Constant (IC.cpp:4-32): velocity = mach * speed_of_sound_ref
    const double sos = std::sqrt(adiabatic_exponent * (kAirMolarMass / m_bar) * temperature);
    double vel[3];
    for (int n = 0; n < 3; ++n) vel[n] = mach[n] * sos;
    const double R = kAirMolarMass / m_bar;
    double fv[NQ], gv[NQ];
    for (int q = 0; q < NQ; ++q) {
        fv[q] = set_equilibrium_value(density, vel, R * temperature, q);
        gv[q] = thermal_g_equilibrium(density, vel, temperature, R, adiabatic_exponent, q);
    }
    if (L->storage != MRB_STORAGE_F64) {
        // reduced-precision storage: build on the host through mrb_upload
        return init_from_state(L, [&](int64_t, int, int, int, double &rho, double *v, double &T, double &Rr,
                                      double &gamma) {
            rho = density;
            v[0] = vel[0];
            v[1] = vel[1];
            v[2] = vel[2];
            T = temperature;
            Rr = R;
            gamma = adiabatic_exponent;
        });
    }
    double *dval = L->d_small + 2 * NQ;
    for (int l = 0; l < 2; ++l) {
        MRB_HIP(hipMemcpy(dval, l == 0 ? fv : gv, sizeof(double) * NQ, hipMemcpyHostToDevice));
        k_init_uniform<<<nblocks(L->total_cells), kBlock>>>((double *)L->pop[l][L->cur], L->d_isf, L->d_boxes,
                                                            L->d_box_base, L->nbox, L->total_cells, dval,
                                                            comp_stride(L), cell_stride(L));
        MRB_HIP(hipGetLastError());
    }
    MRB_HIP(hipDeviceSynchronize());
    return MRB_OK;
}

int mrb_init_fields(
    mrb_level *L, const double *rho, const double *u, const double *v, const double *w, const double *T,
    double adiabatic_exponent, double m_bar)
{
    if (L) {
        L->pops_set = true;
        L->prev_valid = false;
        L->md_src = 0;
    }
    if (!L || !rho || !u || !v || !w || !T) return set_error(MRB_E_ARG, "null argument");
    if (!L->flags_ready) return set_error(MRB_E_STATE, "flags must be set before the initial state");
    if (!(m_bar > 0.0) || !(adiabatic_exponent > 1.0)) return set_error(MRB_E_ARG, "invalid gas parameters");
    // ic::Initializer::initialize (IC.H:460-525) with the per-cell state
    // given for every fab cell (box order, local index x fastest); the
    // equilibria are evaluated on the device with the same expressions
    const int64_t n = L->total_cells;
    double *d = nullptr;
    MRB_HIP(hipMalloc(&d, sizeof(double) * 5 * size_t(n)));
    const double *src[5] = {rho, u, v, w, T};
    hipError_t e = hipSuccess;
    for (int k = 0; k < 5 && e == hipSuccess; ++k)
        e = hipMemcpy(d + k * n, src[k], sizeof(double) * size_t(n), hipMemcpyHostToDevice);
    if (e == hipSuccess && L->storage == MRB_STORAGE_F16_SCALED) {
        // FP64 equilibria into a temporary compact FP64 pair of lattices
        // (component stride 2N, zero shift), then scaled per box into FP16
        const double R = kAirMolarMass / m_bar;
        double *t = nullptr;
        e = hipMalloc(&t, sizeof(double) * (2 * NQ * size_t(n) + 2 * NQ));
        if (e == hipSuccess) e = hipMemset(t + 2 * NQ * n, 0, sizeof(double) * 2 * NQ);
        if (e == hipSuccess) {
            k_init_fields<double><<<nblocks(n), kBlock>>>(t, t + n, L->d_isf, L->d_boxes, L->d_box_base, L->nbox, n,
                                                          d, R, adiabatic_exponent, t + 2 * NQ * n, 2 * n, 1);
            e = hipGetLastError();
        }
        int rc = MRB_OK;
        if (e == hipSuccess) rc = f16_pack(L, 0, t, 2 * n, nullptr);
        if (e == hipSuccess && rc == MRB_OK) rc = f16_pack(L, 1, t + n, 2 * n, nullptr);
        if (e == hipSuccess) e = hipDeviceSynchronize();
        (void)hipFree(t);
        (void)hipFree(d);
        if (e != hipSuccess) return hip_fail(e, "mrb_init_fields");
        L->fresh_depth = 0;
        return rc;
    }
    if (e == hipSuccess) {
        // d_small[0 .. 2NQ) holds the storage shift (zero for FP64)
        const double R = kAirMolarMass / m_bar;
        if (L->storage == MRB_STORAGE_F64)
            k_init_fields<double><<<nblocks(n), kBlock>>>((double *)L->pop[0][L->cur], (double *)L->pop[1][L->cur],
                                                          L->d_isf, L->d_boxes, L->d_box_base, L->nbox, n, d, R,
                                                          adiabatic_exponent, L->d_small, comp_stride(L), cell_stride(L));
        else
            k_init_fields<float><<<nblocks(n), kBlock>>>((float *)L->pop[0][L->cur], (float *)L->pop[1][L->cur],
                                                         L->d_isf, L->d_boxes, L->d_box_base, L->nbox, n, d, R,
                                                         adiabatic_exponent, L->d_small, comp_stride(L), cell_stride(L));
        e = hipGetLastError();
        if (e == hipSuccess) e = hipDeviceSynchronize();
    }
    (void)hipFree(d);
    if (e != hipSuccess) return hip_fail(e, "mrb_init_fields");
    L->fresh_depth = 0;
    if (L->layout == MRB_LAYOUT_FAB) {
        const int rc = fab_fill_boundary_depth(L, L->ng, nullptr);
        if (rc) return rc;
        MRB_HIP(hipDeviceSynchronize());
    }
    return MRB_OK;
}

// ---- reference-layout stages ---------------------------------------------

int mrb_physbc(mrb_level *L, int lattice, void *stream)
{
    if (!L) return set_error(MRB_E_ARG, "null argument");
    if (L->layout != MRB_LAYOUT_FAB) return set_error(MRB_E_STATE, "physbc needs the fab layout");
    if (lattice < 0 || lattice > 1) return set_error(MRB_E_ARG, "lattice must be 0 or 1");
    if (L->n_bcells == 0) return MRB_OK;
    hipStream_t s = as_stream(stream);
    for (size_t p = 1; p < L->bcell_off.size(); ++p) {
        const int64_t n = L->bcell_off[p] - L->bcell_off[p - 1];
        if (n == 0) continue;
        k_physbc_fab<<<nblocks(n), kBlock, 0, s>>>((double *)L->pop[lattice][L->cur], L->d_boxes, L->d_box_base,
                                                   L->nbox, L->d_bcells + L->bcell_off[p - 1], n, 0, 0, 0,
                                                   L->ncell[0] - 1, L->ncell[1] - 1, L->ncell[2] - 1, L->bcdev,
                                                   lattice);
        MRB_HIP(hipGetLastError());
    }
    return MRB_OK;
}

int mrb_fill_ghosts(mrb_level *L, int lattice, void *stream)
{
    if (!L) return set_error(MRB_E_ARG, "null argument");
    if (L->layout != MRB_LAYOUT_FAB) return set_error(MRB_E_STATE, "fill_ghosts needs the fab layout");
    if (lattice < 0 || lattice > 1) return set_error(MRB_E_ARG, "lattice must be 0 or 1");
    // The bounce prefill of FillPatchOps.H:92-110 is overwritten in full by
    // FillBoundary and the boundary functor, so it is not executed.
    // FillBoundary is skipped when all ghost layers already hold the
    // owners' current values (after a collide); the boundary functor
    // always runs.
    if (L->fresh_depth < L->ng) {
        const int rc = fill_boundary(L, L->pop[lattice][L->cur], NQ, false, false, as_stream(stream));
        if (rc) return rc;
        if (lattice == 1) L->fresh_depth = L->ng;
    }
    return mrb_physbc(L, lattice, stream);
}

int mrb_stream(mrb_level *L, int lattice, void *stream)
{
    if (!L) return set_error(MRB_E_ARG, "null argument");
    if (L->layout != MRB_LAYOUT_FAB) return set_error(MRB_E_STATE, "mrb_stream needs the fab layout");
    if (lattice < 0 || lattice > 1) return set_error(MRB_E_ARG, "lattice must be 0 or 1");
    if (!L->flags_ready) return set_error(MRB_E_STATE, "flags not set");
    hipStream_t s = as_stream(stream);
    // the other lattice keeps its buffer index: stream lattice by lattice
    // through the spare buffer and copy back so that cur stays shared
    double *fin = (double *)L->pop[lattice][L->cur];
    double *fout = (double *)L->pop[lattice][1 - L->cur];
    k_stream_fab<<<nblocks(L->total_cells), kBlock, 0, s>>>(fin, fout, L->d_isf, L->d_boxes, L->d_box_base,
                                                            L->nbox, L->total_cells);
    MRB_HIP(hipGetLastError());
    std::swap(L->pop[lattice][0], L->pop[lattice][1]);
    L->fresh_depth = 0;
    if (L->md_src == 2) L->md_src = -1;
    ++L->gen;
    return fill_boundary(L, L->pop[lattice][L->cur], NQ, false, false, s);
}

int mrb_collide(mrb_level *L, void *stream)
{
    if (!L) return set_error(MRB_E_ARG, "null argument");
    if (L->layout != MRB_LAYOUT_FAB) return set_error(MRB_E_STATE, "mrb_collide needs the fab layout");
    if (!L->flags_ready) return set_error(MRB_E_STATE, "flags not set");
    hipStream_t s = as_stream(stream);
    double *f = (double *)L->pop[0][L->cur];
    double *g = (double *)L->pop[1][L->cur];
    const unsigned nb = nblocks(L->total_cells);
    k_moments_fab<<<nb, kBlock, 0, s>>>(f, g, L->d_isf, L->d_md, L->d_boxes, L->d_box_base, L->nbox,
                                        L->total_cells, L->cp);
    MRB_HIP(hipGetLastError());
    int rc = fill_boundary(L, L->d_md, 19, false, true, s);
    if (rc) return rc;
    k_dq_fab<<<nb, kBlock, 0, s>>>(L->d_md, L->d_isf, L->d_dq, L->d_boxes, L->d_box_base, L->nbox,
                                   L->total_cells, L->ncell[0], L->ncell[1], L->ncell[2], L->cp);
    MRB_HIP(hipGetLastError());
    k_eqrelax_fab<<<nb, kBlock, 0, s>>>(f, g, L->d_md, L->d_dq, L->d_isf, L->d_boxes, L->d_box_base,
                                        L->nbox, L->total_cells, L->cp);
    MRB_HIP(hipGetLastError());
    rc = fill_boundary(L, f, NQ, false, false, s);
    if (rc) return rc;
    rc = fill_boundary(L, g, NQ, false, false, s);
    if (rc) return rc;
    L->fresh_depth = L->ng;
    L->md_src = 1;
    return MRB_OK;
}

// ---- hot path --------------------------------------------------------------

int mrb_stream_collide(mrb_level *L, void *stream)
{
    if (!L) return set_error(MRB_E_ARG, "null argument");
    if (!L->flags_ready) return set_error(MRB_E_STATE, "flags not set");
    if (L->layout == MRB_LAYOUT_COMPACT) return compact_stream_collide(L, as_stream(stream));
    // reference layout: ghost fill (FillPatchOps::fillpatch), then stream
    // and collide as one batched step over every box
    int rc;
    if (L->ng < 2 || L->amr_crse) {
        // one ghost layer, or a linked fine level (its uncovered ghosts
        // belong to the two-level schedule): the stages one by one
        for (int l = 0; l < 2; ++l)
            if ((rc = mrb_fill_ghosts(L, l, stream))) return rc;
        for (int l = 0; l < 2; ++l)
            if ((rc = mrb_stream(L, l, stream))) return rc;
        return mrb_collide(L, stream);
    }
    // the batched step pulls ring 1 only: FillBoundary one layer deep
    if (L->fresh_depth < 1 && (rc = fab_fill_boundary_depth(L, 1, as_stream(stream)))) return rc;
    for (int l = 0; l < 2; ++l)
        if ((rc = mrb_physbc(L, l, stream))) return rc;
    return fab_batched_step(L, nullptr, 1, as_stream(stream));
}

int mrb_advance(mrb_level *L, int nsteps, void *stream)
{
    for (int n = 0; n < nsteps; ++n) {
        const int rc = mrb_stream_collide(L, stream);
        if (rc) return rc;
    }
    return MRB_OK;
}

int mrb_advance_timed(mrb_level *L, int nsteps, void *stream, double *kernel_ms, int *nslots)
{
    if (!L || !kernel_ms || !nslots) return set_error(MRB_E_ARG, "null argument");
    if (L->layout != MRB_LAYOUT_COMPACT) return set_error(MRB_E_STATE, "needs the compact layout");
    if (!L->flags_ready) return set_error(MRB_E_STATE, "flags not set");
    hipStream_t s = as_stream(stream);
    for (int k = 0; k < 4; ++k) kernel_ms[k] = 0.0;
    int ns = 0;
    for (int n = 0; n < nsteps; ++n) {
        const int rc = compact_stream_collide_events(L, s, ns);
        if (rc) return rc;
        MRB_HIP(hipEventSynchronize(L->ev[ns]));
        for (int k = 0; k < ns; ++k) {
            float ms = 0.f;
            MRB_HIP(hipEventElapsedTime(&ms, L->ev[k], L->ev[k + 1]));
            kernel_ms[k] += ms;
        }
    }
    *nslots = ns;
    return MRB_OK;
}

int mrb_hot_kernel_count(void) { return kNumTwoPassVariants + kNumFusedVariants; }

const char *mrb_hot_kernel_variant_name(int index) { return compact_variant_name(index); }

int mrb_set_hot_kernel(mrb_level *L, int index)
{
    if (!L) return set_error(MRB_E_ARG, "null argument");
    if (index < 0 || index >= mrb_hot_kernel_count()) return set_error(MRB_E_ARG, "no such kernel variant");
    if (L->layout != MRB_LAYOUT_COMPACT) return set_error(MRB_E_STATE, "needs the compact layout");
    L->hot = index;
    L->hot_default = false;
    L->hot_h32 = -1;
    return MRB_OK;
}

int mrb_autotune_count(const mrb_level *L)
{
    if (!L) return 0;
    return L->layout == MRB_LAYOUT_COMPACT ? mrb_hot_kernel_count() : kNumFabBlockChoices;
}

const char *mrb_autotune_name(const mrb_level *L, int index)
{
    static thread_local std::string name;
    if (!L || index < 0 || index >= mrb_autotune_count(L)) return "";
    if (L->layout == MRB_LAYOUT_COMPACT) return mrb_hot_kernel_variant_name(index);
    const int b = kFabBlockChoices[index];
    name = b == 0 ? std::string("fab_2pass_block_per_tile") : "fab_2pass_persistent_" + std::to_string(b);
    return name.c_str();
}

static int autotune_fab(mrb_level *L, int reps, hipStream_t s, double *ms_out, int *best)
{
    if (L->ng < 2) return set_error(MRB_E_UNSUPPORTED, "the batched fab step needs ng >= 2");
    int arg = 0;
    double best_ms = 1e300;
    for (int v = 0; v < kNumFabBlockChoices; ++v) {
        L->fab_blocks = kFabBlockChoices[v];
        int rc = fab_batched_passes(L, nullptr, s);  // warm-up
        if (rc) return rc;
        MRB_HIP(hipEventRecord(L->ev[3], s));
        for (int r = 0; r < reps; ++r)
            if ((rc = fab_batched_passes(L, nullptr, s))) return rc;
        MRB_HIP(hipEventRecord(L->ev[4], s));
        MRB_HIP(hipEventSynchronize(L->ev[4]));
        float ms = 0.f;
        MRB_HIP(hipEventElapsedTime(&ms, L->ev[3], L->ev[4]));
        const double mean = double(ms) / reps;
        if (ms_out) ms_out[v] = mean;
        if (mean < best_ms) {
            best_ms = mean;
            arg = v;
        }
    }
    L->fab_blocks = kFabBlockChoices[arg];
    ++L->gen;
    if (best) *best = arg;
    return MRB_OK;
}

int mrb_autotune(mrb_level *L, int reps, void *stream, double *ms_out, int *best)
{
    if (!L) return set_error(MRB_E_ARG, "null argument");
    if (reps < 1) return set_error(MRB_E_ARG, "reps must be positive");
    if (L->layout == MRB_LAYOUT_FAB) {
        if (!L->flags_ready) return set_error(MRB_E_STATE, "flags not set");
        return autotune_fab(L, reps, as_stream(stream), ms_out, best);
    }
    if (!L->flags_ready) return set_error(MRB_E_STATE, "flags not set");
    if (reps < 1) return set_error(MRB_E_ARG, "reps must be positive");
    hipStream_t s = as_stream(stream);
    const int n = mrb_hot_kernel_count();
    int arg = L->hot;
    double best_ms = 1e300;
    // reduced-precision levels: the candidates keep the arithmetic of the
    // current step (the FP32- and FP64-arithmetic collides give different
    // results); the FP16 saturation counter is restored afterwards, so that
    // the timing launches do not add to it
    const bool f16 = L->storage == MRB_STORAGE_F16_SCALED;
    const bool reduced = L->storage != MRB_STORAGE_F64;
    uint32_t ovf_saved = 0;
    if (f16) MRB_HIP(hipMemcpyAsync(&ovf_saved, L->d_ovf16, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    if (f16) MRB_HIP(hipStreamSynchronize(s));
    // every candidate is timed in two rounds and keeps its faster mean, so
    // that a clock ramp or a transient on the device does not decide; the
    // current variant runs first as a warm-up of the device
    std::vector<double> mean(size_t(n), -1.0);
    for (int r = 0; r < reps; ++r) {
        const int rc = compact_variant_launch(L, L->hot, s);
        if (rc && rc != MRB_E_UNSUPPORTED) return rc;
    }
    for (int round = 0; round < 2; ++round) {
        for (int v = 0; v < n; ++v) {
            auto launch = [&]() { return compact_variant_launch(L, v, s); };
            if (reduced && fp32_collide_variant(v) != fp32_collide_variant(L->hot)) continue;
            if (round == 1 && mean[size_t(v)] < 0.0) continue;  // not available for this level
            int rc = launch();  // warm-up
            if (rc == MRB_E_UNSUPPORTED) continue;  // storage, boundary types, occupancy
            if (rc) return rc;
            MRB_HIP(hipEventRecord(L->ev[3], s));
            for (int r = 0; r < reps; ++r)
                if ((rc = launch())) return rc;
            MRB_HIP(hipEventRecord(L->ev[4], s));
            MRB_HIP(hipEventSynchronize(L->ev[4]));
            float ms = 0.f;
            MRB_HIP(hipEventElapsedTime(&ms, L->ev[3], L->ev[4]));
            const double m = double(ms) / reps;
            if (mean[size_t(v)] < 0.0 || m < mean[size_t(v)]) mean[size_t(v)] = m;
        }
    }
    for (int v = 0; v < n; ++v) {
        if (ms_out) ms_out[v] = mean[size_t(v)];
        if (mean[size_t(v)] >= 0.0 && mean[size_t(v)] < best_ms) {
            best_ms = mean[size_t(v)];
            arg = v;
        }
    }
    if (f16) {
        MRB_HIP(hipMemcpyAsync(L->d_ovf16, &ovf_saved, sizeof(uint32_t), hipMemcpyHostToDevice, s));
        MRB_HIP(hipStreamSynchronize(s));
    }
    L->hot = arg;
    L->hot_default = false;
    L->hot_h32 = -1;
    if (best) *best = arg;
    return MRB_OK;
}

int mrb_stream_compact(mrb_level *L, void *stream)
{
    if (!L) return set_error(MRB_E_ARG, "null argument");
    if (L->layout != MRB_LAYOUT_COMPACT) return set_error(MRB_E_STATE, "needs the compact layout");
    if (!L->flags_ready) return set_error(MRB_E_STATE, "flags not set");
    return compact_stream_only(L, as_stream(stream));
}

}  // extern "C"

namespace mrb {

unsigned conservation_blocks(const mrb_level *L) { return std::min<unsigned>(kConservationBlocks, nblocks(L->total_cells)); }

int conservation_partials(const mrb_level *L, double *part, hipStream_t s)
{
    const unsigned nb = conservation_blocks(L);
    // shift tables live at d_small[0..2*NQ) (zero for FP64 storage)
    const double *shifts = L->d_small;
    const int64_t nxy = int64_t(L->ncell[0]) * L->ncell[1];
    if (L->storage == MRB_STORAGE_F64) {
        k_conservation<double><<<nb, kBlock, 0, s>>>(
            (const double *)L->pop[0][L->cur], (const double *)L->pop[1][L->cur], L->d_isf, L->d_boxes,
            L->d_box_base, L->nbox, L->total_cells, shifts, shifts + NQ, part, comp_stride(L), cell_stride(L),
            nullptr, nxy);
    } else if (L->storage == MRB_STORAGE_F32_SHIFTED) {
        k_conservation<float><<<nb, kBlock, 0, s>>>(
            (const float *)L->pop[0][L->cur], (const float *)L->pop[1][L->cur], L->d_isf, L->d_boxes,
            L->d_box_base, L->nbox, L->total_cells, shifts, shifts + NQ, part, comp_stride(L), cell_stride(L),
            nullptr, nxy);
    } else {
        k_conservation<_Float16><<<nb, kBlock, 0, s>>>(
            (const _Float16 *)L->pop[0][L->cur], (const _Float16 *)L->pop[1][L->cur], L->d_isf, L->d_boxes,
            L->d_box_base, L->nbox, L->total_cells, shifts, shifts + NQ, part, comp_stride(L), cell_stride(L),
            L->d_e16[L->cur], nxy);
    }
    MRB_HIP(hipGetLastError());
    return MRB_OK;
}

}  // namespace mrb

extern "C" {

int mrb_reduce_conservation(const mrb_level *L, double out[5], void *stream)
{
    if (!L || !out) return set_error(MRB_E_ARG, "null argument");
    hipStream_t s = as_stream(stream);
    const unsigned nb = conservation_blocks(L);
    int rc = conservation_partials(L, L->d_red, s);
    if (rc) return rc;
    std::vector<double> part(size_t(5 * nb));
    MRB_HIP(hipMemcpyAsync(part.data(), L->d_red, sizeof(double) * part.size(), hipMemcpyDeviceToHost, s));
    MRB_HIP(hipStreamSynchronize(s));
    for (int n = 0; n < 5; ++n) out[n] = 0.0;
    for (unsigned b = 0; b < nb; ++b)
        for (int n = 0; n < 5; ++n) out[n] += part[size_t(5 * b + n)];
    return MRB_OK;
}

int mrb_eb_force(const mrb_level *L, double out[3], void *stream)
{
    if (!L || !out) return set_error(MRB_E_ARG, "null argument");
    if (!L->flags_ready) return set_error(MRB_E_STATE, "flags not set");
    hipStream_t s = as_stream(stream);
    const unsigned nb = std::min<unsigned>(4096u, nblocks(L->total_cells));
    const double *shifts = L->d_small;  // zero for FP64 storage
    const int compact = L->layout == MRB_LAYOUT_COMPACT;
    if (L->storage == MRB_STORAGE_F64) {
        k_eb_force<double><<<nb, kBlock, 0, s>>>(
            (const double *)L->pop[0][L->cur], L->d_isf, L->d_boxes, L->d_box_base, L->nbox, L->total_cells,
            shifts, L->d_red, comp_stride(L), cell_stride(L), compact, L->ncell[0], L->ncell[1], L->ncell[2],
            L->periodic[0], L->periodic[1], L->periodic[2]);
    } else if (L->storage == MRB_STORAGE_F32_SHIFTED) {
        k_eb_force<float><<<nb, kBlock, 0, s>>>(
            (const float *)L->pop[0][L->cur], L->d_isf, L->d_boxes, L->d_box_base, L->nbox, L->total_cells,
            shifts, L->d_red, comp_stride(L), cell_stride(L), compact, L->ncell[0], L->ncell[1], L->ncell[2],
            L->periodic[0], L->periodic[1], L->periodic[2]);
    } else {
        return set_error(MRB_E_UNSUPPORTED, "EB force for this storage");
    }
    MRB_HIP(hipGetLastError());
    std::vector<double> part(size_t(5 * nb));
    MRB_HIP(hipMemcpyAsync(part.data(), L->d_red, sizeof(double) * part.size(), hipMemcpyDeviceToHost, s));
    MRB_HIP(hipStreamSynchronize(s));
    for (int n = 0; n < 3; ++n) out[n] = 0.0;
    for (unsigned b = 0; b < nb; ++b)
        for (int n = 0; n < 3; ++n) out[n] += part[size_t(5 * b + n)];
    return MRB_OK;
}

int mrb_storage_saturations(const mrb_level *L, int64_t *count)
{
    if (!L || !count) return set_error(MRB_E_ARG, "null argument");
    *count = 0;
    if (L->storage != MRB_STORAGE_F16_SCALED) return MRB_OK;
    uint32_t v = 0;
    MRB_HIP(hipDeviceSynchronize());
    MRB_HIP(hipMemcpy(&v, L->d_ovf16, sizeof(v), hipMemcpyDeviceToHost));
    *count = v;
    return MRB_OK;
}

int mrb_macro_fields(const mrb_level *L, double *out, void *stream)
{
    if (!L || !out) return set_error(MRB_E_ARG, "null argument");
    if (L->layout != MRB_LAYOUT_COMPACT) return set_error(MRB_E_STATE, "needs the compact layout");
    if (!L->flags_ready) return set_error(MRB_E_STATE, "flags not set");
    if (!L->phys_set) return set_error(MRB_E_STATE, "physics not set");
    return compact_macro(L, out, as_stream(stream));
}

int mrb_vorticity(mrb_level *L, double *out, void *stream)
{
    if (!L || !out) return set_error(MRB_E_ARG, "null argument");
    if (L->layout != MRB_LAYOUT_COMPACT) return set_error(MRB_E_STATE, "needs the compact layout");
    if (!L->flags_ready || !L->phys_set) return set_error(MRB_E_STATE, "flags and physics must be set");
    return compact_vorticity(L, out, as_stream(stream));
}

int mrb_tag_cells(const mrb_level *L, const double *field, int criterion, double value, uint8_t *tags,
                  void *stream)
{
    if (!L || !field || !tags) return set_error(MRB_E_ARG, "null argument");
    if (L->layout != MRB_LAYOUT_COMPACT) return set_error(MRB_E_STATE, "needs the compact layout");
    if (criterion != MRB_TAG_GREATER && criterion != MRB_TAG_LESS && criterion != MRB_TAG_ADJACENT_DIFFERENCE)
        return set_error(MRB_E_ARG, "unknown tagging criterion");
    return compact_tag(L, field, criterion, value, tags, as_stream(stream));
}

int mrb_tag_level(mrb_level *L, const char *field, int criterion, double value, uint8_t *host_tags)
{
    if (!L || !field || !host_tags) return set_error(MRB_E_ARG, "null argument");
    if (criterion != MRB_TAG_GREATER && criterion != MRB_TAG_LESS && criterion != MRB_TAG_ADJACENT_DIFFERENCE)
        return set_error(MRB_E_ARG, "unknown tagging criterion");
    if (!L->flags_ready || !L->pops_set) return set_error(MRB_E_STATE, "flags and populations must be set");
    // LBM::get_field by name: the macrodata and derived plot variables
    static const char *const kNames[26] = {
        "rho", "vel_x", "vel_y", "vel_z", "vel_mag", "two_rho_e", "QCorrX", "QCorrY", "QCorrZ",
        "pxx", "pyy", "pzz", "pxy", "pxz", "pyz", "qx", "qy", "qz", "temperature",
        "vort_x", "vort_y", "vort_z", "vort_mag", "dQCorrX", "dQCorrY", "dQCorrZ"};
    int idx = -1;
    for (int n = 0; n < 26; ++n)
        if (std::string(field) == kNames[n]) idx = n;
    if (idx < 0) return set_error(MRB_E_ARG, std::string("unknown field ") + field);
    const int64_t N = int64_t(L->ncell[0]) * L->ncell[1] * L->ncell[2];
    // ErrorEst tags the levels below the finest, which cover the domain
    int64_t covered = 0;
    for (const BoxDev &B : L->boxes)
        covered += int64_t(B.vhi[0] - B.vlo[0] + 1) * (B.vhi[1] - B.vlo[1] + 1) * (B.vhi[2] - B.vlo[2] + 1);
    if (covered != N) return set_error(MRB_E_UNSUPPORTED, "mrb_tag_level: the level's boxes must cover the domain");
    double *dense = nullptr, *der = nullptr, *md = nullptr;
    uint8_t *tags = nullptr;
    hipError_t e = hipMalloc(&dense, sizeof(double) * size_t(N));
    if (e == hipSuccess) e = hipMalloc(&tags, size_t(N));
    int rc = MRB_OK;
    if (e == hipSuccess) {
        if (L->layout == MRB_LAYOUT_COMPACT) {
            e = hipMalloc(&md, sizeof(double) * 26 * size_t(N));
            if (e == hipSuccess) rc = compact_plot_fields(L, md, md + 19 * N, nullptr);
            if (e == hipSuccess && rc == MRB_OK)
                e = hipMemcpy(dense, md + int64_t(idx) * N, sizeof(double) * size_t(N), hipMemcpyDeviceToDevice);
        } else {
            e = hipMalloc(&der, sizeof(double) * 7 * size_t(L->total_cells));
            if (e == hipSuccess) rc = fab_plot_fields(L, der, nullptr);
            if (e == hipSuccess && rc == MRB_OK) {
                const bool is_md = idx < 19;
                k_fab_to_dense<<<nblocks(L->total_cells), kBlock>>>(is_md ? L->d_md : der, is_md ? 19 : 7,
                                                                    is_md ? idx : idx - 19, dense, L->d_boxes,
                                                                    L->d_box_base, L->nbox, L->total_cells,
                                                                    L->ncell[0], L->ncell[1]);
                e = hipGetLastError();
            }
        }
    }
    if (e == hipSuccess && rc == MRB_OK) rc = compact_tag(L, dense, criterion, value, tags, nullptr);
    if (e == hipSuccess && rc == MRB_OK) e = hipDeviceSynchronize();
    if (e == hipSuccess && rc == MRB_OK) e = hipMemcpy(host_tags, tags, size_t(N), hipMemcpyDeviceToHost);
    (void)hipFree(dense);
    (void)hipFree(tags);
    (void)hipFree(der);
    (void)hipFree(md);
    if (e != hipSuccess) return hip_fail(e, "mrb_tag_level");
    return rc;
}

int mrb_storage_shift(const mrb_level *L, int lattice, double out[27])
{
    if (!L || !out || lattice < 0 || lattice > 1) return set_error(MRB_E_ARG, "bad argument");
    for (int q = 0; q < NQ; ++q)
        out[q] = L->storage == MRB_STORAGE_F64 ? 0.0 : (lattice == 0 ? L->fshift[q] : L->gshift[q]);
    return MRB_OK;
}

const char *mrb_hot_kernel_name(const mrb_level *L)
{
    if (!L) return "";
    return L->layout == MRB_LAYOUT_COMPACT ? compact_kernel_name(L) : "fab-stages";
}

int mrb_population_ptr(const mrb_level *L, int lattice, void **dev_ptr, int64_t *bytes)
{
    if (!L || !dev_ptr || lattice < 0 || lattice > 1) return set_error(MRB_E_ARG, "bad argument");
    *dev_ptr = L->pop[lattice][L->cur];
    if (bytes) *bytes = int64_t(pop_elem_size(L->storage)) * NQ * L->total_cells;
    return MRB_OK;
}

}  // extern "C"
