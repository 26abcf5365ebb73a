// D3Q27 lattice tables and per-cell lattice-Boltzmann math shared by every
// kernel of the engine (FP64 parity path, fused hot path, reduced-precision
// variants and the adjoint).
//
// Every floating-point expression keeps the operation order of the reference
// (NREL/marbles Source/Utilities.H, Source/LBM.cpp) and the library is built
// with -ffp-contract=off, so the FP64 path reproduces the reference
// arithmetic bit for bit.  Per-cell terms that the reference recomputes for
// each of the 27 directions are hoisted; hoisting changes no result because
// the recomputed values are identical.  Terms the reference multiplies by a
// zero lattice-velocity component are skipped: x + (+-0) == x, so the sums
// are unchanged (only the sign of an exact zero could differ).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>
#include <utility>

namespace mrb {

constexpr int NQ = 27;

// Source/Stencil.H:49-79 (velocity order) and 125-134 (opposites).
constexpr int kEV[NQ][3] = {
    {0, 0, 0},   {1, 0, 0},   {-1, 0, 0},  {0, 1, 0},   {0, -1, 0},
    {0, 0, 1},   {0, 0, -1},  {1, 1, 0},   {-1, -1, 0}, {1, -1, 0},
    {-1, 1, 0},  {1, 0, 1},   {-1, 0, -1}, {1, 0, -1},  {-1, 0, 1},
    {0, 1, 1},   {0, -1, -1}, {0, 1, -1},  {0, -1, 1},  {1, 1, 1},
    {-1, -1, -1}, {1, -1, 1}, {-1, 1, -1}, {1, -1, -1}, {-1, 1, 1},
    {1, 1, -1},  {-1, -1, 1}};
constexpr int kBounce[NQ] = {0,  2,  1,  4,  3,  6,  5,  8,  7,  10, 9,  12, 11, 14,
                             13, 16, 15, 18, 17, 20, 19, 22, 21, 24, 23, 26, 25};

constexpr int ev(int q, int d) { return kEV[q][d]; }
constexpr int iabs(int a) { return a < 0 ? -a : a; }
constexpr int csq(int q) { return iabs(kEV[q][0]) + iabs(kEV[q][1]) + iabs(kEV[q][2]); }

// Index of the direction whose component d is reflected (Stencil.H:136-167).
constexpr int reflect(int q, int d)
{
    int r[3] = {kEV[q][0], kEV[q][1], kEV[q][2]};
    r[d] = -r[d];
    for (int p = 0; p < NQ; ++p)
        if (kEV[p][0] == r[0] && kEV[p][1] == r[1] && kEV[p][2] == r[2]) return p;
    return -1;
}

// Index of the direction with lattice velocity (cx, cy, cz).
constexpr int qof(int cx, int cy, int cz)
{
    for (int p = 0; p < NQ; ++p)
        if (kEV[p][0] == cx && kEV[p][1] == cy && kEV[p][2] == cz) return p;
    return -1;
}

// Stencil.H:81-85: WO is computed from the other three weights.
constexpr double kWSC1 = 2.0 / 27.0;
constexpr double kWFCC1 = 1.0 / 54.0;
constexpr double kWBCC1 = 1.0 / 216.0;
constexpr double kWO = 1.0 - (6.0 * (kWSC1) + 12.0 * (kWFCC1) + 8.0 * (kWBCC1));
constexpr double weight(int q)
{
    return csq(q) == 0 ? kWO : csq(q) == 1 ? kWSC1 : csq(q) == 2 ? kWFCC1 : kWBCC1;
}
constexpr double kTheta0 = 1.0 / 3.0;          // Stencil.H:42
constexpr double kAirMolarMass = 28.96;         // Constants.H:65
// Constants.H:57-58: SMALL_NUM = epsilon * 1e10, the sentinel tolerance of
// masked_avgdown
constexpr double kSmallNum = 2.220446049250313080847263336181640625e-16 * 1e10;

// Compile-time loop: f(std::integral_constant<int, I>) for I = 0..N-1.
template <class F, int... Is>
__host__ __device__ __forceinline__ void static_for_impl(F&& f, std::integer_sequence<int, Is...>)
{
    (f(std::integral_constant<int, Is>{}), ...);
}
template <int N, class F>
__host__ __device__ __forceinline__ void static_for(F&& f)
{
    static_for_impl(f, std::make_integer_sequence<int, N>{});
}

// Per-level constants of the collide (LBM.cpp:639-644, 775-777, 817-819).
struct CollideParams {
    double nu, alpha;
    double R;        // specific gas constant m_R_u / m_m_bar
    double cv;       // R / (adiabatic_exponent - 1)
    double dt;       // level dt
    double ms;       // mesh speed
    double inv_dx[3];
};

// Raw moments of the post-stream populations (LBM.cpp:834-862), summed in
// direction order.
// R = the arithmetic type of the sums
template <class R>
struct MomentsR {
    R rho, jx, jy, jz, pxx, pyy, pzz, pxy, pxz, pyz, E, qx, qy, qz;
};
using Moments = MomentsR<double>;

// f, g: anything indexable by q (arrays, pointers, row views of LDS)
template <class FA, class GA, class R>
__host__ __device__ __forceinline__ void accumulate_moments(const FA& f, const GA& g, MomentsR<R>& m)
{
    m.rho = m.jx = m.jy = m.jz = R(0);
    m.pxx = m.pyy = m.pzz = m.pxy = m.pxz = m.pyz = R(0);
    m.E = m.qx = m.qy = m.qz = R(0);
    static_for<NQ>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        constexpr int cx = ev(q, 0), cy = ev(q, 1), cz = ev(q, 2);
        const R fq = R(f[q]);
        const R gq = R(g[q]);
        m.rho += fq;
        if constexpr (cx == 1) m.jx += fq;
        if constexpr (cx == -1) m.jx += -fq;
        if constexpr (cy == 1) m.jy += fq;
        if constexpr (cy == -1) m.jy += -fq;
        if constexpr (cz == 1) m.jz += fq;
        if constexpr (cz == -1) m.jz += -fq;
        if constexpr (cx != 0) m.pxx += fq;
        if constexpr (cy != 0) m.pyy += fq;
        if constexpr (cz != 0) m.pzz += fq;
        if constexpr (cx * cy == 1) m.pxy += fq;
        if constexpr (cx * cy == -1) m.pxy += -fq;
        if constexpr (cx * cz == 1) m.pxz += fq;
        if constexpr (cx * cz == -1) m.pxz += -fq;
        if constexpr (cy * cz == 1) m.pyz += fq;
        if constexpr (cy * cz == -1) m.pyz += -fq;
        m.E += gq;
        if constexpr (cx == 1) m.qx += gq;
        if constexpr (cx == -1) m.qx += -gq;
        if constexpr (cy == 1) m.qy += gq;
        if constexpr (cy == -1) m.qy += -gq;
        if constexpr (cz == 1) m.qz += gq;
        if constexpr (cz == -1) m.qz += -gq;
    });
}

// Only the sums the Q-correction needs (rho, j, 2 rho E).
template <class T, class R>
__host__ __device__ __forceinline__ void accumulate_q_moments(
    const T* __restrict__ f, const T* __restrict__ g, R& rho, R& jx, R& jy, R& jz, R& E)
{
    rho = jx = jy = jz = E = R(0);
    static_for<NQ>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        constexpr int cx = ev(q, 0), cy = ev(q, 1), cz = ev(q, 2);
        const R fq = R(f[q]);
        rho += fq;
        if constexpr (cx == 1) jx += fq;
        if constexpr (cx == -1) jx += -fq;
        if constexpr (cy == 1) jy += fq;
        if constexpr (cy == -1) jy += -fq;
        if constexpr (cz == 1) jz += fq;
        if constexpr (cz == -1) jz += -fq;
        E += R(g[q]);
    });
}

// Macroscopic fields derived in f_to_macrodata (LBM.cpp:863-901).
// R = the arithmetic type (FP64 everywhere except the FP32-arithmetic
// collide of the scaled FP16 storage)
template <class R>
struct MacroR {
    R u, v, w, T, Qx, Qy, Qz;
};
using Macro = MacroR<double>;

template <class R>
__host__ __device__ __forceinline__ MacroR<R> macro_from_t(R rho, R jx, R jy, R jz, R E, const CollideParams& P)
{
    MacroR<R> M;
    M.u = jx * (R(P.ms) / rho);
    M.v = jy * (R(P.ms) / rho);
    M.w = jz * (R(P.ms) / rho);
    // get_temperature, scalar overload (Utilities.H:187-196)
    M.T = (R(0.50) / R(P.cv)) * ((E / rho) - (M.u * M.u + M.v * M.v + M.w * M.w));
    const R k = (R(1.0) - R(3.0) * R(P.R) * M.T);
    M.Qx = rho * M.u * (k - M.u * M.u);
    M.Qy = rho * M.v * (k - M.v * M.v);
    M.Qz = rho * M.w * (k - M.w * M.w);
    return M;
}

__host__ __device__ __forceinline__ Macro macro_from(
    double rho, double jx, double jy, double jz, double E, const CollideParams& P)
{
    return macro_from_t<double>(rho, jx, jy, jz, E, P);
}

// gradient() of Utilities.H:279-312 given the neighbour availability
// (inside the domain box and fluid) and the three Q values.
__host__ __device__ __forceinline__ double one_sided_gradient(
    bool p_ok, bool m_ok, double qp, double qc, double qm, double idx)
{
    double vp = 0.0, vc = 0.0, vm = 0.0;
    if (!p_ok && !m_ok) {
        vp = 0.0;
        vc = 0.0;
        vm = 0.0;
    } else if (!p_ok) {
        vp = 0.0;
        vc = qc;
        vm = -qm;
    } else if (!m_ok) {
        vp = qp;
        vc = -qc;
        vm = 0.0;
    } else {
        vp = 0.5 * qp;
        vc = 0.0;
        vm = -0.5 * qm;
    }
    return (vp + vc + vm) * idx;
}

// phi of the product-form equilibrium for one lattice-velocity component
// c in {-1,0,1} (Utilities.H:53-60).
template <int c, class R>
__host__ __device__ __forceinline__ R phi_of(R u, R P)
{
    return R(c) * R(0.5) * u + R(iabs(c)) * (R(1.50) * P - R(1.0)) - P + R(1.0);
}

// Per-cell coefficients of the f and g equilibria (LBM.cpp:659-758) and the
// relaxation rate (LBM.cpp:795-797).
template <class R>
struct CollideCoeffsR {
    R omega;
    R rphx[3], phy[3], phz[3];  // rho*phi_x, phi_y, phi_z at c+1
    R E, a1x, a1y, a1z;
    R tx[2], ty[2], tz[2];      // (c*c - theta0) * a2_aa for c*c = 0, 1
    R dxy, dxz, dyz;            // 2 * a2_ab
};
using CollideCoeffs = CollideCoeffsR<double>;

template <class R>
__host__ __device__ __forceinline__ CollideCoeffsR<R> collide_coeffs_t(
    const MomentsR<R>& m, const MacroR<R>& M, R dqx, R dqy, R dqz, const CollideParams& P)
{
    CollideCoeffsR<R> C;
    const R Rg = R(P.R), dt = R(P.dt);
    const R rho = m.rho;
    const R T = M.T;
    const R u = M.u, v = M.v, w = M.w;
    const R omega = R(1.0) / (R(P.nu) / (Rg * T * dt) + R(0.5));
    const R omega_one = R(1.0) / (R(P.alpha) / (Rg * T * dt) + R(0.5));
    const R omega_one_by_omega = omega_one / omega;
    const R omega_corr = (R(2.0) - omega) / (R(2.0) * omega * rho);
    C.omega = omega;

    const R pxx_ext = u * u + Rg * T + dt * (omega_corr)*dqx;
    const R pyy_ext = v * v + Rg * T + dt * (omega_corr)*dqy;
    const R pzz_ext = w * w + Rg * T + dt * (omega_corr)*dqz;
    C.rphx[0] = rho * phi_of<-1>(u, pxx_ext);
    C.rphx[1] = rho * phi_of<0>(u, pxx_ext);
    C.rphx[2] = rho * phi_of<1>(u, pxx_ext);
    C.phy[0] = phi_of<-1>(v, pyy_ext);
    C.phy[1] = phi_of<0>(v, pyy_ext);
    C.phy[2] = phi_of<1>(v, pyy_ext);
    C.phz[0] = phi_of<-1>(w, pzz_ext);
    C.phz[1] = phi_of<0>(w, pzz_ext);
    C.phz[2] = phi_of<1>(w, pzz_ext);

    const R E = m.E;
    const R energy = E / (R(2.0) * rho);
    const R Tv = (R(0.50) / R(P.cv)) * ((E / rho) - u * u + v * v + w * w);
    const R p = rho * Rg * Tv;
    const R h = energy + (p / rho);
    const R hp = h + (p / rho);
    R qx_eq = R(2.0) * rho * u * h;
    R qy_eq = R(2.0) * rho * v * h;
    R qz_eq = R(2.0) * rho * w * h;
    const R rxx = R(2.0) * rho * u * u * hp + R(2.0) * p * h;
    const R ryy = R(2.0) * rho * v * v * hp + R(2.0) * p * h;
    const R rzz = R(2.0) * rho * w * w * hp + R(2.0) * p * h;
    const R rxy = R(2.0) * rho * u * v * hp + R(0);
    const R rxz = R(2.0) * rho * u * w * hp + R(0);
    const R ryz = R(2.0) * rho * v * w * hp + R(0);
    qx_eq *= omega_one_by_omega;
    qy_eq *= omega_one_by_omega;
    qz_eq *= omega_one_by_omega;
    qx_eq += (R(1.0) - omega_one_by_omega) *
             (m.qx - R(2.0) * u * m.pxx - R(2.0) * v * m.pxy - R(2.0) * w * m.pxz - u * dt * dqx);
    qy_eq += (R(1.0) - omega_one_by_omega) *
             (m.qy - R(2.0) * u * m.pxy - R(2.0) * v * m.pyy - R(2.0) * w * m.pyz - v * dt * dqy);
    qz_eq += (R(1.0) - omega_one_by_omega) *
             (m.qz - R(2.0) * u * m.pxz - R(2.0) * v * m.pyz - R(2.0) * w * m.pzz - w * dt * dqz);
    const R one_bystheta0 = R(1.0 / kTheta0);
    C.E = E;
    C.a1x = qx_eq * one_bystheta0;
    C.a1y = qy_eq * one_bystheta0;
    C.a1z = qz_eq * one_bystheta0;
    const R a2xx = (rxx - E * R(1.0) * R(kTheta0)) * one_bystheta0 * one_bystheta0;
    const R a2yy = (ryy - E * R(1.0) * R(kTheta0)) * one_bystheta0 * one_bystheta0;
    const R a2zz = (rzz - E * R(1.0) * R(kTheta0)) * one_bystheta0 * one_bystheta0;
    const R a2xy = rxy * one_bystheta0 * one_bystheta0;
    const R a2xz = rxz * one_bystheta0 * one_bystheta0;
    const R a2yz = ryz * one_bystheta0 * one_bystheta0;
    C.tx[0] = R(0 - kTheta0) * a2xx;
    C.tx[1] = R(1 - kTheta0) * a2xx;
    C.ty[0] = R(0 - kTheta0) * a2yy;
    C.ty[1] = R(1 - kTheta0) * a2yy;
    C.tz[0] = R(0 - kTheta0) * a2zz;
    C.tz[1] = R(1 - kTheta0) * a2zz;
    C.dxy = R(2.0) * a2xy;
    C.dxz = R(2.0) * a2xz;
    C.dyz = R(2.0) * a2yz;
    return C;
}

__host__ __device__ __forceinline__ CollideCoeffs collide_coeffs(
    const Moments& m, const Macro& M, double dqx, double dqy, double dqz, const CollideParams& P)
{
    return collide_coeffs_t<double>(m, M, dqx, dqy, dqz, P);
}

// f equilibrium of direction q (Utilities.H:42-63, product order kept)
template <int q, class Co>
__host__ __device__ __forceinline__ auto feq_of(const Co& C)
{
    constexpr int cx = ev(q, 0), cy = ev(q, 1), cz = ev(q, 2);
    return ((C.rphx[cx + 1]) * C.phy[cy + 1]) * C.phz[cz + 1];
}

// g equilibrium of direction q (Utilities.H:147-160, summation order kept)
template <int q, class Co>
__host__ __device__ __forceinline__ auto geq_of(const Co& C)
{
    using R = decltype(C.E);
    constexpr int cx = ev(q, 0), cy = ev(q, 1), cz = ev(q, 2);
    R ge = C.E;
    if constexpr (cx == 1) ge = ge + C.a1x;
    if constexpr (cx == -1) ge = ge + -C.a1x;
    if constexpr (cy == 1) ge = ge + C.a1y;
    if constexpr (cy == -1) ge = ge + -C.a1y;
    if constexpr (cz == 1) ge = ge + C.a1z;
    if constexpr (cz == -1) ge = ge + -C.a1z;
    R S = C.tx[cx * cx] + C.ty[cy * cy];
    if constexpr (cx * cy == 1) S = S + C.dxy;
    if constexpr (cx * cy == -1) S = S + -C.dxy;
    S = S + C.tz[cz * cz];
    if constexpr (cx * cz == 1) S = S + C.dxz;
    if constexpr (cx * cz == -1) S = S + -C.dxz;
    if constexpr (cy * cz == 1) S = S + C.dyz;
    if constexpr (cy * cz == -1) S = S + -C.dyz;
    ge = ge + R(0.5) * S;
    return ge * R(weight(q));
}

// Equilibria of LBM::macrodata_to_equilibrium (LBM.cpp:659-758) hoisted per
// cell, followed by the BGK relaxation of LBM::relax_f_to_equilibrium
// (LBM.cpp:793-801), in arithmetic type R.  fo/go may alias f/g.
template <class R, class FA, class GA, class FOut>
__host__ __device__ __forceinline__ void collide_cell_t(
    const FA& f, const GA& g, const MomentsR<R>& m, const MacroR<R>& M, R dqx,
    R dqy, R dqz, const CollideParams& P, FOut&& out)
{
    const CollideCoeffsR<R> C = collide_coeffs_t<R>(m, M, dqx, dqy, dqz, P);
    const R omega = C.omega;
    static_for<NQ>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        const R fq = R(f[q]);
        const R gq = R(g[q]);
        out(qc, fq + omega * (feq_of<q>(C) - fq), gq + omega * (geq_of<q>(C) - gq));
    });
}

template <class TIn, class FOut>
__host__ __device__ __forceinline__ void collide_cell(
    const TIn* __restrict__ f, const TIn* __restrict__ g, const Moments& m, const Macro& M,
    double dqx, double dqy, double dqz, const CollideParams& P, FOut&& out)
{
    collide_cell_t<double>(f, g, m, M, dqx, dqy, dqz, P, out);
}

// Product-form equilibrium without correction (Utilities.H:11-40): the
// velocity-inlet f ghost value and the initial f.
__host__ __device__ inline double set_equilibrium_value(
    double rho, const double vel[3], double rt, int q)
{
    const double pxx = vel[0] * vel[0] + rt;
    const double pyy = vel[1] * vel[1] + rt;
    const int c0 = kEV[q][0], c1 = kEV[q][1], c2 = kEV[q][2];
    const double phix = c0 * 0.5 * vel[0] + iabs(c0) * (1.50 * pxx - 1.0) - pxx + 1.0;
    const double phiy = c1 * 0.5 * vel[1] + iabs(c1) * (1.50 * pyy - 1.0) - pyy + 1.0;
    const double pzz = vel[2] * vel[2] + rt;
    const double phiz = c2 * 0.5 * vel[2] + iabs(c2) * (1.50 * pzz - 1.0) - pzz + 1.0;
    return rho * phix * phiy * phiz;
}

// g equilibrium of a prescribed state (BC.H:141-189 velocity_bc_g and
// IC.H:497-518): energy, equilibrium moments, Grad expansion.
__host__ __device__ inline double thermal_g_equilibrium(
    double rho, const double vel[3], double T, double R, double gamma, int q)
{
    const double cv = R / (gamma - 1.0);
    // get_energy, RealVect overload (Utilities.H:176-185)
    const double E = rho * (2.0 * cv * T + vel[0] * vel[0] + vel[1] * vel[1] + vel[2] * vel[2]);
    const double energy = E / (2.0 * rho);
    const double Tv = (0.50 / cv) * ((E / rho) - vel[0] * vel[0] + vel[1] * vel[1] + vel[2] * vel[2]);
    const double p = rho * R * Tv;
    const double h = energy + (p / rho);
    const double hp = h + (p / rho);
    const double qeq[3] = {2.0 * rho * vel[0] * h, 2.0 * rho * vel[1] * h, 2.0 * rho * vel[2] * h};
    const double rxx = 2.0 * rho * vel[0] * vel[0] * hp + 2.0 * p * h;
    const double ryy = 2.0 * rho * vel[1] * vel[1] * hp + 2.0 * p * h;
    const double rzz = 2.0 * rho * vel[2] * vel[2] * hp + 2.0 * p * h;
    const double rxy = 2.0 * rho * vel[0] * vel[1] * hp + 0;
    const double rxz = 2.0 * rho * vel[0] * vel[2] * hp + 0;
    const double ryz = 2.0 * rho * vel[1] * vel[2] * hp + 0;
    const double inv = 1.0 / kTheta0;
    const double a1x = qeq[0] * inv, a1y = qeq[1] * inv, a1z = qeq[2] * inv;
    const double a2xx = (rxx - E * 1.0 * kTheta0) * inv * inv;
    const double a2yy = (ryy - E * 1.0 * kTheta0) * inv * inv;
    const double a2zz = (rzz - E * 1.0 * kTheta0) * inv * inv;
    const double a2xy = rxy * inv * inv;
    const double a2xz = rxz * inv * inv;
    const double a2yz = ryz * inv * inv;
    const int c0 = kEV[q][0], c1 = kEV[q][1], c2 = kEV[q][2];
    double fv = E + a1x * c0 + a1y * c1 + a1z * c2;
    fv += 0.5 * ((c0 * c0 - kTheta0) * a2xx + (c1 * c1 - kTheta0) * a2yy +
                 2.0 * (c0 * c1 - 0) * a2xy + (c2 * c2 - kTheta0) * a2zz +
                 2.0 * (c0 * c2 - 0) * a2xz + 2.0 * (c1 * c2 - 0) * a2yz);
    fv *= weight(q);
    return fv;
}

}  // namespace mrb
