// Isothermal f-only variant with partial bounce-back, and its hand-written
// adjoint with sqrt(N) checkpointing.
//
// Forward step (one cell, T frozen at T0, R = 1):
//   pull      f*_q(x) = f_q(x - c_q); y, z periodic; x-lo ghost = the inlet's
//             f_eq(rho_in, u_in, R T0); x-hi ghost = the boundary cell next
//             to it (outflow copy)
//   moments   rho = sum f*, u = j ms / rho, Q_a = rho u_a (1 - 3 R T0 - u_a^2)
//   dQ        one-sided gradient with the domain-box rule (Utilities.H:279-312)
//   collide   w = 1/tau, w_corr = (2 - w)/(2 w rho), P_a = u_a^2 + R T0 +
//             dt w_corr dQ_a, f_eq = ((rho phi_x) phi_y) phi_z,
//             f^c = f* + w (f_eq - f*)   (LBM.cpp:659-697, 765-807, T = T0)
//   partial bounce-back (Walsh, Burwinkle, Saar 2009)
//             f_new_q = (1 - n_s) f^c_q + n_s f*_qbar
// At n_s = 0 this is the reference's fluid update with the temperature held
// at T0; at n_s = 1 the incoming populations are reflected (a solid cell).
//
// Loss J = mean over the outlet plane (x = nx-1) of u_x + 1e-3 sum over
// cells of rho |u|^2 / 2, from the final populations.
//
// Adjoint of one step (lambda = dJ/d populations), two passes:
//   adj: per cell, lambda of the post-collision populations gathered through
//        the transpose of the pull from the previous pass's output, then the
//        transpose of collide + partial bounce-back: lambda at the pulled
//        populations (without the Q-correction path), dJ/d(dQ), dJ/dn_s
//        (accumulated), dJ/dtau (block partial sums)
//   adj_q: per cell, the transpose of dQ (a 3-point gather along each axis)
//        and of Q(rho, u), kept as four coefficients that the next gather
//        adds to each pulled population's lambda
// Checkpoints: the forward sweep writes a checkpoint every `every` steps in
// place and keeps the last segment's states; the reverse sweep recomputes
// each earlier segment from its checkpoint.
// Layout: f[q*N + cell], cell = (z*ny + y)*nx + x, FP64.
#include <algorithm>
#include <cmath>
#include <memory>
#include <vector>

#include "mrb_internal.h"

struct mrb_iso {
    int n[3] = {0, 0, 0};
    int64_t N = 0;
    double T0 = 1.0 / 3.0, rho0 = 1.0, u_in = 0.0, ms = 1.0, dt = 1.0;
    double inlet[mrb::NQ] = {};
    // device work arrays
    double *f[2] = {nullptr, nullptr};
    double *lam[2] = {nullptr, nullptr};  // lambda ping-pong (27 N each)
    double *q = nullptr;                  // 3 N: Q of the current state
    double *dqbar = nullptr;              // 3 N
    double *uo = nullptr;                 // 3 N: u of the pulled state
    double *bu = nullptr;                 // 4 N: shared lambda terms of one adjoint pass
    double *ab = nullptr;                 // 4 N: per-cell gather coefficients (alpha, beta), one 32-byte record per cell
    double *part = nullptr;               // one partial sum per block of the cell grid
    int64_t nparts = 0;
    // checkpoints (one per segment), the states and Q arrays of one segment;
    // kept between gradient calls of the same segment shape
    std::vector<double *> ckpt, seg, segq;
    // the recorded forward sweep that mrb_iso_reverse differentiates
    bool rec_valid = false;
    int rec_nsteps = 0, rec_every = 0;
    double rec_tau = 0.0;
    const double *rec_ns = nullptr, *rec_u0 = nullptr;
};

namespace mrb {
namespace {

struct IsoArgs {
    int nx, ny, nz;
    int64_t N;
    double w, tau, RT, k, ms, dt;
    double inlet[NQ];
};

__device__ __forceinline__ bool iso_cell(const IsoArgs &A, int &x, int &y, int &z, int64_t &cell)
{
    x = int(blockIdx.x * blockDim.x + threadIdx.x);
    y = int(blockIdx.y * blockDim.y + threadIdx.y);
    z = int(blockIdx.z);
    if (x >= A.nx || y >= A.ny) return false;
    cell = (int64_t(z) * A.ny + y) * A.nx + x;
    return true;
}

__device__ __forceinline__ int wrap(int a, int n) { return a < 0 ? a + n : (a >= n ? a - n : a); }

// pull of one cell (see the file comment), without branches: source
// coordinates per velocity component (index c + 1 holds the source of c),
// every load issued unconditionally, the inlet constant selected afterwards
__device__ __forceinline__ void iso_pull(const double *__restrict__ f, const IsoArgs &A, int x, int y, int z,
                                         double fs[NQ])
{
    const uint32_t nxy = uint32_t(A.nx) * uint32_t(A.ny);
    const uint32_t xs[3] = {uint32_t(x + 1 >= A.nx ? A.nx - 1 : x + 1), uint32_t(x), uint32_t(x > 0 ? x - 1 : 0)};
    const uint32_t yo[3] = {uint32_t(wrap(y + 1, A.ny)) * uint32_t(A.nx), uint32_t(y) * uint32_t(A.nx),
                            uint32_t(wrap(y - 1, A.ny)) * uint32_t(A.nx)};
    const uint32_t zo[3] = {uint32_t(wrap(z + 1, A.nz)) * nxy, uint32_t(z) * nxy, uint32_t(wrap(z - 1, A.nz)) * nxy};
    double r[NQ];
    static_for<NQ>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        constexpr int cx = ev(q, 0), cy = ev(q, 1), cz = ev(q, 2);
        const double *fq = f + int64_t(q) * A.N;
        r[q] = fq[zo[cz + 1] + yo[cy + 1] + xs[cx + 1]];
    });
    const bool inlet = x == 0;
    static_for<NQ>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        if constexpr (ev(q, 0) == 1)
            fs[q] = inlet ? A.inlet[q] : r[q];
        else
            fs[q] = r[q];
    });
}

__device__ __forceinline__ void iso_moments(const double fs[NQ], const IsoArgs &A, double &rho, double u[3])
{
    double j[3] = {0.0, 0.0, 0.0};
    rho = 0.0;
    static_for<NQ>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        rho += fs[q];
        if constexpr (ev(q, 0) != 0) j[0] += ev(q, 0) * fs[q];
        if constexpr (ev(q, 1) != 0) j[1] += ev(q, 1) * fs[q];
        if constexpr (ev(q, 2) != 0) j[2] += ev(q, 2) * fs[q];
    });
    for (int a = 0; a < 3; ++a) u[a] = j[a] * (A.ms / rho);
}

__device__ __forceinline__ double iso_phi(int c, double ua, double P)
{
    return ((c * 0.5) * ua + (c != 0 ? 1.0 : 0.0) * (1.5 * P - 1.0)) - P + 1.0;
}

// dQ_a at a cell from the Q array (domain-box rule on every axis); the
// neighbour loads are clamped into the box and issued unconditionally
__device__ __forceinline__ void iso_dq(const double *__restrict__ Q, const IsoArgs &A, int x, int y, int z,
                                       int64_t cell, double dq[3])
{
    const int pos[3] = {x, y, z}, n[3] = {A.nx, A.ny, A.nz};
    const int64_t st[3] = {1, A.nx, int64_t(A.nx) * A.ny};
    double qp[3], q0[3], qm[3];
    for (int a = 0; a < 3; ++a) {
        const double *Qa = Q + a * A.N;
        const bool p_ok = pos[a] + 1 < n[a], m_ok = pos[a] - 1 >= 0;
        qp[a] = Qa[cell + (p_ok ? st[a] : 0)];
        q0[a] = Qa[cell];
        qm[a] = Qa[cell - (m_ok ? st[a] : 0)];
    }
    for (int a = 0; a < 3; ++a) {
        const bool p_ok = pos[a] + 1 < n[a], m_ok = pos[a] - 1 >= 0;
        dq[a] = one_sided_gradient(p_ok, m_ok, p_ok ? qp[a] : 0.0, q0[a], m_ok ? qm[a] : 0.0, 1.0);
    }
}

// f_eq with the correction terms, in the reference's operation order
template <int q>
__device__ __forceinline__ double iso_feq(double rho, const double u[3], const double P[3])
{
    return ((rho * iso_phi(ev(q, 0), u[0], P[0])) * iso_phi(ev(q, 1), u[1], P[1])) * iso_phi(ev(q, 2), u[2], P[2]);
}

__global__ void __launch_bounds__(256) k_iso_init(double *__restrict__ f, const double *__restrict__ u0, IsoArgs A,
                                                  double rho0)
{
    int x, y, z;
    int64_t cell;
    if (!iso_cell(A, x, y, z, cell)) return;
    const double u[3] = {u0[cell], u0[A.N + cell], u0[2 * A.N + cell]};
    const double P[3] = {u[0] * u[0] + A.RT, u[1] * u[1] + A.RT, u[2] * u[2] + A.RT};
    static_for<NQ>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        f[int64_t(q) * A.N + cell] = iso_feq<q>(rho0, u, P);
    });
}

__global__ void __launch_bounds__(256) k_iso_q(const double *__restrict__ f, double *__restrict__ Q, IsoArgs A)
{
    int x, y, z;
    int64_t cell;
    if (!iso_cell(A, x, y, z, cell)) return;
    double fs[NQ], rho, u[3];
    iso_pull(f, A, x, y, z, fs);
    iso_moments(fs, A, rho, u);
    for (int a = 0; a < 3; ++a) Q[a * A.N + cell] = rho * u[a] * (A.k - u[a] * u[a]);
}

__global__ void __launch_bounds__(256) k_iso_step(const double *__restrict__ f, double *__restrict__ fo,
                                                  const double *__restrict__ Q, const double *__restrict__ ns,
                                                  IsoArgs A)
{
    int x, y, z;
    int64_t cell;
    if (!iso_cell(A, x, y, z, cell)) return;
    double fs[NQ], rho, u[3], dq[3];
    iso_pull(f, A, x, y, z, fs);
    iso_moments(fs, A, rho, u);
    iso_dq(Q, A, x, y, z, cell, dq);
    const double wc = (2.0 - A.w) / (2.0 * A.w * rho);
    double P[3];
    for (int a = 0; a < 3; ++a) P[a] = u[a] * u[a] + A.RT + A.dt * wc * dq[a];
    const double s = ns[cell];
    static_for<NQ>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        const double fc = fs[q] + A.w * (iso_feq<q>(rho, u, P) - fs[q]);
        fo[int64_t(q) * A.N + cell] = (1.0 - s) * fc + s * fs[kBounce[q]];
    });
}

// J partial sums: outlet-plane u_x / (ny nz) + 1e-3 rho |u|^2 / 2
__global__ void __launch_bounds__(256) k_iso_loss(const double *__restrict__ f, double *__restrict__ part, IsoArgs A)
{
    __shared__ double sh[256];
    int x, y, z;
    int64_t cell;
    double v = 0.0;
    if (iso_cell(A, x, y, z, cell)) {
        double fc[NQ], rho, u[3];
        for (int q = 0; q < NQ; ++q) fc[q] = f[int64_t(q) * A.N + cell];
        iso_moments(fc, A, rho, u);
        v = 1e-3 * (0.5 * rho * (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]));
        if (x == A.nx - 1) v += u[0] / (double(A.ny) * A.nz);
    }
    const int t = threadIdx.y * blockDim.x + threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) sh[t] += sh[t + w];
        __syncthreads();
    }
    if (t == 0) part[(int64_t(blockIdx.z) * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x] = sh[0];
}

// lambda_N = dJ/df_N
__global__ void __launch_bounds__(256) k_iso_seed(const double *__restrict__ f, double *__restrict__ lam, IsoArgs A)
{
    int x, y, z;
    int64_t cell;
    if (!iso_cell(A, x, y, z, cell)) return;
    double fc[NQ], rho, u[3];
    for (int q = 0; q < NQ; ++q) fc[q] = f[int64_t(q) * A.N + cell];
    iso_moments(fc, A, rho, u);
    const double u2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    const bool outlet = x == A.nx - 1;
    const double inv = 1.0 / (double(A.ny) * A.nz);
    static_for<NQ>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        const double cu = ev(q, 0) * u[0] + ev(q, 1) * u[1] + ev(q, 2) * u[2];
        // d(rho |u|^2 / 2)/df_q = ms c.u - |u|^2 / 2 ; du_x/df_q = (ms c_x - u_x) / rho
        double l = 1e-3 * (A.ms * cu - 0.5 * u2);
        if (outlet) l += (A.ms * ev(q, 0) - u[0]) / rho * inv;
        lam[int64_t(q) * A.N + cell] = l;
    });
}

// lambda of one post-collision population read through the pull by the cell
// t: the pulled populations' lambda there plus the Q-path term
// alpha(t) + beta(t) . c_q (ab holds the record alpha, beta_x, beta_y,
// beta_z of each cell: the coefficients a gather needs from one neighbour
// share one cache line)
template <int q>
__device__ __forceinline__ double lam_at(const double *__restrict__ lams, const double *__restrict__ ab, int64_t N,
                                         int64_t t)
{
    const double *r = ab + 4 * t;
    double v = lams[int64_t(q) * N + t] + r[0];
    static_for<3>([&](auto ac) {
        constexpr int a = decltype(ac)::value;
        if constexpr (ev(q, a) == 1) v += r[a + 1];
        if constexpr (ev(q, a) == -1) v -= r[a + 1];
    });
    return v;
}

// transpose of the pull, as a gather: f_q(x) was read by the cell x + c_q
// (when inside the domain in x) and, on the x-hi boundary for c_x = -1, also
// through the outflow ghost by the boundary cell at (x, y + c_y, z + c_z)
template <int q>
__device__ __forceinline__ double gather_one(const double *__restrict__ lams, const double *__restrict__ ab,
                                             const IsoArgs &A, int x, int y, int z)
{
    constexpr int cx = ev(q, 0), cy = ev(q, 1), cz = ev(q, 2);
    const int ty = wrap(y + cy, A.ny), tz = wrap(z + cz, A.nz);
    const int tx = x + cx;
    const bool in = tx >= 0 && tx < A.nx;
    const int64_t row = (int64_t(tz) * A.ny + ty) * A.nx;
    const double v = lam_at<q>(lams, ab, A.N, row + (tx < 0 ? 0 : (tx >= A.nx ? A.nx - 1 : tx)));
    double l = in ? v : 0.0;
    if constexpr (cx == -1)
        if (x == A.nx - 1) l += lam_at<q>(lams, ab, A.N, row + x);
    return l;
}

__device__ __forceinline__ void iso_gather_lambda(const double *__restrict__ lams, const double *__restrict__ ab,
                                                  const IsoArgs &A, int x, int y, int z, double lo[NQ])
{
    static_for<NQ>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        lo[q] = gather_one<q>(lams, ab, A, x, y, z);
    });
}

// one adjoint step at one cell: lambda_{n+1} (read directly from lin for the
// last step, else gathered from the previous pass's output) through the
// transpose of collide + partial bounce-back.  Per bounce pair (q, qbar
// adjacent in the ordering) the pulled populations' lambda is stored as soon
// as the pair's lambda_{n+1} arrive: (1 - w)(1 - n_s) lambda_q + n_s
// lambda_qbar.  The terms shared by all q (through rho, u and the
// Q-correction) are left to the next gather, as four per-cell coefficients
// in bu (base, ms ubar / rho), together with dJ/d(dQ) times dt w_corr
// (dqbar) and u of the pulled state (uo); dJ/dn_s is accumulated and dJ/dw
// summed per block.
template <bool GATHER>
__global__ void __launch_bounds__(256) k_iso_adj(
    const double *__restrict__ f, const double *__restrict__ Q, const double *__restrict__ lin,
    const double *__restrict__ ab, const double *__restrict__ ns, double *__restrict__ lams,
    double *__restrict__ dqbar, double *__restrict__ uo, double *__restrict__ bu, double *__restrict__ dns,
    double *__restrict__ part, IsoArgs A)
{
    __shared__ double sh[256];
    int x, y, z;
    int64_t cell;
    double gw = 0.0;
    if (iso_cell(A, x, y, z, cell)) {
        double fs[NQ], rho, u[3], dq[3];
        iso_pull(f, A, x, y, z, fs);
        iso_moments(fs, A, rho, u);
        iso_dq(Q, A, x, y, z, cell, dq);
        const double w = A.w;
        const double wc = (2.0 - w) / (2.0 * w * rho);
        double P[3];
        for (int a = 0; a < 3; ++a) P[a] = u[a] * u[a] + A.RT + A.dt * wc * dq[a];
        const double s = ns[cell];
        const double keep = (1.0 - w) * (1.0 - s);
        double rbar = 0.0, ubar[3] = {0.0, 0.0, 0.0}, pbar[3] = {0.0, 0.0, 0.0}, gns = 0.0;
        // contribution of one population, given its lambda_{n+1}
        auto one = [&](auto qc, double lq) {
            constexpr int q = decltype(qc)::value;
            constexpr int cx = ev(q, 0), cy = ev(q, 1), cz = ev(q, 2);
            const double px = iso_phi(cx, u[0], P[0]);
            const double py = iso_phi(cy, u[1], P[1]);
            const double pz = iso_phi(cz, u[2], P[2]);
            const double feq = ((rho * px) * py) * pz;
            const double fc = fs[q] + w * (feq - fs[q]);
            const double lc = (1.0 - s) * lq;
            gns += lq * (fs[kBounce[q]] - fc);
            gw += lc * (feq - fs[q]);
            const double mu = w * lc;
            rbar += mu * (px * py * pz);
            const double mr = mu * rho;
            const double g[3] = {mr * (py * pz), mr * (px * pz), mr * (px * py)};
            // dphi/du = c / 2, dphi/dP = 1/2 (c != 0) or -1 (c = 0)
            static_for<3>([&](auto ac) {
                constexpr int a = decltype(ac)::value;
                constexpr int c = ev(q, a);
                if constexpr (c != 0) {
                    const double h = 0.5 * g[a];
                    ubar[a] += c > 0 ? h : -h;
                    pbar[a] += h;
                } else {
                    pbar[a] -= g[a];
                }
            });
        };
        auto lam_in = [&](auto qc) {
            constexpr int q = decltype(qc)::value;
            if constexpr (GATHER)
                return gather_one<q>(lin, ab, A, x, y, z);
            else
                return lin[int64_t(q) * A.N + cell];
        };
        {
            const double l0 = lam_in(std::integral_constant<int, 0>{});
            one(std::integral_constant<int, 0>{}, l0);
            lams[cell] = keep * l0 + s * l0;
        }
        static_for<(NQ - 1) / 2>([&](auto pc) {
            constexpr int q = 2 * decltype(pc)::value + 1, qb = q + 1;
            static_assert(kBounce[q] == qb, "bounce pairs are adjacent");
            const double lq = lam_in(std::integral_constant<int, q>{});
            const double lb = lam_in(std::integral_constant<int, qb>{});
            one(std::integral_constant<int, q>{}, lq);
            one(std::integral_constant<int, qb>{}, lb);
            lams[int64_t(q) * A.N + cell] = keep * lq + s * lb;
            lams[int64_t(qb) * A.N + cell] = keep * lb + s * lq;
        });
        double wcbar = 0.0;
        for (int a = 0; a < 3; ++a) {
            ubar[a] += 2.0 * u[a] * pbar[a];
            wcbar += pbar[a] * A.dt * dq[a];
            dqbar[a * A.N + cell] = pbar[a] * A.dt * wc;
        }
        rbar += wcbar * (-wc / rho);
        gw += wcbar * (-1.0 / (w * w * rho));
        // du_a/df*_q = (ms c_qa - u_a) / rho: a shared part and a signed sum
        // of ms ubar_a / rho over the components of c_q
        double base = rbar;
        for (int a = 0; a < 3; ++a) {
            const double ub = ubar[a] / rho;
            base -= ub * u[a];
            bu[(a + 1) * A.N + cell] = ub * A.ms;
            uo[a * A.N + cell] = u[a];
        }
        bu[cell] = base;
        dns[cell] += gns;
    }
    const int t = threadIdx.y * blockDim.x + threadIdx.x;
    sh[t] = gw;
    __syncthreads();
    for (int s2 = 128; s2 > 0; s2 >>= 1) {
        if (t < s2) sh[t] += sh[t + s2];
        __syncthreads();
    }
    if (t == 0) part[(int64_t(blockIdx.z) * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x] += sh[0];
}

// transpose of dQ (a 3-point gather of dqbar along each axis) and of
// Q_a(rho, u) = rho u_a (k - u_a^2), whose derivative with respect to the
// pulled population q is 2 u_a^3 + (k - 3 u_a^2) ms c_qa.  Stores the
// per-cell coefficients of the term that the next gather adds to every
// pulled population's lambda, alpha + beta . c_q, with alpha = base +
// sum_a Qbar_a 2 u_a^3 and beta_a = ms ubar_a / rho + Qbar_a (k - 3 u_a^2) ms
// (base and ms ubar / rho from bu).
__global__ void __launch_bounds__(256) k_iso_adj_q(const double *__restrict__ dqbar, const double *__restrict__ uo,
                                                   const double *__restrict__ bu, double *__restrict__ ab, IsoArgs A)
{
    int x, y, z;
    int64_t cell;
    if (!iso_cell(A, x, y, z, cell)) return;
    const int pos[3] = {x, y, z}, n[3] = {A.nx, A.ny, A.nz};
    const int64_t st[3] = {1, A.nx, int64_t(A.nx) * A.ny};
    double alpha = bu[cell];
    double beta[3];
    for (int a = 0; a < 3; ++a) {
        const double *g = dqbar + a * A.N;
        const bool p_ok = pos[a] + 1 < n[a], m_ok = pos[a] - 1 >= 0;
        const double gc = g[cell];
        const double gm = g[cell - (m_ok ? st[a] : 0)];
        const double gp = g[cell + (p_ok ? st[a] : 0)];
        // own gradient: -Q (only p), +Q (only m), 0 otherwise
        double v = (p_ok && !m_ok) ? -gc : ((m_ok && !p_ok) ? gc : 0.0);
        // the neighbour below uses this cell as its p, the one above as its m
        if (m_ok) v += (pos[a] - 2 >= 0 ? 0.5 : 1.0) * gm;
        if (p_ok) v -= (pos[a] + 2 < n[a] ? 0.5 : 1.0) * gp;
        const double ua = uo[a * A.N + cell];
        alpha += v * (2.0 * ua * ua * ua);
        beta[a] = bu[(a + 1) * A.N + cell] + v * ((A.k - 3.0 * ua * ua) * A.ms);
    }
    double2 *r = reinterpret_cast<double2 *>(ab + 4 * cell);
    r[0] = make_double2(alpha, beta[0]);
    r[1] = make_double2(beta[1], beta[2]);
}

// dJ/du0 from lambda_0 (gathered) through the initial equilibrium
__global__ void __launch_bounds__(256) k_iso_u0grad(const double *__restrict__ lams, const double *__restrict__ ab,
                                                    const double *__restrict__ u0, double *__restrict__ du0, IsoArgs A,
                                                    double rho0)
{
    int x, y, z;
    int64_t cell;
    if (!iso_cell(A, x, y, z, cell)) return;
    double lo[NQ];
    iso_gather_lambda(lams, ab, A, x, y, z, lo);
    const double u[3] = {u0[cell], u0[A.N + cell], u0[2 * A.N + cell]};
    const double P[3] = {u[0] * u[0] + A.RT, u[1] * u[1] + A.RT, u[2] * u[2] + A.RT};
    double g[3] = {0.0, 0.0, 0.0};
    static_for<NQ>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        constexpr int c[3] = {ev(q, 0), ev(q, 1), ev(q, 2)};
        const double ph[3] = {iso_phi(c[0], u[0], P[0]), iso_phi(c[1], u[1], P[1]), iso_phi(c[2], u[2], P[2])};
        const double l = lo[q] * rho0;
        for (int a = 0; a < 3; ++a) {
            const double dphi = c[a] * 0.5 + ((c[a] != 0 ? 1.5 : 0.0) - 1.0) * 2.0 * u[a];
            g[a] += l * dphi * ph[(a + 1) % 3] * ph[(a + 2) % 3];
        }
    });
    for (int a = 0; a < 3; ++a) du0[a * A.N + cell] = g[a];
}

IsoArgs iso_args(const mrb_iso *S, double tau)
{
    IsoArgs A{};
    A.nx = S->n[0];
    A.ny = S->n[1];
    A.nz = S->n[2];
    A.N = S->N;
    A.tau = tau;
    A.w = 1.0 / tau;
    A.RT = 1.0 * S->T0;
    A.k = 1.0 - 3.0 * A.RT;
    A.ms = S->ms;
    A.dt = S->dt;
    for (int q = 0; q < NQ; ++q) A.inlet[q] = S->inlet[q];
    return A;
}

dim3 iso_grid(const IsoArgs &A) { return dim3(unsigned((A.nx + 63) / 64), unsigned((A.ny + 3) / 4), unsigned(A.nz)); }

int iso_step(const mrb_iso *S, const IsoArgs &A, const double *f, double *fo, const double *ns, hipStream_t s)
{
    k_iso_q<<<iso_grid(A), dim3(64, 4, 1), 0, s>>>(f, S->q, A);
    MRB_HIP(hipGetLastError());
    k_iso_step<<<iso_grid(A), dim3(64, 4, 1), 0, s>>>(f, fo, S->q, ns, A);
    MRB_HIP(hipGetLastError());
    return MRB_OK;
}

int iso_loss(const mrb_iso *S, const IsoArgs &A, const double *f, hipStream_t s, double *J)
{
    k_iso_loss<<<iso_grid(A), dim3(64, 4, 1), 0, s>>>(f, S->part, A);
    MRB_HIP(hipGetLastError());
    std::vector<double> h(size_t(S->nparts));
    MRB_HIP(hipMemcpyAsync(h.data(), S->part, sizeof(double) * h.size(), hipMemcpyDeviceToHost, s));
    MRB_HIP(hipStreamSynchronize(s));
    double acc = 0.0;
    for (double v : h) acc += v;
    *J = acc;
    return MRB_OK;
}

}  // namespace
}  // namespace mrb

using namespace mrb;

extern "C" {

int mrb_iso_create(const int ncell[3], double T0, double rho0, double u_in, mrb_iso **out)
{
    if (!ncell || !out) return set_error(MRB_E_ARG, "null argument");
    *out = nullptr;
    for (int d = 0; d < 3; ++d)
        if (ncell[d] < 3) return set_error(MRB_E_ARG, "every extent must be at least 3 cells");
    if (!(T0 > 0.0) || !(rho0 > 0.0)) return set_error(MRB_E_ARG, "T0 and rho0 must be positive");
    std::unique_ptr<mrb_iso> S(new mrb_iso());
    for (int d = 0; d < 3; ++d) S->n[d] = ncell[d];
    S->N = int64_t(ncell[0]) * ncell[1] * ncell[2];
    if (S->N >= (int64_t(1) << 31)) return set_error(MRB_E_ARG, "at most 2^31 cells");
    S->T0 = T0;
    S->rho0 = rho0;
    S->u_in = u_in;
    const double vel[3] = {u_in, 0.0, 0.0};
    for (int q = 0; q < NQ; ++q) S->inlet[q] = set_equilibrium_value(rho0, vel, 1.0 * T0, q);
    const size_t pop = sizeof(double) * NQ * size_t(S->N);
    for (int i = 0; i < 2; ++i) {
        MRB_HIP(hipMalloc(&S->f[i], pop));
        MRB_HIP(hipMalloc(&S->lam[i], pop));
    }
    MRB_HIP(hipMalloc(&S->q, sizeof(double) * 3 * size_t(S->N)));
    MRB_HIP(hipMalloc(&S->dqbar, sizeof(double) * 3 * size_t(S->N)));
    MRB_HIP(hipMalloc(&S->uo, sizeof(double) * 3 * size_t(S->N)));
    MRB_HIP(hipMalloc(&S->bu, sizeof(double) * 4 * size_t(S->N)));
    MRB_HIP(hipMalloc(&S->ab, sizeof(double) * 4 * size_t(S->N)));
    {
        IsoArgs A{};
        A.nx = S->n[0];
        A.ny = S->n[1];
        A.nz = S->n[2];
        const dim3 g = iso_grid(A);
        S->nparts = int64_t(g.x) * g.y * g.z;
    }
    MRB_HIP(hipMalloc(&S->part, sizeof(double) * size_t(S->nparts)));
    *out = S.release();
    return MRB_OK;
}

static void iso_free_states(mrb_iso *S)
{
    for (double *p : S->ckpt) (void)hipFree(p);
    for (double *p : S->seg) (void)hipFree(p);
    for (double *p : S->segq) (void)hipFree(p);
    S->ckpt.clear();
    S->seg.clear();
    S->segq.clear();
}

// nseg checkpoints, every - 1 segment states and every Q arrays
static int iso_ensure_states(mrb_iso *S, int nseg, int every)
{
    if (S->ckpt.size() == size_t(nseg) && S->seg.size() == size_t(every - 1) && S->segq.size() == size_t(every))
        return MRB_OK;
    MRB_HIP(hipDeviceSynchronize());
    iso_free_states(S);
    const size_t pop = sizeof(double) * NQ * size_t(S->N);
    S->ckpt.assign(size_t(nseg), nullptr);
    S->seg.assign(size_t(every - 1), nullptr);
    S->segq.assign(size_t(every), nullptr);
    for (auto &p : S->ckpt) MRB_HIP(hipMalloc(&p, pop));
    for (auto &p : S->seg) MRB_HIP(hipMalloc(&p, pop));
    for (auto &p : S->segq) MRB_HIP(hipMalloc(&p, sizeof(double) * 3 * size_t(S->N)));
    return MRB_OK;
}

int mrb_iso_destroy(mrb_iso *S)
{
    if (!S) return MRB_OK;
    (void)hipDeviceSynchronize();
    iso_free_states(S);
    for (int i = 0; i < 2; ++i) {
        (void)hipFree(S->f[i]);
        (void)hipFree(S->lam[i]);
    }
    (void)hipFree(S->q);
    (void)hipFree(S->dqbar);
    (void)hipFree(S->uo);
    (void)hipFree(S->bu);
    (void)hipFree(S->ab);
    (void)hipFree(S->part);
    delete S;
    return MRB_OK;
}

int mrb_iso_forward(mrb_iso *S, const double *ns, double tau, const double *u0, int nsteps, double *J, void *stream)
{
    if (!S || !ns || !u0 || !J) return set_error(MRB_E_ARG, "null argument");
    if (!(tau > 0.5) || nsteps < 0) return set_error(MRB_E_ARG, "tau must exceed 1/2 and nsteps be >= 0");
    hipStream_t s = as_stream(stream);
    S->rec_valid = false;  // the ping-pong pair holds the recording's final state
    const IsoArgs A = iso_args(S, tau);
    k_iso_init<<<iso_grid(A), dim3(64, 4, 1), 0, s>>>(S->f[0], u0, A, S->rho0);
    MRB_HIP(hipGetLastError());
    int cur = 0;
    for (int n = 0; n < nsteps; ++n) {
        const int rc = iso_step(S, A, S->f[cur], S->f[1 - cur], ns, s);
        if (rc) return rc;
        cur = 1 - cur;
    }
    if (cur != 0) MRB_HIP(hipMemcpyAsync(S->f[0], S->f[1], sizeof(double) * NQ * size_t(S->N),
                                         hipMemcpyDeviceToDevice, s));
    return iso_loss(S, A, S->f[0], s, J);
}

// Device bytes of the recording buffers for segments of `every` steps:
// one checkpoint per segment, every - 1 further states and every Q arrays
static double iso_record_bytes(const mrb_iso *S, int nsteps, int every)
{
    const double pop = double(sizeof(double)) * NQ * double(S->N), q3 = double(sizeof(double)) * 3 * double(S->N);
    const int nseg = (nsteps + every - 1) / every;
    return double(nseg + every - 1) * pop + double(every) * q3;
}

// The longest segment whose recording fits half of the device memory that
// is free (counting the recording buffers already held): the whole run
// when it fits, so that the reverse sweep recomputes nothing; never
// shorter than ceil(sqrt(nsteps)), whose buffers are the fewest
static int iso_auto_every(const mrb_iso *S, int nsteps, int sqrt_every)
{
    size_t freeb = 0, totalb = 0;
    if (hipMemGetInfo(&freeb, &totalb) != hipSuccess) return sqrt_every;
    const double pop = double(sizeof(double)) * NQ * double(S->N), q3 = double(sizeof(double)) * 3 * double(S->N);
    const double held = double(S->ckpt.size() + S->seg.size()) * pop + double(S->segq.size()) * q3;
    const double budget = 0.5 * (double(freeb) + held);
    for (int e = nsteps; e > sqrt_every; --e)
        if (iso_record_bytes(S, nsteps, e) <= budget) return e;
    return sqrt_every;
}

int mrb_iso_record(mrb_iso *S, const double *ns, double tau, const double *u0, int nsteps, int every, double *J,
                   void *stream)
{
    if (!S || !ns || !u0 || !J) return set_error(MRB_E_ARG, "null argument");
    if (!(tau > 0.5) || nsteps < 1) return set_error(MRB_E_ARG, "tau must exceed 1/2 and nsteps be >= 1");
    const int sqrt_every = std::max(1, int(std::ceil(std::sqrt(double(nsteps)))));
    if (every == 0) every = sqrt_every;
    if (every < 0) every = iso_auto_every(S, nsteps, std::min(sqrt_every, nsteps));
    every = std::min(every, nsteps);
    S->rec_valid = false;
    hipStream_t s = as_stream(stream);
    const IsoArgs A = iso_args(S, tau);
    const dim3 g = iso_grid(A);
    const dim3 b(64, 4, 1);
    const int nseg = (nsteps + every - 1) / every;
    const int last0 = (nseg - 1) * every;  // first step of the last segment
    {
        const int rc = iso_ensure_states(S, nseg, every);
        if (rc) return rc;
    }
    auto state = [&](int sg, int i) { return i == 0 ? S->ckpt[size_t(sg)] : S->seg[size_t(i - 1)]; };
    // forward sweep: checkpoints written in place, the last segment's states
    // and Q arrays kept, the other states in the ping-pong pair
    auto fwd_state = [&](int n) {
        if (n >= last0) return state(nseg - 1, n - last0);
        if (n % every == 0) return S->ckpt[size_t(n / every)];
        return S->f[n & 1];
    };
    k_iso_init<<<g, b, 0, s>>>(fwd_state(0), u0, A, S->rho0);
    MRB_HIP(hipGetLastError());
    for (int n = 0; n < nsteps; ++n) {
        const double *fn = fwd_state(n);
        double *fo = n + 1 < nsteps ? fwd_state(n + 1) : S->f[nsteps & 1];
        double *qn = n >= last0 ? S->segq[size_t(n - last0)] : S->q;
        k_iso_q<<<g, b, 0, s>>>(fn, qn, A);
        MRB_HIP(hipGetLastError());
        k_iso_step<<<g, b, 0, s>>>(fn, fo, qn, ns, A);
        MRB_HIP(hipGetLastError());
    }
    const int rc = iso_loss(S, A, S->f[nsteps & 1], s, J);
    if (rc) return rc;
    S->rec_valid = true;
    S->rec_nsteps = nsteps;
    S->rec_every = every;
    S->rec_tau = tau;
    S->rec_ns = ns;
    S->rec_u0 = u0;
    return MRB_OK;
}

int mrb_iso_reverse(mrb_iso *S, double *dtau, double *dns, double *du0, void *stream)
{
    if (!S || !dtau || !dns || !du0) return set_error(MRB_E_ARG, "null argument");
    if (!S->rec_valid) return set_error(MRB_E_STATE, "no recorded forward sweep (mrb_iso_record)");
    hipStream_t s = as_stream(stream);
    const int nsteps = S->rec_nsteps, every = S->rec_every;
    const double tau = S->rec_tau;
    const double *ns = S->rec_ns, *u0 = S->rec_u0;
    const IsoArgs A = iso_args(S, tau);
    const dim3 g = iso_grid(A);
    const dim3 b(64, 4, 1);
    const int nseg = (nsteps + every - 1) / every;
    auto state = [&](int sg, int i) { return i == 0 ? S->ckpt[size_t(sg)] : S->seg[size_t(i - 1)]; };
    // reverse sweep: two passes per step
    MRB_HIP(hipMemsetAsync(dns, 0, sizeof(double) * size_t(S->N), s));
    MRB_HIP(hipMemsetAsync(S->part, 0, sizeof(double) * size_t(S->nparts), s));
    k_iso_seed<<<g, b, 0, s>>>(S->f[nsteps & 1], S->lam[0], A);
    MRB_HIP(hipGetLastError());
    int lc = 0;
    bool first = true;
    for (int sg = nseg - 1; sg >= 0; --sg) {
        const int s0 = sg * every, len = std::min(nsteps, s0 + every) - s0;
        if (sg < nseg - 1) {
            // recompute the segment's states and their Q arrays from its checkpoint
            for (int i = 0; i < len; ++i) {
                k_iso_q<<<g, b, 0, s>>>(state(sg, i), S->segq[size_t(i)], A);
                MRB_HIP(hipGetLastError());
                if (i + 1 < len) {
                    k_iso_step<<<g, b, 0, s>>>(state(sg, i), state(sg, i + 1), S->segq[size_t(i)], ns, A);
                    MRB_HIP(hipGetLastError());
                }
            }
        }
        for (int i = len - 1; i >= 0; --i) {
            const double *fn = state(sg, i);
            const double *qn = S->segq[size_t(i)];
            if (first)
                k_iso_adj<false><<<g, b, 0, s>>>(fn, qn, S->lam[lc], S->ab, ns, S->lam[1 - lc], S->dqbar, S->uo,
                                                  S->bu, dns, S->part, A);
            else
                k_iso_adj<true><<<g, b, 0, s>>>(fn, qn, S->lam[lc], S->ab, ns, S->lam[1 - lc], S->dqbar, S->uo,
                                                 S->bu, dns, S->part, A);
            MRB_HIP(hipGetLastError());
            k_iso_adj_q<<<g, b, 0, s>>>(S->dqbar, S->uo, S->bu, S->ab, A);
            MRB_HIP(hipGetLastError());
            lc = 1 - lc;
            first = false;
        }
    }
    // the segment buffers now hold recomputed states of the first segment;
    // the recording is used up
    S->rec_valid = false;
    k_iso_u0grad<<<g, b, 0, s>>>(S->lam[lc], S->ab, u0, du0, A, S->rho0);
    MRB_HIP(hipGetLastError());
    std::vector<double> part(size_t(S->nparts));
    MRB_HIP(hipMemcpyAsync(part.data(), S->part, sizeof(double) * part.size(), hipMemcpyDeviceToHost, s));
    MRB_HIP(hipStreamSynchronize(s));
    double gw = 0.0;
    for (double v : part) gw += v;
    *dtau = gw * (-1.0 / (tau * tau));  // w = 1/tau
    return MRB_OK;
}

int mrb_iso_gradient(mrb_iso *S, const double *ns, double tau, const double *u0, int nsteps, int every, double *J,
                     double *dtau, double *dns, double *du0, void *stream)
{
    if (!S || !ns || !u0 || !J || !dtau || !dns || !du0) return set_error(MRB_E_ARG, "null argument");
    const int rc = mrb_iso_record(S, ns, tau, u0, nsteps, every, J, stream);
    if (rc) return rc;
    return mrb_iso_reverse(S, dtau, dns, du0, stream);
}

int mrb_iso_record_interval(const mrb_iso *S, int *every)
{
    if (!S || !every) return set_error(MRB_E_ARG, "null argument");
    *every = S->rec_every;
    return MRB_OK;
}

int mrb_iso_state(const mrb_iso *S, double *host_f)
{
    if (!S || !host_f) return set_error(MRB_E_ARG, "null argument");
    MRB_HIP(hipDeviceSynchronize());
    MRB_HIP(hipMemcpy(host_f, S->f[0], sizeof(double) * NQ * size_t(S->N), hipMemcpyDeviceToHost));
    return MRB_OK;
}

}  // extern "C"
