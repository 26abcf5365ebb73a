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
// Adjoint of one step (lambda = dJ/d populations), three passes:
//   a: per cell, the transpose of collide + partial bounce-back: lambda at
//      the pulled populations (without the Q-correction term), dJ/d(dQ),
//      dJ/dn_s (accumulated), dJ/dtau (block partial sums)
//   b: per cell, the transpose of dQ (a 3-point gather along each axis) and
//      of Q(rho, u) added to the pulled populations' lambda
//   c: the transpose of the pull (a push, written as a gather)
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
    double *lam[2] = {nullptr, nullptr};
    double *lams = nullptr;   // 27 N: lambda of the pulled populations
    double *q = nullptr;      // 3 N: Q of the current state
    double *dqbar = nullptr;  // 3 N
    double *ru = nullptr;     // 4 N: rho, u of the pulled state
    double *part = nullptr;   // one partial sum per block of the cell grid
    int64_t nparts = 0;
    std::vector<double *> ckpt, seg, segq;
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

// pull of one cell (see the file comment)
__device__ __forceinline__ void iso_pull(const double *__restrict__ f, const IsoArgs &A, int x, int y, int z,
                                         double fs[NQ])
{
    static_for<NQ>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        constexpr int cx = ev(q, 0), cy = ev(q, 1), cz = ev(q, 2);
        const int sy = wrap(y - cy, A.ny), sz = wrap(z - cz, A.nz);
        int sx = x - cx;
        const bool inlet = sx < 0;
        sx = sx >= A.nx ? A.nx - 1 : (sx < 0 ? 0 : sx);
        const double v = f[int64_t(q) * A.N + (int64_t(sz) * A.ny + sy) * A.nx + sx];
        fs[q] = inlet ? A.inlet[q] : v;
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

// dQ_a at a cell from the Q array (domain-box rule on every axis)
__device__ __forceinline__ void iso_dq(const double *__restrict__ Q, const IsoArgs &A, int x, int y, int z,
                                       int64_t cell, double dq[3])
{
    const int pos[3] = {x, y, z}, n[3] = {A.nx, A.ny, A.nz};
    const int64_t st[3] = {1, A.nx, int64_t(A.nx) * A.ny};
    for (int a = 0; a < 3; ++a) {
        const double *Qa = Q + a * A.N;
        const bool p_ok = pos[a] + 1 < n[a], m_ok = pos[a] - 1 >= 0;
        const double qp = p_ok ? Qa[cell + st[a]] : 0.0;
        const double qm = m_ok ? Qa[cell - st[a]] : 0.0;
        dq[a] = one_sided_gradient(p_ok, m_ok, qp, Qa[cell], qm, 1.0);
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

// pass a: transpose of collide + partial bounce-back at one cell
__global__ void __launch_bounds__(256) k_iso_adj_a(
    const double *__restrict__ f, const double *__restrict__ Q, const double *__restrict__ lam_in,
    const double *__restrict__ ns, double *__restrict__ lams, double *__restrict__ dqbar, double *__restrict__ ru,
    double *__restrict__ dns, double *__restrict__ part, IsoArgs A)
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
        double lo[NQ];
        for (int q = 0; q < NQ; ++q) lo[q] = lam_in[int64_t(q) * A.N + cell];
        double rbar = 0.0, ubar[3] = {0.0, 0.0, 0.0}, pbar[3] = {0.0, 0.0, 0.0}, gns = 0.0;
        double ls[NQ];
        static_for<NQ>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            constexpr int cx = ev(q, 0), cy = ev(q, 1), cz = ev(q, 2);
            const double px = iso_phi(cx, u[0], P[0]);
            const double py = iso_phi(cy, u[1], P[1]);
            const double pz = iso_phi(cz, u[2], P[2]);
            const double feq = ((rho * px) * py) * pz;
            const double fc = fs[q] + w * (feq - fs[q]);
            const double lc = (1.0 - s) * lo[q];
            gns += lo[q] * (fs[kBounce[q]] - fc);
            gw += lc * (feq - fs[q]);
            const double mu = w * lc;
            rbar += mu * (px * py * pz);
            const double gx = mu * rho * py * pz, gy = mu * rho * px * pz, gz = mu * rho * px * py;
            ubar[0] += gx * (cx * 0.5);
            ubar[1] += gy * (cy * 0.5);
            ubar[2] += gz * (cz * 0.5);
            pbar[0] += gx * ((cx != 0 ? 1.5 : 0.0) - 1.0);
            pbar[1] += gy * ((cy != 0 ? 1.5 : 0.0) - 1.0);
            pbar[2] += gz * ((cz != 0 ? 1.5 : 0.0) - 1.0);
            ls[q] = (1.0 - w) * lc;
        });
        // partial bounce-back's reflected term: lambda*_q += n_s lambda_qbar
        static_for<NQ>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            ls[q] += s * lo[kBounce[q]];
        });
        double wcbar = 0.0;
        for (int a = 0; a < 3; ++a) {
            ubar[a] += 2.0 * u[a] * pbar[a];
            wcbar += pbar[a] * A.dt * dq[a];
            dqbar[a * A.N + cell] = pbar[a] * A.dt * wc;
        }
        rbar += wcbar * (-wc / rho);
        gw += wcbar * (-1.0 / (w * w * rho));
        static_for<NQ>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            double v = ls[q] + rbar;
            v += ubar[0] * (A.ms * ev(q, 0) - u[0]) / rho;
            v += ubar[1] * (A.ms * ev(q, 1) - u[1]) / rho;
            v += ubar[2] * (A.ms * ev(q, 2) - u[2]) / rho;
            lams[int64_t(q) * A.N + cell] = v;
        });
        ru[cell] = rho;
        ru[A.N + cell] = u[0];
        ru[2 * A.N + cell] = u[1];
        ru[3 * A.N + cell] = u[2];
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

// pass b: transpose of dQ (gather) and of Q(rho, u)
__global__ void __launch_bounds__(256) k_iso_adj_b(double *__restrict__ lams, const double *__restrict__ dqbar,
                                                   const double *__restrict__ ru, IsoArgs A)
{
    int x, y, z;
    int64_t cell;
    if (!iso_cell(A, x, y, z, cell)) return;
    const int pos[3] = {x, y, z}, n[3] = {A.nx, A.ny, A.nz};
    const int64_t st[3] = {1, A.nx, int64_t(A.nx) * A.ny};
    double qbar[3];
    for (int a = 0; a < 3; ++a) {
        const double *g = dqbar + a * A.N;
        const bool p_ok = pos[a] + 1 < n[a], m_ok = pos[a] - 1 >= 0;
        // own gradient: vc = -Q (only p), +Q (only m), 0 otherwise
        double v = (p_ok && !m_ok) ? -g[cell] : ((m_ok && !p_ok) ? g[cell] : 0.0);
        // the neighbour below uses this cell as its p
        if (m_ok) {
            const bool mm_ok = pos[a] - 2 >= 0;
            v += (mm_ok ? 0.5 : 1.0) * g[cell - st[a]];
        }
        // the neighbour above uses this cell as its m
        if (p_ok) {
            const bool pp_ok = pos[a] + 2 < n[a];
            v -= (pp_ok ? 0.5 : 1.0) * g[cell + st[a]];
        }
        qbar[a] = v;  // idx = 1 (unit cells)
    }
    const double u[3] = {ru[A.N + cell], ru[2 * A.N + cell], ru[3 * A.N + cell]};
    double cq[3], du[3];
    for (int a = 0; a < 3; ++a) {
        // dQ_a/df_q = u k - u^3 + (k - 3u^2)(ms c - u)
        cq[a] = u[a] * A.k - u[a] * u[a] * u[a] - (A.k - 3.0 * u[a] * u[a]) * u[a];
        du[a] = (A.k - 3.0 * u[a] * u[a]) * A.ms;
    }
    static_for<NQ>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        double v = 0.0;
        v += qbar[0] * (cq[0] + du[0] * ev(q, 0));
        v += qbar[1] * (cq[1] + du[1] * ev(q, 1));
        v += qbar[2] * (cq[2] + du[2] * ev(q, 2));
        lams[int64_t(q) * A.N + cell] += v;
    });
}

// pass c: transpose of the pull
__global__ void __launch_bounds__(256) k_iso_adj_c(const double *__restrict__ lams, double *__restrict__ lam_out,
                                                   IsoArgs A)
{
    int x, y, z;
    int64_t cell;
    if (!iso_cell(A, x, y, z, cell)) return;
    static_for<NQ>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        constexpr int cx = ev(q, 0), cy = ev(q, 1), cz = ev(q, 2);
        const int ty = wrap(y + cy, A.ny), tz = wrap(z + cz, A.nz);
        const int tx = x + cx;
        double v = 0.0;
        if (tx >= 0 && tx < A.nx) v += lams[int64_t(q) * A.N + (int64_t(tz) * A.ny + ty) * A.nx + tx];
        if (cx == -1 && x == A.nx - 1)  // read again through the outflow ghost
            v += lams[int64_t(q) * A.N + (int64_t(tz) * A.ny + ty) * A.nx + x];
        lam_out[int64_t(q) * A.N + cell] = v;
    });
}

// dJ/du0 from lambda_0 through the initial equilibrium
__global__ void __launch_bounds__(256) k_iso_u0grad(const double *__restrict__ lam, const double *__restrict__ u0,
                                                    double *__restrict__ du0, IsoArgs A, double rho0)
{
    int x, y, z;
    int64_t cell;
    if (!iso_cell(A, x, y, z, cell)) return;
    const double u[3] = {u0[cell], u0[A.N + cell], u0[2 * A.N + cell]};
    const double P[3] = {u[0] * u[0] + A.RT, u[1] * u[1] + A.RT, u[2] * u[2] + A.RT};
    double g[3] = {0.0, 0.0, 0.0};
    static_for<NQ>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        constexpr int c[3] = {ev(q, 0), ev(q, 1), ev(q, 2)};
        const double ph[3] = {iso_phi(c[0], u[0], P[0]), iso_phi(c[1], u[1], P[1]), iso_phi(c[2], u[2], P[2])};
        const double l = lam[int64_t(q) * A.N + cell] * rho0;
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
    MRB_HIP(hipMalloc(&S->lams, pop));
    MRB_HIP(hipMalloc(&S->q, sizeof(double) * 3 * size_t(S->N)));
    MRB_HIP(hipMalloc(&S->dqbar, sizeof(double) * 3 * size_t(S->N)));
    MRB_HIP(hipMalloc(&S->ru, sizeof(double) * 4 * size_t(S->N)));
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

int mrb_iso_destroy(mrb_iso *S)
{
    if (!S) return MRB_OK;
    (void)hipDeviceSynchronize();
    iso_free_states(S);
    for (int i = 0; i < 2; ++i) {
        (void)hipFree(S->f[i]);
        (void)hipFree(S->lam[i]);
    }
    (void)hipFree(S->lams);
    (void)hipFree(S->q);
    (void)hipFree(S->dqbar);
    (void)hipFree(S->ru);
    (void)hipFree(S->part);
    delete S;
    return MRB_OK;
}

int mrb_iso_forward(mrb_iso *S, const double *ns, double tau, const double *u0, int nsteps, double *J, void *stream)
{
    if (!S || !ns || !u0 || !J) return set_error(MRB_E_ARG, "null argument");
    if (!(tau > 0.5) || nsteps < 0) return set_error(MRB_E_ARG, "tau must exceed 1/2 and nsteps be >= 0");
    hipStream_t s = as_stream(stream);
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

int mrb_iso_gradient(mrb_iso *S, const double *ns, double tau, const double *u0, int nsteps, int every, double *J,
                     double *dtau, double *dns, double *du0, void *stream)
{
    if (!S || !ns || !u0 || !J || !dtau || !dns || !du0) return set_error(MRB_E_ARG, "null argument");
    if (!(tau > 0.5) || nsteps < 1) return set_error(MRB_E_ARG, "tau must exceed 1/2 and nsteps be >= 1");
    if (every < 1) every = std::max(1, int(std::ceil(std::sqrt(double(nsteps)))));
    hipStream_t s = as_stream(stream);
    const IsoArgs A = iso_args(S, tau);
    const dim3 g = iso_grid(A);
    const dim3 b(64, 4, 1);
    const size_t pop = sizeof(double) * NQ * size_t(S->N);
    const int nseg = (nsteps + every - 1) / every;
    // checkpoints and one segment of states (sqrt(N) memory)
    iso_free_states(S);
    S->ckpt.assign(size_t(nseg), nullptr);
    S->seg.assign(size_t(every), nullptr);
    S->segq.assign(size_t(every), nullptr);
    for (auto &p : S->ckpt) MRB_HIP(hipMalloc(&p, pop));
    for (auto &p : S->seg) MRB_HIP(hipMalloc(&p, pop));
    for (auto &p : S->segq) MRB_HIP(hipMalloc(&p, sizeof(double) * 3 * size_t(S->N)));
    // forward sweep, keeping every `every`-th state
    k_iso_init<<<g, b, 0, s>>>(S->f[0], u0, A, S->rho0);
    MRB_HIP(hipGetLastError());
    int cur = 0;
    for (int n = 0; n < nsteps; ++n) {
        if (n % every == 0)
            MRB_HIP(hipMemcpyAsync(S->ckpt[size_t(n / every)], S->f[cur], pop, hipMemcpyDeviceToDevice, s));
        const int rc = iso_step(S, A, S->f[cur], S->f[1 - cur], ns, s);
        if (rc) return rc;
        cur = 1 - cur;
    }
    int rc = iso_loss(S, A, S->f[cur], s, J);
    if (rc) return rc;
    // reverse sweep
    MRB_HIP(hipMemsetAsync(dns, 0, sizeof(double) * size_t(S->N), s));
    MRB_HIP(hipMemsetAsync(S->part, 0, sizeof(double) * size_t(S->nparts), s));
    k_iso_seed<<<g, b, 0, s>>>(S->f[cur], S->lam[0], A);
    MRB_HIP(hipGetLastError());
    int lc = 0;
    for (int sg = nseg - 1; sg >= 0; --sg) {
        const int s0 = sg * every, s1 = std::min(nsteps, s0 + every);
        // recompute the segment's states and their Q arrays
        MRB_HIP(hipMemcpyAsync(S->seg[0], S->ckpt[size_t(sg)], pop, hipMemcpyDeviceToDevice, s));
        for (int i = 0; i < s1 - s0; ++i) {
            k_iso_q<<<g, b, 0, s>>>(S->seg[size_t(i)], S->segq[size_t(i)], A);
            MRB_HIP(hipGetLastError());
            if (i + 1 < s1 - s0) {
                k_iso_step<<<g, b, 0, s>>>(S->seg[size_t(i)], S->seg[size_t(i + 1)], S->segq[size_t(i)], ns, A);
                MRB_HIP(hipGetLastError());
            }
        }
        for (int n = s1 - 1; n >= s0; --n) {
            const double *fn = S->seg[size_t(n - s0)];
            const double *qn = S->segq[size_t(n - s0)];
            k_iso_adj_a<<<g, b, 0, s>>>(fn, qn, S->lam[lc], ns, S->lams, S->dqbar, S->ru, dns, S->part, A);
            MRB_HIP(hipGetLastError());
            k_iso_adj_b<<<g, b, 0, s>>>(S->lams, S->dqbar, S->ru, A);
            MRB_HIP(hipGetLastError());
            k_iso_adj_c<<<g, b, 0, s>>>(S->lams, S->lam[1 - lc], A);
            MRB_HIP(hipGetLastError());
            lc = 1 - lc;
        }
    }
    k_iso_u0grad<<<g, b, 0, s>>>(S->lam[lc], u0, du0, A, S->rho0);
    MRB_HIP(hipGetLastError());
    std::vector<double> part(size_t(S->nparts));
    MRB_HIP(hipMemcpyAsync(part.data(), S->part, sizeof(double) * part.size(), hipMemcpyDeviceToHost, s));
    MRB_HIP(hipStreamSynchronize(s));
    double gw = 0.0;
    for (double v : part) gw += v;
    *dtau = gw * (-1.0 / (tau * tau));  // w = 1/tau
    iso_free_states(S);
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
