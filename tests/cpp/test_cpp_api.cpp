// The C++ host surface (include/marbles_amd.hpp) on the GPU, against the
// oracle's restatement (oracle/lbm_oracle.h) of the same steps:
//   - a periodic Taylor-Green box with an EB sphere on a fab level stepped
//     through the reference's stage calls (fillpatch, stream, collide),
//   - the same state on a compact level through the fused advance,
//   - Stencil tables, configuration errors thrown as marbles_amd::Error.
// Prints one line and exits 0 when every check holds.
#include <cmath>
#include <cstdio>
#include <vector>

#include "lbm_oracle.h"
#include "marbles_amd.hpp"

namespace {

int failures = 0;

void expect(bool ok, const char *what)
{
    if (!ok) {
        std::printf("FAILED: %s\n", what);
        ++failures;
    }
}

// valid cells (x fastest) of a fab with ng ghosts and extents n, per comp
std::vector<double> valid_of(const std::vector<double> &fab, const int n[3], int ng, int ncomp)
{
    const int fx = n[0] + 2 * ng, fy = n[1] + 2 * ng, fz = n[2] + 2 * ng;
    std::vector<double> out;
    for (int c = 0; c < ncomp; ++c)
        for (int k = 0; k < n[2]; ++k)
            for (int j = 0; j < n[1]; ++j)
                for (int i = 0; i < n[0]; ++i)
                    out.push_back(fab[size_t(((int64_t(c) * fz + k + ng) * fy + j + ng) * fx + i + ng)]);
    return out;
}

double rel_err(const std::vector<double> &a, const std::vector<double> &b)
{
    double m = 0.0, s = 1e-300;
    for (size_t i = 0; i < a.size(); ++i) {
        m = std::fmax(m, std::fabs(a[i] - b[i]));
        s = std::fmax(s, std::fabs(b[i]));
    }
    return m / s;
}

}  // namespace

int main()
{
    using namespace marbles_amd;
    const int n[3] = {16, 12, 10}, ng = 3;
    const int lo[3] = {0, 0, 0}, hi[3] = {n[0] - 1, n[1] - 1, n[2] - 1}, per[3] = {1, 1, 1};
    const double plo[3] = {0.0, 0.0, 0.0}, phi[3] = {16.0, 12.0, 10.0};
    const double T0 = 0.2, nu = 0.1 * T0;

    orc_level *O = orc_create(lo, hi, per, plo, phi, 1, lo, hi, ng);
    const orc_phys op{nu, nu, 5.0 / 3.0, 28.96, 1.0, 1.0};
    orc_set_phys(O, &op);
    orc_bc ob{};
    orc_set_bc(O, &ob);
    const double c[3] = {8.0, 6.0, 5.0};
    orc_init_flags_sphere(O, c, 2.5, 0);
    const double omega[3] = {2 * M_PI / 16, 2 * M_PI / 12, 2 * M_PI / 10};
    orc_ic_taylorgreen(O, 1.0, 0.03, omega);
    orc_init_finish(O);
    auto fab = [&](int kind) {
        std::vector<double> v(size_t(orc_fab_size(O, kind, 0)));
        orc_get_fab(O, kind, 0, v.data());
        return v;
    };
    std::vector<int32_t> isf(size_t(orc_fab_size(O, ORC_FAB_ISFLUID, 0)));
    orc_get_fab(O, ORC_FAB_ISFLUID, 0, isf.data());
    const std::vector<double> f0 = fab(ORC_FAB_F), g0 = fab(ORC_FAB_G);

    const mrb_phys ph{nu, nu, 5.0 / 3.0, 28.96, 1.0, 1.0};
    mrb_bc bc{};
    const std::array<int, 3> N{n[0], n[1], n[2]}, P{1, 1, 1};
    const std::array<double, 3> PL{0.0, 0.0, 0.0}, PH{16.0, 12.0, 10.0};

    // fab level, the reference's stage calls
    Level E(N, P, PL, PH, {Box{{0, 0, 0}, {n[0] - 1, n[1] - 1, n[2] - 1}}}, ng);
    E.set_phys(ph);
    E.set_bc(bc);
    E.set_is_fluid(isf);
    E.upload(MRB_ARR_F, f0);
    E.upload(MRB_ARR_G, g0);
    const int steps = 6;
    for (int s = 0; s < steps; ++s) {
        for (int lat = 0; lat < 2; ++lat) {
            E.fillpatch(lat);
            E.stream(lat);
        }
        E.collide();
    }
    // compact level, the fused step, from the same initial state
    Level C = Level::compact(N, P, PL, PH);
    C.set_phys(ph);
    C.set_bc(bc);
    {
        std::vector<int32_t> vi;
        const int fx = n[0] + 2 * ng, fy = n[1] + 2 * ng, fz = n[2] + 2 * ng;
        for (int comp = 0; comp < 2; ++comp)
            for (int k = 0; k < n[2]; ++k)
                for (int j = 0; j < n[1]; ++j)
                    for (int i = 0; i < n[0]; ++i)
                        vi.push_back(isf[size_t(((int64_t(comp) * fz + k + ng) * fy + j + ng) * fx + i + ng)]);
        C.set_is_fluid(vi);
    }
    C.upload(MRB_ARR_F, valid_of(f0, n, ng, 27));
    C.upload(MRB_ARR_G, valid_of(g0, n, ng, 27));
    C.advance(steps);

    orc_evolve(O, steps);
    for (int kind : {ORC_FAB_F, ORC_FAB_G}) {
        const std::vector<double> ref = valid_of(fab(kind), n, ng, 27);
        const int arr = kind == ORC_FAB_F ? MRB_ARR_F : MRB_ARR_G;
        const double es = rel_err(valid_of(E.download(arr), n, ng, 27), ref);
        const double ec = rel_err(C.download(arr), ref);
        std::printf("lattice %d: stages rel err %.3e, fused %.3e\n", kind, es, ec);
        expect(es <= 1e-12, "stage calls match the oracle");
        expect(ec <= 1e-12, "fused advance matches the oracle");
    }
    const auto cons = C.conservation();
    expect(std::isfinite(cons[0]) && cons[0] > 0.0, "conservation sums");

    Stencil st;
    double wsum = 0.0;
    bool opposite = true;
    for (int q = 0; q < Stencil::N_MICRO_STATES; ++q) {
        wsum += st.weights[size_t(q)];
        for (int d = 0; d < 3; ++d) opposite = opposite && st.evs[size_t(st.bounce_dirs[size_t(q)])][size_t(d)] == -st.evs[size_t(q)][size_t(d)];
    }
    expect(std::fabs(wsum - 1.0) < 1e-15 && opposite && st.evs[0] == std::array<int, 3>{0, 0, 0}, "stencil tables");

    bool threw = false;
    try {
        mrb_bc bad{};
        bad.bc_type[0] = 2;  // a velocity face on a periodic axis
        E.set_bc(bad);
    } catch (const Error &e) {
        threw = e.code() == MRB_E_ARG;
    }
    expect(threw, "a bad configuration throws marbles_amd::Error");

    orc_destroy(O);
    if (failures) return 1;
    std::printf("cpp api ok: stage calls and fused advance match the oracle (%d steps)\n", steps);
    return 0;
}
