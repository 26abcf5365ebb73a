/* C++17 host surface over libmarbles_amd.so (include/marbles_amd.h),
 * mirroring the reference solver's calls so that a caller of NREL/marbles
 * replaces them one for one:
 *
 *   reference (Source/)                           here
 *   lbm::stencil::Stencil (Stencil.H:41-169)      marbles_amd::Stencil
 *   MultiFab definitions of                       marbles_amd::Level
 *     LBM::MakeNewLevelFromScratch (LBM.cpp:1148-1199)
 *   FillPatchOps::fillpatch / physbc              Level::fillpatch / physbc
 *     (FillPatchOps.H:66-159), one level
 *   LBM::stream (LBM.cpp:558-604)                 Level::stream
 *   LBM::collide (LBM.cpp:607-618)                Level::collide
 *   LBM::advance (LBM.cpp:523-544), fused         Level::advance
 *   average_down_with_ghosts (Utilities.H:353)    TwoLevel::average_down
 *   two-level FillPatchOps::fillpatch (118-129)   TwoLevel::fillpatch_fine
 *   LBM::evolve for finest_level = 1              TwoLevel::advance
 *   LBM::write_checkpoint_file / read_...         write_checkpoint / read_checkpoint
 *   LBM::write_plot_file                          Level::write_plotfile
 *
 * Errors: where the reference calls amrex::Abort on a bad configuration,
 * every call here throws marbles_amd::Error carrying the library's status
 * code and message.  Streams are hipStream_t passed as void* (nullptr: the
 * default stream); calls on a stream are asynchronous unless the C ABI
 * documents them as synchronous.  Header only; link -lmarbles_amd. */
#pragma once

#include <array>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "marbles_amd.h"

namespace marbles_amd {

class Error : public std::runtime_error {
  public:
    Error(int code, const std::string &what) : std::runtime_error(what), code_(code) {}
    int code() const { return code_; }

  private:
    int code_;
};

inline void check(int rc)
{
    if (rc != MRB_OK) {
        const char *m = mrb_last_error();
        throw Error(rc, m ? m : "marbles_amd error");
    }
}

struct Stencil {
    static constexpr int N_MICRO_STATES = 27;
    static constexpr double THETA0 = 1.0 / 3.0;
    std::array<std::array<int, 3>, N_MICRO_STATES> evs{};
    std::array<double, N_MICRO_STATES> weights{};
    std::array<int, N_MICRO_STATES> bounce_dirs{};

    Stencil()
    {
        int e[3 * N_MICRO_STATES];
        check(mrb_stencil(e, weights.data(), bounce_dirs.data()));
        for (int q = 0; q < N_MICRO_STATES; ++q)
            for (int d = 0; d < 3; ++d) evs[q][d] = e[3 * q + d];
    }
};

struct Box {
    std::array<int, 3> lo, hi;  // valid cells, inclusive
};

class Level {
  public:
    Level(const std::array<int, 3> &ncell, const std::array<int, 3> &periodic, const std::array<double, 3> &prob_lo,
          const std::array<double, 3> &prob_hi, const std::vector<Box> &boxes, int ng = 3,
          int layout = MRB_LAYOUT_FAB, int storage = MRB_STORAGE_F64)
        : boxes_(boxes)
    {
        std::vector<int> lo, hi;
        for (const Box &b : boxes)
            for (int d = 0; d < 3; ++d) {
                lo.push_back(b.lo[d]);
                hi.push_back(b.hi[d]);
            }
        mrb_level_desc desc{};
        for (int d = 0; d < 3; ++d) {
            desc.ncell[d] = ncell[d];
            desc.periodic[d] = periodic[d];
            desc.prob_lo[d] = prob_lo[d];
            desc.prob_hi[d] = prob_hi[d];
        }
        desc.nbox = int(boxes.size());
        desc.box_lo = lo.data();
        desc.box_hi = hi.data();
        desc.ng = layout == MRB_LAYOUT_FAB ? ng : 0;
        desc.layout = layout;
        desc.nlattice = 2;
        desc.storage = storage;
        check(mrb_level_create(&desc, &h_));
    }
    /* one box over the domain, compact layout (the fused hot path) */
    static Level compact(const std::array<int, 3> &ncell, const std::array<int, 3> &periodic,
                         const std::array<double, 3> &prob_lo, const std::array<double, 3> &prob_hi,
                         int storage = MRB_STORAGE_F64)
    {
        return Level(ncell, periodic, prob_lo, prob_hi, {Box{{0, 0, 0}, {ncell[0] - 1, ncell[1] - 1, ncell[2] - 1}}},
                     0, MRB_LAYOUT_COMPACT, storage);
    }
    ~Level()
    {
        if (h_) mrb_level_destroy(h_);
    }
    Level(const Level &) = delete;
    Level &operator=(const Level &) = delete;
    Level(Level &&o) noexcept : h_(std::exchange(o.h_, nullptr)), boxes_(std::move(o.boxes_)) {}
    Level &operator=(Level &&o) noexcept
    {
        if (this != &o) {
            if (h_) mrb_level_destroy(h_);
            h_ = std::exchange(o.h_, nullptr);
            boxes_ = std::move(o.boxes_);
        }
        return *this;
    }

    mrb_level *get() const { return h_; }
    const std::vector<Box> &boxes() const { return boxes_; }

    /* LBM physics and boundary plugins */
    void set_phys(const mrb_phys &p) { check(mrb_set_phys(h_, &p)); }
    void set_bc(const mrb_bc &b) { check(mrb_set_bc(h_, &b)); }

    /* LBM::initialize_is_fluid: is_fluid and eb_boundary (2 int32 comps per
     * fab, ghosts included, boxes back to back), or the built-in sphere */
    void set_is_fluid(const std::vector<int32_t> &isf)
    {
        check(mrb_upload(h_, MRB_ARR_ISFLUID, isf.data()));
        check(mrb_finalize_flags(h_));
    }
    void set_sphere(const std::array<double, 3> &center, double radius, bool fluid_inside)
    {
        check(mrb_compute_flags_sphere(h_, center.data(), radius, fluid_inside ? 1 : 0));
    }
    void set_cylinder(const std::array<double, 3> &center, double radius, double height, int direction,
                      bool fluid_inside)
    {
        check(mrb_compute_flags_cylinder(h_, center.data(), radius, height, direction, fluid_inside ? 1 : 0));
    }

    /* populations of MRB_ARR_F / MRB_ARR_G (FP64, AMReX FAB order) */
    std::int64_t elements(int arr) const
    {
        std::int64_t n = 0;
        for (int b = 0; b < int(boxes_.size()); ++b) n += mrb_fab_size(h_, arr, b);
        return n;
    }
    std::vector<double> download(int arr) const
    {
        std::vector<double> v(size_t(elements(arr)));
        check(mrb_download(h_, arr, v.data()));
        return v;
    }
    void upload(int arr, const std::vector<double> &v)
    {
        if (std::int64_t(v.size()) != elements(arr)) throw Error(MRB_E_ARG, "upload: wrong element count");
        check(mrb_upload(h_, arr, v.data()));
    }

    /* LBM::advance stages (LBM.cpp:535-543) and the fused step */
    void fillpatch(int lattice, void *stream = nullptr) { check(mrb_fill_ghosts(h_, lattice, stream)); }
    void physbc(int lattice, void *stream = nullptr) { check(mrb_physbc(h_, lattice, stream)); }
    void stream(int lattice, void *stream = nullptr) { check(mrb_stream(h_, lattice, stream)); }
    void collide(void *stream = nullptr) { check(mrb_collide(h_, stream)); }
    void advance(int nsteps, void *stream = nullptr) { check(mrb_advance(h_, nsteps, stream)); }

    /* FP64 sums of rho, rho u, rho E over the valid fluid cells */
    std::array<double, 5> conservation(void *stream = nullptr) const
    {
        std::array<double, 5> out{};
        check(mrb_reduce_conservation(h_, out.data(), stream));
        return out;
    }
    /* LBM::compute_eb_forces (LBM.cpp:994-1044) */
    std::array<double, 3> eb_force(void *stream = nullptr) const
    {
        std::array<double, 3> out{};
        check(mrb_eb_force(h_, out.data(), stream));
        return out;
    }
    void write_plotfile(const std::string &dir, double time, int step, int flags = 0)
    {
        check(mrb_write_plotfile(dir.c_str(), h_, time, step, flags));
    }

  private:
    mrb_level *h_ = nullptr;
    std::vector<Box> boxes_;
};

/* Two levels, refinement ratio 2 (LBM::time_step recursion for
 * finest_level = 1); the levels must outlive this object. */
class TwoLevel {
  public:
    TwoLevel(Level &crse, Level &fine) : c_(crse), f_(fine) { check(mrb_amr_link(c_.get(), f_.get())); }
    void average_down(void *stream = nullptr) { check(mrb_coalesce(f_.get(), c_.get(), stream)); }
    void fillpatch_fine(int lattice, void *stream = nullptr)
    {
        check(mrb_explode(c_.get(), f_.get(), lattice, stream));
    }
    void advance(int coarse_steps, bool graph = false, void *stream = nullptr)
    {
        check(graph ? mrb_amr_advance_graph(c_.get(), f_.get(), coarse_steps, stream)
                    : mrb_amr_advance(c_.get(), f_.get(), coarse_steps, stream));
    }

  private:
    Level &c_;
    Level &f_;
};

inline void write_checkpoint(const std::string &dir, const std::vector<Level *> &levels, const std::vector<int> &steps,
                             const std::vector<double> &dts, const std::vector<double> &times)
{
    std::vector<mrb_level *> h;
    for (Level *l : levels) h.push_back(l->get());
    if (steps.size() != h.size() || dts.size() != h.size() || times.size() != h.size())
        throw Error(MRB_E_ARG, "write_checkpoint: one step, dt and time per level");
    check(mrb_write_checkpoint(dir.c_str(), h.data(), int(h.size()), steps.data(), dts.data(), times.data()));
}

/* the checkpoint's per-level boxes, for creating the levels of a restart */
inline std::vector<std::vector<Box>> checkpoint_boxes(const std::string &dir)
{
    int nl = 0;
    check(mrb_checkpoint_header(dir.c_str(), &nl, nullptr, nullptr, nullptr, 0));
    std::vector<std::vector<Box>> out(static_cast<size_t>(nl));
    for (int l = 0; l < nl; ++l) {
        int nb = 0;
        check(mrb_checkpoint_boxes(dir.c_str(), l, &nb, nullptr, nullptr, 0));
        std::vector<int> lo(size_t(3 * nb)), hi(size_t(3 * nb));
        check(mrb_checkpoint_boxes(dir.c_str(), l, &nb, lo.data(), hi.data(), nb));
        for (int b = 0; b < nb; ++b)
            out[size_t(l)].push_back(Box{{lo[size_t(3 * b)], lo[size_t(3 * b + 1)], lo[size_t(3 * b + 2)]},
                                         {hi[size_t(3 * b)], hi[size_t(3 * b + 1)], hi[size_t(3 * b + 2)]}});
    }
    return out;
}

inline void read_checkpoint(const std::string &dir, const std::vector<Level *> &levels)
{
    std::vector<mrb_level *> h;
    for (Level *l : levels) h.push_back(l->get());
    check(mrb_read_checkpoint(dir.c_str(), h.data(), int(h.size())));
}

}  // namespace marbles_amd
