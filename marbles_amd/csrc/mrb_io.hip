// Checkpoint files in AMReX's on-disk formats (host code).
//
// LBM::write_checkpoint_file (Source/LBM.cpp:1684-1750) writes
//   <dir>/Header                 "Checkpoint file for LBM", the finest level,
//                                the steps, time steps and times of every
//                                level, then every level's BoxArray
//   <dir>/Level_<l>/f_00_H, _D_00000, g_00_H, _D_00000
//                                the f and g MultiFabs (ghost cells
//                                included) through amrex::VisMF::Write
// and LBM::read_checkpoint_file (LBM.cpp:1752-1860) reads them back, then
// zeroes the populations of solid cells (fill_f_inside_eb) and refills the
// ghost cells.  The VisMF header is AMReX's version-1 text header (version,
// how, component count, ghost width, BoxArray, FabOnDisk entries, per-fab
// minima and maxima); each FAB in the data file is a text line "FAB <real
// descriptor><box> <ncomp>" followed by the values, component-major, x
// fastest, as little-endian IEEE doubles.
#include <sys/stat.h>

#include <algorithm>
#include <array>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <limits>
#include <sstream>
#include <string>
#include <vector>

#include "mrb_internal.h"

namespace mrb {
namespace {

// RealDescriptor of little-endian IEEE-754 double (FABio_binary)
constexpr const char *kRealDescriptor = "((8, (64 11 52 0 1 12 0 1023)),(8, (8 7 6 5 4 3 2 1)))";

std::string box_text(const int lo[3], const int hi[3])
{
    std::ostringstream os;
    os << "((" << lo[0] << ',' << lo[1] << ',' << lo[2] << ") (" << hi[0] << ',' << hi[1] << ',' << hi[2]
       << ") (0,0,0))";
    return os.str();
}

int make_dirs(const std::string &path)
{
    std::string cur;
    std::istringstream is(path);
    std::string part;
    if (!path.empty() && path[0] == '/') cur = "/";
    while (std::getline(is, part, '/')) {
        if (part.empty()) continue;
        cur += part;
        if (mkdir(cur.c_str(), 0755) != 0 && errno != EEXIST)
            return set_error(MRB_E_ARG, "cannot create directory " + cur + ": " + std::strerror(errno));
        cur += '/';
    }
    return MRB_OK;
}

// The valid boxes of a level and the box its stored fab covers (grown by
// the ghost width; the compact layout stores the domain box without ghosts)
struct LevelBoxes {
    std::vector<std::array<int, 6>> valid;  // lo[3], hi[3]
    int ng = 0;
};

LevelBoxes level_boxes(const mrb_level *L)
{
    LevelBoxes B;
    B.ng = L->layout == MRB_LAYOUT_FAB ? L->ng : 0;
    for (int b = 0; b < L->nbox; ++b) {
        const BoxDev &d = L->boxes[size_t(b)];
        B.valid.push_back({d.vlo[0], d.vlo[1], d.vlo[2], d.vhi[0], d.vhi[1], d.vhi[2]});
    }
    return B;
}

// amrex::VisMF::Write of one MultiFab: <prefix>_H and <prefix>_D_00000.
// data: every fab back to back, fab b = ncomp components of ncell[b] values
int write_multifab(const std::string &prefix, const LevelBoxes &B, int ncomp, const double *data,
                   const std::vector<int64_t> &ncell)
{
    const std::string base = prefix.substr(prefix.find_last_of('/') + 1);
    const std::string dname = base + "_D_00000";
    std::ofstream d(prefix + "_D_00000", std::ios::binary | std::ios::trunc);
    if (!d) return set_error(MRB_E_ARG, "cannot write " + prefix + "_D_00000");
    const size_t nfab = B.valid.size();
    std::vector<int64_t> offset(nfab);
    std::vector<std::vector<double>> vmin(nfab, std::vector<double>(size_t(ncomp))), vmax = vmin;
    const double *p = data;
    for (size_t b = 0; b < nfab; ++b) {
        const auto &v = B.valid[b];
        const int lo[3] = {v[0] - B.ng, v[1] - B.ng, v[2] - B.ng};
        const int hi[3] = {v[3] + B.ng, v[4] + B.ng, v[5] + B.ng};
        const int n0 = hi[0] - lo[0] + 1, n1 = hi[1] - lo[1] + 1;
        offset[b] = int64_t(d.tellp());
        d << "FAB " << kRealDescriptor << box_text(lo, hi) << ' ' << ncomp << '\n';
        d.write(reinterpret_cast<const char *>(p), std::streamsize(sizeof(double) * size_t(ncomp) * ncell[b]));
        // minima and maxima over the valid box
        for (int c = 0; c < ncomp; ++c) {
            double mn = std::numeric_limits<double>::max(), mx = -mn;
            const double *pc = p + int64_t(c) * ncell[b];
            for (int k = v[2]; k <= v[5]; ++k)
                for (int j = v[1]; j <= v[4]; ++j)
                    for (int i = v[0]; i <= v[3]; ++i) {
                        const double x = pc[(int64_t(k - lo[2]) * n1 + (j - lo[1])) * n0 + (i - lo[0])];
                        mn = std::min(mn, x);
                        mx = std::max(mx, x);
                    }
            vmin[b][size_t(c)] = mn;
            vmax[b][size_t(c)] = mx;
        }
        p += int64_t(ncomp) * ncell[b];
    }
    if (!d) return set_error(MRB_E_ARG, "write failed: " + prefix + "_D_00000");
    std::ofstream h(prefix + "_H", std::ios::trunc);
    if (!h) return set_error(MRB_E_ARG, "cannot write " + prefix + "_H");
    h.setf(std::ios::scientific, std::ios::floatfield);
    h.precision(15);
    h << 1 << '\n' << 0 << '\n' << ncomp << '\n' << B.ng << '\n';
    h << '(' << nfab << " 0" << '\n';
    for (const auto &v : B.valid) {
        const int lo[3] = {v[0], v[1], v[2]}, hi[3] = {v[3], v[4], v[5]};
        h << box_text(lo, hi) << '\n';
    }
    h << ')' << '\n';
    h << nfab << '\n';
    for (size_t b = 0; b < nfab; ++b) h << "FabOnDisk: " << dname << ' ' << offset[b] << '\n';
    h << '\n';
    for (const auto *mm : {&vmin, &vmax}) {
        h << nfab << ',' << ncomp << '\n';
        for (size_t b = 0; b < nfab; ++b) {
            for (int c = 0; c < ncomp; ++c) h << (*mm)[b][size_t(c)] << ',';
            h << '\n';
        }
        h << '\n';
    }
    if (!h) return set_error(MRB_E_ARG, "write failed: " + prefix + "_H");
    return MRB_OK;
}

// "((a,b,c) (d,e,f) (t,u,v))" -> lo, hi
bool parse_box(const std::string &s, int lo[3], int hi[3])
{
    int t[3];
    return std::sscanf(s.c_str(), " ((%d,%d,%d) (%d,%d,%d) (%d,%d,%d))", &lo[0], &lo[1], &lo[2], &hi[0], &hi[1],
                       &hi[2], &t[0], &t[1], &t[2]) == 9;
}

// a BoxArray as BoxArray::writeOn prints it: "(n 0", n boxes, ")"
bool read_box_array(std::istream &is, std::vector<std::array<int, 6>> &out)
{
    std::string line;
    while (std::getline(is, line) && line.find('(') == std::string::npos) {
    }
    int n = 0, z = 0;
    if (std::sscanf(line.c_str(), " (%d %d", &n, &z) != 2 || n < 0) return false;
    out.clear();
    for (int b = 0; b < n; ++b) {
        if (!std::getline(is, line)) return false;
        int lo[3], hi[3];
        if (!parse_box(line, lo, hi)) return false;
        out.push_back({lo[0], lo[1], lo[2], hi[0], hi[1], hi[2]});
    }
    return bool(std::getline(is, line));  // ")"
}

// amrex::VisMF::Read of one MultiFab into `data` (fabs back to back), the
// boxes, ghost width and component count checked against the level's
int read_multifab(const std::string &prefix, const LevelBoxes &B, int ncomp, double *data,
                  const std::vector<int64_t> &ncell)
{
    std::ifstream h(prefix + "_H");
    if (!h) return set_error(MRB_E_ARG, "cannot read " + prefix + "_H");
    int vers = 0, how = 0, nc = 0;
    std::string ngs;
    h >> vers >> how >> nc >> ngs;
    int ng = -1;
    if (!ngs.empty() && ngs[0] == '(') {
        int a, b, c;
        if (std::sscanf(ngs.c_str(), "(%d,%d,%d)", &a, &b, &c) == 3 && a == b && b == c) ng = a;
    } else {
        ng = std::atoi(ngs.c_str());
    }
    std::vector<std::array<int, 6>> ba;
    if (!h || !read_box_array(h, ba)) return set_error(MRB_E_ARG, "malformed header " + prefix + "_H");
    if (nc != ncomp || ng != B.ng || ba != B.valid)
        return set_error(MRB_E_ARG, "checkpoint " + prefix + " does not match the level (boxes, ghosts, comps)");
    size_t nfod = 0;
    h >> nfod;
    if (nfod != ba.size()) return set_error(MRB_E_ARG, "malformed FabOnDisk list in " + prefix + "_H");
    const std::string dir = prefix.substr(0, prefix.find_last_of('/') + 1);
    double *p = data;
    for (size_t b = 0; b < nfod; ++b) {
        std::string tag, name;
        int64_t off = 0;
        h >> tag >> name >> off;
        if (tag != "FabOnDisk:") return set_error(MRB_E_ARG, "malformed FabOnDisk entry in " + prefix + "_H");
        std::ifstream d(dir + name, std::ios::binary);
        if (!d) return set_error(MRB_E_ARG, "cannot read " + dir + name);
        d.seekg(off);
        std::string line;
        std::getline(d, line);
        const size_t at = line.find(")))");
        if (line.compare(0, 4, "FAB ") != 0 || at == std::string::npos ||
            line.substr(4, at + 3 - 4) != kRealDescriptor)
            return set_error(MRB_E_UNSUPPORTED, "FAB in " + name + " is not little-endian IEEE double");
        int lo[3], hi[3];
        const std::string rest = line.substr(at + 3);
        if (!parse_box(rest, lo, hi)) return set_error(MRB_E_ARG, "malformed FAB header in " + name);
        const auto &v = B.valid[b];
        for (int k = 0; k < 3; ++k)
            if (lo[k] != v[k] - B.ng || hi[k] != v[3 + k] + B.ng)
                return set_error(MRB_E_ARG, "FAB box in " + name + " does not match the level");
        d.read(reinterpret_cast<char *>(p), std::streamsize(sizeof(double) * size_t(ncomp) * ncell[b]));
        if (!d) return set_error(MRB_E_ARG, "short FAB data in " + name);
        p += int64_t(ncomp) * ncell[b];
    }
    return MRB_OK;
}

std::vector<int64_t> fab_cells(const mrb_level *L)
{
    std::vector<int64_t> n;
    for (int b = 0; b < L->nbox; ++b) n.push_back(L->boxes[size_t(b)].ncell);
    return n;
}

std::string level_dir(const char *dir, int l) { return std::string(dir) + "/Level_" + std::to_string(l); }

// LBM's plot variable names (LBM.cpp:14-82): macrodata, then f_00..f_26 and
// g_00..g_26 (save_streaming), then the derived fields (save_derived), then
// is_fluid and eb_boundary
std::vector<std::string> plot_names(bool streaming, bool derived)
{
    std::vector<std::string> v = {"rho",  "vel_x", "vel_y", "vel_z", "vel_mag", "two_rho_e", "QCorrX",
                                  "QCorrY", "QCorrZ", "pxx", "pyy", "pzz", "pxy", "pxz", "pyz", "qx", "qy", "qz",
                                  "temperature"};
    if (streaming) {
        for (const char *p : {"f_", "g_"})
            for (int q = 0; q < NQ; ++q) {
                char b[8];
                std::snprintf(b, sizeof b, "%s%02d", p, q);
                v.push_back(b);
            }
    }
    if (derived)
        for (const char *n : {"vort_x", "vort_y", "vort_z", "vort_mag", "dQCorrX", "dQCorrY", "dQCorrZ"})
            v.push_back(n);
    v.push_back("is_fluid");
    v.push_back("eb_boundary");
    return v;
}

}  // namespace
}  // namespace mrb

using namespace mrb;

extern "C" {

int mrb_write_checkpoint(const char *dir, mrb_level *const *levels, int nlevels, const int *steps,
                         const double *dts, const double *times)
{
    if (!dir || !levels || nlevels < 1 || !steps || !dts || !times) return set_error(MRB_E_ARG, "null argument");
    for (int l = 0; l < nlevels; ++l)
        if (!levels[l]) return set_error(MRB_E_ARG, "null level");
    int rc = make_dirs(dir);
    if (rc) return rc;
    {
        std::ofstream h(std::string(dir) + "/Header", std::ios::trunc | std::ios::binary);
        if (!h) return set_error(MRB_E_ARG, std::string("cannot write ") + dir + "/Header");
        h.precision(17);
        h << "Checkpoint file for LBM\n";
        h << nlevels - 1 << "\n";
        for (int l = 0; l < nlevels; ++l) h << steps[l] << " ";
        h << "\n";
        for (int l = 0; l < nlevels; ++l) h << dts[l] << " ";
        h << "\n";
        for (int l = 0; l < nlevels; ++l) h << times[l] << " ";
        h << "\n";
        for (int l = 0; l < nlevels; ++l) {
            const LevelBoxes B = level_boxes(levels[l]);
            h << '(' << B.valid.size() << ' ' << 0 << '\n';
            for (const auto &v : B.valid) {
                const int lo[3] = {v[0], v[1], v[2]}, hi[3] = {v[3], v[4], v[5]};
                h << box_text(lo, hi) << '\n';
            }
            h << ')' << '\n';
        }
        if (!h) return set_error(MRB_E_ARG, std::string("write failed: ") + dir + "/Header");
    }
    for (int l = 0; l < nlevels; ++l) {
        const mrb_level *L = levels[l];
        rc = make_dirs(level_dir(dir, l));
        if (rc) return rc;
        const LevelBoxes B = level_boxes(L);
        const std::vector<int64_t> nc = fab_cells(L);
        std::vector<double> host(size_t(NQ) * size_t(L->total_cells));
        for (int lat = 0; lat < 2; ++lat) {
            rc = mrb_download(L, lat == 0 ? MRB_ARR_F : MRB_ARR_G, host.data());
            if (rc) return rc;
            rc = write_multifab(level_dir(dir, l) + (lat == 0 ? "/f_00" : "/g_00"), B, NQ, host.data(), nc);
            if (rc) return rc;
        }
    }
    return MRB_OK;
}

int mrb_checkpoint_header(const char *dir, int *nlevels, int *steps, double *dts, double *times, int max_levels)
{
    if (!dir || !nlevels) return set_error(MRB_E_ARG, "null argument");
    std::ifstream h(std::string(dir) + "/Header");
    if (!h) return set_error(MRB_E_ARG, std::string("cannot read ") + dir + "/Header");
    std::string line;
    std::getline(h, line);
    int finest = -1;
    h >> finest;
    std::getline(h, line);
    if (finest < 0) return set_error(MRB_E_ARG, "malformed checkpoint header");
    *nlevels = finest + 1;
    if (!steps && !dts && !times) return MRB_OK;
    if (max_levels < finest + 1) return set_error(MRB_E_ARG, "arrays shorter than the checkpoint's level count");
    for (int row = 0; row < 3; ++row) {
        std::getline(h, line);
        std::istringstream ls(line);
        for (int l = 0; l <= finest; ++l) {
            double v = 0.0;
            ls >> v;
            if (row == 0 && steps) steps[l] = int(v);
            if (row == 1 && dts) dts[l] = v;
            if (row == 2 && times) times[l] = v;
        }
        if (!ls) return set_error(MRB_E_ARG, "malformed checkpoint header");
    }
    return MRB_OK;
}

int mrb_checkpoint_boxes(const char *dir, int level, int *nbox, int *box_lo, int *box_hi, int max_boxes)
{
    if (!dir || !nbox) return set_error(MRB_E_ARG, "null argument");
    std::ifstream h(std::string(dir) + "/Header");
    if (!h) return set_error(MRB_E_ARG, std::string("cannot read ") + dir + "/Header");
    std::string line;
    for (int i = 0; i < 5; ++i) std::getline(h, line);
    std::vector<std::array<int, 6>> ba;
    for (int l = 0; l <= level; ++l)
        if (!read_box_array(h, ba)) return set_error(MRB_E_ARG, "checkpoint has no such level");
    *nbox = int(ba.size());
    if (!box_lo && !box_hi) return MRB_OK;
    if (!box_lo || !box_hi || max_boxes < int(ba.size())) return set_error(MRB_E_ARG, "box arrays too short");
    for (size_t b = 0; b < ba.size(); ++b)
        for (int k = 0; k < 3; ++k) {
            box_lo[3 * b + k] = ba[b][size_t(k)];
            box_hi[3 * b + k] = ba[b][size_t(3 + k)];
        }
    return MRB_OK;
}

int mrb_read_checkpoint(const char *dir, mrb_level *const *levels, int nlevels)
{
    if (!dir || !levels || nlevels < 1) return set_error(MRB_E_ARG, "null argument");
    int nl = 0;
    int rc = mrb_checkpoint_header(dir, &nl, nullptr, nullptr, nullptr, 0);
    if (rc) return rc;
    if (nl != nlevels) return set_error(MRB_E_ARG, "level count differs from the checkpoint's");
    for (int l = 0; l < nlevels; ++l) {
        mrb_level *L = levels[l];
        if (!L) return set_error(MRB_E_ARG, "null level");
        if (!L->flags_ready) return set_error(MRB_E_STATE, "flags not set (restart recomputes them first)");
        const LevelBoxes B = level_boxes(L);
        const std::vector<int64_t> nc = fab_cells(L);
        std::vector<int32_t> isf(size_t(2) * size_t(L->total_cells));
        rc = mrb_download(L, MRB_ARR_ISFLUID, isf.data());
        if (rc) return rc;
        std::vector<double> host(size_t(NQ) * size_t(L->total_cells));
        for (int lat = 0; lat < 2; ++lat) {
            rc = read_multifab(level_dir(dir, l) + (lat == 0 ? "/f_00" : "/g_00"), B, NQ, host.data(), nc);
            if (rc) return rc;
            // fill_f_inside_eb (LBM.cpp:1278-1298): zero in solid cells
            int64_t base = 0;
            for (int b = 0; b < L->nbox; ++b) {
                const int64_t n = nc[size_t(b)];
                for (int64_t c = 0; c < n; ++c)
                    if (isf[size_t(2 * base + c)] == 0)
                        for (int q = 0; q < NQ; ++q) host[size_t(base * NQ + q * n + c)] = 0.0;
                base += n;
            }
            rc = mrb_upload(L, lat == 0 ? MRB_ARR_F : MRB_ARR_G, host.data());
            if (rc) return rc;
        }
    }
    return MRB_OK;
}

// the plot variables of a fab level: per box, every variable over the
// valid cells, boxes back to back
static int fab_plot(mrb_level *L, double *out, bool streaming, bool derived)
{
    const int64_t T = L->total_cells;
    double *der = nullptr;
    MRB_HIP(hipMalloc(&der, sizeof(double) * 7 * size_t(T)));
    int rc = fab_plot_fields(L, der, nullptr);
    hipError_t e = rc == MRB_OK ? hipDeviceSynchronize() : hipSuccess;
    std::vector<double> md(size_t(19) * size_t(T)), dv(size_t(7) * size_t(T));
    if (rc == MRB_OK && e == hipSuccess) e = hipMemcpy(md.data(), L->d_md, sizeof(double) * md.size(), hipMemcpyDeviceToHost);
    if (rc == MRB_OK && e == hipSuccess) e = hipMemcpy(dv.data(), der, sizeof(double) * dv.size(), hipMemcpyDeviceToHost);
    (void)hipFree(der);
    if (e != hipSuccess) return hip_fail(e, "plot fields");
    if (rc) return rc;
    std::vector<double> f, g;
    if (streaming) {
        f.resize(size_t(NQ) * size_t(T));
        g.resize(f.size());
        if ((rc = mrb_download(L, MRB_ARR_F, f.data()))) return rc;
        if ((rc = mrb_download(L, MRB_ARR_G, g.data()))) return rc;
    }
    std::vector<int32_t> isf(size_t(2) * size_t(T));
    if ((rc = mrb_download(L, MRB_ARR_ISFLUID, isf.data()))) return rc;
    double *o = out;
    for (int b = 0; b < L->nbox; ++b) {
        const BoxDev &B = L->boxes[size_t(b)];
        auto emit = [&](const double *src, int64_t base_comps, int comp) {
            const double *p = src + B.base * base_comps + int64_t(comp) * B.ncell;
            for (int k = B.vlo[2]; k <= B.vhi[2]; ++k)
                for (int j = B.vlo[1]; j <= B.vhi[1]; ++j)
                    for (int i = B.vlo[0]; i <= B.vhi[0]; ++i)
                        *o++ = p[(int64_t(k - B.lo[2]) * B.n[1] + (j - B.lo[1])) * B.n[0] + (i - B.lo[0])];
        };
        for (int c = 0; c < 19; ++c) emit(md.data(), 19, c);
        if (streaming) {
            for (int q = 0; q < NQ; ++q) emit(f.data(), NQ, q);
            for (int q = 0; q < NQ; ++q) emit(g.data(), NQ, q);
        }
        if (derived)
            for (int c = 0; c < 7; ++c) emit(dv.data(), 7, c);
        for (int c = 0; c < 2; ++c) {
            const int32_t *p = isf.data() + B.base * 2 + int64_t(c) * B.ncell;
            for (int k = B.vlo[2]; k <= B.vhi[2]; ++k)
                for (int j = B.vlo[1]; j <= B.vhi[1]; ++j)
                    for (int i = B.vlo[0]; i <= B.vhi[0]; ++i)
                        *o++ = double(p[(int64_t(k - B.lo[2]) * B.n[1] + (j - B.lo[1])) * B.n[0] + (i - B.lo[0])]);
        }
    }
    return MRB_OK;
}

int mrb_plot_fields(mrb_level *L, double *out, int flags)
{
    if (!L || !out) return set_error(MRB_E_ARG, "null argument");
    if (!L->flags_ready || !L->pops_set) return set_error(MRB_E_STATE, "flags and populations must be set");
    const bool streaming = flags & MRB_PLOT_STREAMING, derived = flags & MRB_PLOT_DERIVED;
    if (L->layout == MRB_LAYOUT_FAB) return fab_plot(L, out, streaming, derived);
    const int64_t N = L->total_cells;
    double *d = nullptr;
    MRB_HIP(hipMalloc(&d, sizeof(double) * 26 * size_t(N)));
    int rc = compact_plot_fields(L, d, d + 19 * N, nullptr);
    hipError_t e = rc == MRB_OK ? hipDeviceSynchronize() : hipSuccess;
    double *o = out;
    if (rc == MRB_OK && e == hipSuccess) e = hipMemcpy(o, d, sizeof(double) * 19 * size_t(N), hipMemcpyDeviceToHost);
    o += 19 * N;
    if (rc == MRB_OK && e == hipSuccess && derived) {
        const int64_t skip = streaming ? 2 * NQ * N : 0;
        e = hipMemcpy(o + skip, d + 19 * N, sizeof(double) * 7 * size_t(N), hipMemcpyDeviceToHost);
    }
    (void)hipFree(d);
    if (e != hipSuccess) return hip_fail(e, "plot fields");
    if (rc) return rc;
    if (streaming) {
        if ((rc = mrb_download(L, MRB_ARR_F, o))) return rc;
        if ((rc = mrb_download(L, MRB_ARR_G, o + NQ * N))) return rc;
        o += 2 * NQ * N;
    }
    if (derived) o += 7 * N;
    std::vector<int32_t> isf(size_t(2) * size_t(N));
    if ((rc = mrb_download(L, MRB_ARR_ISFLUID, isf.data()))) return rc;
    for (int64_t c = 0; c < 2 * N; ++c) o[c] = double(isf[size_t(c)]);
    return MRB_OK;
}

// LBM::open_forces_file / output_forces_file (LBM.cpp:1925-1969), columns
// of width DATWIDTH = 24 and DATPRECISION = 16 significant digits
// (Constants.H:63-64)
int mrb_forces_file_open(const char *path, int initialize)
{
    if (!path) return set_error(MRB_E_ARG, "null argument");
    std::ifstream probe(path);
    if (probe.good() && !initialize) return MRB_OK;  // appended to
    std::ofstream o(path, std::ios::out | std::ios::trunc);
    if (!o) return set_error(MRB_E_ARG, std::string("cannot write ") + path);
    o << std::setw(24) << "          time";
    for (const char *c : {"          fx", "          fy", "          fz"}) o << std::setw(24) << c;
    o << std::endl;
    return o ? MRB_OK : set_error(MRB_E_ARG, std::string("write failed: ") + path);
}

int mrb_forces_file_append(const char *path, double time, const double force[3])
{
    if (!path || !force) return set_error(MRB_E_ARG, "null argument");
    std::ofstream o(path, std::ios::app);
    if (!o) return set_error(MRB_E_ARG, std::string("cannot append to ") + path);
    o << std::setw(24) << std::setprecision(16) << time;
    for (int d = 0; d < 3; ++d) o << std::setw(24) << std::setprecision(16) << force[d];
    o << std::endl;
    return o ? MRB_OK : set_error(MRB_E_ARG, std::string("write failed: ") + path);
}

int mrb_plot_nvars(int flags)
{
    return int(plot_names(flags & MRB_PLOT_STREAMING, flags & MRB_PLOT_DERIVED).size());
}

int mrb_write_plotfile_levels(const char *dir, mrb_level *const *levels, int nlevels, double time, const int *steps,
                              int flags)
{
    if (!dir || !levels || nlevels < 1 || !steps) return set_error(MRB_E_ARG, "null argument");
    for (int l = 0; l < nlevels; ++l) {
        if (!levels[l]) return set_error(MRB_E_ARG, "null level");
        if (l > 0)
            for (int d = 0; d < 3; ++d)
                if (levels[l]->ncell[d] != 2 * levels[l - 1]->ncell[d])
                    return set_error(MRB_E_ARG, "plotfile levels must refine by 2");
    }
    const std::vector<std::string> names = plot_names(flags & MRB_PLOT_STREAMING, flags & MRB_PLOT_DERIVED);
    const int nv = int(names.size());
    // amrex::WriteMultiLevelPlotfile -> WriteGenericPlotfileHeader
    const mrb_level *L0 = levels[0];
    int rc = make_dirs(dir);
    if (rc) return rc;
    std::ofstream h(std::string(dir) + "/Header", std::ios::trunc);
    if (!h) return set_error(MRB_E_ARG, std::string("cannot write ") + dir + "/Header");
    h.precision(17);
    h << "HyperCLaw-V1.1" << '\n' << nv << '\n';
    for (const auto &n : names) h << n << '\n';
    h << 3 << '\n' << time << '\n' << nlevels - 1 << '\n';
    for (int k = 0; k < 3; ++k) h << L0->prob_lo[k] << ' ';
    h << '\n';
    for (int k = 0; k < 3; ++k) h << L0->prob_hi[k] << ' ';
    h << '\n';
    for (int l = 0; l + 1 < nlevels; ++l) h << 2 << ' ';
    h << '\n';
    for (int l = 0; l < nlevels; ++l) {
        const int dlo[3] = {0, 0, 0};
        const int dhi[3] = {levels[l]->ncell[0] - 1, levels[l]->ncell[1] - 1, levels[l]->ncell[2] - 1};
        h << box_text(dlo, dhi) << ' ';
    }
    h << '\n';
    for (int l = 0; l < nlevels; ++l) h << steps[l] << ' ';
    h << '\n';
    for (int l = 0; l < nlevels; ++l) {
        for (int k = 0; k < 3; ++k) h << levels[l]->dx[k] << ' ';
        h << '\n';
    }
    h << 0 << '\n' << 0 << '\n';
    std::vector<LevelBoxes> LB(static_cast<size_t>(nlevels));
    for (int l = 0; l < nlevels; ++l) {
        const mrb_level *L = levels[l];
        LB[size_t(l)] = level_boxes(L);
        LB[size_t(l)].ng = 0;
        h << l << ' ' << LB[size_t(l)].valid.size() << ' ' << time << '\n' << steps[l] << '\n';
        for (const auto &v : LB[size_t(l)].valid)
            for (int k = 0; k < 3; ++k)
                h << L->prob_lo[k] + v[size_t(k)] * L->dx[k] << ' '
                  << L->prob_lo[k] + (v[size_t(3 + k)] + 1) * L->dx[k] << '\n';
        h << "Level_" << l << "/Cell" << '\n';
    }
    if (!h) return set_error(MRB_E_ARG, std::string("write failed: ") + dir + "/Header");
    h.close();
    for (int l = 0; l < nlevels; ++l) {
        std::vector<int64_t> nvalid;
        int64_t N = 0;
        for (const auto &v : LB[size_t(l)].valid) {
            nvalid.push_back(int64_t(v[3] - v[0] + 1) * (v[4] - v[1] + 1) * (v[5] - v[2] + 1));
            N += nvalid.back();
        }
        std::vector<double> data(size_t(nv) * size_t(N));
        if ((rc = mrb_plot_fields(levels[l], data.data(), flags))) return rc;
        if ((rc = make_dirs(level_dir(dir, l)))) return rc;
        if ((rc = write_multifab(level_dir(dir, l) + "/Cell", LB[size_t(l)], nv, data.data(), nvalid))) return rc;
    }
    return MRB_OK;
}

int mrb_write_plotfile(const char *dir, mrb_level *lvl, double time, int step, int flags)
{
    mrb_level *const one[1] = {lvl};
    return mrb_write_plotfile_levels(dir, one, 1, time, &step, flags);
}

}  // extern "C"
