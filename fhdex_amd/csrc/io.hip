// io.hip -- on-disk formats of the compressible path: the AMReX plotfile of WritePlotFile
// (src_compressible/writePlotFile.cpp:7-190) and the checkpoint of WriteCheckPoint /
// ReadCheckPoint (src_compressible/checkpoint.cpp:22-255), for the single-box, single-level
// domain this library steps.
//
// The byte layouts are AMReX's (an external dependency of the reference, not part of its
// source tree):
//   plotfile  <base><step:9 digits>/Header       text: "HyperCLaw-V1.1", variable names,
//             dimension, time, level data, box extents, "Level_0/Cell"
//             <dir>/Level_0/Cell_H, Cell_D_00000  one VisMF MultiFab, no ghost cells
//   checkpoint chk<step:9>/Header                text written by checkpoint.cpp:77-97
//             <dir>/Level_0/<name>_H, _D_00000     one VisMF MultiFab per field, ngc ghosts
//   VisMF _H  version 1, how 1 (NFiles), ncomp, ngrow, BoxArray, FabOnDisk list, per-FAB
//             minima and maxima; _D holds per FAB the line
//             "FAB ((8, (64 11 52 0 1 12 0 1023)),(8, (8 7 6 5 4 3 2 1)))<box> <ncomp>"
//             (IEEE-754 binary64, little-endian) followed by the FAB's values, x fastest,
//             component slowest -- the same order as the device arrays.
// The checkpoint Header gets one line after the BoxArray that the reference's reader never
// reaches ("fhdx_philox_seed <seed>"): with counter-based noise keyed by (seed, step), the seed
// and the step number are all a restart needs to continue the run bit for bit.
#include <hip/hip_runtime.h>

#include <sys/stat.h>
#include <sys/types.h>

#include <cerrno>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <limits>
#include <sstream>
#include <string>
#include <vector>

#include "common.hpp"

using namespace fhdx;

namespace {

constexpr const char* kRealDescriptor = "((8, (64 11 52 0 1 12 0 1023)),(8, (8 7 6 5 4 3 2 1)))";

struct Box3 {
    int lo[3], hi[3];
    long size(int d) const { return (long)hi[d] - lo[d] + 1; }
    long cells() const { return size(0) * size(1) * size(2); }
    bool contains(const Box3& b) const
    {
        for (int d = 0; d < 3; ++d)
            if (b.lo[d] < lo[d] || b.hi[d] > hi[d]) return false;
        return true;
    }
};

std::string box_text(const Box3& b)
{
    std::ostringstream s;
    s << "((" << b.lo[0] << ',' << b.lo[1] << ',' << b.lo[2] << ") (" << b.hi[0] << ',' << b.hi[1] << ','
      << b.hi[2] << ") (0,0,0))";
    return s.str();
}

// BoxArray::writeOn of a one-box array
void write_boxarray(std::ostream& os, const Box3& b) { os << "(1 0\n" << box_text(b) << "\n)"; }

// Parses "((a,b,c) (d,e,f) (t,u,v))" (whitespace tolerant); cell-centred boxes only.
bool read_box(std::istream& is, Box3& b)
{
    char c;
    int t[3];
    auto expect = [&](char want) { return (is >> c) && c == want; };
    if (!expect('(') || !expect('(')) return false;
    for (int d = 0; d < 3; ++d)
        if (!(is >> b.lo[d]) || !expect(d < 2 ? ',' : ')')) return false;
    if (!expect('(')) return false;
    for (int d = 0; d < 3; ++d)
        if (!(is >> b.hi[d]) || !expect(d < 2 ? ',' : ')')) return false;
    if (!expect('(')) return false;
    for (int d = 0; d < 3; ++d)
        if (!(is >> t[d]) || !expect(d < 2 ? ',' : ')')) return false;
    if (!expect(')')) return false;
    return t[0] == 0 && t[1] == 0 && t[2] == 0;
}

// BoxArray::readFrom of a one-box array: "(1 <hash>" box ")"
bool read_boxarray(std::istream& is, Box3& b)
{
    char c;
    long n, hash;
    if (!(is >> c) || c != '(' || !(is >> n >> hash) || n != 1) return false;
    if (!read_box(is, b)) return false;
    return (is >> c) && c == ')';
}

std::string step_name(const char* base, int step)
{
    std::ostringstream s;
    s << base << std::setw(9) << std::setfill('0') << step;  // amrex::Concatenate(base, step, 9)
    return s.str();
}

bool make_dir(const std::string& d)
{
    if (mkdir(d.c_str(), 0755) == 0 || errno == EEXIST) return true;
    set_error("cannot create directory " + d + ": " + std::strerror(errno));
    return false;
}

Box3 fab_box(const fhdx_fab& f)
{
    Box3 b;
    for (int d = 0; d < 3; ++d) {
        b.lo[d] = f.lo[d];
        b.hi[d] = f.hi[d];
    }
    return b;
}

Box3 valid_box(const fhdx_params& p)
{
    Box3 b;
    for (int d = 0; d < 3; ++d) {
        b.lo[d] = 0;
        b.hi[d] = p.n_cells[d] - 1;
    }
    return b;
}

Box3 grown_box(const fhdx_params& p)
{
    Box3 b = valid_box(p);
    for (int d = 0; d < 3; ++d) {
        b.lo[d] -= p.ngc;
        b.hi[d] += p.ngc;
    }
    return b;
}

// Component c of a FAB (device or host memory) into host memory.
struct Mover {
    bool host;
    hipStream_t s;
    int get(const fhdx_fab& f, int c, std::vector<double>& out) const
    {
        const long n = fab_box(f).cells();
        out.resize(n);
        const double* src = f.p + (long)c * n;
        if (host) {
            std::memcpy(out.data(), src, sizeof(double) * n);
            return FHDX_OK;
        }
        FHDX_HIP_TRY(hipMemcpyAsync(out.data(), src, sizeof(double) * n, hipMemcpyDefault, s));
        FHDX_HIP_TRY(hipStreamSynchronize(s));
        return FHDX_OK;
    }
    int put(const fhdx_fab& f, int c, const std::vector<double>& in) const
    {
        const long n = fab_box(f).cells();
        double* dst = f.p + (long)c * n;
        if (host) {
            std::memcpy(dst, in.data(), sizeof(double) * n);
            return FHDX_OK;
        }
        FHDX_HIP_TRY(hipMemcpyAsync(dst, in.data(), sizeof(double) * n, hipMemcpyDefault, s));
        FHDX_HIP_TRY(hipStreamSynchronize(s));
        return FHDX_OK;
    }
};

// Copies the cells of `region` from an array over box `from` into an array over box `to`.
void copy_region(const std::vector<double>& src, const Box3& from, std::vector<double>& dst, const Box3& to,
                 const Box3& region)
{
    const long fy = from.size(0), fz = fy * from.size(1);
    const long ty = to.size(0), tz = ty * to.size(1);
    const long nx = region.size(0);
    for (int k = region.lo[2]; k <= region.hi[2]; ++k)
        for (int j = region.lo[1]; j <= region.hi[1]; ++j) {
            const long a = (k - from.lo[2]) * fz + (j - from.lo[1]) * fy + (region.lo[0] - from.lo[0]);
            const long b = (k - to.lo[2]) * tz + (j - to.lo[1]) * ty + (region.lo[0] - to.lo[0]);
            std::memcpy(&dst[b], &src[a], sizeof(double) * nx);
        }
}

// One field of a VisMF MultiFab: a FAB and its first `ncomp` components.
struct Part {
    const fhdx_fab* fab;
    int comp0, ncomp;
};

// VisMF::Write of a one-box MultiFab over `box` (the valid box grown by `ngrow`) whose
// components are the concatenation of `parts`.  Streams one component at a time.
int vismf_write(const std::string& dir, const std::string& name, const Box3& valid, int ngrow,
                const std::vector<Part>& parts, const Mover& mv)
{
    Box3 box = valid;
    for (int d = 0; d < 3; ++d) {
        box.lo[d] -= ngrow;
        box.hi[d] += ngrow;
    }
    int ncomp = 0;
    for (const Part& q : parts) ncomp += q.ncomp;
    const std::string dname = name + "_D_00000";
    std::ofstream D(dir + "/" + dname, std::ios::binary | std::ios::trunc);
    if (!D) {
        set_error("cannot open " + dir + "/" + dname + " for writing");
        return FHDX_ERR_ARG;
    }
    D << "FAB " << kRealDescriptor << box_text(box) << ' ' << ncomp << '\n';
    std::vector<double> mn, mx, comp, out(box.cells());
    for (const Part& q : parts)
        for (int c = q.comp0; c < q.comp0 + q.ncomp; ++c) {
            int r = mv.get(*q.fab, c, comp);
            if (r) return r;
            const Box3 fb = fab_box(*q.fab);
            copy_region(comp, fb, out, box, box);
            double lo = std::numeric_limits<double>::infinity(), hi = -lo;
            for (double v : out) {
                lo = v < lo ? v : lo;
                hi = v > hi ? v : hi;
            }
            mn.push_back(lo);
            mx.push_back(hi);
            D.write(reinterpret_cast<const char*>(out.data()), (std::streamsize)(sizeof(double) * out.size()));
        }
    D.close();
    if (!D) {
        set_error("write failed: " + dir + "/" + dname);
        return FHDX_ERR_ARG;
    }
    std::ofstream H(dir + "/" + name + "_H", std::ios::trunc);
    if (!H) {
        set_error("cannot open " + dir + "/" + name + "_H for writing");
        return FHDX_ERR_ARG;
    }
    H << 1 << '\n' << 1 << '\n' << ncomp << '\n' << ngrow << '\n';
    write_boxarray(H, valid);
    H << '\n' << 1 << '\n' << "FabOnDisk: " << dname << ' ' << 0 << '\n' << '\n';
    H << std::scientific << std::setprecision(16);
    for (const std::vector<double>* v : {&mn, &mx}) {
        H << 1 << ',' << ncomp << '\n';
        for (double x : *v) H << x << ',';
        H << '\n' << '\n';
    }
    H.close();
    if (!H) {
        set_error("write failed: " + dir + "/" + name + "_H");
        return FHDX_ERR_ARG;
    }
    return FHDX_OK;
}

// VisMF::Read of a one-box MultiFab with `ncomp` components into `fab` (components 0..ncomp-1):
// the cells both boxes hold are copied, the others keep their values.
int vismf_read(const std::string& dir, const std::string& name, const fhdx_fab& fab, int ncomp, const Mover& mv)
{
    std::ifstream H(dir + "/" + name + "_H");
    if (!H) {
        set_error("cannot open " + dir + "/" + name + "_H");
        return FHDX_ERR_ARG;
    }
    int vers, how, nc, ngrow, nfab;
    Box3 valid;
    std::string tag, dname;
    long offset;
    if (!(H >> vers >> how >> nc >> ngrow) || vers != 1 || !read_boxarray(H, valid) || !(H >> nfab) || nfab != 1 ||
        !(H >> tag >> dname >> offset) || tag != "FabOnDisk:") {
        set_error("unsupported or malformed VisMF header " + dir + "/" + name + "_H");
        return FHDX_ERR_ARG;
    }
    if (nc != ncomp) {
        set_error("VisMF " + name + ": " + std::to_string(nc) + " components, expected " + std::to_string(ncomp));
        return FHDX_ERR_ARG;
    }
    std::ifstream D(dir + "/" + dname, std::ios::binary);
    if (!D || !D.seekg(offset)) {
        set_error("cannot open " + dir + "/" + dname);
        return FHDX_ERR_ARG;
    }
    std::string line;
    std::getline(D, line);
    const std::string head = std::string("FAB ") + kRealDescriptor;
    if (line.compare(0, head.size(), head) != 0) {
        set_error("FAB " + dname + " is not little-endian IEEE binary64 (only that layout is read)");
        return FHDX_ERR_UNSUPPORTED;
    }
    std::istringstream ls(line.substr(head.size()));
    Box3 fbox;
    int fnc;
    if (!read_box(ls, fbox) || !(ls >> fnc) || fnc != ncomp) {
        set_error("malformed FAB header in " + dname);
        return FHDX_ERR_ARG;
    }
    const Box3 tb = fab_box(fab);
    Box3 region;
    for (int d = 0; d < 3; ++d) {
        region.lo[d] = fbox.lo[d] > tb.lo[d] ? fbox.lo[d] : tb.lo[d];
        region.hi[d] = fbox.hi[d] < tb.hi[d] ? fbox.hi[d] : tb.hi[d];
        if (region.lo[d] > region.hi[d]) {
            set_error("checkpoint box does not overlap the array of " + name);
            return FHDX_ERR_ARG;
        }
    }
    std::vector<double> comp(fbox.cells()), cur;
    for (int c = 0; c < ncomp; ++c) {
        D.read(reinterpret_cast<char*>(comp.data()), (std::streamsize)(sizeof(double) * comp.size()));
        if (!D) {
            set_error("short read in " + dname);
            return FHDX_ERR_ARG;
        }
        int r = mv.get(fab, c, cur);
        if (!r) {
            copy_region(comp, fbox, cur, tb, region);
            r = mv.put(fab, c, cur);
        }
        if (r) return r;
    }
    return FHDX_OK;
}

bool check_fab(const fhdx_params& p, const fhdx_fab& f, int ncomp, const char* what)
{
    const Box3 g = grown_box(p);
    bool ok = f.p != nullptr && f.ncomp >= ncomp;
    for (int d = 0; d < 3; ++d) ok = ok && f.lo[d] == g.lo[d] && f.hi[d] == g.hi[d];
    if (!ok) set_error(std::string("bad cell array shape: ") + what);
    return ok;
}

struct Names {
    std::vector<std::string> v;
    void add(const std::string& s) { v.push_back(s); }
    void species(const std::string& s, int ns)
    {
        for (int i = 0; i < ns; ++i) v.push_back(s + char('0' + i));
    }
};

}  // namespace

extern "C" {

int fhdx_write_plotfile(const fhdx_params* p, const char* plot_base_name, int step, double time,
                        const double prob_lo[3], const fhdx_fields* f, int plot_means, int plot_vars, int flags,
                        void* stream)
{
    if (!p || !plot_base_name || !prob_lo || !f || step < 0) return FHDX_ERR_ARG;
    int r = fhdx_params_check(p);
    if (r) return r;
    const int ns = p->nspecies, nv = 5 + ns, np = 6 + 2 * ns;
    if (!check_fab(*p, f->cu, nv, "cu") || !check_fab(*p, f->prim, np, "prim") ||
        !check_fab(*p, f->spatialCross, 8, "spatialCross") || !check_fab(*p, f->eta, 1, "eta") ||
        !check_fab(*p, f->kappa, 1, "kappa"))
        return FHDX_ERR_ARG;
    if (plot_means == 1 && (!check_fab(*p, f->cuMeans, 5, "cuMeans") || !check_fab(*p, f->primMeans, 6, "primMeans")))
        return FHDX_ERR_ARG;
    if (plot_vars == 1 && (!check_fab(*p, f->cuVars, 5, "cuVars") || !check_fab(*p, f->primVars, 5, "primVars")))
        return FHDX_ERR_ARG;
    // components and names in the order of writePlotFile.cpp:47-180
    std::vector<Part> parts{{&f->cu, 0, nv}, {&f->prim, 0, np}};
    Names n;
    n.add("rhoInstant");
    n.add("jxInstant");
    n.add("jyInstant");
    n.add("jzInstant");
    n.add("rhoEInstant");
    n.species("rhoYkInstant_", ns);
    n.add("rhoInstant");
    n.add("uxInstant");
    n.add("uyInstant");
    n.add("uzInstant");
    n.add("tInstant");
    n.add("pInstant");
    n.species("YkInstant_", ns);
    n.species("XkInstant_", ns);
    if (plot_means == 1) {
        parts.push_back({&f->cuMeans, 0, 5});
        parts.push_back({&f->primMeans, 0, 6});
        for (const char* s : {"rhoMean", "jxMean", "jyMean", "jzMean", "rhoEMean", "rhoMean", "uxMean", "uyMean",
                              "uzMean", "tMean", "pMean"})
            n.add(s);
    }
    if (plot_vars == 1) {
        parts.push_back({&f->cuVars, 0, 5});
        parts.push_back({&f->primVars, 0, 5});
        for (const char* s :
             {"rhoVar", "jxVar", "jyVar", "jzVar", "rhoEVar", "rhoVar", "uxVar", "uyVar", "uzVar", "tVar"})
            n.add(s);
    }
    parts.push_back({&f->spatialCross, 0, 8});
    for (const char* s : {"Tyzaveragedcross", "Tyzaveraged", "T*T", "delT*delT", "delT*delrho", "delu*delrho",
                          "deljx*delrho", "delrho*delrhoE"})
        n.add(s);
    parts.push_back({&f->eta, 0, 1});
    parts.push_back({&f->kappa, 0, 1});
    n.add("eta");
    n.add("kappa");

    const std::string dir = step_name(plot_base_name, step);
    if (!make_dir(dir) || !make_dir(dir + "/Level_0")) return FHDX_ERR_ARG;
    const Box3 vb = valid_box(*p);
    const Mover mv{(flags & FHDX_IO_HOST_ARRAYS) != 0, reinterpret_cast<hipStream_t>(stream)};
    if ((r = vismf_write(dir + "/Level_0", "Cell", vb, 0, parts, mv))) return r;

    std::ofstream H(dir + "/Header", std::ios::trunc);
    if (!H) {
        set_error("cannot open " + dir + "/Header");
        return FHDX_ERR_ARG;
    }
    H.precision(17);
    H << "HyperCLaw-V1.1\n" << n.v.size() << '\n';
    for (const std::string& s : n.v) H << s << '\n';
    H << 3 << '\n' << time << '\n' << 0 << '\n';
    double hi[3];
    for (int d = 0; d < 3; ++d) hi[d] = prob_lo[d] + p->n_cells[d] * p->dx[d];
    for (int d = 0; d < 3; ++d) H << prob_lo[d] << ' ';
    H << '\n';
    for (int d = 0; d < 3; ++d) H << hi[d] << ' ';
    H << '\n' << '\n';                     // no refinement ratios on one level
    H << box_text(vb) << ' ' << '\n';      // domain of level 0
    H << step << ' ' << '\n';              // level steps
    for (int d = 0; d < 3; ++d) H << p->dx[d] << ' ';
    H << '\n' << 0 << '\n' << 0 << '\n';   // Cartesian coordinates, boundary width 0
    H << 0 << ' ' << 1 << ' ' << time << '\n' << step << '\n';
    for (int d = 0; d < 3; ++d)
        H << prob_lo[d] + p->dx[d] * vb.lo[d] << ' ' << prob_lo[d] + p->dx[d] * (vb.hi[d] + 1) << '\n';
    H << "Level_0/Cell\n";
    H.close();
    if (!H) {
        set_error("write failed: " + dir + "/Header");
        return FHDX_ERR_ARG;
    }
    return FHDX_OK;
}

// The fields of checkpoint.cpp:107-133 with their component counts (main_driver.cpp:205-238).
static std::vector<std::pair<const char*, std::pair<const fhdx_fab*, int>>> checkpoint_fields(const fhdx_params& p,
                                                                                              const fhdx_fields& f)
{
    const int nv = 5 + p.nspecies, np = 6 + 2 * p.nspecies;
    return {{"cu", {&f.cu, nv}},
            {"cuMeans", {&f.cuMeans, nv}},
            {"cuVars", {&f.cuVars, nv}},
            {"prim", {&f.prim, np}},
            {"primMeans", {&f.primMeans, np}},
            {"primVars", {&f.primVars, np + 5}},
            {"spatialCross", {&f.spatialCross, 8}},
            {"miscStats", {&f.miscStats, 10}},
            {"eta", {&f.eta, 1}},
            {"kappa", {&f.kappa, 1}}};
}

int fhdx_write_checkpoint(const fhdx_params* p, const char* chk_base_name, int step, double time, int statsCount,
                          uint64_t seed, const fhdx_fields* f, int flags, void* stream)
{
    if (!p || !chk_base_name || !f || step < 0) return FHDX_ERR_ARG;
    int r = fhdx_params_check(p);
    if (r) return r;
    const auto fields = checkpoint_fields(*p, *f);
    for (const auto& q : fields)
        if (!check_fab(*p, *q.second.first, q.second.second, q.first)) return FHDX_ERR_ARG;
    const std::string dir = step_name(chk_base_name, step);
    if (!make_dir(dir) || !make_dir(dir + "/Level_0")) return FHDX_ERR_ARG;
    {
        std::ofstream H(dir + "/Header", std::ios::trunc | std::ios::binary);
        if (!H) {
            set_error("cannot open " + dir + "/Header");
            return FHDX_ERR_ARG;
        }
        H.precision(17);
        H << "Checkpoint file for FHDeX/compress\n" << step << '\n' << time << '\n' << statsCount << '\n';
        write_boxarray(H, valid_box(*p));
        H << '\n' << "fhdx_philox_seed " << seed << '\n';
        H.close();
        if (!H) {
            set_error("write failed: " + dir + "/Header");
            return FHDX_ERR_ARG;
        }
    }
    const Mover mv{(flags & FHDX_IO_HOST_ARRAYS) != 0, reinterpret_cast<hipStream_t>(stream)};
    const Box3 vb = valid_box(*p);
    for (const auto& q : fields)
        if ((r = vismf_write(dir + "/Level_0", q.first, vb, p->ngc, {{q.second.first, 0, q.second.second}}, mv)))
            return r;
    return FHDX_OK;
}

int fhdx_read_checkpoint(const fhdx_params* p, const char* chk_base_name, int restart, int reset_stats, int* step,
                         double* time, int* statsCount, uint64_t* seed, int* has_seed, const fhdx_fields* f,
                         int flags, void* stream)
{
    if (!p || !chk_base_name || !f || !step || !time || !statsCount || restart < 0) return FHDX_ERR_ARG;
    int r = fhdx_params_check(p);
    if (r) return r;
    const auto fields = checkpoint_fields(*p, *f);
    for (const auto& q : fields)
        if (!check_fab(*p, *q.second.first, q.second.second, q.first)) return FHDX_ERR_ARG;
    const std::string dir = step_name(chk_base_name, restart);
    std::ifstream H(dir + "/Header");
    if (!H) {
        set_error("cannot open " + dir + "/Header");
        return FHDX_ERR_ARG;
    }
    std::string title;
    int st, sc;
    double t;
    Box3 ba;
    std::getline(H, title);
    if (!(H >> st) || !(H >> t) || !(H >> sc) || !read_boxarray(H, ba)) {
        set_error("malformed checkpoint header " + dir + "/Header");
        return FHDX_ERR_ARG;
    }
    const Box3 vb = valid_box(*p);
    for (int d = 0; d < 3; ++d)
        if (ba.lo[d] != vb.lo[d] || ba.hi[d] != vb.hi[d]) {
            set_error("checkpoint domain " + box_text(ba) + " differs from the run's " + box_text(vb));
            return FHDX_ERR_ARG;
        }
    std::string tag;
    uint64_t sd;
    const bool found = (H >> tag >> sd) && tag == "fhdx_philox_seed";
    if (found && seed) *seed = sd;
    if (has_seed) *has_seed = found ? 1 : 0;
    const Mover mv{(flags & FHDX_IO_HOST_ARRAYS) != 0, reinterpret_cast<hipStream_t>(stream)};
    // checkpoint.cpp:207-252: cu, prim, eta, kappa always; the statistics unless reset_stats
    for (const auto& q : fields) {
        const std::string name = q.first;
        const bool stat = !(name == "cu" || name == "prim" || name == "eta" || name == "kappa");
        if (!stat || reset_stats == 0) {
            r = vismf_read(dir + "/Level_0", name, *q.second.first, q.second.second, mv);
        } else if (reset_stats == 1) {
            const fhdx_fab& fab = *q.second.first;
            std::vector<double> zero(fab_box(fab).cells(), 0.0);
            for (int c = 0; c < q.second.second && !r; ++c) r = mv.put(fab, c, zero);
        }
        if (r) return r;
    }
    *step = st + 1;  // the reference resumes with the step after the checkpoint (checkpoint.cpp:172-173)
    *time = t;
    *statsCount = reset_stats == 1 ? 1 : sc;
    return FHDX_OK;
}

}  // extern "C"
