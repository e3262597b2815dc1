// ose_host_sim.cpp -- the simulation object behind the ose3d driver: the
// parameter set of the reference's main() (ose_3d.cu:75-199) as SU-style
// key=value pairs and presets, model and shot construction, and the time
// loop through include/ose_api.h.
#include <algorithm>
#include <cerrno>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iterator>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "ose_host.h"

namespace {

struct ShotSpec {
    int x, y, z;   // interior coordinates
};

struct ModelSpec {
    int nx, ny, nz;
    int kind;          // oseh_model kind, or -1 for a velocity file
    uint64_t seed;
    std::vector<ShotSpec> shots;
};

struct BuiltModel {
    ModelSpec spec;
    int bl, br, bf, bb, bu, bd, nxb, nyb, nzb;
    std::vector<float> v;   // padded
    double vmax, R, z;
    int K;
    std::vector<double> coef;
    std::vector<float> gx, gy, gz;
    int id = -1;
    std::vector<int> shot_ids;
    std::string adjusted;   // extent=smooth changes, as JSON entries
};

struct ShotRec {
    int model;
    int src[3];              // padded
    std::vector<int> rec;    // listed receivers, interior (x, y, z) triples; empty: the plane rec_z
};

}  // namespace

struct oseh_sim {
    std::map<std::string, std::string> p;
    std::vector<ModelSpec> specs;
    std::vector<BuiltModel> models;
    std::vector<ShotRec> shots;
    std::vector<float> s, h;
    ose_ctx* ctx = nullptr;
    std::string err;
    int steps_done = 0;
    // streamed snapshots: output directory and the first write error
    std::string snap_dir;
    std::mutex snap_mu;
    std::string snap_err;

    int fail(int code, const std::string& m)
    {
        err = m;
        return code;
    }
    std::string get(const std::string& k, const std::string& d) const
    {
        auto it = p.find(k);
        return it == p.end() ? d : it->second;
    }
    // Numeric keys parse strictly: a value that is not a whole finite number
    // (in int range for geti) gives the default and is remembered in `bad`,
    // which oseh_sim_build reports as an error.
    mutable std::string bad;
    double getd(const std::string& k, double d) const
    {
        auto it = p.find(k);
        if (it == p.end()) return d;
        const char* c = it->second.c_str();
        char* end = nullptr;
        errno = 0;
        const double v = std::strtod(c, &end);
        if (end == c || *end || errno == ERANGE || !std::isfinite(v)) {
            if (bad.empty()) bad = k + "=" + it->second;
            return d;
        }
        return v;
    }
    int geti(const std::string& k, int d) const
    {
        auto it = p.find(k);
        if (it == p.end()) return d;
        const char* c = it->second.c_str();
        char* end = nullptr;
        errno = 0;
        const long v = std::strtol(c, &end, 10);
        if (end == c || *end || errno == ERANGE || v < INT_MIN || v > INT_MAX) {
            if (bad.empty()) bad = k + "=" + it->second;
            return d;
        }
        return (int)v;
    }
};

// first failed write of a streamed snapshot, reported once the run returns
static int snap_error(oseh_sim* s)
{
    std::lock_guard<std::mutex> lk(s->snap_mu);
    return s->snap_err.empty() ? OSE_OK : s->fail(OSE_EINVAL, s->snap_err);
}

namespace {

void parse_into(std::map<std::string, std::string>& p, const std::string& s)
{
    std::istringstream is(s);
    std::string tok;
    while (is >> tok) {
        const size_t eq = tok.find('=');
        if (eq == std::string::npos) continue;
        p[tok.substr(0, eq)] = tok.substr(eq + 1);
    }
}

// Benchmark configurations.  Sizes, velocities, sources and receivers follow
// the configuration texts; unstated spacing and time step are 10 m and 1 ms.
const char* preset_text(const std::string& name)
{
    if (name == "ref")   // ose_3d.cu:75-119 as shipped
        return "nx=101 ny=101 nz=101 boundary=25 nt=300 fm=40 dt=0.001 amp=100 dx=10 dy=10 dz=10 "
               "model=twolayer v0=3000 v1=4000 iface=30 sx=50 sy=50 sz=0 rec_z=0 korder=ref "
               "taper=ref hilbert=unnorm extent=exact delay=40";
    if (name == "c0")
        return "nx=100 ny=100 nz=100 boundary=20 nt=1000 fm=25 dt=0.001 amp=100 dx=10 dy=10 dz=10 "
               "model=twolayer v0=2500 v1=3500 iface=50 sx=50 sy=50 sz=10 rec_z=5 korder=tail "
               "ktol=1e-7 taper=cerjan alpha=0.015 hilbert=unnorm extent=exact delay=40";
    if (name == "c1")
        return "nx=256 ny=256 nz=256 boundary=40 nt=2000 fm=20 dt=0.001 amp=100 dx=10 dy=10 dz=10 "
               "model=layered8 seed=2 sx=128 sy=128 sz=0 rec_z=0 korder=tail ktol=1e-7 taper=ref "
               "hilbert=unnorm extent=smooth delay=40";
    if (name == "c2")
        return "batch=c2 boundary=40 nt=1500 fm=20 dt=0.001 amp=100 dx=10 dy=10 dz=10 rec_z=0 "
               "korder=tail ktol=1e-7 taper=ref hilbert=unnorm extent=smooth delay=40";
    if (name == "c3")
        return "nx=256 ny=256 nz=256 boundary=8 nt=2000 fm=20 dt=0.001 amp=100 dx=10 dy=10 dz=10 "
               "model=layered8 seed=2000 sx=128 sy=128 sz=0 rec_z=0 korder=tail ktol=1e-7 taper=ref "
               "hilbert=unnorm extent=smooth delay=40";
    if (name == "c4")
        return "nx=768 ny=768 nz=768 boundary=40 nt=3000 fm=20 dt=0.001 amp=100 dx=10 dy=10 dz=10 "
               "model=salt12 seed=5 batch=c4 rec_z=0 korder=tail ktol=1e-7 taper=ref hilbert=unnorm "
               "extent=smooth delay=40";
    return nullptr;
}

int model_kind(const std::string& m)
{
    if (m == "const") return 0;
    if (m == "twolayer") return 1;
    if (m == "layered8") return 2;
    if (m == "salt12") return 3;
    if (m == "train") return 4;
    if (m == "file") return -1;
    return -2;
}

int build_specs(oseh_sim* s)
{
    s->specs.clear();
    const std::string batch = s->get("batch", "");
    if (batch == "c2") {
        ModelSpec a{512, 512, 384, 2, 3, {}};
        for (int iy = 0; iy < 4; ++iy)
            for (int ix = 0; ix < 4; ++ix) a.shots.push_back({64 + 128 * ix, 64 + 128 * iy, 0});
        s->specs.push_back(a);
        const int dims[4][3] = {{192, 192, 192}, {256, 256, 200}, {320, 240, 160}, {288, 160, 224}};
        for (int k = 0; k < 4; ++k) {
            ModelSpec m{dims[k][0], dims[k][1], dims[k][2], 2, (uint64_t)(31 + k), {}};
            m.shots.push_back({dims[k][0] / 2, dims[k][1] / 2, 0});
            s->specs.push_back(m);
        }
    } else {
        ModelSpec m{s->geti("nx", 64), s->geti("ny", 64), s->geti("nz", 64),
                    model_kind(s->get("model", "layered8")), (uint64_t)s->geti("seed", 1), {}};
        if (s->get("vfile", "") != "") m.kind = -1;
        if (m.kind == -2) return s->fail(OSE_EINVAL, "unknown model " + s->get("model", ""));
        if (batch == "c4") {
            for (int iy = 0; iy < 2; ++iy)
                for (int ix = 0; ix < 4; ++ix)
                    m.shots.push_back({(int)(m.nx * (0.2 + 0.2 * ix)), (int)(m.ny * (0.35 + 0.3 * iy)), 0});
        } else if (s->p.count("srcgrid")) {
            const int g = s->geti("srcgrid", 1);
            if (g < 1 || g > 256) return s->fail(OSE_EINVAL, "srcgrid must be in [1, 256]");
            for (int iy = 0; iy < g; ++iy)
                for (int ix = 0; ix < g; ++ix)
                    m.shots.push_back({(int)((ix + 0.5) * m.nx / g), (int)((iy + 0.5) * m.ny / g), s->geti("sz", 0)});
        } else {
            // shots on a line: (sx + is*jsx, sy + is*jsy, sz + is*jsz), sg_init_host ose_3d_lib.cu:598-608
            const int ns = s->geti("nshot", 1);
            if (ns < 1 || ns > 65536) return s->fail(OSE_EINVAL, "nshot must be in [1, 65536]");
            for (int k = 0; k < ns; ++k)
                m.shots.push_back({s->geti("sx", m.nx / 2) + k * s->geti("jsx", 1),
                                   s->geti("sy", m.ny / 2) + k * s->geti("jsy", 1),
                                   s->geti("sz", 0) + k * s->geti("jsz", 1)});
        }
        s->specs.push_back(m);
    }
    const int cap = s->geti("max_shots", 0);
    if (cap > 0)
        for (auto& m : s->specs)
            if ((int)m.shots.size() > cap) m.shots.resize(cap);
    return OSE_OK;
}

// pads of one axis: before/after, the odd rule (judge_odd, ose_3d_lib.cu:36-44),
// then the extent rule.  extent=exact (the default) keeps the reference's
// padded extent, which every size up to 4096 supports; extent=smooth widens
// the after-pad to the next extent compiled in autotuned launch shapes
// (faster transforms, but a different periodic domain and taper than the
// reference's grid), and says so on stderr and in the info record.
int axis_pads(oseh_sim* s, const char* axis, int n, int b, int& before, int& after, int& nb, std::string& adjusted)
{
    if (n < 1 || n > 4096 || b < 0 || b > 4096)
        return s->fail(OSE_EINVAL, std::string(axis) + ": interior cells in [1, 4096] and pad cells in [0, 4096]");
    before = b;
    after = b;
    nb = before + n + after;
    oseh_judge_odd(&after, &nb);
    const std::string ext = s->get("extent", "exact");
    if (ext != "exact" && ext != "smooth") return s->fail(OSE_EINVAL, "extent must be exact or smooth");
    if (ext == "smooth" && ose_extent_kind(nb) != 1) {   // not compiled in autotuned launch shapes
        const int t = ose_supported_extent(nb);
        if (t > 0) {
            std::fprintf(stderr, "ose: extent=smooth widens the %s padded extent %d to %d (after-pad %d -> %d cells)\n",
                         axis, nb, t, after, after + t - nb);
            adjusted += std::string(adjusted.empty() ? "" : ", ") + "{\"axis\": \"" + axis +
                        "\", \"from\": " + std::to_string(nb) + ", \"to\": " + std::to_string(t) + "}";
            after += t - nb;   // the taper is extended over the extra cells
            nb = t;
        }
    }
    if (!ose_extent_is_supported(nb))
        return s->fail(OSE_EUNSUPPORTED, "padded extent " + std::to_string(nb) + " is outside [2, 4096]");
    if (ose_extent_kind(nb) == 2) {
        const int t = ose_supported_extent(nb);
        std::fprintf(stderr, "ose: the %s padded extent %d has no compiled transform kernels and runs on the "
                             "slower runtime-length path%s%s\n",
                     axis, nb, t > 0 ? "; extent=smooth would pad it to " : "",
                     t > 0 ? std::to_string(t).c_str() : "");
    }
    return OSE_OK;
}

int build_model(oseh_sim* s, const ModelSpec& spec, BuiltModel& m)
{
    m.spec = spec;
    const int b = s->geti("boundary", 20);
    int rc;
    m.adjusted.clear();
    if ((rc = axis_pads(s, "x", spec.nx, s->geti("bl", b), m.bl, m.br, m.nxb, m.adjusted))) return rc;
    if ((rc = axis_pads(s, "y", spec.ny, s->geti("bf", b), m.bf, m.bb, m.nyb, m.adjusted))) return rc;
    if ((rc = axis_pads(s, "z", spec.nz, s->geti("bu", b), m.bu, m.bd, m.nzb, m.adjusted))) return rc;
    std::vector<float> vi((size_t)spec.nx * spec.ny * spec.nz);
    if (spec.kind == -1) {
        rc = oseh_read_velocity_xyz(s->get("vfile", "").c_str(), spec.nx, spec.ny, spec.nz, vi.data());
        if (rc) return s->fail(rc, "cannot read velocity file " + s->get("vfile", ""));
    } else {
        rc = oseh_model(spec.kind, spec.nx, spec.ny, spec.nz, spec.seed, s->getd("v0", 2000.0),
                        s->getd("v1", 3500.0), s->geti("iface", spec.nz / 2), vi.data());
        if (rc) return s->fail(rc, "model generation failed");
    }
    m.v.resize((size_t)m.nxb * m.nyb * m.nzb);
    oseh_pad_velocity(vi.data(), spec.nx, spec.ny, spec.nz, m.bl, m.br, m.bf, m.bb, m.bu, m.bd, m.v.data());
    m.vmax = *std::max_element(m.v.begin(), m.v.end());   // find_max, ose_3d_lib.cu:544-555
    const double dx = s->getd("dx", 10), dy = s->getd("dy", 10), dz = s->getd("dz", 10), dt = s->getd("dt", 1e-3);
    if (!(dx > 0) || !(dy > 0) || !(dz > 0) || !(dt > 0) || !std::isfinite(dx * dy * dz * dt))
        return s->fail(OSE_EINVAL, "dx, dy, dz and dt must be positive");
    if (!(m.vmax > 0) || !std::isfinite(m.vmax)) return s->fail(OSE_EINVAL, "velocities must be positive and finite");
    m.R = oseh_spectral_radius(m.vmax, dx, dy, dz);
    m.z = m.R * dt;
    const std::string ko = s->get("korder", "tail");
    if (ko == "ref")
        m.K = oseh_order_reference(m.z);
    else if (ko == "tail")
        m.K = oseh_order_tail(m.z, s->getd("ktol", 1e-7));
    else
        m.K = std::atoi(ko.c_str());
    if (m.K < 2 || m.K > 64) return s->fail(OSE_EINVAL, "order K out of range");
    m.coef.resize(m.K + 1);
    if ((rc = oseh_coefficients(m.z, m.K, m.coef.data()))) return s->fail(rc, "Bessel coefficients failed");
    const std::string tp = s->get("taper", "ref");
    const int mode = tp == "ref" ? OSEH_TAPER_REF : tp == "cerjan" ? OSEH_TAPER_CERJAN : OSEH_TAPER_NONE;
    const double alpha = s->getd("alpha", 0.015);
    m.gx.resize(m.nxb);
    m.gy.resize(m.nyb);
    m.gz.resize(m.nzb);
    oseh_taper(m.gx.data(), m.bl, spec.nx, m.br, mode, alpha);
    oseh_taper(m.gy.data(), m.bf, spec.ny, m.bb, mode, alpha);
    oseh_taper(m.gz.data(), m.bu, spec.nz, m.bd, mode, alpha);
    return OSE_OK;
}

// Boundary-network parameters: a .npy file holding one little-endian
// float32 array (version 1-3 header, C order), or raw float32 otherwise.
int read_params_f32(const std::string& path, std::vector<float>& out, std::string& err)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) {
        err = "cannot open " + path;
        return OSE_EINVAL;
    }
    std::vector<char> all((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    size_t off = 0;
    if (all.size() >= 10 && std::memcmp(all.data(), "\x93NUMPY", 6) == 0) {
        const int major = (unsigned char)all[6];
        size_t hlen = 0, hoff = 0;
        if (major == 1) {
            hlen = (unsigned char)all[8] | ((size_t)(unsigned char)all[9] << 8);
            hoff = 10;
        } else if (all.size() >= 12) {
            hlen = 0;
            for (int k = 0; k < 4; ++k) hlen |= (size_t)(unsigned char)all[8 + k] << (8 * k);
            hoff = 12;
        }
        if (!hoff || hoff + hlen > all.size()) {
            err = path + ": bad .npy header";
            return OSE_EINVAL;
        }
        const std::string h(all.data() + hoff, hlen);
        if (h.find("'<f4'") == std::string::npos || h.find("'fortran_order': False") == std::string::npos) {
            err = path + ": expected a C-order little-endian float32 array";
            return OSE_EINVAL;
        }
        off = hoff + hlen;
    }
    const size_t n = (all.size() - off) / sizeof(float);
    if (n == 0 || (all.size() - off) % sizeof(float)) {
        err = path + ": size is not a whole number of float32 values";
        return OSE_EINVAL;
    }
    out.resize(n);
    std::memcpy(out.data(), all.data() + off, n * sizeof(float));
    return OSE_OK;
}

}  // namespace

extern "C" {

int oseh_sim_create(oseh_sim** out, const char* params)
{
    if (!out) return OSE_EINVAL;
    *out = nullptr;
    auto* s = new (std::nothrow) oseh_sim();
    if (!s) return OSE_ENOMEM;
    std::map<std::string, std::string> user;
    parse_into(user, params ? params : "");
    auto pr = user.find("preset");
    if (pr != user.end()) {
        const char* t = preset_text(pr->second);
        if (!t) {
            delete s;
            return OSE_EINVAL;
        }
        parse_into(s->p, t);
    }
    for (auto& kv : user) s->p[kv.first] = kv.second;
    if (s->get("boundary", "") == "" && !s->p.count("bl")) s->p["boundary"] = "20";
    *out = s;
    return OSE_OK;
}

const char* oseh_sim_error(oseh_sim* s) { return s ? s->err.c_str() : "null simulation"; }

// Receiver geometry of every shot.  Without ng/ngx/ngy the shot records the
// whole interior plane rec_z (what rec_shotdata_3d does, ose_3d_kernel.cu:
// 425-436).  ng=<n>: a line of n receivers at (gxbeg + r*jgx, gybeg + r*jgy,
// gzbeg + r*jgz), the geometry sg_init_device builds from the reference's
// gxbeg..jgz (ose_3d.cu:105-114, ose_3d_kernel.cu:202-212); defaults
// jgx = 1, jgy = jgz = 0.  ngx/ngy=<n>: an ngx x ngy grid at depth gzbeg,
// x fastest, spacing jgx, jgy (default 1).  csdgather=1 moves the spread with
// the source: shot k's receivers shift by its source's offset from the
// model's first source (a common-shot gather with a fixed source-receiver
// layout).  Coordinates are interior cells; a receiver outside the interior
// is an error.
static int build_receivers(oseh_sim* s)
{
    const int ng = s->geti("ng", 0), ngx = s->geti("ngx", 0), ngy = s->geti("ngy", 0);
    if (ng < 0 || ngx < 0 || ngy < 0) return s->fail(OSE_EINVAL, "ng, ngx and ngy must not be negative");
    const bool grid = ngx > 0 || ngy > 0;
    if (ng == 0 && !grid) return OSE_OK;
    if (ng > 0 && grid) return s->fail(OSE_EINVAL, "give either ng (a receiver line) or ngx/ngy (a grid)");
    const int gx0 = s->geti("gxbeg", 0), gy0 = s->geti("gybeg", 0), gz0 = s->geti("gzbeg", 0);
    const int jgx = s->geti("jgx", 1), jgy = s->geti("jgy", grid ? 1 : 0), jgz = s->geti("jgz", 0);
    const bool csd = s->geti("csdgather", 0) != 0;
    std::vector<int> first(s->models.size(), -1);
    for (size_t k = 0; k < s->shots.size(); ++k) {
        ShotRec& sh = s->shots[k];
        const BuiltModel& m = s->models[sh.model];
        if (first[sh.model] < 0) first[sh.model] = (int)k;
        const ShotRec& f = s->shots[first[sh.model]];
        const int ox = csd ? sh.src[0] - f.src[0] : 0, oy = csd ? sh.src[1] - f.src[1] : 0,
                  oz = csd ? sh.src[2] - f.src[2] : 0;
        sh.rec.clear();
        auto add = [&](int x, int y, int z) {
            x += ox, y += oy, z += oz;
            if (x < 0 || x >= m.spec.nx || y < 0 || y >= m.spec.ny || z < 0 || z >= m.spec.nz)
                return s->fail(OSE_EINVAL, "receiver " + std::to_string(sh.rec.size() / 3) + " of shot " +
                                               std::to_string(k) + " at (" + std::to_string(x) + ", " +
                                               std::to_string(y) + ", " + std::to_string(z) +
                                               ") is outside the interior");
            sh.rec.insert(sh.rec.end(), {x, y, z});
            return (int)OSE_OK;
        };
        int rc;
        if (grid) {
            for (int iy = 0; iy < std::max(ngy, 1); ++iy)
                for (int ix = 0; ix < std::max(ngx, 1); ++ix)
                    if ((rc = add(gx0 + ix * jgx, gy0 + iy * jgy, gz0))) return rc;
        } else {
            for (int r = 0; r < ng; ++r)
                if ((rc = add(gx0 + r * jgx, gy0 + r * jgy, gz0 + r * jgz))) return rc;
        }
    }
    return OSE_OK;
}

static int sim_build(oseh_sim* s);

int oseh_sim_build(oseh_sim* s)
{
    if (!s) return OSE_EINVAL;
    if (!s->models.empty()) return OSE_OK;
    int rc;
    try {
        rc = sim_build(s);
    } catch (const std::bad_alloc&) {
        s->models.clear();
        return s->fail(OSE_ENOMEM, "out of host memory building the models");
    } catch (const std::exception& e) {
        s->models.clear();
        return s->fail(OSE_EINVAL, std::string("model build failed: ") + e.what());
    }
    if (rc) {
        s->models.clear();
        return rc;
    }
    if (!s->bad.empty()) {
        s->models.clear();
        return s->fail(OSE_EINVAL, "parameter " + s->bad + " is not a number in range");
    }
    return OSE_OK;
}

static int sim_build(oseh_sim* s)
{
    int rc;
    if ((rc = build_specs(s))) return rc;
    // source wavelets (Alloc_wavelet, ose_3d_lib.cu:557-586)
    const int nt = s->geti("nt", 300);
    if (nt <= 0) return s->fail(OSE_EINVAL, "nt must be positive");
    std::vector<double> w(nt), hh(nt);
    oseh_ricker(w.data(), nt, s->getd("fm", 25.0), s->getd("dt", 1e-3), s->getd("amp", 100.0),
                s->geti("delay", 40));
    // hilbert=unnorm (the reference: nt H[s]), norm (H[s]) or analytic (-H[s]:
    // with the exp(-i L dt) step, s - i H[s] is the analytic signal that radiates)
    const std::string hm = s->get("hilbert", "unnorm");
    if (hm != "unnorm" && hm != "norm" && hm != "analytic")
        return s->fail(OSE_EINVAL, "hilbert must be unnorm, norm or analytic");
    oseh_hilbert(w.data(), hh.data(), nt, hm == "unnorm" ? 0 : 1);
    if (hm == "analytic")
        for (double& v : hh) v = -v;
    s->s.assign(w.begin(), w.end());
    s->h.assign(hh.begin(), hh.end());
    s->models.clear();
    for (auto& spec : s->specs) {
        BuiltModel m;
        if ((rc = build_model(s, spec, m))) return rc;
        s->models.push_back(std::move(m));
    }
    s->shots.clear();
    for (size_t mi = 0; mi < s->models.size(); ++mi) {
        const BuiltModel& m = s->models[mi];
        for (auto& sh : m.spec.shots) {
            if (sh.x < 0 || sh.x >= m.spec.nx || sh.y < 0 || sh.y >= m.spec.ny || sh.z < 0 || sh.z >= m.spec.nz)
                return s->fail(OSE_EINVAL, "source outside the interior");
            s->shots.push_back(ShotRec{(int)mi, {m.bl + sh.x, m.bf + sh.y, m.bu + sh.z}, {}});
        }
    }
    return build_receivers(s);
}

int oseh_gaussian_profile(double sigma, int center, int n, int* start, float* out, int cap)
{
    if (!start || !out || center < 0 || center >= n || sigma < 0.0) return -1;
    const int r = sigma > 0.0 ? (int)std::ceil(4.0 * sigma) : 0;
    const int a = std::max(0, center - r), b = std::min(n - 1, center + r);
    if (b - a + 1 > cap) return -1;
    double sum = 0.0;
    std::vector<double> g(b - a + 1);
    for (int j = a; j <= b; ++j) {
        const double t = sigma > 0.0 ? (j - center) / sigma : 0.0;
        g[j - a] = std::exp(-0.5 * t * t);
        sum += g[j - a];
    }
    for (int j = a; j <= b; ++j) out[j - a] = (float)(g[j - a] / sum);
    *start = a;
    return b - a + 1;
}

int oseh_sim_setup(oseh_sim* s)
{
    if (!s) return OSE_EINVAL;
    if (s->ctx) return s->fail(OSE_ESTATE, "already set up");
    int rc;
    if ((rc = oseh_sim_build(s))) return rc;
    const int nt = s->geti("nt", 300);
    if ((rc = ose_create(&s->ctx, s->geti("device", 0)))) return s->fail(rc, "ose_create failed");
    ose_set_options(s->ctx, s->geti("graph", 1), 0);
    // streams=1|2: one stream, or two concurrent lanes for several models
    if ((rc = ose_set_streams(s->ctx, s->geti("streams", 2))))
        return s->fail(rc, "streams must be 1 or 2");
    const int rec_z = s->geti("rec_z", 0);
    for (size_t mi = 0; mi < s->models.size(); ++mi) {
        BuiltModel& m = s->models[mi];
        ose_model_desc d{};
        d.nxb = m.nxb;
        d.nyb = m.nyb;
        d.nzb = m.nzb;
        d.bl = m.bl;
        d.bf = m.bf;
        d.bu = m.bu;
        d.nx = m.spec.nx;
        d.ny = m.spec.ny;
        d.nz = m.spec.nz;
        d.dx = s->getd("dx", 10);
        d.dy = s->getd("dy", 10);
        d.dz = s->getd("dz", 10);
        d.dt = s->getd("dt", 1e-3);
        d.R = m.R;
        d.K = m.K;
        d.coef = m.coef.data();
        d.taper_x = m.gx.data();
        d.taper_y = m.gy.data();
        d.taper_z = m.gz.data();
        d.rec_z = rec_z;
        if ((rc = ose_add_model(s->ctx, &d, m.v.data(), &m.id)))
            return s->fail(rc, std::string("ose_add_model: ") + ose_last_error(s->ctx));
    }
    // shots in context order (ose_add_shot ids equal the shot indices)
    for (const ShotRec& r : s->shots) {
        BuiltModel& m = s->models[r.model];
        int sid = -1;
        if ((rc = ose_add_shot(s->ctx, m.id, r.src[0], r.src[1], r.src[2], s->s.data(), s->h.data(), nt, &sid)))
            return s->fail(rc, std::string("ose_add_shot: ") + ose_last_error(s->ctx));
        m.shot_ids.push_back(sid);
        // extended source: src_sigma=<cells> or auto (spectrum e^{-k^2 sigma^2 / 2}
        // below 1e-3 at the temporal Nyquist pi / (v_src dt): sigma = 3.72 v dt / pi)
        const std::string ss = s->get("src_sigma", "0");
        if (ss != "0" && ss != "0.0") {
            const double d3[3] = {s->getd("dx", 10), s->getd("dy", 10), s->getd("dz", 10)};
            const int n3[3] = {m.nxb, m.nyb, m.nzb};
            const double vs = m.v[((size_t)r.src[2] * m.nyb + r.src[1]) * m.nxb + r.src[0]];
            int box[6];
            float prof[3][64];
            for (int a = 0; a < 3; ++a) {
                const double sig = ss == "auto" ? 3.72 * vs * s->getd("dt", 1e-3) / (M_PI * d3[a]) : std::atof(ss.c_str());
                const int w = oseh_gaussian_profile(sig, r.src[a], n3[a], &box[a], prof[a], 64);
                if (w < 1) return s->fail(OSE_EINVAL, "src_sigma " + ss + ": source profile wider than 64 cells");
                box[3 + a] = w;
            }
            if ((rc = ose_set_source_profile(s->ctx, sid, box, prof[0], prof[1], prof[2])))
                return s->fail(rc, std::string("ose_set_source_profile: ") + ose_last_error(s->ctx));
        }
        if (!r.rec.empty()) {
            std::vector<int> xyz(r.rec.size());
            for (size_t i = 0; i < r.rec.size(); i += 3) {
                xyz[i] = r.rec[i] + m.bl;
                xyz[i + 1] = r.rec[i + 1] + m.bf;
                xyz[i + 2] = r.rec[i + 2] + m.bu;
            }
            if ((rc = ose_set_receivers(s->ctx, sid, xyz.data(), (int)(xyz.size() / 3))))
                return s->fail(rc, std::string("ose_set_receivers: ") + ose_last_error(s->ctx));
        }
    }
    // learned thin pad: bnet=<parameters file> [guard=4] [bnet_taper=replace|keep]
    const std::string bn = s->get("bnet", "");
    if (!bn.empty()) {
        std::vector<float> par;
        std::string err;
        if ((rc = read_params_f32(bn, par, err))) return s->fail(rc, err);
        const int replace = s->get("bnet_taper", "replace") == "keep" ? 0 : 1;
        for (BuiltModel& m : s->models)
            if ((rc = ose_set_boundary_net(s->ctx, m.id, par.data(), par.size(), s->getd("guard", 4.0), replace)))
                return s->fail(rc, std::string("ose_set_boundary_net: ") + ose_last_error(s->ctx));
    }
    if (!s->bad.empty()) return s->fail(OSE_EINVAL, "parameter " + s->bad + " is not a number in range");
    return OSE_OK;
}

ose_ctx* oseh_sim_ctx(oseh_sim* s) { return s ? s->ctx : nullptr; }

int oseh_read_params(const char* path, float* out, size_t cap, size_t* n)
{
    if (!path || !n) return OSE_EINVAL;
    std::vector<float> p;
    std::string err;
    int rc = read_params_f32(path, p, err);
    if (rc) return rc;
    *n = p.size();
    if (out) {
        if (cap < p.size()) return OSE_EINVAL;
        std::memcpy(out, p.data(), p.size() * sizeof(float));
    }
    return OSE_OK;
}

int oseh_sim_run(oseh_sim* s, int nsteps, double* ms)
{
    if (!s) return OSE_EINVAL;
    if (!s->ctx) return s->fail(OSE_ESTATE, "not set up");
    double t = 0.0;
    int rc = ose_timed_run(s->ctx, nsteps, &t);
    if (rc) return s->fail(rc, std::string("ose_timed_run: ") + ose_last_error(s->ctx));
    s->steps_done += nsteps;
    if (ms) *ms = t;
    return snap_error(s);
}

int oseh_sim_info(oseh_sim* s, char* buf, size_t cap)
{
    if (!s || !buf) return OSE_EINVAL;
    std::ostringstream o;
    double interior = 0, padded = 0;
    o << "{\"models\": [";
    for (size_t i = 0; i < s->models.size(); ++i) {
        const BuiltModel& m = s->models[i];
        const double ns = (double)m.spec.shots.size();
        interior += ns * m.spec.nx * m.spec.ny * m.spec.nz;
        padded += ns * (double)m.nxb * m.nyb * m.nzb;
        o << (i ? ", " : "") << "{\"interior\": [" << m.spec.nx << ", " << m.spec.ny << ", " << m.spec.nz
          << "], \"padded\": [" << m.nxb << ", " << m.nyb << ", " << m.nzb << "], \"pads\": [" << m.bl << ", "
          << m.br << ", " << m.bf << ", " << m.bb << ", " << m.bu << ", " << m.bd << "], \"K\": " << m.K
          << ", \"R\": " << m.R << ", \"z\": " << m.z << ", \"vmax\": " << m.vmax
          << ", \"shots\": " << m.spec.shots.size() << ", \"extent_adjusted\": [" << m.adjusted << "]}";
    }
    o << "], \"nshots\": " << s->shots.size() << ", \"interior_cells\": " << (long long)interior
      << ", \"padded_cells\": " << (long long)padded << ", \"nt\": " << s->geti("nt", 300)
      << ", \"taper\": \"" << s->get("taper", "ref") << "\", \"korder\": \"" << s->get("korder", "tail")
      << "\", \"bnet\": \"" << s->get("bnet", "") << "\"}";
    const std::string t = o.str();
    if (t.size() + 1 > cap) return OSE_EINVAL;
    std::memcpy(buf, t.c_str(), t.size() + 1);
    return OSE_OK;
}

// interior (x, y, z) of every receiver of shot k: its list, or the plane
// rec_z with x fastest
static std::vector<int> receiver_cells(const oseh_sim* s, size_t k)
{
    const ShotRec& sh = s->shots[k];
    if (!sh.rec.empty()) return sh.rec;
    const ModelSpec& m = s->models[sh.model].spec;
    const int rz = s->geti("rec_z", 0);
    std::vector<int> xyz;
    xyz.reserve((size_t)3 * m.nx * m.ny);
    for (int iy = 0; iy < m.ny; ++iy)
        for (int ix = 0; ix < m.nx; ++ix) xyz.insert(xyz.end(), {ix, iy, rz});
    return xyz;
}

int oseh_sim_receivers(oseh_sim* s, int shot, int* xyz, int* n)
{
    if (!s || !n) return OSE_EINVAL;
    if (s->models.empty()) return s->fail(OSE_ESTATE, "not built");
    if (shot < 0 || shot >= (int)s->shots.size()) return s->fail(OSE_EINVAL, "bad shot index");
    const std::vector<int> c = receiver_cells(s, (size_t)shot);
    const int cnt = (int)(c.size() / 3);
    if (xyz && *n >= cnt) std::copy(c.begin(), c.end(), xyz);
    const int cap = *n;
    *n = cnt;
    return (xyz && cap < cnt) ? OSE_EINVAL : OSE_OK;
}

int oseh_sim_write(oseh_sim* s, const char* dir)
{
    if (!s || !dir) return OSE_EINVAL;
    if (!s->ctx) return s->fail(OSE_ESTATE, "not set up");
    int rc;
    if ((rc = ose_sync(s->ctx))) return s->fail(rc, std::string("ose_sync: ") + ose_last_error(s->ctx));
    if ((rc = snap_error(s))) return rc;
    const int nt = std::min(s->steps_done, s->geti("nt", 300));
    const double dx = s->getd("dx", 10), dy = s->getd("dy", 10), dz = s->getd("dz", 10), dt = s->getd("dt", 1e-3);
    for (size_t k = 0; k < s->shots.size(); ++k) {
        const BuiltModel& m = s->models[s->shots[k].model];
        const std::vector<int> cells = receiver_cells(s, k);
        const int nrec = (int)(cells.size() / 3);
        std::vector<float> tr((size_t)nrec * nt);
        if ((rc = ose_get_traces(s->ctx, (int)k, 0, nt, tr.data())))
            return s->fail(rc, std::string("ose_get_traces: ") + ose_last_error(s->ctx));
        const std::string base = std::string(dir) + "/shot" + std::to_string(k);
        if ((rc = oseh_write_raw((base + ".bin").c_str(), tr.data(), tr.size())))
            return s->fail(rc, "cannot write " + base + ".bin");
        std::vector<double> rx(nrec), ry(nrec), rz(nrec);
        for (int r = 0; r < nrec; ++r) {
            rx[r] = cells[3 * r] * dx;
            ry[r] = cells[3 * r + 1] * dy;
            rz[r] = cells[3 * r + 2] * dz;
        }
        const double sx = (s->shots[k].src[0] - m.bl) * dx, sy = (s->shots[k].src[1] - m.bf) * dy,
                     sz = (s->shots[k].src[2] - m.bu) * dz;
        if (nt <= 65535 && (rc = oseh_write_su((base + ".su").c_str(), tr.data(), nrec, nt, dt, rx.data(),
                                               ry.data(), rz.data(), sx, sy, sz, (int)k)))
            return s->fail(rc, "cannot write " + base + ".su");
        if (s->geti("segy", 0) && (rc = oseh_write_segy((base + ".segy").c_str(), tr.data(), nrec, nt, dt, rx.data(),
                                                         ry.data(), rz.data(), sx, sy, sz, (int)k)))
            return s->fail(rc, "cannot write " + base + ".segy");
    }
    if (s->geti("snap", 0) && (rc = oseh_sim_snapshot(s, dir))) return rc;
    return OSE_OK;
}

void oseh_sim_destroy(oseh_sim* s)
{
    if (!s) return;
    if (s->ctx) ose_destroy(s->ctx);
    delete s;
}

// Writer of streamed snapshots (runs on the library's snapshot thread): per
// lane p-<step>-shot<k>.bin in the reference's [y][x][z] order for float32,
// else p-<step>-shot<k>.osnap with the encoded lane.
static void sim_snapshot_cb(void* user, const ose_snapshot_info* in, const void* data)
{
    auto* s = static_cast<oseh_sim*>(user);
    const int nl = (in->lanes & 1) + ((in->lanes >> 1) & 1);
    std::vector<char> lane(oseh_snapshot_lane_bytes(in->format, in->nx, in->ny, in->nz));
    int l = 0;
    for (int b = 0; b < 2; ++b) {
        if (!((in->lanes >> b) & 1)) continue;
        const std::string base = s->snap_dir + "/" + (b == 0 ? "p-" : "q-") + std::to_string(in->step) + "-shot" +
                                 std::to_string(in->shot_id);
        int rc = oseh_snapshot_lane(in->format, nl, l++, in->nx, in->ny, in->nz, data, lane.data());
        if (!rc && in->format == OSE_SNAP_F32) {
            rc = oseh_write_snapshot((base + ".bin").c_str(), (const float*)lane.data(), in->nx, in->ny, in->nz);
        } else if (!rc) {
            const int box[6] = {in->x0, in->y0, in->z0, in->nx, in->ny, in->nz};
            rc = oseh_write_osnap((base + ".osnap").c_str(), in->format, b, in->step, in->shot_id, box, lane.data());
        }
        if (rc) {
            std::lock_guard<std::mutex> lk(s->snap_mu);
            if (s->snap_err.empty()) s->snap_err = "cannot write snapshot " + base;
        }
    }
}

int oseh_sim_stream_snapshots(oseh_sim* s, const char* dir, int every)
{
    if (!s || every < 0 || (every > 0 && !dir)) return OSE_EINVAL;
    if (!s->ctx) return s->fail(OSE_ESTATE, "not set up");
    const std::string f = s->get("snap_format", "f32"), r = s->get("snap_region", "padded"),
                      l = s->get("snap_lanes", "pq");
    const int format = f == "f32" ? OSE_SNAP_F32 : f == "bf16" ? OSE_SNAP_BF16 : f == "q16" ? OSE_SNAP_Q16
                                                                 : f == "q8" ? OSE_SNAP_Q8 : -1;
    const int lanes = l == "pq" ? 3 : l == "p" ? 1 : l == "q" ? 2 : 0;
    const int interior = r == "interior" ? 1 : r == "padded" ? 0 : -1;
    if (format < 0 || lanes == 0 || interior < 0)
        return s->fail(OSE_EINVAL, "snap_format must be f32|bf16|q16|q8, snap_lanes pq|p|q, snap_region padded|interior");
    int rc;
    if ((rc = ose_set_snapshots(s->ctx, 0, 0, 0, 0, nullptr, nullptr)))
        return s->fail(rc, std::string("ose_set_snapshots: ") + ose_last_error(s->ctx));
    s->snap_dir = every > 0 ? dir : "";
    if (every > 0 && (rc = ose_set_snapshots(s->ctx, every, lanes, format, interior, sim_snapshot_cb, s)))
        return s->fail(rc, std::string("ose_set_snapshots: ") + ose_last_error(s->ctx));
    return snap_error(s);
}

int oseh_sim_snapshot(oseh_sim* s, const char* dir)
{
    if (!s || !dir) return OSE_EINVAL;
    if (!s->ctx) return s->fail(OSE_ESTATE, "not set up");
    int rc;
    for (size_t k = 0; k < s->shots.size(); ++k) {
        const BuiltModel& m = s->models[s->shots[k].model];
        const size_t n = (size_t)m.nxb * m.nyb * m.nzb;
        std::vector<float> re(n), im(n);
        if ((rc = ose_get_wavefield(s->ctx, (int)k, re.data(), im.data())))
            return s->fail(rc, std::string("ose_get_wavefield: ") + ose_last_error(s->ctx));
        // ./wf/{p,q}-<it>-single.bin of ose_3d.cu:261-269, one pair per shot
        const std::string tag = std::to_string(s->steps_done) + "-shot" + std::to_string(k) + ".bin";
        if ((rc = oseh_write_snapshot((std::string(dir) + "/p-" + tag).c_str(), re.data(), m.nxb, m.nyb, m.nzb)) ||
            (rc = oseh_write_snapshot((std::string(dir) + "/q-" + tag).c_str(), im.data(), m.nxb, m.nyb, m.nzb)))
            return s->fail(rc, "cannot write snapshot");
    }
    return OSE_OK;
}

// Checkpoint: ckpt.txt (step, shot count, per shot the padded extents and
// receiver count), and per shot the wavefield (re, im: float32 [z][y][x])
// and the traces recorded so far ([nrec][step] float32).
// FNV-1a (64-bit) over the bytes that fix a run: every model's grid, pads,
// velocity, order, coefficients and tapers, every shot's source cell and
// receiver count, the wavelets, the spacing and the time step.  A checkpoint
// records it and a restore refuses a checkpoint of any other configuration.
static uint64_t run_fingerprint(const oseh_sim* s)
{
    uint64_t h = 1469598103934665603ULL;
    auto mix = [&h](const void* p, size_t n) {
        const unsigned char* c = static_cast<const unsigned char*>(p);
        for (size_t i = 0; i < n; ++i) h = (h ^ c[i]) * 1099511628211ULL;
    };
    for (const BuiltModel& m : s->models) {
        const int dims[9] = {m.bl, m.br, m.bf, m.bb, m.bu, m.bd, m.nxb, m.nyb, m.nzb};
        mix(dims, sizeof dims);
        mix(m.v.data(), m.v.size() * sizeof(float));
        mix(&m.K, sizeof m.K);
        mix(m.coef.data(), m.coef.size() * sizeof(double));
        mix(m.gx.data(), m.gx.size() * sizeof(float));
        mix(m.gy.data(), m.gy.size() * sizeof(float));
        mix(m.gz.data(), m.gz.size() * sizeof(float));
    }
    for (size_t k = 0; k < s->shots.size(); ++k) {
        mix(s->shots[k].src, sizeof s->shots[k].src);
        const size_t nrec = receiver_cells(s, k).size();
        mix(&nrec, sizeof nrec);
    }
    mix(s->s.data(), s->s.size() * sizeof(float));
    mix(s->h.data(), s->h.size() * sizeof(float));
    const double dd[4] = {s->getd("dx", 10), s->getd("dy", 10), s->getd("dz", 10), s->getd("dt", 1e-3)};
    mix(dd, sizeof dd);
    return h;
}

static std::string ckpt_base(const std::string& dir, int step, size_t k)
{
    return dir + "/ckpt-s" + std::to_string(step) + "-shot" + std::to_string(k);
}

// Writes the state of every shot after `step` steps into files named for the
// step, then replaces ckpt.txt through a rename (atomic within a
// directory), then removes the files of the checkpoint it replaced: a crash
// at any point leaves a ckpt.txt whose files are all present.
int oseh_sim_checkpoint(oseh_sim* s, const char* dir)
{
    if (!s || !dir) return OSE_EINVAL;
    if (!s->ctx) return s->fail(OSE_ESTATE, "not set up");
    int rc, step = 0;
    if ((rc = ose_get_step(s->ctx, &step))) return s->fail(rc, std::string("ose_get_step: ") + ose_last_error(s->ctx));
    const std::string d(dir);
    int old_step = -1;
    {
        std::ifstream f(d + "/ckpt.txt");
        std::string tag, key;
        int version = 0;
        if (f >> tag >> version >> key >> old_step && !(tag == "ose-checkpoint" && version == 2 && key == "step"))
            old_step = -1;
    }
    std::ostringstream meta;
    char fp[32];
    std::snprintf(fp, sizeof fp, "%016llx", (unsigned long long)run_fingerprint(s));
    meta << "ose-checkpoint 2\nstep " << step << "\nfingerprint " << fp << "\nshots " << s->shots.size() << "\n";
    for (size_t k = 0; k < s->shots.size(); ++k) {
        const BuiltModel& m = s->models[s->shots[k].model];
        const size_t n = (size_t)m.nxb * m.nyb * m.nzb;
        const int nrec = (int)(receiver_cells(s, k).size() / 3);
        meta << "shot " << k << " " << m.nxb << " " << m.nyb << " " << m.nzb << " " << nrec << "\n";
        std::vector<float> re(n), im(n), tr((size_t)nrec * step);
        if ((rc = ose_get_wavefield(s->ctx, (int)k, re.data(), im.data())) ||
            (step > 0 && (rc = ose_get_traces(s->ctx, (int)k, 0, step, tr.data()))))
            return s->fail(rc, std::string("checkpoint read-back: ") + ose_last_error(s->ctx));
        const std::string base = ckpt_base(d, step, k);
        if ((rc = oseh_write_raw((base + "-re.bin").c_str(), re.data(), n)) ||
            (rc = oseh_write_raw((base + "-im.bin").c_str(), im.data(), n)) ||
            (rc = oseh_write_raw((base + "-traces.bin").c_str(), tr.data(), tr.size())))
            return s->fail(rc, "cannot write checkpoint " + base);
    }
    // the learned pad's state per model: whether its guard has tripped
    for (size_t mi = 0; mi < s->models.size(); ++mi) {
        int active = 0, tripped = 0;
        if ((rc = ose_boundary_status(s->ctx, s->models[mi].id, &active, &tripped)))
            return s->fail(rc, std::string("ose_boundary_status: ") + ose_last_error(s->ctx));
        meta << "bnet " << mi << " " << active << " " << tripped << "\n";
    }
    {
        std::ofstream f(d + "/ckpt.txt.tmp");
        f << meta.str();
        f.flush();
        if (!f) return s->fail(OSE_EINVAL, "cannot write " + d + "/ckpt.txt.tmp");
    }
    if (std::rename((d + "/ckpt.txt.tmp").c_str(), (d + "/ckpt.txt").c_str()))
        return s->fail(OSE_EINVAL, "cannot replace " + d + "/ckpt.txt");
    if (old_step >= 0 && old_step != step)
        for (size_t k = 0; k < s->shots.size(); ++k)
            for (const char* suf : {"-re.bin", "-im.bin", "-traces.bin"})
                std::remove((ckpt_base(d, old_step, k) + suf).c_str());
    return OSE_OK;
}

static int read_floats(const std::string& path, std::vector<float>& v)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) return OSE_EINVAL;
    f.read(reinterpret_cast<char*>(v.data()), (std::streamsize)(v.size() * sizeof(float)));
    return (size_t)f.gcount() == v.size() * sizeof(float) && f.peek() == EOF ? OSE_OK : OSE_EINVAL;
}

int oseh_sim_restore(oseh_sim* s, const char* dir)
{
    if (!s || !dir) return OSE_EINVAL;
    if (!s->ctx) return s->fail(OSE_ESTATE, "not set up");
    const std::string d(dir);
    std::ifstream f(d + "/ckpt.txt");
    std::string tag, key, fp;
    int version = 0, step = -1;
    size_t nshots = 0;
    if (!(f >> tag >> version) || tag != "ose-checkpoint" || version != 2 || !(f >> key >> step) || key != "step" ||
        !(f >> key >> fp) || key != "fingerprint" || !(f >> key >> nshots) || key != "shots")
        return s->fail(OSE_EINVAL, "not a checkpoint: " + d + "/ckpt.txt");
    char want[32];
    std::snprintf(want, sizeof want, "%016llx", (unsigned long long)run_fingerprint(s));
    if (fp != want)
        return s->fail(OSE_EINVAL, "checkpoint " + d + " was written by another configuration (model, order, "
                                   "coefficients, tapers, sources, receivers, wavelets or spacing differ)");
    if (nshots != s->shots.size() || step < 0 || step > s->geti("nt", 300))
        return s->fail(OSE_EINVAL, "checkpoint does not match this configuration");
    int rc;
    for (size_t k = 0; k < nshots; ++k) {
        const BuiltModel& m = s->models[s->shots[k].model];
        size_t idx = 0;
        int nxb = 0, nyb = 0, nzb = 0, nrec = 0;
        if (!(f >> key >> idx >> nxb >> nyb >> nzb >> nrec) || key != "shot" || idx != k || nxb != m.nxb ||
            nyb != m.nyb || nzb != m.nzb || (size_t)nrec * 3 != receiver_cells(s, k).size())
            return s->fail(OSE_EINVAL, "checkpoint shot " + std::to_string(k) + " does not match this configuration");
        const size_t n = (size_t)nxb * nyb * nzb;
        std::vector<float> re(n), im(n), tr((size_t)nrec * step);
        const std::string base = ckpt_base(d, step, k);
        if (read_floats(base + "-re.bin", re) || read_floats(base + "-im.bin", im) ||
            read_floats(base + "-traces.bin", tr))
            return s->fail(OSE_EINVAL, "cannot read checkpoint " + base);
        if ((rc = ose_set_wavefield(s->ctx, (int)k, re.data(), im.data())) ||
            (step > 0 && (rc = ose_set_traces(s->ctx, (int)k, 0, step, tr.data()))))
            return s->fail(rc, std::string("checkpoint upload: ") + ose_last_error(s->ctx));
    }
    // a learned pad whose guard had tripped continues on the plain taper
    for (size_t mi = 0; mi < s->models.size(); ++mi) {
        size_t idx = 0;
        int active = 0, tripped = 0;
        if (!(f >> key >> idx >> active >> tripped) || key != "bnet" || idx != mi)
            return s->fail(OSE_EINVAL, "checkpoint lacks the boundary-network state of model " + std::to_string(mi));
        int now_active = 0, now_tripped = 0;
        if ((rc = ose_boundary_status(s->ctx, s->models[mi].id, &now_active, &now_tripped)))
            return s->fail(rc, std::string("ose_boundary_status: ") + ose_last_error(s->ctx));
        if (tripped && !now_tripped && (now_active || active) &&
            (rc = ose_trip_boundary_net(s->ctx, s->models[mi].id)))
            return s->fail(rc, std::string("ose_trip_boundary_net: ") + ose_last_error(s->ctx));
    }
    if ((rc = ose_set_step(s->ctx, step))) return s->fail(rc, std::string("ose_set_step: ") + ose_last_error(s->ctx));
    s->steps_done = step;
    return OSE_OK;
}

int oseh_sim_nmodels(oseh_sim* s) { return s ? (int)s->models.size() : 0; }

int oseh_sim_model_dims(oseh_sim* s, int mi, int* d)
{
    if (!s || !d || mi < 0 || mi >= (int)s->models.size()) return OSE_EINVAL;
    const BuiltModel& m = s->models[mi];
    const int v[12] = {m.spec.nx, m.spec.ny, m.spec.nz, m.bl, m.br, m.bf, m.bb, m.bu, m.bd, m.nxb, m.nyb, m.nzb};
    std::memcpy(d, v, sizeof v);
    return OSE_OK;
}

int oseh_sim_model_velocity(oseh_sim* s, int mi, float* v)
{
    if (!s || !v || mi < 0 || mi >= (int)s->models.size()) return OSE_EINVAL;
    std::copy(s->models[mi].v.begin(), s->models[mi].v.end(), v);
    return OSE_OK;
}

int oseh_sim_model_params(oseh_sim* s, int mi, double* R, int* K, double* coef, float* gx, float* gy,
                          float* gz, int* rec_z)
{
    if (!s || mi < 0 || mi >= (int)s->models.size()) return OSE_EINVAL;
    const BuiltModel& m = s->models[mi];
    if (R) *R = m.R;
    if (K) *K = m.K;
    if (coef) std::copy(m.coef.begin(), m.coef.end(), coef);
    if (gx) std::copy(m.gx.begin(), m.gx.end(), gx);
    if (gy) std::copy(m.gy.begin(), m.gy.end(), gy);
    if (gz) std::copy(m.gz.begin(), m.gz.end(), gz);
    if (rec_z) *rec_z = s->geti("rec_z", 0);
    return OSE_OK;
}

int oseh_sim_nshots(oseh_sim* s) { return s ? (int)s->shots.size() : 0; }

int oseh_sim_shot(oseh_sim* s, int k, int* model, int* src, int* nt, float* sw, float* hw)
{
    if (!s || k < 0 || k >= (int)s->shots.size()) return OSE_EINVAL;
    if (model) *model = s->shots[k].model;
    if (src) std::memcpy(src, s->shots[k].src, sizeof(int) * 3);
    if (nt) *nt = (int)s->s.size();
    if (sw) std::copy(s->s.begin(), s->s.end(), sw);
    if (hw) std::copy(s->h.begin(), s->h.end(), hw);
    return OSE_OK;
}

}  // extern "C"
