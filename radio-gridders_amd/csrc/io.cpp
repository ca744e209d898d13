// io.cpp -- visibility / W-kernel files in the reference's HDF5 layouts and
// the file-level wrappers (wproj_CUDA_flat & co., grid_common.cuh:45-64;
// image_dft_host_flat, dft_common.cuh:12-18).
//
// libhdf5 (>= 1.10, 64-bit hid_t) is loaded at run time with dlopen, so the
// library itself has no link-time HDF5 dependency.  Layouts:
//   visibilities  vis/<a1>/<a2>/{frequency[F], time[T], uvw[T][3], vis[T][F][1]}
//                 (load_vis / load_vis_group, src/common/hdf5.c:64-369), or the
//                 legacy flat group with attribute type="Visibility" and
//                 antenna1/antenna2 datasets (hdf5.c:200-290);
//   W-kernels     wkern/<theta %g>/<w>/kern[R][R][S][S] (load_wkern_CUDA,
//                 hdf5.c:1362-1520);
//   complex       compound {r: f64, i: f64} (init_dtype_cpx, hdf5.c:21-31).
#include <dlfcn.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "rg_internal.hpp"

namespace rg {

namespace h5 {

typedef int64_t hid_t;
typedef int herr_t;
typedef unsigned long long hsize_t;

struct H5G_info_t {
    int storage_type;
    hsize_t nlinks;
    int64_t max_corder;
    int mounted; // hbool_t
};

struct Api {
    bool ok = false;
    void *lib = nullptr;
    herr_t (*open)(void);
    herr_t (*get_libversion)(unsigned *, unsigned *, unsigned *);
    herr_t (*eset_auto2)(hid_t, void *, void *);
    hid_t (*fopen)(const char *, unsigned, hid_t);
    hid_t (*fcreate)(const char *, unsigned, hid_t, hid_t);
    herr_t (*fclose)(hid_t);
    hid_t (*gopen2)(hid_t, const char *, hid_t);
    hid_t (*gcreate2)(hid_t, const char *, hid_t, hid_t, hid_t);
    herr_t (*gclose)(hid_t);
    herr_t (*gget_info)(hid_t, H5G_info_t *);
    long (*lget_name_by_idx)(hid_t, const char *, int, int, hsize_t, char *, size_t, hid_t);
    int (*lexists)(hid_t, const char *, hid_t);
    hid_t (*dopen2)(hid_t, const char *, hid_t);
    hid_t (*dcreate2)(hid_t, const char *, hid_t, hid_t, hid_t, hid_t, hid_t);
    herr_t (*dclose)(hid_t);
    hid_t (*dget_space)(hid_t);
    hid_t (*dget_type)(hid_t);
    herr_t (*dread)(hid_t, hid_t, hid_t, hid_t, hid_t, void *);
    herr_t (*dwrite)(hid_t, hid_t, hid_t, hid_t, hid_t, const void *);
    int (*sget_simple_extent_ndims)(hid_t);
    int (*sget_simple_extent_dims)(hid_t, hsize_t *, hsize_t *);
    hid_t (*screate_simple)(int, const hsize_t *, const hsize_t *);
    herr_t (*sclose)(hid_t);
    size_t (*tget_size)(hid_t);
    int (*tget_class)(hid_t);
    hid_t (*tcreate)(int, size_t);
    herr_t (*tinsert)(hid_t, const char *, size_t, hid_t);
    herr_t (*tclose)(hid_t);
    int (*aexists)(hid_t, const char *);
    hid_t (*aopen)(hid_t, const char *, hid_t);
    hid_t (*aget_type)(hid_t);
    herr_t (*aread)(hid_t, hid_t, void *);
    herr_t (*aclose)(hid_t);
    hid_t (*pcreate)(hid_t);
    herr_t (*pset_create_intermediate_group)(hid_t, unsigned);
    herr_t (*pclose)(hid_t);
    hid_t *t_ieee_f64le = nullptr, *t_std_i64le = nullptr, *p_cls_link_create = nullptr;
    hid_t cpx = -1;
};

static Api g_api;
static bool g_tried = false;

template <typename F> static bool sym(void *lib, const char *name, F &out)
{
    out = reinterpret_cast<F>(dlsym(lib, name));
    return out != nullptr;
}

static Api *api()
{
    if (g_tried)
        return g_api.ok ? &g_api : nullptr;
    g_tried = true;
    const char *cands[] = {getenv("RG_HDF5_LIB"), "libhdf5.so", "libhdf5.so.103",
                           "/opt/conda/lib/libhdf5.so", "/usr/lib/x86_64-linux-gnu/libhdf5_serial.so"};
    void *lib = nullptr;
    for (const char *c : cands) {
        if (!c)
            continue;
        lib = dlopen(c, RTLD_NOW | RTLD_LOCAL);
        if (lib)
            break;
    }
    if (!lib)
        return nullptr;
    Api &a = g_api;
    a.lib = lib;
    bool ok = sym(lib, "H5open", a.open) && sym(lib, "H5get_libversion", a.get_libversion) &&
              sym(lib, "H5Eset_auto2", a.eset_auto2) && sym(lib, "H5Fopen", a.fopen) &&
              sym(lib, "H5Fcreate", a.fcreate) && sym(lib, "H5Fclose", a.fclose) &&
              sym(lib, "H5Gopen2", a.gopen2) && sym(lib, "H5Gcreate2", a.gcreate2) &&
              sym(lib, "H5Gclose", a.gclose) && sym(lib, "H5Gget_info", a.gget_info) &&
              sym(lib, "H5Lget_name_by_idx", a.lget_name_by_idx) &&
              sym(lib, "H5Lexists", a.lexists) && sym(lib, "H5Dopen2", a.dopen2) &&
              sym(lib, "H5Dcreate2", a.dcreate2) && sym(lib, "H5Dclose", a.dclose) &&
              sym(lib, "H5Dget_space", a.dget_space) && sym(lib, "H5Dget_type", a.dget_type) &&
              sym(lib, "H5Dread", a.dread) && sym(lib, "H5Dwrite", a.dwrite) &&
              sym(lib, "H5Sget_simple_extent_ndims", a.sget_simple_extent_ndims) &&
              sym(lib, "H5Sget_simple_extent_dims", a.sget_simple_extent_dims) &&
              sym(lib, "H5Screate_simple", a.screate_simple) && sym(lib, "H5Sclose", a.sclose) &&
              sym(lib, "H5Tget_size", a.tget_size) && sym(lib, "H5Tget_class", a.tget_class) &&
              sym(lib, "H5Tcreate", a.tcreate) && sym(lib, "H5Tinsert", a.tinsert) &&
              sym(lib, "H5Tclose", a.tclose) && sym(lib, "H5Aexists", a.aexists) &&
              sym(lib, "H5Aopen", a.aopen) && sym(lib, "H5Aget_type", a.aget_type) &&
              sym(lib, "H5Aread", a.aread) && sym(lib, "H5Aclose", a.aclose) &&
              sym(lib, "H5Pcreate", a.pcreate) &&
              sym(lib, "H5Pset_create_intermediate_group", a.pset_create_intermediate_group) &&
              sym(lib, "H5Pclose", a.pclose);
    a.t_ieee_f64le = (hid_t *)dlsym(lib, "H5T_IEEE_F64LE_g");
    a.t_std_i64le = (hid_t *)dlsym(lib, "H5T_STD_I64LE_g");
    a.p_cls_link_create = (hid_t *)dlsym(lib, "H5P_CLS_LINK_CREATE_ID_g");
    if (!ok || !a.t_ieee_f64le || !a.t_std_i64le || !a.p_cls_link_create)
        return nullptr;
    unsigned maj = 0, min = 0, rel = 0;
    if (a.open() < 0 || a.get_libversion(&maj, &min, &rel) < 0 || maj != 1 || min < 10)
        return nullptr;
    a.eset_auto2(0 /* H5E_DEFAULT */, nullptr, nullptr);
    // H5T_COMPOUND = 6
    a.cpx = a.tcreate(6, 16);
    a.tinsert(a.cpx, "r", 0, *a.t_ieee_f64le);
    a.tinsert(a.cpx, "i", 8, *a.t_ieee_f64le);
    a.ok = true;
    return &a;
}

constexpr unsigned ACC_RDONLY = 0x0000u, ACC_TRUNC = 0x0002u;
constexpr hid_t P_DEFAULT = 0, S_ALL = 0;
constexpr int T_STRING = 3;

} // namespace h5

// Host-side hierarchical visibilities (struct bl_data / vis_data, hdf5.cuh:41-65).
struct Baseline {
    int a1, a2, T, F;
    std::vector<double> time, freq, uvw, vis; // uvw [T][3] metres, vis [T][F] (re, im)
};

// load_vis_group (hdf5.c:64-143): shape checks and the baseline-length filter
// on the first uvw sample, in metres, [min_len, max_len).  Returns 1 when the
// group is kept, 0 when it is filtered out or holds no time samples (the
// reference reads the first uvw sample of an empty group out of bounds), and
// a negative RG_ERR_* code when its datasets are malformed or a read fails.
static int read_vis_group(h5::Api *a, h5::hid_t g, Baseline &bl, double min_len, double max_len,
                          bool filter, int64_t *total)
{
    using namespace h5;
    hid_t fds = a->dopen2(g, "frequency", P_DEFAULT), tds = a->dopen2(g, "time", P_DEFAULT);
    hid_t uds = a->dopen2(g, "uvw", P_DEFAULT), vds = a->dopen2(g, "vis", P_DEFAULT);
    int st = (fds >= 0 && tds >= 0 && uds >= 0 && vds >= 0) ? 1 : RG_ERR_SHAPE;
    hsize_t fdim = 0, tdim = 0, udim[2] = {0, 0}, vdim[3] = {0, 0, 0};
    auto shape = [&](hid_t ds, int nd, hsize_t *dims, size_t esz) {
        hid_t sp = a->dget_space(ds), ty = a->dget_type(ds);
        bool r = a->sget_simple_extent_ndims(sp) == nd && a->tget_size(ty) == esz &&
                 a->sget_simple_extent_dims(sp, dims, nullptr) >= 0;
        a->sclose(sp);
        a->tclose(ty);
        return r;
    };
    if (st == 1 &&
        !(shape(fds, 1, &fdim, 8) && shape(tds, 1, &tdim, 8) && shape(uds, 2, udim, 8) &&
          udim[0] == tdim && udim[1] == 3 && shape(vds, 3, vdim, 16) && vdim[0] == tdim &&
          vdim[1] == fdim && vdim[2] == 1))
        st = RG_ERR_SHAPE;
    if (st == 1 && (tdim == 0 || fdim == 0))
        st = 0; // nothing to grid
    if (st == 1) {
        if (total)
            *total += (int64_t)(tdim * fdim);
        bl.T = (int)tdim;
        bl.F = (int)fdim;
        bl.uvw.resize(tdim * 3);
        if (a->dread(uds, *a->t_ieee_f64le, S_ALL, S_ALL, P_DEFAULT, bl.uvw.data()) < 0) {
            st = RG_ERR_IO;
        } else {
            const double len = std::sqrt(bl.uvw[0] * bl.uvw[0] + bl.uvw[1] * bl.uvw[1]);
            if (filter && (len < min_len || len >= max_len)) {
                st = 0;
            } else {
                bl.time.resize(tdim);
                bl.freq.resize(fdim);
                bl.vis.resize(tdim * fdim * 2);
                if (a->dread(tds, *a->t_ieee_f64le, S_ALL, S_ALL, P_DEFAULT, bl.time.data()) < 0 ||
                    a->dread(fds, *a->t_ieee_f64le, S_ALL, S_ALL, P_DEFAULT, bl.freq.data()) < 0 ||
                    a->dread(vds, a->cpx, S_ALL, S_ALL, P_DEFAULT, bl.vis.data()) < 0)
                    st = RG_ERR_IO;
            }
        }
    }
    if (fds >= 0) a->dclose(fds);
    if (tds >= 0) a->dclose(tds);
    if (uds >= 0) a->dclose(uds);
    if (vds >= 0) a->dclose(vds);
    return st;
}

static int load_vis(const char *filename, double min_len, double max_len,
                    std::vector<Baseline> &out)
{
    using namespace h5;
    Api *a = api();
    if (!a)
        return fail(RG_ERR_UNSUPPORTED, "libhdf5 (>= 1.10) could not be loaded");
    hid_t f = a->fopen(filename, ACC_RDONLY, P_DEFAULT);
    if (f < 0)
        return fail(RG_ERR_IO, "could not open visibility file %s", filename);
    hid_t g = a->gopen2(f, "vis", P_DEFAULT);
    if (g < 0) {
        a->fclose(f);
        return fail(RG_ERR_IO, "could not open 'vis' group in visibility file %s", filename);
    }
    out.clear();
    int status = RG_OK;
    bool legacy = false;
    if (a->aexists(g, "type") > 0) {
        hid_t at = a->aopen(g, "type", P_DEFAULT);
        hid_t ty = at >= 0 ? a->aget_type(at) : -1;
        if (ty >= 0 && a->tget_class(ty) == T_STRING) {
            std::vector<char> buf(a->tget_size(ty) + 1, 0);
            char *vl = nullptr;
            // variable-length strings read as char*, fixed-length into the buffer
            if (a->tget_size(ty) == sizeof(char *)) {
                if (a->aread(at, ty, &vl) >= 0 && vl)
                    legacy = std::string(vl) == "Visibility";
            } else if (a->aread(at, ty, buf.data()) >= 0) {
                legacy = std::string(buf.data()) == "Visibility";
            }
        }
        if (ty >= 0) a->tclose(ty);
        if (at >= 0) a->aclose(at);
    }
    if (legacy) {
        // hdf5.c:200-290: one group of all visibilities, split into one
        // single-sample baseline per row, filter per row.
        Baseline data;
        int64_t total = 0;
        const int rs = read_vis_group(a, g, data, -DBL_MAX, DBL_MAX, false, &total);
        if (rs < 0) {
            status = fail(rs, "malformed or unreadable legacy visibility group in %s", filename);
        } else if (rs == 0) {
            // no rows: no visibilities
        } else {
            hid_t d1 = a->dopen2(g, "antenna1", P_DEFAULT), d2 = a->dopen2(g, "antenna2", P_DEFAULT);
            std::vector<int64_t> a1(data.T), a2(data.T);
            if (d1 < 0 || d2 < 0 ||
                a->dread(d1, *a->t_std_i64le, S_ALL, S_ALL, P_DEFAULT, a1.data()) < 0 ||
                a->dread(d2, *a->t_std_i64le, S_ALL, S_ALL, P_DEFAULT, a2.data()) < 0) {
                status = fail(RG_ERR_SHAPE, "antenna datasets missing in %s", filename);
            } else {
                for (int i = 0; i < data.T; ++i) {
                    const double *uvw = &data.uvw[3 * i];
                    const double len = std::sqrt(uvw[0] * uvw[0] + uvw[1] * uvw[1]);
                    if (len < min_len || len >= max_len)
                        continue;
                    Baseline b;
                    b.a1 = (int)a1[i];
                    b.a2 = (int)a2[i];
                    b.T = 1;
                    b.F = data.F;
                    b.time = {data.time[i]};
                    b.freq = data.freq;
                    b.uvw = {uvw[0], uvw[1], uvw[2]};
                    b.vis.assign(data.vis.begin() + (size_t)i * data.F * 2,
                                 data.vis.begin() + (size_t)(i + 1) * data.F * 2);
                    out.push_back(std::move(b));
                }
            }
            if (d1 >= 0) a->dclose(d1);
            if (d2 >= 0) a->dclose(d2);
        }
    } else {
        // hdf5.c:292-340: antenna count = number of children + 1, groups a1/a2
        H5G_info_t info;
        if (a->gget_info(g, &info) < 0 || info.nlinks == 0) {
            status = fail(RG_ERR_IO, "found no antenna data in visibility file %s", filename);
        } else {
            const int nant = (int)info.nlinks + 1;
            for (int a1 = 0; a1 < nant - 1 && status == RG_OK; ++a1) {
                char n1[32];
                snprintf(n1, sizeof(n1), "%d", a1);
                if (a->lexists(g, n1, P_DEFAULT) <= 0)
                    continue;
                hid_t g1 = a->gopen2(g, n1, P_DEFAULT);
                if (g1 < 0)
                    continue;
                for (int a2 = a1 + 1; a2 < nant; ++a2) {
                    char n2[32];
                    snprintf(n2, sizeof(n2), "%d", a2);
                    if (a->lexists(g1, n2, P_DEFAULT) <= 0)
                        continue;
                    hid_t g2 = a->gopen2(g1, n2, P_DEFAULT);
                    if (g2 < 0)
                        continue;
                    Baseline b;
                    b.a1 = a1;
                    b.a2 = a2;
                    // a group of the wrong shape is skipped as the reference
                    // does (hdf5.c:88-95); a failed read is an error
                    const int rs = read_vis_group(a, g2, b, min_len, max_len, true, nullptr);
                    if (rs > 0)
                        out.push_back(std::move(b));
                    else if (rs < 0 && rs != RG_ERR_SHAPE)
                        status = fail(rs, "reading baseline %d-%d of %s failed", a1, a2,
                                      filename);
                    a->gclose(g2);
                }
                a->gclose(g1);
            }
        }
    }
    a->gclose(g);
    a->fclose(f);
    return status;
}

// flatten_visibilities_CUDA (hdf5.c:827-872) with the index written before it
// is advanced: u = uvw * f / c, order baseline -> time -> frequency.
static void flatten(const std::vector<Baseline> &bls, std::vector<double> &u, std::vector<double> &v,
                    std::vector<double> &w, std::vector<double> &vis)
{
    const double c = 299792458.0;
    size_t n = 0;
    for (const auto &b : bls)
        n += (size_t)b.T * b.F;
    u.resize(n);
    v.resize(n);
    w.resize(n);
    vis.resize(2 * n);
    size_t k = 0;
    for (const auto &b : bls)
        for (int t = 0; t < b.T; ++t)
            for (int f = 0; f < b.F; ++f) {
                u[k] = b.uvw[3 * t + 0] * b.freq[f] / c;
                v[k] = b.uvw[3 * t + 1] * b.freq[f] / c;
                w[k] = b.uvw[3 * t + 2] * b.freq[f] / c;
                vis[2 * k] = b.vis[2 * ((size_t)t * b.F + f)];
                vis[2 * k + 1] = b.vis[2 * ((size_t)t * b.F + f) + 1];
                ++k;
            }
}

struct HostKern {
    int P = 0, R = 0, S = 0;
    double w_min = 0, w_step = 0;
    std::vector<double> data; // [P][R][R][S][S] complex, indexed by w
};

// load_wkern_CUDA (hdf5.c:1362-1520): planes named by their w value, w_min and
// w_max initialised from the names, w_step = (w_max - w_min) / (P - 1), plane i
// is the kernel whose w is nearest to w_min + i * w_step.
static int load_wkern(const char *filename, double theta, HostKern &k)
{
    using namespace h5;
    Api *a = api();
    if (!a)
        return fail(RG_ERR_UNSUPPORTED, "libhdf5 (>= 1.10) could not be loaded");
    hid_t f = a->fopen(filename, ACC_RDONLY, P_DEFAULT);
    if (f < 0)
        return fail(RG_ERR_IO, "could not open w kernel file %s", filename);
    char gname[64];
    snprintf(gname, sizeof(gname), "wkern/%g", theta);
    hid_t g = a->gopen2(f, gname, P_DEFAULT);
    if (g < 0) {
        a->fclose(f);
        return fail(RG_ERR_IO, "could not open '%s' group in w kernel file %s", gname, filename);
    }
    H5G_info_t info;
    int status = RG_OK;
    if (a->gget_info(g, &info) < 0 || info.nlinks == 0)
        status = fail(RG_ERR_IO, "found no w-kernels in w-kernel file %s", filename);
    std::vector<double> ws;
    std::vector<std::vector<double>> planes;
    for (hsize_t i = 0; status == RG_OK && i < info.nlinks; ++i) {
        char name[256];
        // H5_INDEX_NAME = 0, H5_ITER_INC = 0
        if (a->lget_name_by_idx(g, ".", 0, 0, i, name, sizeof(name), P_DEFAULT) < 0) {
            status = fail(RG_ERR_IO, "cannot list w-kernel planes in %s", filename);
            break;
        }
        const double w = atof(name);
        std::string dn = std::string(name) + "/kern";
        hid_t ds = a->dopen2(g, dn.c_str(), P_DEFAULT);
        if (ds < 0) {
            status = fail(RG_ERR_SHAPE, "missing dataset %s/%s", gname, dn.c_str());
            break;
        }
        hid_t sp = a->dget_space(ds), ty = a->dget_type(ds);
        hsize_t dims[4] = {0, 0, 0, 0};
        const bool fmt = a->sget_simple_extent_ndims(sp) == 4 && a->tget_size(ty) == 16 &&
                         a->sget_simple_extent_dims(sp, dims, nullptr) >= 0;
        a->sclose(sp);
        a->tclose(ty);
        if (fmt) {
            if (k.R == 0) {
                k.R = (int)dims[0];
                k.S = (int)dims[3];
                if (dims[2] != dims[3])
                    status = fail(RG_ERR_SHAPE, "non-square W-kernels are not supported");
            }
            if ((int)dims[0] == k.R && (int)dims[1] == k.R && (int)dims[2] == k.S &&
                (int)dims[3] == k.S) {
                std::vector<double> d((size_t)k.R * k.R * k.S * k.S * 2);
                if (a->dread(ds, a->cpx, S_ALL, S_ALL, P_DEFAULT, d.data()) < 0) {
                    status = fail(RG_ERR_IO, "cannot read %s/%s", gname, dn.c_str());
                    a->dclose(ds);
                    break;
                }
                ws.push_back(w);
                planes.push_back(std::move(d));
            } else {
                fprintf(stderr, "kernel %s has inconsistent dimensions - ignored!\n", dn.c_str());
            }
        }
        a->dclose(ds);
    }
    a->gclose(g);
    a->fclose(f);
    if (status != RG_OK)
        return status;
    if (planes.empty() || k.R <= 0 || k.S <= 0)
        return fail(RG_ERR_SHAPE, "invalid dimensions in w-kernel file %s", filename);
    k.P = (int)planes.size();
    k.w_min = *std::min_element(ws.begin(), ws.end());
    const double w_max = *std::max_element(ws.begin(), ws.end());
    k.w_step = k.P > 1 ? (w_max - k.w_min) / (k.P - 1) : 1.0;
    const size_t pe = planes[0].size();
    k.data.resize(pe * k.P);
    for (int i = 0; i < k.P; ++i) {
        const double w = k.w_min + i * k.w_step;
        int best = 0;
        for (int j = 1; j < k.P; ++j)
            if (std::fabs(ws[j] - w) < std::fabs(ws[best] - w))
                best = j;
        std::copy(planes[best].begin(), planes[best].end(), k.data.begin() + pe * i);
    }
    return RG_OK;
}

// ---- device staging for the file-level wrappers ------------------------------------------

struct DeviceInputs {
    DevBuf u, v, w, vis, kern;
    rg_flat_vis fv{};
    rg_wkern wk{};
    ~DeviceInputs()
    {
        u.release();
        v.release();
        w.release();
        vis.release();
        kern.release();
    }
};

static int upload_vis(const std::vector<double> &u, const std::vector<double> &v,
                      const std::vector<double> &w, const std::vector<double> &vis,
                      DeviceInputs &d)
{
    const size_t n = u.size();
    RG_TRY(d.u.ensure(8 * n + 8));
    RG_TRY(d.v.ensure(8 * n + 8));
    RG_TRY(d.w.ensure(8 * n + 8));
    RG_TRY(d.vis.ensure(16 * n + 16));
    RG_HIP(hipMemcpy(d.u.p, u.data(), 8 * n, hipMemcpyHostToDevice));
    RG_HIP(hipMemcpy(d.v.p, v.data(), 8 * n, hipMemcpyHostToDevice));
    RG_HIP(hipMemcpy(d.w.p, w.data(), 8 * n, hipMemcpyHostToDevice));
    RG_HIP(hipMemcpy(d.vis.p, vis.data(), 16 * n, hipMemcpyHostToDevice));
    d.fv.u = d.u.as<double>();
    d.fv.v = d.v.as<double>();
    d.fv.w = d.w.as<double>();
    d.fv.vis = d.vis.p;
    d.fv.n = (int64_t)n;
    d.fv.vis_type = RG_C128;
    return RG_OK;
}

static int upload_kern(const HostKern &k, DeviceInputs &d)
{
    RG_TRY(d.kern.ensure(8 * k.data.size()));
    RG_HIP(hipMemcpy(d.kern.p, k.data.data(), 8 * k.data.size(), hipMemcpyHostToDevice));
    d.wk.data = d.kern.p;
    d.wk.type = RG_C128;
    d.wk.plane_count = k.P;
    d.wk.oversampling = k.R;
    d.wk.size = k.S;
    d.wk.w_min = k.w_min;
    d.wk.w_step = k.w_step;
    return RG_OK;
}

static int32_t g_file_mode = RG_MODE_BATCHED;
static int32_t g_image_conv = RG_IMAGE_GPU;

static int file_grid(const char *visfile, const char *wkernfile, void *grid, int grid_size,
                     double theta, double bl_min, double bl_max, int subgrid, int margin,
                     double winc)
{
    if (!visfile || !wkernfile || !grid)
        return fail(RG_ERR_PARAM, "NULL argument");
    std::vector<Baseline> bls;
    RG_TRY(load_vis(visfile, bl_min, bl_max, bls));
    HostKern hk;
    RG_TRY(load_wkern(wkernfile, theta, hk));
    std::vector<double> u, v, w, vis;
    flatten(bls, u, v, w, vis);
    int dev = 0;
    RG_HIP(hipGetDevice(&dev));
    rg_context *ctx = nullptr;
    RG_TRY(rg_context_create(&ctx, dev));
    int st;
    {
        DeviceInputs d;
        st = upload_vis(u, v, w, vis, d);
        if (st == RG_OK)
            st = upload_kern(hk, d);
        if (st == RG_OK) {
            rg_grid_params p{};
            p.grid_size = grid_size;
            p.theta = theta;
            p.mode = g_file_mode;
            p.subgrid_size = subgrid;
            p.subgrid_margin = margin;
            p.winc = winc;
            p.compensated = 1;
            if (g_image_conv == RG_IMAGE_CPU) // weight before gridding (main.c:151-154)
                st = rg_weight(ctx, d.fv.u, d.fv.v, d.fv.n, d.vis.p, RG_C128, grid_size, theta,
                               nullptr, nullptr);
            if (st == RG_OK)
                st = rg_grid(ctx, &p, &d.fv, &d.wk, grid, RG_C128, nullptr);
        }
        if (st == RG_OK) {
            rg_grid_stats gs;
            st = rg_context_stats(ctx, &gs);
            if (st == RG_OK && gs.n_bad_plane > 0)
                st = fail(RG_ERR_RANGE, "%lld visibilities fall outside the W-kernel table's w range",
                          (long long)gs.n_bad_plane);
        }
        if (st == RG_OK) {
            if (g_image_conv == RG_IMAGE_CPU) {
                // make_hermitian, shift, forward FFT, shift (main.c:258-281)
                st = rg_make_hermitian(grid, RG_C128, grid_size, nullptr);
                if (st == RG_OK)
                    st = rg_fft_shift(grid, RG_C128, grid_size, nullptr);
                if (st == RG_OK)
                    st = rg_fft2d(ctx, grid, RG_C128, grid_size, 1, -1, nullptr);
                if (st == RG_OK)
                    st = rg_fft_shift(grid, RG_C128, grid_size, nullptr);
            } else {
                st = rg_grid_to_image(ctx, grid, RG_C128, grid_size, nullptr);
            }
        }
        if (st == RG_OK) {
            hipError_t e = hipDeviceSynchronize();
            if (e != hipSuccess)
                st = fail(RG_ERR_HIP, "%s", hipGetErrorString(e));
        }
    }
    rg_context_destroy(ctx);
    return st;
}

static int file_dft(const char *visfile, void *grid_host, void *grid_dev, int grid_size,
                    double lambda, double bl_min, double bl_max, int blocks, int threads)
{
    if (!visfile || !grid_host || !grid_dev)
        return fail(RG_ERR_PARAM, "NULL argument");
    std::vector<Baseline> bls;
    RG_TRY(load_vis(visfile, bl_min, bl_max, bls));
    std::vector<double> u, v, w, vis;
    flatten(bls, u, v, w, vis);
    int dev = 0;
    RG_HIP(hipGetDevice(&dev));
    rg_context *ctx = nullptr;
    RG_TRY(rg_context_create(&ctx, dev));
    int st;
    {
        DeviceInputs d;
        st = upload_vis(u, v, w, vis, d);
        if (st == RG_OK)
            st = rg_dft_image(ctx, &d.fv, grid_size, lambda, grid_dev, blocks, threads,
                              RG_DFT_FP64, nullptr);
        if (st == RG_OK) {
            hipError_t e = hipMemcpy(grid_host, grid_dev, 16 * (size_t)grid_size * grid_size,
                                     hipMemcpyDeviceToHost);
            if (e != hipSuccess)
                st = fail(RG_ERR_HIP, "%s", hipGetErrorString(e));
        }
    }
    rg_context_destroy(ctx);
    return st;
}

} // namespace rg

using namespace rg;

extern "C" {

int rg_set_file_mode(int32_t mode)
{
    if (mode != RG_MODE_PLAIN && mode != RG_MODE_BATCHED)
        return fail(RG_ERR_PARAM, "unknown mode %d", mode);
    g_file_mode = mode;
    return RG_OK;
}

int rg_set_image_convention(int32_t convention)
{
    if (convention != RG_IMAGE_GPU && convention != RG_IMAGE_CPU)
        return fail(RG_ERR_PARAM, "unknown image convention %d", convention);
    g_image_conv = convention;
    return RG_OK;
}

int rg_hdf5_available(void) { return h5::api() ? 1 : 0; }

int rg_wtowers_flat(const char *visfile, const char *wkernfile, void *grid, void *gridh,
                    int grid_size, double theta, double lambda, double bl_min, double bl_max,
                    int subgrid_size, int subgrid_margin, double winc)
{
    (void)gridh;
    (void)lambda;
    if (subgrid_size <= 0)
        return fail(RG_ERR_PARAM, "W-towers needs a positive subgrid size");
    return file_grid(visfile, wkernfile, grid, grid_size, theta, bl_min, bl_max, subgrid_size,
                     subgrid_margin, winc);
}

int rg_wtowers(const char *visfile, const char *wkernfile, void *grid, void *gridh, int grid_size,
               double theta, double lambda, double bl_min, double bl_max, int subgrid_size,
               int subgrid_margin, double winc)
{
    return rg_wtowers_flat(visfile, wkernfile, grid, gridh, grid_size, theta, lambda, bl_min,
                           bl_max, subgrid_size, subgrid_margin, winc);
}

int rg_wproj_flat(const char *visfile, const char *wkernfile, void *grid, void *gridh,
                  int grid_size, double theta, double lambda, double bl_min, double bl_max,
                  int threads_per_block)
{
    (void)gridh;
    (void)lambda;
    (void)threads_per_block;
    return file_grid(visfile, wkernfile, grid, grid_size, theta, bl_min, bl_max, 0, 0, 0.0);
}

int rg_wproj(const char *visfile, const char *wkernfile, void *grid, void *gridh, int grid_size,
             double theta, double lambda, double bl_min, double bl_max, int threads_per_block)
{
    return rg_wproj_flat(visfile, wkernfile, grid, gridh, grid_size, theta, lambda, bl_min, bl_max,
                         threads_per_block);
}

int rg_image_dft_host_flat(const char *visfile, void *grid_host, void *grid_dev, int grid_size,
                           double theta, double lambda, double bl_min, double bl_max, int blocks,
                           int threads_block)
{
    (void)theta;
    return file_dft(visfile, grid_host, grid_dev, grid_size, lambda, bl_min, bl_max, blocks,
                    threads_block);
}

int rg_image_dft_host(const char *visfile, void *grid_host, void *grid_dev, int grid_size,
                      double theta, double lambda, double bl_min, double bl_max, int blocks,
                      int threads_block)
{
    return rg_image_dft_host_flat(visfile, grid_host, grid_dev, grid_size, theta, lambda, bl_min,
                                  bl_max, blocks, threads_block);
}

int rg_load_flat_vis(const char *visfile, double bl_min, double bl_max, int64_t *n, double **u,
                     double **v, double **w, double **vis)
{
    if (!visfile || !n || !u || !v || !w || !vis)
        return fail(RG_ERR_PARAM, "NULL argument");
    std::vector<Baseline> bls;
    RG_TRY(load_vis(visfile, bl_min, bl_max, bls));
    std::vector<double> hu, hv, hw, hvis;
    flatten(bls, hu, hv, hw, hvis);
    *n = (int64_t)hu.size();
    auto dup = [](const std::vector<double> &x) {
        double *p = (double *)malloc(sizeof(double) * (x.size() + 1));
        std::copy(x.begin(), x.end(), p);
        return p;
    };
    *u = dup(hu);
    *v = dup(hv);
    *w = dup(hw);
    *vis = dup(hvis);
    return RG_OK;
}

int rg_load_wkern(const char *wkernfile, double theta, int32_t *planes, int32_t *R, int32_t *S,
                  double *w_min, double *w_step, double **data)
{
    if (!wkernfile || !planes || !R || !S || !w_min || !w_step || !data)
        return fail(RG_ERR_PARAM, "NULL argument");
    HostKern k;
    RG_TRY(load_wkern(wkernfile, theta, k));
    *planes = k.P;
    *R = k.R;
    *S = k.S;
    *w_min = k.w_min;
    *w_step = k.w_step;
    *data = (double *)malloc(sizeof(double) * k.data.size());
    std::copy(k.data.begin(), k.data.end(), *data);
    return RG_OK;
}

// load_sep_kern / load_sep_kern_T (src/common/hdf5.c:372-500): dataset
// "sepkern/kern" (2-D, FP64) and the file attributes du, dw, x0.  The data are
// read in file order; the two loaders differ only in which dimension they
// call the oversampling.
int rg_load_sep_kern(const char *filename, int32_t transposed, int32_t *size,
                     int32_t *oversampling, double *du, double *dw, double *x0, double **data)
{
    using namespace h5;
    if (!filename || !size || !oversampling || !du || !dw || !x0 || !data)
        return fail(RG_ERR_PARAM, "NULL argument");
    Api *a = api();
    if (!a)
        return fail(RG_ERR_UNSUPPORTED, "libhdf5 (>= 1.10) could not be loaded");
    hid_t f = a->fopen(filename, ACC_RDONLY, P_DEFAULT);
    if (f < 0)
        return fail(RG_ERR_IO, "could not open separable kernel file %s", filename);
    hid_t ds = a->lexists(f, "sepkern", P_DEFAULT) > 0 ? a->dopen2(f, "sepkern/kern", P_DEFAULT) : -1;
    if (ds < 0) {
        a->fclose(f);
        return fail(RG_ERR_IO, "'sepkern/kern' dataset could not be opened from file %s", filename);
    }
    hid_t sp = a->dget_space(ds), ty = a->dget_type(ds);
    hsize_t dims[2] = {0, 0};
    const bool fmt = a->sget_simple_extent_ndims(sp) == 2 && a->tget_size(ty) == sizeof(double) &&
                     a->sget_simple_extent_dims(sp, dims, nullptr) >= 0 && dims[0] > 0 &&
                     dims[1] > 0;
    a->sclose(sp);
    a->tclose(ty);
    int st = RG_OK;
    std::vector<double> buf;
    if (!fmt) {
        st = fail(RG_ERR_SHAPE, "'sepkern/kern' dataset has wrong format in file %s", filename);
    } else {
        buf.resize((size_t)dims[0] * dims[1]);
        if (a->dread(ds, *a->t_ieee_f64le, S_ALL, S_ALL, P_DEFAULT, buf.data()) < 0)
            st = fail(RG_ERR_IO, "failed to read separable kernel data from %s", filename);
    }
    a->dclose(ds);
    auto attr = [&](const char *name) {
        double v = std::nan("");
        if (a->aexists(f, name) > 0) {
            hid_t at = a->aopen(f, name, P_DEFAULT);
            if (at >= 0) {
                if (a->aread(at, *a->t_ieee_f64le, &v) < 0)
                    v = std::nan("");
                a->aclose(at);
            }
        }
        return v;
    };
    const double vdu = attr("du"), vdw = attr("dw"), vx0 = attr("x0");
    a->fclose(f);
    if (st != RG_OK)
        return st;
    *size = (int32_t)(transposed ? dims[0] : dims[1]);
    *oversampling = (int32_t)(transposed ? dims[1] : dims[0]);
    *du = vdu;
    *dw = vdw;
    *x0 = vx0;
    *data = (double *)malloc(sizeof(double) * buf.size());
    std::copy(buf.begin(), buf.end(), *data);
    return RG_OK;
}

void rg_free_host(void *p) { free(p); }

int rg_write_vis_hdf5(const char *filename, int32_t n_bl, const int32_t *a1, const int32_t *a2,
                      int32_t T, int32_t F, const double *time, const double *freq,
                      const double *uvw, const double *vis)
{
    using namespace h5;
    Api *a = api();
    if (!a)
        return fail(RG_ERR_UNSUPPORTED, "libhdf5 (>= 1.10) could not be loaded");
    if (!filename || n_bl < 0 || T < 0 || F <= 0 || (n_bl > 0 && (!a1 || !a2 || !uvw || !vis)) ||
        !time || !freq)
        return fail(RG_ERR_PARAM, "invalid arguments");
    hid_t f = a->fcreate(filename, ACC_TRUNC, P_DEFAULT, P_DEFAULT);
    if (f < 0)
        return fail(RG_ERR_IO, "could not create %s", filename);
    hid_t lcpl = a->pcreate(*a->p_cls_link_create);
    a->pset_create_intermediate_group(lcpl, 1);
    int st = RG_OK;
    auto write = [&](hid_t loc, const char *name, int nd, const hsize_t *dims, hid_t ty,
                     const void *data) {
        hid_t sp = a->screate_simple(nd, dims, nullptr);
        hid_t ds = a->dcreate2(loc, name, ty, sp, lcpl, P_DEFAULT, P_DEFAULT);
        bool ok = ds >= 0 && a->dwrite(ds, ty, S_ALL, S_ALL, P_DEFAULT, data) >= 0;
        if (ds >= 0) a->dclose(ds);
        a->sclose(sp);
        return ok;
    };
    for (int b = 0; b < n_bl && st == RG_OK; ++b) {
        char base[96];
        snprintf(base, sizeof(base), "vis/%d/%d/", a1[b], a2[b]);
        const hsize_t df = (hsize_t)F, dt = (hsize_t)T, du[2] = {(hsize_t)T, 3},
                      dv[3] = {(hsize_t)T, (hsize_t)F, 1};
        std::string p = base;
        if (!write(f, (p + "frequency").c_str(), 1, &df, *a->t_ieee_f64le, freq) ||
            !write(f, (p + "time").c_str(), 1, &dt, *a->t_ieee_f64le, time) ||
            !write(f, (p + "uvw").c_str(), 2, du, *a->t_ieee_f64le, uvw + (size_t)b * T * 3) ||
            !write(f, (p + "vis").c_str(), 3, dv, a->cpx, vis + (size_t)b * T * F * 2))
            st = fail(RG_ERR_IO, "writing baseline %d-%d failed", a1[b], a2[b]);
    }
    a->pclose(lcpl);
    a->fclose(f);
    return st;
}

int rg_write_wkern_hdf5(const char *filename, double theta, int32_t planes, int32_t R, int32_t S,
                        double w_min, double w_step, const double *data)
{
    using namespace h5;
    Api *a = api();
    if (!a)
        return fail(RG_ERR_UNSUPPORTED, "libhdf5 (>= 1.10) could not be loaded");
    if (!filename || !data || planes <= 0 || R <= 0 || S <= 0)
        return fail(RG_ERR_PARAM, "invalid arguments");
    hid_t f = a->fcreate(filename, ACC_TRUNC, P_DEFAULT, P_DEFAULT);
    if (f < 0)
        return fail(RG_ERR_IO, "could not create %s", filename);
    hid_t lcpl = a->pcreate(*a->p_cls_link_create);
    a->pset_create_intermediate_group(lcpl, 1);
    int st = RG_OK;
    const hsize_t dims[4] = {(hsize_t)R, (hsize_t)R, (hsize_t)S, (hsize_t)S};
    const size_t pe = (size_t)R * R * S * S * 2;
    for (int p = 0; p < planes && st == RG_OK; ++p) {
        char name[128];
        snprintf(name, sizeof(name), "wkern/%g/%.17g/kern", theta, w_min + p * w_step);
        hid_t sp = a->screate_simple(4, dims, nullptr);
        hid_t ds = a->dcreate2(f, name, a->cpx, sp, lcpl, P_DEFAULT, P_DEFAULT);
        if (ds < 0 || a->dwrite(ds, a->cpx, S_ALL, S_ALL, P_DEFAULT, data + pe * p) < 0)
            st = fail(RG_ERR_IO, "writing %s failed", name);
        if (ds >= 0) a->dclose(ds);
        a->sclose(sp);
    }
    a->pclose(lcpl);
    a->fclose(f);
    return st;
}

} // extern "C"
