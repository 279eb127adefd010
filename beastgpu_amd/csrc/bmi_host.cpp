// Host side of libbmi: the C ABI, mesh and space preprocessing (patches,
// patch colouring, first-touch flags), plans and the launch sequence of one
// assembly.  See include/bmi.h for the contract of each entry point.
#include "../../include/bmi.h"
#include "bmi_internal.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <numeric>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

bmi_status fail(bmi_status s, const std::string &msg) {
    g_err = msg;
    return s;
}

#define BMI_HIP(expr)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess) return fail(BMI_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <class T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    hipError_t alloc(size_t count) {
        release();
        n = count;
        return hipMalloc((void **)&p, sizeof(T) * (count ? count : 1));
    }
    hipError_t upload(const std::vector<T> &h) {
        hipError_t e = alloc(h.size());
        if (e != hipSuccess) return e;
        if (h.empty()) return hipSuccess;
        return hipMemcpy(p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice);
    }
};

constexpr double kPi = 3.14159265358979323846;

// Symmetric triangle rules on the reference simplex (weights sum to 1/2):
// order 1 centroid, order 2 Strang-Fix interior 3-point, order 3 the
// 4-point rule with a negative centroid weight.
bool far_rule(int order, int &n, double l[BMI_MAX_RULE][3], double w[BMI_MAX_RULE]) {
    if (order == 1) {
        n = 1;
        l[0][0] = l[0][1] = l[0][2] = 1.0 / 3.0;
        w[0] = 0.5;
        return true;
    }
    if (order == 2) {
        n = 3;
        for (int i = 0; i < 3; ++i) {
            for (int j = 0; j < 3; ++j) l[i][j] = (i == j) ? 2.0 / 3.0 : 1.0 / 6.0;
            w[i] = 1.0 / 6.0;
        }
        return true;
    }
    if (order == 3) {
        n = 4;
        l[0][0] = l[0][1] = l[0][2] = 1.0 / 3.0;
        w[0] = -27.0 / 96.0;
        for (int i = 0; i < 3; ++i) {
            for (int j = 0; j < 3; ++j) l[i + 1][j] = (i == j) ? 3.0 / 5.0 : 1.0 / 5.0;
            w[i + 1] = 25.0 / 96.0;
        }
        return true;
    }
    return false;
}

// Gauss-Legendre on [0,1] by Newton iteration on the Legendre recurrence.
void gauss_legendre01(int n, double *x, double *w) {
    for (int i = 0; i < n; ++i) {
        double z = std::cos(kPi * (i + 0.75) / (n + 0.5));
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = z;
            for (int k = 2; k <= n; ++k) {
                const double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            if (n == 1) { p1 = z; p0 = 1.0; }
            const double dp = n * (z * p1 - p0) / (z * z - 1.0);
            const double dz = p1 / dp;
            z -= dz;
            if (std::fabs(dz) < 1e-16) break;
        }
        double p0 = 1.0, p1 = z;
        for (int k = 2; k <= n; ++k) {
            const double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
            p0 = p1;
            p1 = p2;
        }
        const double dp = n * (z * p1 - p0) / (z * z - 1.0);
        x[i] = 0.5 * (1.0 - z);
        w[i] = 1.0 / ((1.0 - z * z) * dp * dp);
    }
}

uint64_t spread3(uint64_t v) {
    v &= 0x1fffff;
    v = (v | v << 32) & 0x1f00000000ffffULL;
    v = (v | v << 16) & 0x1f0000ff0000ffULL;
    v = (v | v << 8) & 0x100f00f00f00f00fULL;
    v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
    v = (v | v << 2) & 0x1249249249249249ULL;
    return v;
}

}  // namespace

struct bmi_context {
    int device = 0;
    hipStream_t stream = nullptr;
};

struct bmi_mesh {
    bmi_context *ctx = nullptr;
    int64_t nv = 0, nt = 0;
    std::vector<double> vtx, cen, area;  // per cell
    double hmax = 0.0;
    DevBuf<double> d_vtx, d_cen, d_area;
};

struct bmi_space {
    bmi_mesh *mesh = nullptr;
    int kind = 0;
    int64_t nfun = 0;
    int npatch = 0;
    int ncolour = 0;
    bool uncovered = false;  // some function has no local shape
    std::vector<int> patch_tri, row_ptr, row_fn, row_slot_ptr, patch_colour;
    std::vector<unsigned char> row_first, slot_kl;
    std::vector<double> slot_cf;
    std::vector<std::vector<int>> by_colour;
    DevBuf<int> d_patch_tri, d_row_ptr, d_row_fn, d_row_slot_ptr, d_by_colour;
    DevBuf<unsigned char> d_row_first, d_slot_kl;
    DevBuf<double> d_slot_cf;
    std::vector<int> colour_off;  // offsets of each colour's list in d_by_colour
};

struct bmi_matrix {
    bmi_context *ctx = nullptr;
    int64_t rows = 0, cols = 0;
    int dtype = BMI_C64;
    void *d = nullptr;
};

struct bmi_plan {
    bmi_context *ctx = nullptr;
    const bmi_space *X = nullptr, *Y = nullptr;
    bmi_operator op{};
    bmi_options opts{};
    int op_kind = 0;
    OpConst opc{};
    RuleConst rule{};
    DevBuf<double> d_xpts, d_ypts;
    int npa = 0, npb = 0;
    long long nblocks = 0;
    DevBuf<int> d_blk_count, d_blk_off, d_partials, d_total;
    int capacity = 0;
    DevBuf<int> d_spec_K, d_spec_L;
    DevBuf<unsigned char> d_spec_cls, d_spec_loc;
    DevBuf<double2> d_spec_val;
    SpecialArgs sargs{};
    ClassifyArgs cargs{};
    int64_t pairs[6] = {0, 0, 0, 0, 0, 0};
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    bmi_stats last{};
};

namespace {

DevSpace dev_view(const bmi_space *s, const double *pts) {
    DevSpace d{};
    d.nt = (int)s->mesh->nt;
    d.npatch = s->npatch;
    d.vtx = s->mesh->d_vtx.p;
    d.cen = s->mesh->d_cen.p;
    d.area = s->mesh->d_area.p;
    d.pts = pts;
    d.patch_tri = s->d_patch_tri.p;
    d.row_ptr = s->d_row_ptr.p;
    d.row_fn = s->d_row_fn.p;
    d.row_first = s->d_row_first.p;
    d.row_slot_ptr = s->d_row_slot_ptr.p;
    d.slot_kl = s->d_slot_kl.p;
    d.slot_cf = s->d_slot_cf.p;
    return d;
}

// Patches: runs of BMI_PATCH cells in Morton order of their centroids.
// Rows of a patch: the functions its cells carry, ascending.  Patches that
// share a function get different colours (greedy); the first-touch flag of a
// (patch, function) row is set on the patch with the smallest colour.
bmi_status build_patches(bmi_space *s, const int64_t *gidx, const double *coeff) {
    const bmi_mesh *m = s->mesh;
    const int64_t nt = m->nt;
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int64_t k = 0; k < nt; ++k)
        for (int d = 0; d < 3; ++d) {
            lo[d] = std::min(lo[d], m->cen[3 * k + d]);
            hi[d] = std::max(hi[d], m->cen[3 * k + d]);
        }
    double ext = 0.0;
    for (int d = 0; d < 3; ++d) ext = std::max(ext, hi[d] - lo[d]);
    if (ext <= 0.0) ext = 1.0;
    std::vector<uint64_t> key(nt);
    for (int64_t k = 0; k < nt; ++k) {
        uint64_t c[3];
        for (int d = 0; d < 3; ++d) {
            double u = (m->cen[3 * k + d] - lo[d]) / ext;
            c[d] = (uint64_t)std::min(2097151.0, std::max(0.0, u * 2097151.0));
        }
        key[k] = spread3(c[0]) | spread3(c[1]) << 1 | spread3(c[2]) << 2;
    }
    std::vector<int> order(nt);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return key[a] < key[b]; });

    const int np = (int)((nt + BMI_PATCH - 1) / BMI_PATCH);
    s->npatch = np;
    s->patch_tri.assign((size_t)np * BMI_PATCH, -1);
    for (int64_t i = 0; i < nt; ++i) s->patch_tri[i] = order[i];

    s->row_ptr.assign(np + 1, 0);
    s->row_fn.clear();
    s->row_slot_ptr.clear();
    s->slot_kl.clear();
    s->slot_cf.clear();
    std::vector<char> covered(s->nfun, 0);
    struct Slot { int64_t fn; int kl; double cf; };
    std::vector<Slot> sl;
    for (int p = 0; p < np; ++p) {
        sl.clear();
        for (int k = 0; k < BMI_PATCH; ++k) {
            const int K = s->patch_tri[(size_t)p * BMI_PATCH + k];
            if (K < 0) continue;
            for (int i = 0; i < 3; ++i) {
                const int64_t fn = gidx[3 * (int64_t)K + i];
                if (fn < 0) continue;
                sl.push_back({fn, 3 * k + i, coeff[3 * (int64_t)K + i]});
                covered[fn] = 1;
            }
        }
        std::stable_sort(sl.begin(), sl.end(), [](const Slot &a, const Slot &b) {
            return a.fn != b.fn ? a.fn < b.fn : a.kl < b.kl;
        });
        for (size_t u = 0; u < sl.size(); ++u) {
            if (u == 0 || sl[u].fn != sl[u - 1].fn) {
                s->row_fn.push_back((int)sl[u].fn);
                s->row_slot_ptr.push_back((int)s->slot_kl.size());
            }
            s->slot_kl.push_back((unsigned char)sl[u].kl);
            s->slot_cf.push_back(sl[u].cf);
        }
        s->row_ptr[p + 1] = (int)s->row_fn.size();
    }
    s->row_slot_ptr.push_back((int)s->slot_kl.size());
    s->uncovered = std::find(covered.begin(), covered.end(), 0) != covered.end();

    // function -> patches
    std::vector<int> fcount(s->nfun + 1, 0);
    for (int fn : s->row_fn) fcount[fn + 1]++;
    for (int64_t f = 0; f < s->nfun; ++f) fcount[f + 1] += fcount[f];
    std::vector<int> fpat(s->row_fn.size()), fpos(fcount.begin(), fcount.end() - 1);
    for (int p = 0; p < np; ++p)
        for (int r = s->row_ptr[p]; r < s->row_ptr[p + 1]; ++r) fpat[fpos[s->row_fn[r]]++] = p;

    s->patch_colour.assign(np, -1);
    int ncol = 0;
    std::vector<int> mark(64, -1);
    for (int p = 0; p < np; ++p) {
        for (int r = s->row_ptr[p]; r < s->row_ptr[p + 1]; ++r) {
            const int fn = s->row_fn[r];
            for (int e = fcount[fn]; e < fcount[fn + 1]; ++e) {
                const int o = fpat[e];
                const int c = s->patch_colour[o];
                if (o != p && c >= 0) {
                    if (c >= (int)mark.size()) mark.resize(c + 1, -1);
                    mark[c] = p;
                }
            }
        }
        int c = 0;
        while (c < (int)mark.size() && mark[c] == p) ++c;
        s->patch_colour[p] = c;
        ncol = std::max(ncol, c + 1);
    }
    s->ncolour = ncol;
    std::vector<int> mincol(s->nfun, 1 << 30);
    for (int p = 0; p < np; ++p)
        for (int r = s->row_ptr[p]; r < s->row_ptr[p + 1]; ++r)
            mincol[s->row_fn[r]] = std::min(mincol[s->row_fn[r]], s->patch_colour[p]);
    s->row_first.assign(s->row_fn.size(), 0);
    for (int p = 0; p < np; ++p)
        for (int r = s->row_ptr[p]; r < s->row_ptr[p + 1]; ++r)
            s->row_first[r] = mincol[s->row_fn[r]] == s->patch_colour[p];
    s->by_colour.assign(ncol, {});
    for (int p = 0; p < np; ++p) s->by_colour[s->patch_colour[p]].push_back(p);
    std::vector<int> flat;
    s->colour_off.assign(ncol + 1, 0);
    for (int c = 0; c < ncol; ++c) {
        flat.insert(flat.end(), s->by_colour[c].begin(), s->by_colour[c].end());
        s->colour_off[c + 1] = (int)flat.size();
    }

    hipError_t e;
    if ((e = s->d_patch_tri.upload(s->patch_tri)) != hipSuccess ||
        (e = s->d_row_ptr.upload(s->row_ptr)) != hipSuccess || (e = s->d_row_fn.upload(s->row_fn)) != hipSuccess ||
        (e = s->d_row_slot_ptr.upload(s->row_slot_ptr)) != hipSuccess ||
        (e = s->d_row_first.upload(s->row_first)) != hipSuccess ||
        (e = s->d_slot_kl.upload(s->slot_kl)) != hipSuccess || (e = s->d_slot_cf.upload(s->slot_cf)) != hipSuccess ||
        (e = s->d_by_colour.upload(flat)) != hipSuccess)
        return fail(BMI_EDEVICE, std::string("space upload: ") + hipGetErrorString(e));
    return BMI_OK;
}

std::vector<double> rule_points(const bmi_mesh *m, int n, const double l[BMI_MAX_RULE][3]) {
    std::vector<double> pts((size_t)m->nt * 3 * BMI_MAX_RULE, 0.0);
    for (int64_t k = 0; k < m->nt; ++k) {
        const double *v = &m->vtx[9 * k];
        for (int I = 0; I < n; ++I)
            for (int d = 0; d < 3; ++d) {
                // x = v3 + b1 (v1 - v3) + b2 (v2 - v3), relative to the centroid
                const double x = v[6 + d] + l[I][0] * (v[d] - v[6 + d]) + l[I][1] * (v[3 + d] - v[6 + d]);
                pts[(size_t)k * 3 * BMI_MAX_RULE + 3 * I + d] = x - m->cen[3 * k + d];
            }
    }
    return pts;
}

}  // namespace

extern "C" {

const char *bmi_version(void) { return "bmi 0.1 (gfx950)"; }

const char *bmi_last_error(void) { return g_err.c_str(); }

void bmi_options_default(bmi_options *o) {
    if (!o) return;
    o->far_test_order = 2;
    o->far_trial_order = 3;
    o->ss_order[0] = o->ss_order[1] = o->ss_order[2] = 5;
    o->near_order = 4;
    o->precision = BMI_PREC_FP64;
    o->near_factor = 2.0;
}

bmi_status bmi_context_create(int32_t device, bmi_context **out) {
    if (!out) return fail(BMI_EINVAL, "out is NULL");
    int n = 0;
    BMI_HIP(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(BMI_EINVAL, "no such device");
    BMI_HIP(hipSetDevice(device));
    auto ctx = std::make_unique<bmi_context>();
    ctx->device = device;
    BMI_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    *out = ctx.release();
    return BMI_OK;
}

bmi_status bmi_context_destroy(bmi_context *ctx) {
    if (!ctx) return BMI_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return BMI_OK;
}

bmi_status bmi_mesh_create(bmi_context *ctx, const double *xyz, int64_t nv, const int32_t *tri, int64_t nt,
                           bmi_mesh **out) {
    if (!ctx || !xyz || !tri || !out) return fail(BMI_EINVAL, "NULL argument");
    if (nv < 3 || nt < 1 || nt > (1LL << 30)) return fail(BMI_EINVAL, "bad mesh size");
    BMI_HIP(hipSetDevice(ctx->device));
    auto m = std::make_unique<bmi_mesh>();
    m->ctx = ctx;
    m->nv = nv;
    m->nt = nt;
    m->vtx.resize(9 * nt);
    m->cen.resize(3 * nt);
    m->area.resize(nt);
    for (int64_t i = 0; i < 3 * nv; ++i)
        if (!std::isfinite(xyz[i])) return fail(BMI_EINVAL, "non-finite vertex coordinate");
    double hmax = 0.0;
    for (int64_t k = 0; k < nt; ++k) {
        for (int i = 0; i < 3; ++i) {
            const int32_t v = tri[3 * k + i];
            if (v < 0 || v >= nv) return fail(BMI_EINVAL, "vertex index out of range in cell " + std::to_string(k));
            for (int d = 0; d < 3; ++d) m->vtx[9 * k + 3 * i + d] = xyz[3 * (int64_t)v + d];
        }
        const double *p = &m->vtx[9 * k];
        double t1[3], t2[3];
        for (int d = 0; d < 3; ++d) {
            t1[d] = p[d] - p[6 + d];
            t2[d] = p[3 + d] - p[6 + d];
            m->cen[3 * k + d] = (p[d] + p[3 + d] + p[6 + d]) / 3.0;
        }
        const double n0 = t1[1] * t2[2] - t1[2] * t2[1], n1 = t1[2] * t2[0] - t1[0] * t2[2],
                     n2 = t1[0] * t2[1] - t1[1] * t2[0];
        m->area[k] = 0.5 * std::sqrt(n0 * n0 + n1 * n1 + n2 * n2);
        if (!(m->area[k] > 0.0)) return fail(BMI_EINVAL, "degenerate cell " + std::to_string(k));
        for (int i = 0; i < 3; ++i) {
            const double *a = &p[3 * i], *b = &p[3 * ((i + 1) % 3)];
            const double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
            hmax = std::max(hmax, std::sqrt(dx * dx + dy * dy + dz * dz));
        }
    }
    m->hmax = hmax;
    hipError_t e;
    if ((e = m->d_vtx.upload(m->vtx)) != hipSuccess || (e = m->d_cen.upload(m->cen)) != hipSuccess ||
        (e = m->d_area.upload(m->area)) != hipSuccess)
        return fail(BMI_EDEVICE, std::string("mesh upload: ") + hipGetErrorString(e));
    *out = m.release();
    return BMI_OK;
}

bmi_status bmi_mesh_free(bmi_mesh *m) {
    if (m) {
        (void)hipSetDevice(m->ctx->device);
        delete m;
    }
    return BMI_OK;
}

bmi_status bmi_space_create(bmi_mesh *mesh, int32_t kind, const int64_t *gidx, const double *coeff, int64_t nfun,
                            bmi_space **out) {
    if (!mesh || !gidx || !coeff || !out) return fail(BMI_EINVAL, "NULL argument");
    if (kind != BMI_SPACE_LAGRANGE_C0D1 && kind != BMI_SPACE_RWG) return fail(BMI_EUNSUPPORTED, "unknown space kind");
    if (nfun < 1 || nfun > (1LL << 31) - 1) return fail(BMI_EINVAL, "bad number of functions");
    for (int64_t i = 0; i < 3 * mesh->nt; ++i) {
        if (gidx[i] >= nfun) return fail(BMI_EINVAL, "function index out of range");
        if (gidx[i] >= 0 && !std::isfinite(coeff[i])) return fail(BMI_EINVAL, "non-finite coefficient");
    }
    BMI_HIP(hipSetDevice(mesh->ctx->device));
    auto s = std::make_unique<bmi_space>();
    s->mesh = mesh;
    s->kind = kind;
    s->nfun = nfun;
    bmi_status st = build_patches(s.get(), gidx, coeff);
    if (st != BMI_OK) return st;
    *out = s.release();
    return BMI_OK;
}

bmi_status bmi_space_free(bmi_space *s) {
    if (s) {
        (void)hipSetDevice(s->mesh->ctx->device);
        delete s;
    }
    return BMI_OK;
}

bmi_status bmi_matrix_create(bmi_context *ctx, int64_t rows, int64_t cols, int32_t dtype, bmi_matrix **out) {
    if (!ctx || !out) return fail(BMI_EINVAL, "NULL argument");
    if (rows < 1 || cols < 1) return fail(BMI_EINVAL, "bad matrix size");
    if (dtype != BMI_C64 && dtype != BMI_C32) return fail(BMI_EINVAL, "bad dtype");
    BMI_HIP(hipSetDevice(ctx->device));
    const size_t es = dtype == BMI_C64 ? 16 : 8;
    const size_t bytes = (size_t)rows * (size_t)cols * es;
    size_t freeb = 0, totalb = 0;
    BMI_HIP(hipMemGetInfo(&freeb, &totalb));
    if (bytes > freeb) return fail(BMI_ENOMEM, "matrix of " + std::to_string(bytes) + " bytes exceeds free device memory");
    auto m = std::make_unique<bmi_matrix>();
    m->ctx = ctx;
    m->rows = rows;
    m->cols = cols;
    m->dtype = dtype;
    hipError_t e = hipMalloc(&m->d, bytes);
    if (e != hipSuccess) return fail(BMI_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    *out = m.release();
    return BMI_OK;
}

bmi_status bmi_matrix_free(bmi_matrix *m) {
    if (m) {
        (void)hipSetDevice(m->ctx->device);
        if (m->d) (void)hipFree(m->d);
        delete m;
    }
    return BMI_OK;
}

bmi_status bmi_matrix_device_ptr(const bmi_matrix *m, void **ptr, int64_t *ld) {
    if (!m) return fail(BMI_EINVAL, "NULL matrix");
    if (ptr) *ptr = m->d;
    if (ld) *ld = m->rows;
    return BMI_OK;
}

bmi_status bmi_plan_create(bmi_context *ctx, const bmi_operator *op, const bmi_space *test, const bmi_space *trial,
                           const bmi_options *opts, bmi_plan **out) {
    if (!ctx || !op || !test || !trial || !out) return fail(BMI_EINVAL, "NULL argument");
    if (test->mesh->ctx != ctx || trial->mesh->ctx != ctx) return fail(BMI_EINVAL, "spaces belong to another context");
    bmi_options o;
    bmi_options_default(&o);
    if (opts) o = *opts;
    if (op->kind == BMI_OP_HH3D_SINGLELAYER) {
        if (test->kind != BMI_SPACE_LAGRANGE_C0D1 || trial->kind != BMI_SPACE_LAGRANGE_C0D1)
            return fail(BMI_EUNSUPPORTED, "the Helmholtz single layer is assembled on lagrangec0d1 spaces");
    } else if (op->kind == BMI_OP_MW_SINGLELAYER) {
        if (test->kind != BMI_SPACE_RWG || trial->kind != BMI_SPACE_RWG)
            return fail(BMI_EUNSUPPORTED, "the Maxwell single layer is assembled on RWG spaces");
    } else {
        return fail(BMI_EUNSUPPORTED, "unknown operator kind");
    }
    for (int i = 0; i < 3; ++i)
        if (o.ss_order[i] < 1 || o.ss_order[i] > BMI_MAX_GL) return fail(BMI_EINVAL, "ss_order must be 1..8");
    if (o.near_order < 1 || o.near_order > 8) return fail(BMI_EINVAL, "near_order must be 1..8");
    if (o.precision != BMI_PREC_FP64 && o.precision != BMI_PREC_MIXED) return fail(BMI_EINVAL, "bad precision");
    if (!(o.near_factor >= 0.0)) return fail(BMI_EINVAL, "bad near_factor");
    const double vals[6] = {op->alpha[0], op->alpha[1], op->beta[0], op->beta[1], op->gamma[0], op->gamma[1]};
    for (double v : vals)
        if (!std::isfinite(v)) return fail(BMI_EINVAL, "non-finite operator coefficient");

    BMI_HIP(hipSetDevice(ctx->device));
    auto P = std::make_unique<bmi_plan>();
    P->ctx = ctx;
    P->X = test;
    P->Y = trial;
    P->op = *op;
    P->opts = o;
    P->op_kind = op->kind == BMI_OP_MW_SINGLELAYER ? OPK_MW : OPK_HH3D;
    P->opc.alpha_re = op->alpha[0];
    P->opc.alpha_im = op->alpha[1];
    P->opc.beta_re = op->beta[0];
    P->opc.beta_im = op->beta[1];
    P->opc.gamma_re = op->gamma[0];
    P->opc.k2pi = op->gamma[1] / (2.0 * kPi);
    P->opc.lossy = op->gamma[0] != 0.0;

    double lt[BMI_MAX_RULE][3] = {}, lb[BMI_MAX_RULE][3] = {}, wt[BMI_MAX_RULE] = {}, wb[BMI_MAX_RULE] = {};
    int nt = 0, nb = 0;
    if (!far_rule(o.far_test_order, nt, lt, wt) || !far_rule(o.far_trial_order, nb, lb, wb))
        return fail(BMI_EINVAL, "far rule orders must be 1, 2 or 3");
    RuleConst &R = P->rule;
    R.nt = nt;
    R.nb = nb;
    for (int I = 0; I < BMI_MAX_RULE; ++I)
        for (int d = 0; d < 3; ++d) {
            R.lt[I][d] = lt[I][d];
            R.lb[I][d] = lb[I][d];
        }
    for (int I = 0; I < BMI_MAX_RULE; ++I)
        for (int J = 0; J < BMI_MAX_RULE; ++J) R.w4pi[I][J] = wt[I] * wb[J] / (4.0 * kPi);

    hipError_t e;
    if ((e = P->d_xpts.upload(rule_points(test->mesh, nt, lt))) != hipSuccess ||
        (e = P->d_ypts.upload(rule_points(trial->mesh, nb, lb))) != hipSuccess)
        return fail(BMI_EDEVICE, std::string("rule points: ") + hipGetErrorString(e));

    P->npa = test->npatch;
    P->npb = trial->npatch;
    P->nblocks = (long long)P->npa * P->npb;
    if (P->nblocks >= (1LL << 31) - 1) return fail(BMI_EUNSUPPORTED, "too many patch blocks");
    const int nbk = (int)P->nblocks;
    if ((e = P->d_blk_count.alloc(nbk)) != hipSuccess || (e = P->d_blk_off.alloc(nbk + 1)) != hipSuccess ||
        (e = P->d_partials.alloc(scan_partials_needed(nbk) + 1)) != hipSuccess ||
        (e = P->d_total.alloc(1)) != hipSuccess)
        return fail(BMI_ENOMEM, std::string("plan buffers: ") + hipGetErrorString(e));

    ClassifyArgs &C = P->cargs;
    C.X = dev_view(test, P->d_xpts.p);
    C.Y = dev_view(trial, P->d_ypts.p);
    C.npa = P->npa;
    C.npb = P->npb;
    C.k2x16 = 16.0 * (op->gamma[0] * op->gamma[0] + op->gamma[1] * op->gamma[1]);
    const double h = std::max(test->mesh->hmax, trial->mesh->hmax);
    C.near_d2 = o.precision == BMI_PREC_MIXED ? (o.near_factor * h) * (o.near_factor * h) : -1.0;
    C.blk_count = P->d_blk_count.p;
    C.blk_off = P->d_blk_off.p;

    // first classification: size the special-pair buffers
    hipStream_t s = ctx->stream;
    BMI_HIP(launch_classify(C, 0, s));
    BMI_HIP(launch_scan(P->d_blk_count.p, P->d_blk_off.p, nbk, P->d_partials.p, P->d_total.p, s));
    int total = 0;
    BMI_HIP(hipMemcpyAsync(&total, P->d_total.p, sizeof(int), hipMemcpyDeviceToHost, s));
    BMI_HIP(hipStreamSynchronize(s));
    P->capacity = total;
    const int cap = std::max(total, 1);
    if ((e = P->d_spec_K.alloc(cap)) != hipSuccess || (e = P->d_spec_L.alloc(cap)) != hipSuccess ||
        (e = P->d_spec_cls.alloc(cap)) != hipSuccess || (e = P->d_spec_loc.alloc(cap)) != hipSuccess ||
        (e = P->d_spec_val.alloc((size_t)cap * 9)) != hipSuccess)
        return fail(BMI_ENOMEM, std::string("special-pair buffers: ") + hipGetErrorString(e));
    C.spec_K = P->d_spec_K.p;
    C.spec_L = P->d_spec_L.p;
    C.spec_cls = P->d_spec_cls.p;
    C.spec_loc = P->d_spec_loc.p;
    C.capacity = total;
    BMI_HIP(launch_classify(C, 1, s));
    std::vector<unsigned char> cls(cap, 0);
    if (total > 0) BMI_HIP(hipMemcpyAsync(cls.data(), P->d_spec_cls.p, total, hipMemcpyDeviceToHost, s));
    BMI_HIP(hipStreamSynchronize(s));
    for (int i = 0; i < total; ++i) P->pairs[cls[i] < 6 ? cls[i] : 0]++;
    int64_t special = 0;
    for (int c = 1; c < 6; ++c) special += P->pairs[c];
    P->pairs[0] = test->mesh->nt * trial->mesh->nt - special;

    SpecialArgs &S = P->sargs;
    S.X = C.X;
    S.Y = C.Y;
    S.spec_K = P->d_spec_K.p;
    S.spec_L = P->d_spec_L.p;
    S.spec_cls = P->d_spec_cls.p;
    S.n_special = P->d_total.p;
    S.capacity = total;
    S.spec_val = P->d_spec_val.p;
    S.op = P->opc;
    S.rule = P->rule;
    for (int g = 0; g < 3; ++g) {
        S.ss_n[g] = o.ss_order[g];
        gauss_legendre01(o.ss_order[g], S.glx[g], S.glw[g]);
    }
    {
        double x[8], w[8];
        gauss_legendre01(o.near_order, x, w);
        int n = 0;
        for (int i = 0; i < o.near_order; ++i)
            for (int j = 0; j < o.near_order; ++j) {
                // Duffy-collapsed product rule on the reference simplex
                S.nb1[n] = x[i];
                S.nb2[n] = x[j] * (1.0 - x[i]);
                S.nw[n] = w[i] * w[j] * (1.0 - x[i]);
                ++n;
            }
        S.near_np = n;
    }
    for (int i = 0; i < 4; ++i) BMI_HIP(hipEventCreate(&P->ev[i]));
    *out = P.release();
    return BMI_OK;
}

bmi_status bmi_plan_free(bmi_plan *P) {
    if (P) {
        (void)hipSetDevice(P->ctx->device);
        for (auto &e : P->ev)
            if (e) (void)hipEventDestroy(e);
        delete P;
    }
    return BMI_OK;
}

bmi_status bmi_assemble(bmi_plan *P, bmi_matrix *out) {
    if (!P || !out) return fail(BMI_EINVAL, "NULL argument");
    if (out->ctx != P->ctx) return fail(BMI_EINVAL, "matrix belongs to another context");
    if (out->rows != P->X->nfun || out->cols != P->Y->nfun)
        return fail(BMI_EINVAL, "matrix is " + std::to_string(out->rows) + " x " + std::to_string(out->cols) +
                                    ", the plan assembles " + std::to_string(P->X->nfun) + " x " +
                                    std::to_string(P->Y->nfun));
    BMI_HIP(hipSetDevice(P->ctx->device));
    hipStream_t s = P->ctx->stream;
    const int nbk = (int)P->nblocks;
    BMI_HIP(hipEventRecord(P->ev[0], s));
    BMI_HIP(launch_classify(P->cargs, 0, s));
    BMI_HIP(launch_scan(P->d_blk_count.p, P->d_blk_off.p, nbk, P->d_partials.p, P->d_total.p, s));
    BMI_HIP(launch_classify(P->cargs, 1, s));
    BMI_HIP(hipEventRecord(P->ev[1], s));
    const int mixed = P->opts.precision == BMI_PREC_MIXED;
    BMI_HIP(launch_special(P->sargs, P->capacity, mixed, P->op_kind, s));
    BMI_HIP(hipEventRecord(P->ev[2], s));
    if (P->X->uncovered || P->Y->uncovered)
        BMI_HIP(hipMemsetAsync(out->d, 0, (size_t)out->rows * out->cols * (out->dtype == BMI_C64 ? 16 : 8), s));

    BlockArgs B{};
    B.X = P->cargs.X;
    B.Y = P->cargs.Y;
    B.npb = P->npb;
    B.blk_off = P->d_blk_off.p;
    B.spec_loc = P->d_spec_loc.p;
    B.spec_val = P->d_spec_val.p;
    B.capacity = P->capacity;
    B.Z = out->d;
    B.ld = out->rows;
    B.op = P->opc;
    B.rule = P->rule;
    int launches = 0;
    for (int ct = 0; ct < P->X->ncolour; ++ct) {
        const int ntp = P->X->colour_off[ct + 1] - P->X->colour_off[ct];
        for (int cb = 0; cb < P->Y->ncolour; ++cb) {
            B.tp_list = P->X->d_by_colour.p + P->X->colour_off[ct];
            B.bq_list = P->Y->d_by_colour.p + P->Y->colour_off[cb];
            B.nbq = P->Y->colour_off[cb + 1] - P->Y->colour_off[cb];
            BMI_HIP(launch_blocks(B, ntp, P->op_kind, mixed, out->dtype == BMI_C32, s));
            ++launches;
        }
    }
    BMI_HIP(hipEventRecord(P->ev[3], s));
    BMI_HIP(hipStreamSynchronize(s));
    float t01 = 0, t12 = 0, t23 = 0, t03 = 0;
    BMI_HIP(hipEventElapsedTime(&t01, P->ev[0], P->ev[1]));
    BMI_HIP(hipEventElapsedTime(&t12, P->ev[1], P->ev[2]));
    BMI_HIP(hipEventElapsedTime(&t23, P->ev[2], P->ev[3]));
    BMI_HIP(hipEventElapsedTime(&t03, P->ev[0], P->ev[3]));
    P->last.classify_ms = t01;
    P->last.special_ms = t12;
    P->last.far_ms = t23;
    P->last.total_ms = t03;
    for (int i = 0; i < 6; ++i) P->last.pairs[i] = P->pairs[i];
    P->last.far_launches = launches;
    P->last.n_colours_test = P->X->ncolour;
    P->last.n_colours_trial = P->Y->ncolour;
    return BMI_OK;
}

bmi_status bmi_plan_stats(const bmi_plan *P, bmi_stats *st) {
    if (!P || !st) return fail(BMI_EINVAL, "NULL argument");
    *st = P->last;
    for (int i = 0; i < 6; ++i) st->pairs[i] = P->pairs[i];
    st->n_colours_test = P->X->ncolour;
    st->n_colours_trial = P->Y->ncolour;
    return BMI_OK;
}

bmi_status bmi_matrix_to_host(const bmi_matrix *m, void *dst, int32_t dtype) {
    if (!m || !dst) return fail(BMI_EINVAL, "NULL argument");
    if (dtype != BMI_C64 && dtype != BMI_C32) return fail(BMI_EINVAL, "bad dtype");
    BMI_HIP(hipSetDevice(m->ctx->device));
    const size_t n = (size_t)m->rows * (size_t)m->cols;
    if (dtype == m->dtype) {
        BMI_HIP(hipMemcpy(dst, m->d, n * (dtype == BMI_C64 ? 16 : 8), hipMemcpyDeviceToHost));
        return BMI_OK;
    }
    const size_t chunk = 1 << 22;  // complex values per staging copy
    const size_t es = m->dtype == BMI_C64 ? 16 : 8;
    std::vector<unsigned char> stage(chunk * es);
    for (size_t o = 0; o < n; o += chunk) {
        const size_t c = std::min(chunk, n - o);
        BMI_HIP(hipMemcpy(stage.data(), (const unsigned char *)m->d + o * es, c * es, hipMemcpyDeviceToHost));
        if (m->dtype == BMI_C64) {
            const double *src = (const double *)stage.data();
            float *d = (float *)dst + 2 * o;
            for (size_t i = 0; i < 2 * c; ++i) d[i] = (float)src[i];
        } else {
            const float *src = (const float *)stage.data();
            double *d = (double *)dst + 2 * o;
            for (size_t i = 0; i < 2 * c; ++i) d[i] = (double)src[i];
        }
    }
    return BMI_OK;
}

}  // extern "C"
