// Host side of libbmi: the C ABI, mesh and space preprocessing (patches,
// patch colouring, first-touch flags), plans and the launch sequence of one
// assembly.  See include/bmi.h for the contract of each entry point.
#include "../../include/bmi.h"
#include "bmi_internal.h"

#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <functional>
#include <memory>
#include <mutex>
#include <numeric>
#include <string>
#include <system_error>
#include <thread>
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

}  // namespace

bmi_status set_error(bmi_status s, const std::string &msg) { return fail(s, msg); }

// Status for an exception caught at the C boundary (called inside a catch
// block): out of host memory or threads -> BMI_ENOMEM, else BMI_EINTERNAL.
bmi_status fail_exception() noexcept {
    try {
        try {
            throw;
        } catch (const std::bad_alloc &) {
            return fail(BMI_ENOMEM, "host memory allocation failed");
        } catch (const std::system_error &e) {
            return fail(BMI_ENOMEM, std::string("host resources: ") + e.what());
        } catch (const std::exception &e) {
            return fail(BMI_EINTERNAL, std::string("internal error: ") + e.what());
        } catch (...) {
            return fail(BMI_EINTERNAL, "internal error");
        }
    } catch (...) {
        return BMI_ENOMEM;  // recording the message failed too
    }
}

// Device memory comes from the device's stream-ordered pool, allocated and
// freed on a library stream of its own per device (it carries nothing else):
// a free does not wait for the whole device to go idle, as hipFree does, and
// an allocation waits only for that stream, which matters when several
// contexts work on one device at once.  Every entry point finishes its
// device work before it returns, so a buffer is idle when it is freed.
namespace {
std::mutex g_alloc_mu;
hipStream_t g_alloc_stream[64] = {};

hipStream_t alloc_stream(int *dev) {
    if (hipGetDevice(dev) != hipSuccess || *dev < 0 || *dev >= 64) return nullptr;
    std::lock_guard<std::mutex> lk(g_alloc_mu);
    if (!g_alloc_stream[*dev]) {
        hipStream_t s = nullptr;
        if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return nullptr;
        g_alloc_stream[*dev] = s;
    }
    return g_alloc_stream[*dev];
}

hipMemPool_t device_pool(int dev) {
    hipMemPool_t pool = nullptr;
    if (hipDeviceGetDefaultMemPool(&pool, dev) != hipSuccess) return nullptr;
    return pool;
}
}  // namespace

// Bytes the pool holds for reuse (reserved but not in use).
size_t pool_cached_bytes(int dev) {
    hipMemPool_t pool = device_pool(dev);
    if (!pool) return 0;
    uint64_t reserved = 0, used = 0;
    if (hipMemPoolGetAttribute(pool, hipMemPoolAttrReservedMemCurrent, &reserved) != hipSuccess ||
        hipMemPoolGetAttribute(pool, hipMemPoolAttrUsedMemCurrent, &used) != hipSuccess)
        return 0;
    return reserved > used ? (size_t)(reserved - used) : 0;
}

namespace {
// stream of the entry point running on this thread (its context's), if any
thread_local hipStream_t t_work_stream = nullptr;
}  // namespace

WorkStream::WorkStream(hipStream_t s) : prev_(t_work_stream) { t_work_stream = s; }
WorkStream::~WorkStream() { t_work_stream = prev_; }

hipError_t dev_alloc(void **p, size_t bytes) {
    *p = nullptr;
    int dev = 0;
    hipStream_t s = alloc_stream(&dev);
    if (!s) return hipErrorInvalidDevice;
    // allocate where the caller's work runs: the allocation stream is shared
    // by every context of the device and may sit in a hardware queue behind
    // another context's kernels
    if (t_work_stream) s = t_work_stream;
    hipError_t e = hipMallocAsync(p, bytes ? bytes : 1, s);
    if (e == hipErrorOutOfMemory) {
        // give the pool's cached blocks back and try once more
        (void)hipGetLastError();
        (void)hipStreamSynchronize(s);
        if (hipMemPool_t pool = device_pool(dev)) (void)hipMemPoolTrimTo(pool, 0);
        e = hipMallocAsync(p, bytes ? bytes : 1, s);
    }
    if (e != hipSuccess) {
        *p = nullptr;
        return e;
    }
    return hipStreamSynchronize(s);
}

// Host-to-device copy on the caller's stream (a plain hipMemcpy goes through
// the null stream and waits for other contexts' work).
hipError_t dev_upload(void *dst, const void *src, size_t bytes) {
    int dev = 0;
    hipStream_t s = alloc_stream(&dev);
    if (!s) return hipErrorInvalidDevice;
    if (t_work_stream) s = t_work_stream;
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(s);
}

void dev_free(void *p) {
    if (!p) return;
    int dev = 0;
    hipStream_t s = alloc_stream(&dev);
    if (s) (void)hipFreeAsync(p, s);
}

namespace {

template <class T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void release() {
        dev_free(p);
        p = nullptr;
        n = 0;
    }
    hipError_t alloc(size_t count) {
        release();
        n = count;
        return dev_alloc((void **)&p, sizeof(T) * (count ? count : 1));
    }
    hipError_t upload(const std::vector<T> &h) {
        hipError_t e = alloc(h.size());
        if (e != hipSuccess) return e;
        if (h.empty()) return hipSuccess;
        return dev_upload(p, h.data(), sizeof(T) * h.size());
    }
};

template <class T>
size_t buf_bytes(const DevBuf<T> &b) {
    return b.p ? sizeof(T) * (b.n ? b.n : 1) : 0;
}

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
    // grow-only device workspace of the apply kernels
    void *ws = nullptr;
    size_t ws_bytes = 0;
    // pinned staging ring for device-to-host copies of matrices
    void *stage[2] = {nullptr, nullptr};
    hipEvent_t stage_ev[2] = {nullptr, nullptr};
    // streamed assembly into host memory: copies run on their own stream
    // (default priority) beside the next block's kernels
    hipStream_t copy_stream = nullptr;
    hipEvent_t blk_ev[2] = {nullptr, nullptr};
};

struct bmi_mesh {
    bmi_context *ctx = nullptr;
    int device = 0;  // kept here so that freeing needs no other object
    int64_t nv = 0, nt = 0;
    std::vector<double> vtx, cen, area;  // per cell
    double hmax = 0.0;
    DevBuf<double> d_vtx, d_cen, d_area;
};

struct bmi_space {
    bmi_mesh *mesh = nullptr;
    int device = 0;
    int kind = 0;
    int64_t nfun = 0;
    int npatch = 0;
    int ncolour = 0;
    bool uncovered = false;  // some function has no local shape
    std::vector<int> patch_tri, row_ptr, row_fn, row_slot_ptr, patch_colour;
    std::vector<double> patch_sph, patch_amax;
    std::vector<unsigned char> row_first, slot_kl;
    std::vector<double> slot_cf;
    std::vector<std::vector<int>> by_colour;
    DevBuf<int> d_patch_tri, d_by_colour;
    DevBuf<double> d_patch_sph, d_patch_amax;
    DevBuf<int> d_home_patch, d_home_sb;  // cell -> its patch / its first sub-batch (-1: none)
    std::vector<int> colour_off;  // offsets of each colour's list in d_by_colour
    // test role: panels of prows functions in sub-batches of BMI_PATCH cells;
    // trial role: patches of at most maxcols functions (Lagrange: wide panels,
    // narrow patches)
    int prows = BMI_PANEL_ROWS, maxcols = BMI_MAX_COLS;
    int npanel = 0, nsb = 0;
    std::vector<int> prow;  // npanel + 1: first row (function) of each panel
    DevBuf<int> d_prow;
    std::vector<int> sb_off, sb_cells, sb_row_ptr;
    std::vector<unsigned char> sb_slot_kl;
    std::vector<double> sb_slot_cf, sb_sph, sb_amax;
    DevBuf<int> d_sb_off, d_sb_cells;
    DevBuf<double> d_sb_sph, d_sb_amax;
    // fixed-size records: test role (per sub-batch) and trial role (per patch)
    DevBuf<int> d_trs, d_bnc, d_bfn, d_bcs, d_tslot, d_bslot;
    DevBuf<double> d_tscf, d_bscf;
    int rounds_t = 1, rounds_b = 1, levels_b = 1, ncmax = 1;
    // rounds of each trial patch: npb by slot rank, then npb by slot level
    DevBuf<unsigned char> d_brounds;
    DevBuf<unsigned char> d_tkl, d_bfirst, d_bkl;
    DevBuf<double> d_tcf, d_bcf;
};

struct bmi_matrix {
    bmi_context *ctx = nullptr;
    int device = 0;
    int64_t rows = 0, cols = 0;
    int dtype = BMI_C64;
    void *d = nullptr;
    bool owned = true;  // false: d belongs to the caller (matrix_wrap)
};

struct bmi_plan {
    bmi_context *ctx = nullptr;
    int device = 0;
    const bmi_space *X = nullptr, *Y = nullptr;
    bmi_operator op{};
    bmi_options opts{};
    int op_kind = 0;
    OpConst opc{};
    RuleConst rule{};
    DevBuf<double> d_xpts, d_ypts, d_tgeo, d_bgeo;
    int npa = 0, npb = 0;
    long long nblocks = 0;
    DevBuf<int> d_blk_cnt, d_blk_off, d_total;  // d_total: the classifier's entry counter
    int capacity = 0;
    DevBuf<int> d_spec_K, d_spec_L;
    DevBuf<unsigned char> d_spec_cls, d_spec_loc;
    DevBuf<double2> d_spec_val;
    SpecialArgs sargs{};
    ClassifyArgs cargs{};
    // device copies of the arguments and one-problem prefix tables of the
    // classify / special kernels (which also take whole batches)
    DevBuf<ClassifyArgs> d_cargs;
    DevBuf<SpecialArgs> d_sargs;
    DevBuf<int> d_gbase, d_cbase, d_ebase;
    int cs = 256, nchunks = 0;
    int64_t pairs[6] = {0, 0, 0, 0, 0, 0};
    bool pairs_ready = false;
    DevBuf<double2> d_sc;  // the 64-entry sine/cosine table of the FP64 Green evaluations
    bool rule_default = true;  // the far rule is the default 3 x 4 point rule
    mutable bool filled = false;  // the special-pair entries have been written
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
    d.patch_sph = s->d_patch_sph.p;
    d.patch_amax = s->d_patch_amax.p;
    d.home = s->d_home_patch.p;
    return d;
}

// The test role of a space: cells grouped by panel sub-batch.
DevSpace dev_view_test(const bmi_space *s, const double *pts) {
    DevSpace d = dev_view(s, pts);
    d.npatch = s->nsb;
    d.patch_tri = s->d_sb_cells.p;
    d.patch_sph = s->d_sb_sph.p;
    d.patch_amax = s->d_sb_amax.p;
    d.home = s->d_home_sb.p;
    return d;
}

// Bounding sphere (centre of the vertex bounding box) and largest area of a
// group of cells.
void group_sphere(const bmi_mesh *m, const int *cells, int n, double *sph, double *amax) {
    double blo[3] = {1e300, 1e300, 1e300}, bhi[3] = {-1e300, -1e300, -1e300};
    *amax = 0.0;
    for (int k = 0; k < n; ++k) {
        const int K = cells[k];
        if (K < 0) continue;
        *amax = std::max(*amax, m->area[K]);
        for (int i = 0; i < 3; ++i)
            for (int d = 0; d < 3; ++d) {
                blo[d] = std::min(blo[d], m->vtx[9 * (size_t)K + 3 * i + d]);
                bhi[d] = std::max(bhi[d], m->vtx[9 * (size_t)K + 3 * i + d]);
            }
    }
    double r2 = 0.0;
    for (int d = 0; d < 3; ++d) sph[d] = 0.5 * (blo[d] + bhi[d]);
    for (int k = 0; k < n; ++k) {
        const int K = cells[k];
        if (K < 0) continue;
        for (int i = 0; i < 3; ++i) {
            double q2 = 0.0;
            for (int d = 0; d < 3; ++d) {
                const double x = m->vtx[9 * (size_t)K + 3 * i + d] - sph[d];
                q2 += x * x;
            }
            r2 = std::max(r2, q2);
        }
    }
    sph[3] = std::sqrt(r2) * (1.0 + 1e-12);
}

// Host worker threads, started once for the process (up to 7, beside the
// calling thread): a call hands out the indices 0 .. n - 1 in chunks and
// returns when all are done.  Calls from several threads at once (serving
// workers preparing requests) share the workers; each caller also works on
// its own call, so a call always completes.  Starting threads per call cost
// about 0.35 ms each time (a request's space creation made four such calls).
class HostPool {
  public:
    static HostPool &get() {
        static HostPool *pool = new HostPool();  // never destroyed: workers may still wait at exit
        return *pool;
    }
    int threads() const { return (int)workers_ + 1; }
    void run(int n, int chunk, const std::function<void(int)> &f) {
        auto job = std::make_shared<Job>();
        job->f = &f;
        job->n = n;
        job->chunk = chunk;
        {
            std::lock_guard<std::mutex> lk(mu_);
            jobs_.push_back(job);
        }
        cv_.notify_all();
        work(*job);
        while (job->done.load(std::memory_order_acquire) < n) std::this_thread::yield();
        std::lock_guard<std::mutex> lk(mu_);
        for (auto it = jobs_.begin(); it != jobs_.end(); ++it)
            if (*it == job) {
                jobs_.erase(it);
                break;
            }
    }

  private:
    struct Job {
        const std::function<void(int)> *f = nullptr;
        int n = 0, chunk = 1;
        std::atomic<int> next{0}, done{0};
    };
    HostPool() {
        workers_ = std::min<unsigned>(7u, std::max(1u, std::thread::hardware_concurrency()) - 1u);
        for (unsigned t = 0; t < workers_; ++t) std::thread([this] { loop(); }).detach();
    }
    static void work(Job &j) {
        for (int i; (i = j.next.fetch_add(j.chunk)) < j.n;) {
            const int e = std::min(j.n, i + j.chunk);
            for (int k = i; k < e; ++k) (*j.f)(k);
            j.done.fetch_add(e - i, std::memory_order_release);
        }
    }
    void loop() {
        for (;;) {
            std::shared_ptr<Job> job;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [this] {
                    for (auto &j : jobs_)
                        if (j->next.load() < j->n) return true;
                    return false;
                });
                for (auto &j : jobs_)
                    if (j->next.load() < j->n) {
                        job = j;
                        break;
                    }
            }
            if (job) work(*job);
        }
    }
    unsigned workers_ = 0;
    std::mutex mu_;
    std::condition_variable cv_;
    std::deque<std::shared_ptr<Job>> jobs_;
};

// f(0) ... f(n - 1) on the host pool, in chunks of 8 indices taken in turn.
template <class F>
void parallel_for(int n, F f) {
    if (n < 64) {
        for (int i = 0; i < n; ++i) f(i);
        return;
    }
    const std::function<void(int)> g = [&](int i) { f(i); };
    HostPool::get().run(n, 8, g);
}

// Panels: runs of at most prows consecutive functions (rows) and every cell
// carrying one of them, ordered by Morton rank, cut into sub-batches of
// BMI_PATCH cells.  For each sub-batch and each row of its panel: the slots
// (local cell * 3 + shape, coefficient) of that row's function on the
// sub-batch's cells.
bmi_status build_panels(bmi_space *s, const int64_t *gidx, const double *coeff, const std::vector<int> &morton_rank) {
    const bmi_mesh *m = s->mesh;
    const int64_t nt = m->nt;
    const int R = s->prows;
    const int64_t nf = s->nfun;
    // panel cuts: runs of at most R consecutive functions.  An RWG panel also
    // ends before its cells would outnumber four sub-batches (a 64-row
    // panel carries about 60 cells where the numbering is compact, more
    // where it is not: the cut keeps its sub-batches full)
    s->prow.assign(1, 0);
    std::vector<int> pan_of((size_t)nf, 0);
    {
        const int cap = s->kind == BMI_SPACE_RWG ? 4 * BMI_PATCH : INT32_MAX;
        std::vector<int> fptr((size_t)nf + 1, 0), fcell;
        for (int64_t K = 0; K < nt; ++K)
            for (int i = 0; i < 3; ++i)
                if (gidx[3 * K + i] >= 0) fptr[(size_t)gidx[3 * K + i] + 1]++;
        for (int64_t f = 0; f < nf; ++f) fptr[(size_t)f + 1] += fptr[(size_t)f];
        fcell.resize((size_t)fptr[(size_t)nf]);
        {
            std::vector<int> pos(fptr.begin(), fptr.end() - 1);
            for (int64_t K = 0; K < nt; ++K)
                for (int i = 0; i < 3; ++i)
                    if (gidx[3 * K + i] >= 0) fcell[(size_t)pos[(size_t)gidx[3 * K + i]]++] = (int)K;
        }
        std::vector<int> stamp((size_t)nt, -1);
        int64_t r = 0;
        int P = 0;
        while (r < nf) {
            const int64_t r_first = r;
            int ncell = 0;
            while (r < nf && r - r_first < R) {
                int add = 0;
                for (int u = fptr[(size_t)r]; u < fptr[(size_t)r + 1]; ++u) add += stamp[(size_t)fcell[(size_t)u]] != P;
                if (r > r_first && ncell + add > cap) break;
                for (int u = fptr[(size_t)r]; u < fptr[(size_t)r + 1]; ++u) stamp[(size_t)fcell[(size_t)u]] = P;
                ncell += add;
                pan_of[(size_t)r] = P;
                ++r;
            }
            s->prow.push_back((int)r);
            ++P;
        }
    }
    const int np = (int)s->prow.size() - 1;
    s->npanel = np;
    std::vector<int> cnt(np + 1, 0);
    for (int64_t K = 0; K < nt; ++K) {
        int seen[3] = {-1, -1, -1};
        for (int i = 0; i < 3; ++i) {
            const int64_t f = gidx[3 * K + i];
            if (f < 0) continue;
            const int P = pan_of[(size_t)f];
            if (P != seen[0] && P != seen[1] && P != seen[2]) {
                seen[i] = P;
                cnt[P + 1]++;
            }
        }
    }
    for (int P = 0; P < np; ++P) cnt[P + 1] += cnt[P];
    std::vector<int> cells(cnt[np]), pos(cnt.begin(), cnt.end() - 1);
    for (int64_t K = 0; K < nt; ++K) {
        int seen[3] = {-1, -1, -1};
        for (int i = 0; i < 3; ++i) {
            const int64_t f = gidx[3 * K + i];
            if (f < 0) continue;
            const int P = pan_of[(size_t)f];
            if (P != seen[0] && P != seen[1] && P != seen[2]) {
                seen[i] = P;
                cells[pos[P]++] = (int)K;
            }
        }
    }
    // per panel (in parallel): its sub-batches' cells and their slot lists
    // by row; then concatenated in panel order
    struct Slot { int r, kl; double cf; };
    struct PanelOut {
        std::vector<int> cells, row_ptr;  // row_ptr relative to the panel's slots
        std::vector<unsigned char> kl;
        std::vector<double> cf;
    };
    std::vector<PanelOut> outs(np);
    parallel_for(np, [&](int P) {
        PanelOut &o = outs[P];
        std::sort(cells.begin() + cnt[P], cells.begin() + cnt[P + 1],
                  [&](int a, int b) { return morton_rank[a] < morton_rank[b]; });
        // Round-robin packing: the panel's cells, in Morton order, are dealt
        // over ceil(n/16) sub-batches (neighbours land in different ones);
        // a cell that would share a function of the panel with a cell already
        // in its sub-batch goes to the next sub-batch that has room and no
        // such conflict (one more sub-batch if none has).  So every row has at
        // most one cell per sub-batch (its rank in the sub-batch is 0).
        const int ncell = cnt[P + 1] - cnt[P];
        // rows of the panel as a 128-bit mask (two words)
        std::vector<std::vector<int>> bins;
        std::vector<std::array<uint64_t, 2>> used;
        for (int nb = std::max(1, (ncell + BMI_PATCH - 1) / BMI_PATCH);; ++nb) {
            bins.assign(nb, {});
            used.assign(nb, {0, 0});
            bool ok = true;
            for (int i = 0; i < ncell && ok; ++i) {
                const int K = cells[cnt[P] + i];
                std::array<uint64_t, 2> rows = {0, 0};
                for (int j = 0; j < 3; ++j) {
                    const int64_t f = gidx[3 * (int64_t)K + j];
                    if (f >= 0 && pan_of[(size_t)f] == P) {
                        const int r = (int)(f - s->prow[P]);
                        rows[r >> 6] |= 1ull << (r & 63);
                    }
                }
                ok = false;
                for (int d = 0; d < nb; ++d) {
                    const int b = (i + d) % nb;
                    if ((int)bins[b].size() < BMI_PATCH && !(used[b][0] & rows[0]) && !(used[b][1] & rows[1])) {
                        bins[b].push_back(K);
                        used[b][0] |= rows[0];
                        used[b][1] |= rows[1];
                        ok = true;
                        break;
                    }
                }
            }
            if (ok) break;
        }
        if (ncell == 0) bins.clear();
        std::vector<Slot> sl;
        for (size_t b = 0; b < bins.size(); ++b) {
            sl.clear();
            for (int k = 0; k < BMI_PATCH; ++k) {
                const int K = k < (int)bins[b].size() ? bins[b][k] : -1;
                o.cells.push_back(K);
                if (K < 0) continue;
                for (int i = 0; i < 3; ++i) {
                    const int64_t f = gidx[3 * (int64_t)K + i];
                    if (f < 0 || pan_of[(size_t)f] != P) continue;
                    sl.push_back({(int)(f - s->prow[P]), 3 * k + i, coeff[3 * (int64_t)K + i]});
                }
            }
            std::stable_sort(sl.begin(), sl.end(), [](const Slot &x, const Slot &y) {
                return x.r != y.r ? x.r < y.r : x.kl < y.kl;
            });
            size_t u = 0;
            for (int r = 0; r <= R; ++r) {
                o.row_ptr.push_back((int)o.kl.size());
                while (u < sl.size() && sl[u].r == r) {
                    o.kl.push_back((unsigned char)sl[u].kl);
                    o.cf.push_back(sl[u].cf);
                    ++u;
                }
            }
        }
    });
    s->sb_off.assign(np + 1, 0);
    s->sb_cells.clear();
    s->sb_row_ptr.clear();
    s->sb_slot_kl.clear();
    s->sb_slot_cf.clear();
    for (int P = 0; P < np; ++P) {
        const PanelOut &o = outs[P];
        const int base = (int)s->sb_slot_kl.size();
        s->sb_cells.insert(s->sb_cells.end(), o.cells.begin(), o.cells.end());
        for (int v : o.row_ptr) s->sb_row_ptr.push_back(base + v);
        s->sb_slot_kl.insert(s->sb_slot_kl.end(), o.kl.begin(), o.kl.end());
        s->sb_slot_cf.insert(s->sb_slot_cf.end(), o.cf.begin(), o.cf.end());
        s->sb_off[P + 1] = s->sb_off[P] + (int)(o.cells.size() / BMI_PATCH);
    }
    s->nsb = s->sb_off[np];
    s->sb_sph.assign((size_t)s->nsb * 4, 0.0);
    s->sb_amax.assign(s->nsb, 0.0);
    parallel_for(s->nsb, [&](int b) {
        group_sphere(m, &s->sb_cells[(size_t)b * BMI_PATCH], BMI_PATCH, &s->sb_sph[4 * b], &s->sb_amax[b]);
    });
    // fixed-size slot records per sub-batch
    std::vector<int> trs((size_t)s->nsb * (R + 1), 0);
    std::vector<unsigned char> tkl((size_t)s->nsb * BMI_MAX_ROWS, 0);
    std::vector<double> tcf((size_t)s->nsb * BMI_MAX_ROWS, 0.0);
    // rank-round records: row of each cell slot in its panel and its rank
    // among the sub-batch's slots of that row
    std::vector<int> tslot((size_t)s->nsb * BMI_MAX_ROWS, 0xff);
    std::vector<double> tscf((size_t)s->nsb * BMI_MAX_ROWS, 0.0);
    std::vector<int> sb_panel(s->nsb), sb_rounds(s->nsb, 1);
    for (int P = 0; P < np; ++P)
        for (int b = s->sb_off[P]; b < s->sb_off[P + 1]; ++b) sb_panel[b] = P;
    std::atomic<int> too_many_slots{0}, too_many_cells{0};
    parallel_for(s->nsb, [&](int b) {
        const int *rp = &s->sb_row_ptr[(size_t)b * (R + 1)];
        const int base = rp[0];
        if (rp[R] - base > BMI_MAX_ROWS) {
            too_many_slots = 1;
            return;
        }
        for (int r = 0; r <= R; ++r) trs[(size_t)b * (R + 1) + r] = rp[r] - base;
        for (int u = base; u < rp[R]; ++u) {
            tkl[(size_t)b * BMI_MAX_ROWS + (u - base)] = s->sb_slot_kl[u];
            tcf[(size_t)b * BMI_MAX_ROWS + (u - base)] = s->sb_slot_cf[u];
        }
        const int P = sb_panel[b];
        int seen[BMI_PANEL_ROWS_WIDE] = {0};
        for (int k = 0; k < BMI_PATCH; ++k) {
            const int K = s->sb_cells[(size_t)b * BMI_PATCH + k];
            if (K < 0) continue;
            for (int i = 0; i < 3; ++i) {
                const int64_t f = gidx[3 * (int64_t)K + i];
                if (f < 0 || pan_of[(size_t)f] != P) continue;
                const int r = (int)(f - s->prow[P]);
                const int rank = seen[r]++;
                if (rank > 255) too_many_cells = 1;
                tslot[(size_t)b * BMI_MAX_ROWS + 3 * k + i] = r | ((rank & 0xff) << 8);
                tscf[(size_t)b * BMI_MAX_ROWS + 3 * k + i] = coeff[3 * (int64_t)K + i];
                sb_rounds[b] = std::max(sb_rounds[b], rank + 1);
            }
        }
    });
    if (too_many_slots) return fail(BMI_EINVAL, "more than 48 slots in a sub-batch");
    if (too_many_cells) return fail(BMI_EUNSUPPORTED, "function with more than 256 cells");
    s->rounds_t = 1;
    for (int b = 0; b < s->nsb; ++b) s->rounds_t = std::max(s->rounds_t, sb_rounds[b]);
    // home sub-batch of each cell: the first one that holds it
    std::vector<int> home(nt, -1);
    for (int b = 0; b < s->nsb; ++b)
        for (int k = 0; k < BMI_PATCH; ++k) {
            const int K = s->sb_cells[(size_t)b * BMI_PATCH + k];
            if (K >= 0 && home[K] < 0) home[K] = b;
        }
    hipError_t e;
    if ((e = s->d_tslot.upload(tslot)) != hipSuccess || (e = s->d_tscf.upload(tscf)) != hipSuccess ||
        (e = s->d_home_sb.upload(home)) != hipSuccess)
        return fail(BMI_EDEVICE, std::string("panel upload: ") + hipGetErrorString(e));
    if ((e = s->d_prow.upload(s->prow)) != hipSuccess)
        return fail(BMI_EDEVICE, std::string("panel upload: ") + hipGetErrorString(e));
    if ((e = s->d_sb_off.upload(s->sb_off)) != hipSuccess || (e = s->d_sb_cells.upload(s->sb_cells)) != hipSuccess ||
        (e = s->d_sb_sph.upload(s->sb_sph)) != hipSuccess || (e = s->d_sb_amax.upload(s->sb_amax)) != hipSuccess ||
        (e = s->d_trs.upload(trs)) != hipSuccess || (e = s->d_tkl.upload(tkl)) != hipSuccess ||
        (e = s->d_tcf.upload(tcf)) != hipSuccess)
        return fail(BMI_EDEVICE, std::string("panel upload: ") + hipGetErrorString(e));
    return BMI_OK;
}

// Patches: runs of BMI_PATCH cells in Morton order of their centroids.
// Rows of a patch: the functions its cells carry, ascending.  Patches that
// share a function get different colours (greedy); the first-touch flag of a
// (patch, function) row is set on the patch with the smallest colour.
// Cells in Morton order of their centroids (ties by index): order[i] is the
// i-th cell, rank its inverse.
void morton_order(const bmi_mesh *m, std::vector<int> &order, std::vector<int> &morton_rank) {
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
    // (Morton key, cell) pairs sorted: the cells by key, ties by index
    std::vector<std::pair<uint64_t, int>> key(nt);
    for (int64_t k = 0; k < nt; ++k) {
        uint64_t c[3];
        for (int d = 0; d < 3; ++d) {
            double u = (m->cen[3 * k + d] - lo[d]) / ext;
            c[d] = (uint64_t)std::min(2097151.0, std::max(0.0, u * 2097151.0));
        }
        key[k] = {spread3(c[0]) | spread3(c[1]) << 1 | spread3(c[2]) << 2, (int)k};
    }
    std::sort(key.begin(), key.end());
    order.assign(nt, 0);
    morton_rank.assign(nt, 0);
    for (int64_t i = 0; i < nt; ++i) {
        order[i] = key[i].second;
        morton_rank[order[i]] = (int)i;
    }
}

bmi_status build_patches(bmi_space *s, const int64_t *gidx, const double *coeff, const std::vector<int> &order) {
    const bmi_mesh *m = s->mesh;
    const int64_t nt = m->nt;

    // runs of at most BMI_PATCH consecutive cells carrying at most
    // s->maxcols functions (bounds the LDS tile of the far kernel)
    s->patch_tri.clear();
    {
        // in_run[f] = number of the run that holds function f (runs count from 1)
        std::vector<int> in_run((size_t)s->nfun, 0);
        int run = 0, nfs = 0;
        int64_t i = 0;
        while (i < nt) {
            ++run;
            nfs = 0;
            int n = 0;
            while (i < nt && n < BMI_PATCH) {
                const int K = order[i];
                // the functions this cell would add to the run
                int64_t add[3];
                int na = 0;
                for (int j = 0; j < 3; ++j) {
                    const int64_t f = gidx[3 * (int64_t)K + j];
                    if (f >= 0 && in_run[f] != run && std::find(add, add + na, f) == add + na) add[na++] = f;
                }
                if (n > 0 && nfs + na > s->maxcols) break;
                for (int j = 0; j < na; ++j) in_run[add[j]] = run;
                nfs += na;
                s->patch_tri.push_back(K);
                ++n;
                ++i;
            }
            while (n++ < BMI_PATCH) s->patch_tri.push_back(-1);
        }
    }
    const int np = (int)(s->patch_tri.size() / BMI_PATCH);
    s->npatch = np;

    s->patch_sph.assign((size_t)np * 4, 0.0);
    s->patch_amax.assign(np, 0.0);
    parallel_for(np, [&](int p) {
        group_sphere(m, &s->patch_tri[(size_t)p * BMI_PATCH], BMI_PATCH, &s->patch_sph[4 * p], &s->patch_amax[p]);
    });

    // per patch (in parallel): its slots sorted by function; then
    // concatenated in patch order
    struct Slot { int64_t fn; int kl; double cf; };
    std::vector<std::vector<Slot>> psl(np);
    parallel_for(np, [&](int p) {
        std::vector<Slot> &sl = psl[p];
        for (int k = 0; k < BMI_PATCH; ++k) {
            const int K = s->patch_tri[(size_t)p * BMI_PATCH + k];
            if (K < 0) continue;
            for (int i = 0; i < 3; ++i) {
                const int64_t fn = gidx[3 * (int64_t)K + i];
                if (fn >= 0) sl.push_back({fn, 3 * k + i, coeff[3 * (int64_t)K + i]});
            }
        }
        std::stable_sort(sl.begin(), sl.end(), [](const Slot &a, const Slot &b) {
            return a.fn != b.fn ? a.fn < b.fn : a.kl < b.kl;
        });
    });
    s->row_ptr.assign(np + 1, 0);
    s->row_fn.clear();
    s->row_slot_ptr.clear();
    s->slot_kl.clear();
    s->slot_cf.clear();
    std::vector<char> covered(s->nfun, 0);
    for (int p = 0; p < np; ++p) {
        const std::vector<Slot> &sl = psl[p];
        for (size_t u = 0; u < sl.size(); ++u) {
            if (u == 0 || sl[u].fn != sl[u - 1].fn) {
                s->row_fn.push_back((int)sl[u].fn);
                s->row_slot_ptr.push_back((int)s->slot_kl.size());
                covered[sl[u].fn] = 1;
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

    std::vector<int> home(nt, -1);
    for (int p = 0; p < np; ++p)
        for (int k = 0; k < BMI_PATCH; ++k) {
            const int K = s->patch_tri[(size_t)p * BMI_PATCH + k];
            if (K >= 0) home[K] = p;
        }
    hipError_t e;
    if ((e = s->d_patch_tri.upload(s->patch_tri)) != hipSuccess || (e = s->d_home_patch.upload(home)) != hipSuccess ||
        (e = s->d_by_colour.upload(flat)) != hipSuccess || (e = s->d_patch_sph.upload(s->patch_sph)) != hipSuccess ||
        (e = s->d_patch_amax.upload(s->patch_amax)) != hipSuccess)
        return fail(BMI_EDEVICE, std::string("space upload: ") + hipGetErrorString(e));
    // fixed-size column records per patch
    std::vector<int> bnc(np, 0), bfn((size_t)np * BMI_MAX_ROWS, -1), bcs((size_t)np * (BMI_MAX_ROWS + 1), 0);
    std::vector<unsigned char> bfirst((size_t)np * BMI_MAX_ROWS, 0), bkl((size_t)np * BMI_MAX_ROWS, 0);
    std::vector<double> bcf((size_t)np * BMI_MAX_ROWS, 0.0);
    parallel_for(np, [&](int p) {
        const int r0 = s->row_ptr[p], nr = s->row_ptr[p + 1] - r0;
        const int base = s->row_slot_ptr[r0];
        bnc[p] = nr;
        for (int r = 0; r < nr; ++r) {
            bfn[(size_t)p * BMI_MAX_ROWS + r] = s->row_fn[r0 + r];
            bfirst[(size_t)p * BMI_MAX_ROWS + r] = s->row_first[r0 + r];
            bcs[(size_t)p * (BMI_MAX_ROWS + 1) + r] = s->row_slot_ptr[r0 + r] - base;
        }
        for (int r = nr; r <= BMI_MAX_ROWS; ++r)
            bcs[(size_t)p * (BMI_MAX_ROWS + 1) + r] = s->row_slot_ptr[r0 + nr] - base;
        for (int u = base; u < s->row_slot_ptr[r0 + nr]; ++u) {
            bkl[(size_t)p * BMI_MAX_ROWS + (u - base)] = s->slot_kl[u];
            bcf[(size_t)p * BMI_MAX_ROWS + (u - base)] = s->slot_cf[u];
        }
    });
    // rank-round records: column of each cell slot in its patch (bits 0-7),
    // its rank among the patch's slots of that column (bits 8-15) and its
    // level, the rank among the slots of that column with the same local
    // shape index (bits 16-23).  Rounds by rank touch an entry at most once
    // per round; rounds by level may touch it once per local shape index, in
    // shape order.  Lagrange functions span many cells of a patch with
    // different local indices, so they need far fewer levels than ranks.
    std::vector<int> bslot((size_t)np * BMI_MAX_ROWS, 0xff);
    std::vector<double> bscf((size_t)np * BMI_MAX_ROWS, 0.0);
    std::vector<unsigned char> brounds(2 * (size_t)np, 1);
    parallel_for(np, [&](int p) {
        const int r0 = s->row_ptr[p], nr = s->row_ptr[p + 1] - r0;
        int seen[BMI_MAX_ROWS] = {0}, seen_j[BMI_MAX_ROWS][3] = {};
        int rk = 1, lv = 1;
        for (int l = 0; l < BMI_PATCH; ++l) {
            const int L = s->patch_tri[(size_t)p * BMI_PATCH + l];
            if (L < 0) continue;
            for (int j = 0; j < 3; ++j) {
                const int64_t f = gidx[3 * (int64_t)L + j];
                if (f < 0) continue;
                const int c = (int)(std::lower_bound(s->row_fn.begin() + r0, s->row_fn.begin() + r0 + nr, (int)f) -
                                    (s->row_fn.begin() + r0));
                const int rank = seen[c]++, level = seen_j[c][j]++;
                bslot[(size_t)p * BMI_MAX_ROWS + 3 * l + j] = c | (rank << 8) | (level << 16);
                bscf[(size_t)p * BMI_MAX_ROWS + 3 * l + j] = coeff[3 * (int64_t)L + j];
                rk = std::max(rk, rank + 1);
                lv = std::max(lv, level + 1);
            }
        }
        brounds[p] = (unsigned char)rk;
        brounds[np + p] = (unsigned char)lv;
    });
    s->rounds_b = 1;
    s->levels_b = 1;
    s->ncmax = 1;
    for (int p = 0; p < np; ++p) {
        s->ncmax = std::max(s->ncmax, s->row_ptr[p + 1] - s->row_ptr[p]);
        s->rounds_b = std::max(s->rounds_b, (int)brounds[p]);
        s->levels_b = std::max(s->levels_b, (int)brounds[np + p]);
    }
    if ((e = s->d_bslot.upload(bslot)) != hipSuccess || (e = s->d_bscf.upload(bscf)) != hipSuccess ||
        (e = s->d_brounds.upload(brounds)) != hipSuccess)
        return fail(BMI_EDEVICE, std::string("space upload: ") + hipGetErrorString(e));
    if ((e = s->d_bnc.upload(bnc)) != hipSuccess || (e = s->d_bfn.upload(bfn)) != hipSuccess ||
        (e = s->d_bcs.upload(bcs)) != hipSuccess || (e = s->d_bfirst.upload(bfirst)) != hipSuccess ||
        (e = s->d_bkl.upload(bkl)) != hipSuccess || (e = s->d_bcf.upload(bcf)) != hipSuccess)
        return fail(BMI_EDEVICE, std::string("space upload: ") + hipGetErrorString(e));
    return BMI_OK;
}

}  // namespace

hipStream_t context_stream(const bmi_context *ctx) { return ctx->stream; }

// y = Z x on run's stream, with run's workspace (grown when too small).
bmi_status matrix_apply_on(bmi_context *run, const bmi_matrix *m, const double *x, double *y, int32_t nrhs) {
    if (!run || !m || !x || !y) return fail(BMI_EINVAL, "NULL argument");
    if (nrhs < 1 || nrhs > BMI_MAX_RHS) return fail(BMI_EINVAL, "nrhs must be 1..8");
    if (m->rows > (1LL << 31) - 1 || m->cols > (1LL << 31) - 1) return fail(BMI_EUNSUPPORTED, "matrix too large");
    if (run->device != m->ctx->device) return fail(BMI_EINVAL, "matrix and context are on different devices");
    BMI_HIP(hipSetDevice(run->device));
    hipStream_t s = run->stream;
    WorkStream wsc(s);
    const int rows = (int)m->rows, cols = (int)m->cols;
    int chunk = 0;
    const int nchunk = apply_chunks(rows, cols, &chunk);
    const size_t nx = (size_t)cols * nrhs, ny = (size_t)rows * nrhs, np = (size_t)nchunk * rows * nrhs;
    const size_t need = sizeof(double2) * (nx + ny + np);
    if (need > run->ws_bytes) {
        dev_free(run->ws);
        run->ws = nullptr;
        run->ws_bytes = 0;
        hipError_t e = dev_alloc(&run->ws, need);
        if (e != hipSuccess) return fail(BMI_ENOMEM, std::string("apply workspace: ") + hipGetErrorString(e));
        run->ws_bytes = need;
    }
    double2 *dx = (double2 *)run->ws, *dy = dx + nx, *part = dy + ny;
    BMI_HIP(hipMemcpyAsync(dx, x, sizeof(double2) * nx, hipMemcpyHostToDevice, s));
    BMI_HIP(launch_apply(m->d, m->dtype == BMI_C32, m->rows, rows, cols, dx, nrhs, part, dy, s));
    BMI_HIP(hipMemcpyAsync(y, dy, sizeof(double2) * ny, hipMemcpyDeviceToHost, s));
    BMI_HIP(hipStreamSynchronize(s));
    return BMI_OK;
}

MatrixView matrix_view(const bmi_matrix *m) {
    return MatrixView{m->d, m->rows, m->cols, m->dtype, m->device, m->ctx->stream};
}

// A matrix over caller-owned device memory of at least rows*cols elements.
bmi_status matrix_wrap(bmi_context *ctx, void *mem, int64_t rows, int64_t cols, int32_t dtype, bmi_matrix **out) {
    if (!ctx || !mem || !out) return fail(BMI_EINVAL, "NULL argument");
    auto m = std::make_unique<bmi_matrix>();
    m->ctx = ctx;
    m->device = ctx->device;
    m->rows = rows;
    m->cols = cols;
    m->dtype = dtype;
    m->d = mem;
    m->owned = false;
    *out = m.release();
    return BMI_OK;
}

// The matrix must have the plan's shape and context.
bmi_status check_target(const bmi_plan *P, const bmi_matrix *out) {
    if (out->ctx != P->ctx) return fail(BMI_EINVAL, "matrix belongs to another context");
    if (out->rows != P->X->nfun || out->cols != P->Y->nfun)
        return fail(BMI_EINVAL, "matrix is " + std::to_string(out->rows) + " x " + std::to_string(out->cols) +
                                    ", the plan assembles " + std::to_string(P->X->nfun) + " x " +
                                    std::to_string(P->Y->nfun));
    return BMI_OK;
}

// 1 if the far kernels' compile-time rule applies: default 3 x 4 point
// rule, lossless operator, alpha != 0 (those kernels accumulate entries
// divided by alpha).
int fast_rule(const bmi_plan *P) {
    return P->rule_default && !P->opc.lossy && (P->opc.alpha_re != 0.0 || P->opc.alpha_im != 0.0);
}

// 1 if the plan's Sauter-Schwab rules all have 5 points per dimension and the
// same nodes and weights (the compile-time-order kernel takes one table).
int gl5(const bmi_plan *P) {
    const SpecialArgs &S = P->sargs;
    for (int g = 0; g < 3; ++g) {
        if (S.ss_n[g] != 5) return 0;
        for (int i = 0; i < 5; ++i)
            if (S.glx[g][i] != S.glx[0][i] || S.glw[g][i] != S.glw[0][i]) return 0;
    }
    return 1;
}

// The default far rules as tables: 3 points on test cells, 4 on trial cells
// (in the order of the kernels' compile-time rule).
bool rules_are_default(const bmi_rules &r) {
    int n = 0, m = 0;
    double l[BMI_MAX_RULE][3], w[BMI_MAX_RULE], lb[BMI_MAX_RULE][3], wb[BMI_MAX_RULE];
    far_rule(2, n, l, w);
    far_rule(3, m, lb, wb);
    if (r.test_points != n || r.trial_points != m) return false;
    for (int i = 0; i < n; ++i)
        if (r.test_uv[i][0] != l[i][0] || r.test_uv[i][1] != l[i][1] || r.test_weight[i] != w[i]) return false;
    for (int i = 0; i < m; ++i)
        if (r.trial_uv[i][0] != lb[i][0] || r.trial_uv[i][1] != lb[i][1] || r.trial_weight[i] != wb[i]) return false;
    return true;
}

// true if a vertex of mesh a lies inside a cell of mesh b (within a relative
// tolerance, on its plane and in its closure) without being one of that
// cell's vertices: the meshes overlap, as when one refines the other.
// Candidate cells come from a uniform grid over the cells' bounding boxes.
bool vertex_inside_other(const bmi_mesh *a, const bmi_mesh *b) {
    const int64_t nt = b->nt;
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int64_t k = 0; k < nt; ++k)
        for (int i = 0; i < 3; ++i)
            for (int d = 0; d < 3; ++d) {
                lo[d] = std::min(lo[d], b->vtx[9 * k + 3 * i + d]);
                hi[d] = std::max(hi[d], b->vtx[9 * k + 3 * i + d]);
            }
    const double h = std::max(b->hmax, 1e-300);
    const double tol = 1e-9 * h;
    int64_t dims[3];
    for (int d = 0; d < 3; ++d) dims[d] = std::max<int64_t>(1, std::min<int64_t>(1 << 10, (int64_t)((hi[d] - lo[d]) / h) + 1));
    auto bin = [&](double x, int d) {
        return std::max<int64_t>(0, std::min<int64_t>(dims[d] - 1, (int64_t)((x - lo[d]) / h)));
    };
    std::vector<std::pair<int64_t, int64_t>> ent;  // (bin, cell)
    for (int64_t k = 0; k < nt; ++k) {
        int64_t b0[3], b1[3];
        for (int d = 0; d < 3; ++d) {
            double mn = 1e300, mx = -1e300;
            for (int i = 0; i < 3; ++i) {
                mn = std::min(mn, b->vtx[9 * k + 3 * i + d]);
                mx = std::max(mx, b->vtx[9 * k + 3 * i + d]);
            }
            b0[d] = bin(mn - tol, d);
            b1[d] = bin(mx + tol, d);
        }
        for (int64_t x = b0[0]; x <= b1[0]; ++x)
            for (int64_t y = b0[1]; y <= b1[1]; ++y)
                for (int64_t z = b0[2]; z <= b1[2]; ++z) ent.push_back({(x * dims[1] + y) * dims[2] + z, k});
    }
    std::sort(ent.begin(), ent.end());
    for (int64_t K = 0; K < a->nt; ++K)
        for (int i = 0; i < 3; ++i) {
            const double *p = &a->vtx[9 * K + 3 * i];
            if (p[0] < lo[0] - tol || p[0] > hi[0] + tol || p[1] < lo[1] - tol || p[1] > hi[1] + tol ||
                p[2] < lo[2] - tol || p[2] > hi[2] + tol)
                continue;
            const int64_t key = (bin(p[0], 0) * dims[1] + bin(p[1], 1)) * dims[2] + bin(p[2], 2);
            auto it = std::lower_bound(ent.begin(), ent.end(), std::make_pair(key, (int64_t)-1));
            for (; it != ent.end() && it->first == key; ++it) {
                const double *v = &b->vtx[9 * it->second];
                bool at_vertex = false;
                for (int j = 0; j < 3; ++j) {
                    const double dx = p[0] - v[3 * j], dy = p[1] - v[3 * j + 1], dz = p[2] - v[3 * j + 2];
                    at_vertex |= dx * dx + dy * dy + dz * dz <= tol * tol;
                }
                if (at_vertex) continue;
                // barycentric coordinates of p's projection, distance to the plane
                double e1[3], e2[3], q[3];
                for (int d = 0; d < 3; ++d) {
                    e1[d] = v[d] - v[6 + d];
                    e2[d] = v[3 + d] - v[6 + d];
                    q[d] = p[d] - v[6 + d];
                }
                const double g11 = e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2];
                const double g12 = e1[0] * e2[0] + e1[1] * e2[1] + e1[2] * e2[2];
                const double g22 = e2[0] * e2[0] + e2[1] * e2[1] + e2[2] * e2[2];
                const double r1 = q[0] * e1[0] + q[1] * e1[1] + q[2] * e1[2];
                const double r2 = q[0] * e2[0] + q[1] * e2[1] + q[2] * e2[2];
                const double det = g11 * g22 - g12 * g12;
                if (!(det > 0.0)) continue;
                const double u = (g22 * r1 - g12 * r2) / det, w = (g11 * r2 - g12 * r1) / det;
                double dist2 = 0.0;
                for (int d = 0; d < 3; ++d) {
                    const double o = q[d] - u * e1[d] - w * e2[d];
                    dist2 += o * o;
                }
                const double bt = 1e-9;
                if (dist2 <= tol * tol && u >= -bt && w >= -bt && u + w <= 1.0 + bt) return true;
            }
        }
    return false;
}

// Classification, compaction and special pairs of one assembly on stream s,
// no host synchronisation (graph-capturable).  ev_mid (optional) is recorded
// after the classification.
bmi_status enqueue_prologue(bmi_plan *P, bmi_matrix *out, hipStream_t s, hipEvent_t ev_mid) {
    BMI_HIP(launch_zero_ints(P->d_total.p, 1, s));
    BMI_HIP(launch_classify(P->d_cargs.p, P->d_gbase.p, 1, P->npa, s));
    P->filled = true;
    if (ev_mid) BMI_HIP(hipEventRecord(ev_mid, s));
    const int mixed = P->opts.precision == BMI_PREC_MIXED;
    BMI_HIP(launch_special(P->d_sargs.p, P->d_cbase.p, P->d_ebase.p, 1, P->nchunks, P->capacity, P->cs, mixed,
                           P->op_kind, gl5(P), fast_rule(P), s));
    // functions without local shapes keep zero rows/columns (a fill kernel:
    // memset nodes misbehave in replayed graphs); out == nullptr: the caller
    // fills its own target
    if (out && (P->X->uncovered || P->Y->uncovered))
        BMI_HIP(launch_fill_zero(out->d, (size_t)out->rows * out->cols * (out->dtype == BMI_C64 ? 16 : 8), s));
    return BMI_OK;
}

// Arguments of the far-pair kernels of plan P writing into the column-major
// matrix Z with leading dimension ld (global row r of column c at
// Z[r + c ld]).
PanelArgs panel_args(const bmi_plan *P, void *Z, long long ld) {
    PanelArgs B{};
    B.npanel = P->X->npanel;
    B.panel0 = 0;
    B.nrows = (int)P->X->nfun;
    B.prow = P->X->d_prow.p;
    B.sb_off = P->X->d_sb_off.p;
    B.tgeo = P->d_tgeo.p;
    B.trs = P->X->d_trs.p;
    B.tkl = P->X->d_tkl.p;
    B.tcf = P->X->d_tcf.p;
    B.bgeo = P->d_bgeo.p;
    B.bnc = P->Y->d_bnc.p;
    B.bfn = P->Y->d_bfn.p;
    B.bfirst = P->Y->d_bfirst.p;
    B.bcs = P->Y->d_bcs.p;
    B.bkl = P->Y->d_bkl.p;
    B.bcf = P->Y->d_bcf.p;
    B.tslot = P->X->d_tslot.p;
    B.tscf = P->X->d_tscf.p;
    B.bslot = P->Y->d_bslot.p;
    B.bscf = P->Y->d_bscf.p;
    B.rounds_t = P->X->rounds_t;
    B.fast = fast_rule(P);
    // trial rounds by level: the Helmholtz kernels and the kernels with the
    // compile-time rule; the others by rank (must match panel_block)
    const bool by_level = P->op_kind == OPK_HH3D || B.fast;
    B.rounds_b = by_level ? P->Y->levels_b : P->Y->rounds_b;
    B.brounds = P->Y->d_brounds.p + (by_level ? P->Y->npatch : 0);
    B.ncm = P->Y->ncmax;
    // Lagrange test spaces and panels of other heights than 64 rows: only
    // the rank-round kernels handle them
    B.rr = P->X->kind == BMI_SPACE_LAGRANGE_C0D1 || P->X->prows != BMI_PANEL_ROWS ||
           P->X->rounds_t * B.rounds_b <= 6;
    B.npb = P->npb;
    B.bq_all = P->Y->d_by_colour.p;
    B.blk_off = P->d_blk_off.p;
    B.blk_cnt = P->d_blk_cnt.p;
    B.spec_loc = P->d_spec_loc.p;
    B.spec_val = P->d_spec_val.p;
    B.capacity = P->capacity;
    B.sc = P->d_sc.p;
    B.Z = Z;
    B.ld = ld;
    B.op = P->opc;
    B.rule = P->rule;
    return B;
}

PanelArgs panel_args(const bmi_plan *P, const bmi_matrix *out) { return panel_args(P, out->d, out->rows); }

// The far-pair launches of test panels [p0, p1) of one assembly, one per
// trial colour, on stream s, into Z (element type dtype, leading dimension
// ld, global rows).
bmi_status enqueue_far_range(bmi_plan *P, void *Z, long long ld, int dtype, int p0, int p1, hipStream_t s,
                             int *launches) {
    PanelArgs B = panel_args(P, Z, ld);
    B.panel0 = p0;
    B.npanel = p1 - p0;
    const int mixed = P->opts.precision == BMI_PREC_MIXED;
    for (int cb = 0; cb < P->Y->ncolour; ++cb) {
        B.bq_list = P->Y->d_by_colour.p + P->Y->colour_off[cb];
        B.nbq = P->Y->colour_off[cb + 1] - P->Y->colour_off[cb];
        BMI_HIP(launch_panels(B, P->op_kind, mixed, dtype == BMI_C32, s));
        ++*launches;
    }
    return BMI_OK;
}

// The far-pair launches of one assembly, one per trial colour, on stream s.
bmi_status enqueue_far(bmi_plan *P, bmi_matrix *out, hipStream_t s, int *launches) {
    *launches = 0;
    return enqueue_far_range(P, out->d, out->rows, out->dtype, 0, P->X->npanel, s, launches);
}

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

bmi_status bmi_context_create(int32_t device, bmi_context **out) try {
    if (!out) return fail(BMI_EINVAL, "out is NULL");
    int n = 0;
    BMI_HIP(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(BMI_EINVAL, "no such device");
    BMI_HIP(hipSetDevice(device));
    auto ctx = std::make_unique<bmi_context>();
    ctx->device = device;
    BMI_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    // once per device: the address of the allocation-free apply's scratch
    {
        const bmi_status st = scratch_setup(device);
        if (st != BMI_OK) {
            (void)hipStreamDestroy(ctx->stream);
            return st;
        }
    }
    // keep freed pool memory cached for the next allocation
    hipMemPool_t pool = nullptr;
    if (hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess && pool) {
        uint64_t keep = UINT64_MAX;
        (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
    }
    *out = ctx.release();
    return BMI_OK;
} catch (...) {
    return fail_exception();
}

bmi_status bmi_device_memory(const bmi_context *ctx, int64_t *free_bytes, int64_t *total_bytes) try {
    if (!ctx || !free_bytes || !total_bytes) return fail(BMI_EINVAL, "NULL argument");
    BMI_HIP(hipSetDevice(ctx->device));
    size_t f = 0, t = 0;
    BMI_HIP(hipMemGetInfo(&f, &t));
    f += pool_cached_bytes(ctx->device);  // freed blocks the pool keeps for reuse
    *free_bytes = (int64_t)f;
    *total_bytes = (int64_t)t;
    return BMI_OK;
} catch (...) {
    return fail_exception();
}

bmi_status bmi_context_destroy(bmi_context *ctx) try {
    if (!ctx) return BMI_OK;
    (void)hipSetDevice(ctx->device);
    for (int i = 0; i < 2; ++i) {
        if (ctx->stage[i]) (void)hipHostFree(ctx->stage[i]);
        if (ctx->stage_ev[i]) (void)hipEventDestroy(ctx->stage_ev[i]);
        if (ctx->blk_ev[i]) (void)hipEventDestroy(ctx->blk_ev[i]);
    }
    if (ctx->copy_stream) (void)hipStreamDestroy(ctx->copy_stream);
    dev_free(ctx->ws);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return BMI_OK;
} catch (...) {
    return fail_exception();
}

bmi_status bmi_mesh_create(bmi_context *ctx, const double *xyz, int64_t nv, const int32_t *tri, int64_t nt,
                           bmi_mesh **out) try {
    if (!ctx || !xyz || !tri || !out) return fail(BMI_EINVAL, "NULL argument");
    if (nv < 3 || nt < 1 || nt > (1LL << 30)) return fail(BMI_EINVAL, "bad mesh size");
    BMI_HIP(hipSetDevice(ctx->device));
    WorkStream ws(ctx->stream);
    auto m = std::make_unique<bmi_mesh>();
    m->ctx = ctx;
    m->device = ctx->device;
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
} catch (...) {
    return fail_exception();
}

bmi_status bmi_mesh_free(bmi_mesh *m) try {
    if (m) {
        (void)hipSetDevice(m->device);
        delete m;
    }
    return BMI_OK;
} catch (...) {
    return fail_exception();
}

bmi_status bmi_space_create(bmi_mesh *mesh, int32_t kind, const int64_t *gidx, const double *coeff, int64_t nfun,
                            bmi_space **out) {
    return bmi_space_create_shaped(mesh, kind, gidx, coeff, nfun, 0, 0, out);
}

bmi_status bmi_space_create_shaped(bmi_mesh *mesh, int32_t kind, const int64_t *gidx, const double *coeff,
                                   int64_t nfun, int32_t panel_rows, int32_t patch_cols, bmi_space **out) try {
    if (!mesh || !gidx || !coeff || !out) return fail(BMI_EINVAL, "NULL argument");
    if (kind != BMI_SPACE_LAGRANGE_C0D1 && kind != BMI_SPACE_RWG) return fail(BMI_EUNSUPPORTED, "unknown space kind");
    // the far kernels' tile: rows up to the kernel's panel height, columns up
    // to its widest patch
    const bool lagrange = kind == BMI_SPACE_LAGRANGE_C0D1;
    const int max_rows = lagrange ? BMI_PANEL_ROWS_WIDE : BMI_PANEL_ROWS;
    const int max_cols = lagrange ? BMI_MAX_COLS_WIDE : BMI_MAX_COLS;
    if (panel_rows < 0 || panel_rows > max_rows)
        return fail(BMI_EINVAL, "panel_rows out of range (1.." + std::to_string(max_rows) + ", 0 = default)");
    if (patch_cols != 0 && (patch_cols < 3 || patch_cols > max_cols))
        return fail(BMI_EINVAL, "patch_cols out of range (3.." + std::to_string(max_cols) + ", 0 = default)");
    if (nfun < 1 || nfun > (1LL << 31) - 1) return fail(BMI_EINVAL, "bad number of functions");
    for (int64_t i = 0; i < 3 * mesh->nt; ++i) {
        if (gidx[i] >= nfun) return fail(BMI_EINVAL, "function index out of range");
        if (gidx[i] >= 0 && !std::isfinite(coeff[i])) return fail(BMI_EINVAL, "non-finite coefficient");
    }
    BMI_HIP(hipSetDevice(mesh->ctx->device));
    WorkStream ws(mesh->ctx->stream);
    auto s = std::make_unique<bmi_space>();
    s->mesh = mesh;
    s->device = mesh->device;
    s->kind = kind;
    s->nfun = nfun;
    s->prows = panel_rows ? panel_rows : max_rows;
    s->maxcols = patch_cols ? patch_cols : max_cols;
    // the test-role panels and the trial-role patches are built side by side
    // (they share the Morton order and write disjoint parts of the space)
    std::vector<int> order, rank;
    morton_order(mesh, order, rank);
    bmi_status st_pan = BMI_OK;
    std::string msg_pan;
    const hipStream_t stream = mesh->ctx->stream;
    const int dev = mesh->ctx->device;
    std::thread panels([&] {
        try {
            (void)hipSetDevice(dev);
            WorkStream w(stream);
            st_pan = build_panels(s.get(), gidx, coeff, rank);
        } catch (...) {
            st_pan = fail_exception();
        }
        if (st_pan != BMI_OK) msg_pan = bmi_last_error();
    });
    bmi_status st = BMI_OK;
    try {
        st = build_patches(s.get(), gidx, coeff, order);
    } catch (...) {
        panels.join();
        throw;
    }
    panels.join();
    if (st != BMI_OK) return st;
    if (st_pan != BMI_OK) return set_error(st_pan, msg_pan);
    *out = s.release();
    return BMI_OK;
} catch (...) {
    return fail_exception();
}

bmi_status bmi_space_free(bmi_space *s) try {
    if (s) {
        (void)hipSetDevice(s->device);
        delete s;
    }
    return BMI_OK;
} catch (...) {
    return fail_exception();
}

bmi_status bmi_matrix_create(bmi_context *ctx, int64_t rows, int64_t cols, int32_t dtype, bmi_matrix **out) try {
    if (!ctx || !out) return fail(BMI_EINVAL, "NULL argument");
    if (rows < 1 || cols < 1) return fail(BMI_EINVAL, "bad matrix size");
    if (dtype != BMI_C64 && dtype != BMI_C32) return fail(BMI_EINVAL, "bad dtype");
    BMI_HIP(hipSetDevice(ctx->device));
    WorkStream ws(ctx->stream);
    const size_t es = dtype == BMI_C64 ? 16 : 8;
    const size_t bytes = (size_t)rows * (size_t)cols * es;
    auto m = std::make_unique<bmi_matrix>();
    m->ctx = ctx;
    m->device = ctx->device;
    m->rows = rows;
    m->cols = cols;
    m->dtype = dtype;
    hipError_t e = dev_alloc(&m->d, bytes);
    if (e != hipSuccess) return fail(BMI_ENOMEM, std::string("matrix allocation: ") + hipGetErrorString(e));
    *out = m.release();
    return BMI_OK;
} catch (...) {
    return fail_exception();
}

bmi_status bmi_matrix_free(bmi_matrix *m) try {
    if (m) {
        (void)hipSetDevice(m->device);
        if (m->owned) dev_free(m->d);
        delete m;
    }
    return BMI_OK;
} catch (...) {
    return fail_exception();
}

bmi_status bmi_matrix_device_ptr(const bmi_matrix *m, void **ptr, int64_t *ld) try {
    if (!m) return fail(BMI_EINVAL, "NULL matrix");
    if (ptr) *ptr = m->d;
    if (ld) *ld = m->rows;
    return BMI_OK;
} catch (...) {
    return fail_exception();
}

static bmi_status plan_create(bmi_context *ctx, const bmi_operator *op, const bmi_space *test,
                              const bmi_space *trial, const bmi_options *opts, const bmi_rules *rules,
                              bmi_plan **out);

bmi_status bmi_plan_create(bmi_context *ctx, const bmi_operator *op, const bmi_space *test, const bmi_space *trial,
                           const bmi_options *opts, bmi_plan **out) try {
    return plan_create(ctx, op, test, trial, opts, nullptr, out);
} catch (...) {
    return fail_exception();
}

bmi_status bmi_plan_create_rules(bmi_context *ctx, const bmi_operator *op, const bmi_space *test,
                                 const bmi_space *trial, const bmi_options *opts, const bmi_rules *rules,
                                 bmi_plan **out) try {
    if (!rules) return fail(BMI_EINVAL, "NULL rules");
    return plan_create(ctx, op, test, trial, opts, rules, out);
} catch (...) {
    return fail_exception();
}

static bmi_status plan_create(bmi_context *ctx, const bmi_operator *op, const bmi_space *test,
                              const bmi_space *trial, const bmi_options *opts, const bmi_rules *rules,
                              bmi_plan **out) {
    if (!ctx || !op || !test || !trial || !out) return fail(BMI_EINVAL, "NULL argument");
    if (test->mesh->ctx != ctx || trial->mesh->ctx != ctx) return fail(BMI_EINVAL, "spaces belong to another context");
    bmi_options o;
    bmi_options_default(&o);
    if (opts) o = *opts;
    if (op->kind == BMI_OP_HH3D_SINGLELAYER) {
        if (test->kind != BMI_SPACE_LAGRANGE_C0D1 || trial->kind != BMI_SPACE_LAGRANGE_C0D1)
            return fail(BMI_EUNSUPPORTED, "the Helmholtz single layer is assembled on lagrangec0d1 spaces");
    } else if (op->kind == BMI_OP_MW_SINGLELAYER || op->kind == BMI_OP_MW_DOUBLELAYER) {
        if (test->kind != BMI_SPACE_RWG || trial->kind != BMI_SPACE_RWG)
            return fail(BMI_EUNSUPPORTED, "the Maxwell operators are assembled on RWG spaces");
    } else {
        return fail(BMI_EUNSUPPORTED, "unknown operator kind");
    }
    if (rules) {
        if (rules->test_points < 1 || rules->test_points > BMI_MAX_RULE || rules->trial_points < 1 ||
            rules->trial_points > BMI_MAX_RULE)
            return fail(BMI_EINVAL, "rule points must be 1..4 per cell");
        for (int i = 0; i < BMI_MAX_RULE; ++i) {
            const double v[6] = {rules->test_uv[i][0], rules->test_uv[i][1], rules->test_weight[i],
                                 rules->trial_uv[i][0], rules->trial_uv[i][1], rules->trial_weight[i]};
            for (int j = 0; j < 6; ++j)
                if ((j < 3 ? i < rules->test_points : i < rules->trial_points) && !std::isfinite(v[j]))
                    return fail(BMI_EINVAL, "non-finite rule point or weight");
        }
        for (int g = 0; g < 3; ++g) {
            if (rules->gl_points[g] < 1 || rules->gl_points[g] > BMI_MAX_GL)
                return fail(BMI_EINVAL, "Gauss-Legendre points must be 1..8");
            for (int i = 0; i < rules->gl_points[g]; ++i)
                if (!(rules->gl_node[g][i] >= 0.0 && rules->gl_node[g][i] <= 1.0) ||
                    !std::isfinite(rules->gl_weight[g][i]))
                    return fail(BMI_EINVAL, "Gauss-Legendre nodes must lie in [0, 1] with finite weights");
        }
    } else {
        for (int i = 0; i < 3; ++i)
            if (o.ss_order[i] < 1 || o.ss_order[i] > BMI_MAX_GL) return fail(BMI_EINVAL, "ss_order must be 1..8");
    }
    if (o.near_order < 1 || o.near_order > 8) return fail(BMI_EINVAL, "near_order must be 1..8");
    if (o.precision != BMI_PREC_FP64 && o.precision != BMI_PREC_MIXED) return fail(BMI_EINVAL, "bad precision");
    if (!(o.near_factor >= 0.0)) return fail(BMI_EINVAL, "bad near_factor");
    const double vals[6] = {op->alpha[0], op->alpha[1], op->beta[0], op->beta[1], op->gamma[0], op->gamma[1]};
    for (double v : vals)
        if (!std::isfinite(v)) return fail(BMI_EINVAL, "non-finite operator coefficient");

    if (test->mesh != trial->mesh &&
        (vertex_inside_other(test->mesh, trial->mesh) || vertex_inside_other(trial->mesh, test->mesh)))
        return fail(BMI_EUNSUPPORTED, "the test and trial meshes overlap (a vertex of one lies inside a cell of the "
                                      "other): refinement-aware quadrature is not supported");
    BMI_HIP(hipSetDevice(ctx->device));
    WorkStream ws(ctx->stream);
    auto P = std::make_unique<bmi_plan>();
    P->ctx = ctx;
    P->device = ctx->device;
    P->X = test;
    P->Y = trial;
    P->op = *op;
    P->opts = o;
    P->op_kind = op->kind == BMI_OP_MW_SINGLELAYER ? OPK_MW : (op->kind == BMI_OP_MW_DOUBLELAYER ? OPK_DL : OPK_HH3D);
    P->opc.alpha_re = op->alpha[0];
    P->opc.alpha_im = op->alpha[1];
    P->opc.beta_re = op->beta[0];
    P->opc.beta_im = op->beta[1];
    P->opc.gamma_re = op->gamma[0];
    P->opc.k2pi = op->gamma[1] / (2.0 * kPi);
    P->opc.lossy = op->gamma[0] != 0.0;
    {
        const double an = op->alpha[0] * op->alpha[0] + op->alpha[1] * op->alpha[1];
        if (an > 0.0) {
            P->opc.ia_re = op->alpha[0] / an;
            P->opc.ia_im = -op->alpha[1] / an;
            // 4 beta / alpha = 4 beta conj(alpha) / |alpha|^2
            P->opc.bp_re = 4.0 * (op->beta[0] * op->alpha[0] + op->beta[1] * op->alpha[1]) / an;
            P->opc.bp_im = 4.0 * (op->beta[1] * op->alpha[0] - op->beta[0] * op->alpha[1]) / an;
        }
    }

    double lt[BMI_MAX_RULE][3] = {}, lb[BMI_MAX_RULE][3] = {}, wt[BMI_MAX_RULE] = {}, wb[BMI_MAX_RULE] = {};
    int nt = 0, nb = 0;
    if (rules) {
        nt = rules->test_points;
        nb = rules->trial_points;
        for (int i = 0; i < nt; ++i) {
            lt[i][0] = rules->test_uv[i][0];
            lt[i][1] = rules->test_uv[i][1];
            lt[i][2] = 1.0 - lt[i][0] - lt[i][1];
            wt[i] = rules->test_weight[i];
        }
        for (int i = 0; i < nb; ++i) {
            lb[i][0] = rules->trial_uv[i][0];
            lb[i][1] = rules->trial_uv[i][1];
            lb[i][2] = 1.0 - lb[i][0] - lb[i][1];
            wb[i] = rules->trial_weight[i];
        }
        P->rule_default = rules_are_default(*rules);
    } else {
        if (!far_rule(o.far_test_order, nt, lt, wt) || !far_rule(o.far_trial_order, nb, lb, wb))
            return fail(BMI_EINVAL, "far rule orders must be 1, 2 or 3");
        P->rule_default = o.far_test_order == 2 && o.far_trial_order == 3;
    }
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
    hipStream_t s = ctx->stream;
    // rule points of every cell and the packed geometry records of the test
    // sub-batches and trial patches, built on the device from the meshes
    if ((e = P->d_xpts.alloc((size_t)test->mesh->nt * 3 * BMI_MAX_RULE)) != hipSuccess ||
        (e = P->d_ypts.alloc((size_t)trial->mesh->nt * 3 * BMI_MAX_RULE)) != hipSuccess ||
        (e = P->d_tgeo.alloc((size_t)test->nsb * BMI_GEO)) != hipSuccess ||
        (e = P->d_bgeo.alloc((size_t)trial->npatch * BMI_GEO)) != hipSuccess)
        return fail(BMI_ENOMEM, std::string("rule points: ") + hipGetErrorString(e));
    {
        const bmi_mesh *mt = test->mesh, *mb = trial->mesh;
        BMI_HIP(launch_rule_points(mt->d_vtx.p, mt->d_cen.p, mt->nt, nt, lt, P->d_xpts.p, s));
        BMI_HIP(launch_rule_points(mb->d_vtx.p, mb->d_cen.p, mb->nt, nb, lb, P->d_ypts.p, s));
        BMI_HIP(launch_geo_records(P->d_xpts.p, mt->d_vtx.p, mt->d_cen.p, mt->d_area.p, test->d_sb_cells.p, test->nsb,
                                   P->d_tgeo.p, s));
        BMI_HIP(launch_geo_records(P->d_ypts.p, mb->d_vtx.p, mb->d_cen.p, mb->d_area.p, trial->d_patch_tri.p,
                                   trial->npatch, P->d_bgeo.p, s));
    }

    P->npa = test->nsb;
    P->npb = trial->npatch;
    P->nblocks = (long long)P->npa * P->npb;
    if (P->nblocks >= (1LL << 31) - 1) return fail(BMI_EUNSUPPORTED, "too many patch blocks");
    const int nbk = (int)P->nblocks;
    if ((e = P->d_blk_cnt.alloc(nbk)) != hipSuccess || (e = P->d_blk_off.alloc(nbk)) != hipSuccess ||
        (e = P->d_total.alloc(1)) != hipSuccess || (e = P->d_sc.alloc(64)) != hipSuccess)
        return fail(BMI_ENOMEM, std::string("plan buffers: ") + hipGetErrorString(e));
    BMI_HIP(launch_sincos_table(P->d_sc.p, s));

    ClassifyArgs &C = P->cargs;
    C.X = dev_view_test(test, P->d_xpts.p);
    C.Y = dev_view(trial, P->d_ypts.p);
    C.npa = P->npa;
    C.npb = P->npb;
    C.k2x16 = 16.0 * (op->gamma[0] * op->gamma[0] + op->gamma[1] * op->gamma[1]);
    const double h = std::max(test->mesh->hmax, trial->mesh->hmax);
    C.near_d2 = o.precision == BMI_PREC_MIXED ? (o.near_factor * h) * (o.near_factor * h) : -1.0;
    C.near_d = o.precision == BMI_PREC_MIXED ? o.near_factor * h : -1.0;
    C.blk_cnt = P->d_blk_cnt.p;
    C.blk_off = P->d_blk_off.p;
    C.counter = P->d_total.p;

    // first classification: size the special-pair buffers
    {
        const std::vector<int> gb = {0, P->npa};
        const std::vector<ClassifyArgs> ca = {C};
        if ((e = P->d_gbase.upload(gb)) != hipSuccess || (e = P->d_cargs.upload(ca)) != hipSuccess)
            return fail(BMI_EDEVICE, std::string("plan arguments: ") + hipGetErrorString(e));
    }
    // counting run: no entry buffer yet (capacity 0), the counter gives the size
    BMI_HIP(launch_zero_ints(P->d_total.p, 1, s));
    BMI_HIP(launch_classify(P->d_cargs.p, P->d_gbase.p, 1, P->npa, s));
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
    {
        const std::vector<ClassifyArgs> ca = {C};
        if ((e = P->d_cargs.upload(ca)) != hipSuccess)
            return fail(BMI_EDEVICE, std::string("plan arguments: ") + hipGetErrorString(e));
    }
    // the entries themselves are written by each assembly's classification
    // (or by bmi_plan_stats if it comes first)
    SpecialArgs &S = P->sargs;
    S.X = C.X;
    S.Y = C.Y;
    S.spec_K = P->d_spec_K.p;
    S.spec_L = P->d_spec_L.p;
    S.spec_cls = P->d_spec_cls.p;
    S.spec_loc = P->d_spec_loc.p;
    S.n_special = P->d_total.p;
    S.blk_off = P->d_blk_off.p;
    S.blk_cnt = P->d_blk_cnt.p;
    S.npb = P->npb;
    S.capacity = total;
    S.sc = P->d_sc.p;
    S.spec_val = P->d_spec_val.p;
    S.op = P->opc;
    S.rule = P->rule;
    S.sym = P->X == P->Y && (P->op_kind == OPK_MW || P->op_kind == OPK_HH3D);
    S.cf_zero = P->op_kind == OPK_DL;
    for (int g = 0; g < 3; ++g) {
        if (rules) {
            S.ss_n[g] = rules->gl_points[g];
            for (int i = 0; i < S.ss_n[g]; ++i) {
                S.glx[g][i] = rules->gl_node[g][i];
                S.glw[g][i] = rules->gl_weight[g][i];
            }
        } else {
            S.ss_n[g] = o.ss_order[g];
            gauss_legendre01(o.ss_order[g], S.glx[g], S.glw[g]);
        }
        P->opts.ss_order[g] = S.ss_n[g];
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
    P->cs = special_chunk(P->capacity);
    P->nchunks = (P->capacity + P->cs - 1) / P->cs;
    {
        const std::vector<SpecialArgs> sa = {S};
        const std::vector<int> cb = {0, P->nchunks}, eb = {0, P->capacity};
        if ((e = P->d_sargs.upload(sa)) != hipSuccess || (e = P->d_cbase.upload(cb)) != hipSuccess ||
            (e = P->d_ebase.upload(eb)) != hipSuccess)
            return fail(BMI_EDEVICE, std::string("plan arguments: ") + hipGetErrorString(e));
    }
    for (int i = 0; i < 4; ++i) BMI_HIP(hipEventCreate(&P->ev[i]));
    *out = P.release();
    return BMI_OK;
}

bmi_status bmi_plan_free(bmi_plan *P) try {
    if (P) {
        (void)hipSetDevice(P->device);
        for (auto &e : P->ev)
            if (e) (void)hipEventDestroy(e);
        delete P;
    }
    return BMI_OK;
} catch (...) {
    return fail_exception();
}

bmi_status bmi_assemble(bmi_plan *P, bmi_matrix *out) try {
    if (!P || !out) return fail(BMI_EINVAL, "NULL argument");
    bmi_status st = check_target(P, out);
    if (st != BMI_OK) return st;
    BMI_HIP(hipSetDevice(P->ctx->device));
    hipStream_t s = P->ctx->stream;
    BMI_HIP(hipEventRecord(P->ev[0], s));
    if ((st = enqueue_prologue(P, out, s, P->ev[1])) != BMI_OK) return st;
    BMI_HIP(hipEventRecord(P->ev[2], s));
    int launches = 0;
    if ((st = enqueue_far(P, out, s, &launches)) != BMI_OK) return st;
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
    P->last.far_launches = launches;
    P->last.n_colours_test = 1;
    P->last.n_colours_trial = P->Y->ncolour;
    return BMI_OK;
} catch (...) {
    return fail_exception();
}

bmi_status bmi_plan_stats(const bmi_plan *Pc, bmi_stats *st) try {
    if (!Pc || !st) return fail(BMI_EINVAL, "NULL argument");
    bmi_plan *P = const_cast<bmi_plan *>(Pc);
    if (!P->pairs_ready) {
        // pair counts, on first request: each (test cell, trial cell) pair
        // has one owner entry (the other entries of a cell repeated in
        // several panels carry SPEC_NOT_OWNER); the rest are far pairs
        BMI_HIP(hipSetDevice(P->ctx->device));
        if (!P->filled) {
            BMI_HIP(launch_zero_ints(P->d_total.p, 1, P->ctx->stream));
            BMI_HIP(launch_classify(P->d_cargs.p, P->d_gbase.p, 1, P->npa, P->ctx->stream));
            P->filled = true;
        }
        std::vector<unsigned char> cls(std::max(P->capacity, 1), 0);
        if (P->capacity > 0) {
            BMI_HIP(hipMemcpyAsync(cls.data(), P->d_spec_cls.p, P->capacity, hipMemcpyDeviceToHost, P->ctx->stream));
            BMI_HIP(hipStreamSynchronize(P->ctx->stream));
        }
        for (int c = 0; c < 6; ++c) P->pairs[c] = 0;
        for (int i = 0; i < P->capacity; ++i)
            if (!(cls[i] & SPEC_NOT_OWNER) && cls[i] >= 1 && cls[i] < 6) P->pairs[cls[i]]++;
        int64_t special = 0;
        for (int c = 1; c < 6; ++c) special += P->pairs[c];
        P->pairs[0] = P->X->mesh->nt * P->Y->mesh->nt - special;
        P->pairs_ready = true;
    }
    *st = P->last;
    for (int i = 0; i < 6; ++i) st->pairs[i] = P->pairs[i];
    st->n_colours_test = 1;
    st->n_colours_trial = P->Y->ncolour;
    st->plan_bytes = (int64_t)(buf_bytes(P->d_xpts) + buf_bytes(P->d_ypts) + buf_bytes(P->d_tgeo) +
                               buf_bytes(P->d_bgeo) + buf_bytes(P->d_blk_cnt) + buf_bytes(P->d_blk_off) +
                               buf_bytes(P->d_total) + buf_bytes(P->d_spec_K) + buf_bytes(P->d_spec_L) +
                               buf_bytes(P->d_spec_cls) + buf_bytes(P->d_spec_loc) + buf_bytes(P->d_spec_val) +
                               buf_bytes(P->d_cargs) + buf_bytes(P->d_sargs) + buf_bytes(P->d_gbase) +
                               buf_bytes(P->d_cbase) + buf_bytes(P->d_ebase) + buf_bytes(P->d_sc));
    return BMI_OK;
} catch (...) {
    return fail_exception();
}

// Host copy of a source chunk into the destination, converting the element
// type if needed, split over a few host threads.
static void consume_chunk(const void *src, int src32, void *dst, int dst32, size_t first, size_t count) {
    const unsigned nth = count > (1u << 20) ? 4u : 1u;
    auto work = [&](size_t a, size_t b) {
        if (src32 == dst32) {
            const size_t es = src32 ? 8 : 16;
            std::memcpy((unsigned char *)dst + (first + a) * es, (const unsigned char *)src + a * es, (b - a) * es);
        } else if (!src32) {
            const double *s = (const double *)src;
            float *d = (float *)dst + 2 * first;
            for (size_t i = 2 * a; i < 2 * b; ++i) d[i] = (float)s[i];
        } else {
            const float *s = (const float *)src;
            double *d = (double *)dst + 2 * first;
            for (size_t i = 2 * a; i < 2 * b; ++i) d[i] = (double)s[i];
        }
    };
    if (nth == 1) {
        work(0, count);
        return;
    }
    const size_t per = (count + nth - 1) / nth;
    const std::function<void(int)> g = [&](int t) {
        const size_t a = std::min(count, (size_t)t * per), b = std::min(count, a + per);
        if (a < b) work(a, b);
    };
    HostPool::get().run((int)nth, 1, g);
}

// Bytes per pinned staging buffer of the device-to-host copies.
constexpr size_t kStage = 64u << 20;

static bmi_status ensure_staging(bmi_context *ctx) {
    for (int i = 0; i < 2; ++i) {
        if (!ctx->stage[i]) BMI_HIP(hipHostMalloc(&ctx->stage[i], kStage, hipHostMallocDefault));
        if (!ctx->stage_ev[i]) BMI_HIP(hipEventCreateWithFlags(&ctx->stage_ev[i], hipEventDisableTiming));
    }
    return BMI_OK;
}

// Host copy of elements [o, o + c) of a column-major block of nrb rows (global
// rows r0 .. r0 + nrb - 1 of every column) into the column-major rows-row host
// matrix dst; element size es.  Column runs are copied whole, over a few host
// threads for large pieces.
static void scatter_rows_piece(const unsigned char *src, size_t es, size_t o, size_t c, int64_t nrb, int64_t r0,
                               int64_t rows, unsigned char *dst) {
    auto work = [&](size_t a, size_t b) {
        size_t e = o + a;
        const size_t end = o + b;
        while (e < end) {
            const size_t col = e / (size_t)nrb, row = e % (size_t)nrb;
            const size_t run = std::min(end - e, (size_t)nrb - row);
            std::memcpy(dst + ((size_t)r0 + row + col * (size_t)rows) * es, src + (e - o) * es, run * es);
            e += run;
        }
    };
    const unsigned nth = c > (1u << 20) ? 4u : 1u;
    if (nth == 1) {
        work(0, c);
        return;
    }
    const size_t per = (c + nth - 1) / nth;
    const std::function<void(int)> g = [&](int t) {
        const size_t a = std::min(c, (size_t)t * per), b = std::min(c, a + per);
        if (a < b) work(a, b);
    };
    HostPool::get().run((int)nth, 1, g);
}

bmi_status bmi_assemble_to_host(bmi_plan *P, void *dst, int32_t dtype, int64_t device_budget) try {
    if (!P || !dst) return fail(BMI_EINVAL, "NULL argument");
    if (dtype != BMI_C64 && dtype != BMI_C32) return fail(BMI_EINVAL, "bad dtype");
    if (device_budget <= 0) return fail(BMI_EINVAL, "device budget must be positive");
    bmi_context *ctx = P->ctx;
    BMI_HIP(hipSetDevice(ctx->device));
    WorkStream ws(ctx->stream);
    hipStream_t s = ctx->stream;
    const int64_t rows = P->X->nfun, cols = P->Y->nfun;
    const size_t es = dtype == BMI_C64 ? 16 : 8;
    const int PR = P->X->prows;
    const int npanel = P->X->npanel;
    // panels per block: two staging slots within the budget; at most an
    // eighth of the panels, so that the copies of a block overlap the
    // assembly of the next ones
    const double panel_bytes = (double)PR * (double)cols * (double)es;
    int64_t ppb = (int64_t)((double)device_budget / (2.0 * panel_bytes));
    if (ppb < 1)
        return fail(BMI_ENOMEM, "the device budget holds less than two blocks of one test panel (" +
                                    std::to_string((int64_t)(2.0 * panel_bytes)) + " bytes)");
    ppb = std::min<int64_t>(ppb, std::max<int64_t>(1, (npanel + 7) / 8));
    const int nblk = (int)((npanel + ppb - 1) / ppb);
    const int nslot = nblk > 1 ? 2 : 1;
    // the tallest block's rows (panels are at most PR rows)
    int64_t max_rows = 0;
    for (int64_t p0 = 0; p0 < npanel; p0 += ppb)
        max_rows = std::max<int64_t>(max_rows, P->X->prow[std::min<int64_t>(npanel, p0 + ppb)] - P->X->prow[p0]);
    const size_t slot_elems = (size_t)max_rows * (size_t)cols;
    DevBuf<unsigned char> slot[2];
    for (int i = 0; i < nslot; ++i) {
        hipError_t e = slot[i].alloc(slot_elems * es);
        if (e != hipSuccess) return fail(BMI_ENOMEM, std::string("staging block: ") + hipGetErrorString(e));
    }
    bmi_status st = ensure_staging(ctx);
    if (st != BMI_OK) return st;
    if (!ctx->copy_stream) BMI_HIP(hipStreamCreateWithFlags(&ctx->copy_stream, hipStreamNonBlocking));
    for (int i = 0; i < 2; ++i)
        if (!ctx->blk_ev[i]) BMI_HIP(hipEventCreateWithFlags(&ctx->blk_ev[i], hipEventDisableTiming));
    hipStream_t cs = ctx->copy_stream;

    BMI_HIP(hipEventRecord(P->ev[0], s));
    if ((st = enqueue_prologue(P, nullptr, s, P->ev[1])) != BMI_OK) return st;
    BMI_HIP(hipEventRecord(P->ev[2], s));
    int launches = 0;
    auto block_rows = [&](int b, int64_t &r0, int64_t &nrb) {
        const int64_t p0 = (int64_t)b * ppb, p1 = std::min<int64_t>(npanel, p0 + ppb);
        r0 = P->X->prow[p0];
        nrb = P->X->prow[p1] - r0;
    };
    auto issue_block = [&](int b) -> bmi_status {
        int64_t r0, nrb;
        block_rows(b, r0, nrb);
        unsigned char *base = slot[b % nslot].p;
        if (P->X->uncovered || P->Y->uncovered) BMI_HIP(launch_fill_zero(base, (size_t)nrb * cols * es, s));
        // global row r of column c lands at base[(r - r0) + c nrb]
        const int p0 = (int)((int64_t)b * ppb), p1 = (int)std::min<int64_t>(npanel, (int64_t)(b + 1) * ppb);
        const bmi_status r = enqueue_far_range(P, base - (size_t)r0 * es, nrb, dtype, p0, p1, s, &launches);
        if (r != BMI_OK) return r;
        BMI_HIP(hipEventRecord(ctx->blk_ev[b % nslot], s));
        if (b == nblk - 1) BMI_HIP(hipEventRecord(P->ev[3], s));
        return BMI_OK;
    };
    if ((st = issue_block(0)) != BMI_OK) return st;
    const size_t ce = kStage / es;
    for (int b = 0; b < nblk; ++b) {
        // the other slot was copied out by the previous iteration: the next
        // block's kernels may overwrite it while this one is copied
        if (b + 1 < nblk && (st = issue_block(b + 1)) != BMI_OK) return st;
        int64_t r0, nrb;
        block_rows(b, r0, nrb);
        BMI_HIP(hipStreamWaitEvent(cs, ctx->blk_ev[b % nslot], 0));
        const unsigned char *base = slot[b % nslot].p;
        const size_t n = (size_t)nrb * (size_t)cols;
        const size_t nk = (n + ce - 1) / ce;
        auto issue = [&](size_t k) -> hipError_t {
            const size_t o = k * ce, c = std::min(ce, n - o);
            hipError_t e = hipMemcpyAsync(ctx->stage[k & 1], base + o * es, c * es, hipMemcpyDeviceToHost, cs);
            if (e != hipSuccess) return e;
            return hipEventRecord(ctx->stage_ev[k & 1], cs);
        };
        // two-deep pipeline: piece k+1 is in flight while piece k is scattered
        BMI_HIP(issue(0));
        for (size_t k = 0; k < nk; ++k) {
            if (k + 1 < nk) BMI_HIP(issue(k + 1));
            BMI_HIP(hipEventSynchronize(ctx->stage_ev[k & 1]));
            const size_t o = k * ce, c = std::min(ce, n - o);
            scatter_rows_piece((const unsigned char *)ctx->stage[k & 1], es, o, c, nrb, r0, rows,
                               (unsigned char *)dst);
        }
    }
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
    P->last.far_launches = launches;
    P->last.n_colours_test = 1;
    P->last.n_colours_trial = P->Y->ncolour;
    return BMI_OK;
} catch (...) {
    return fail_exception();
}

bmi_status bmi_matrix_to_host(const bmi_matrix *m, void *dst, int32_t dtype) try {
    if (!m || !dst) return fail(BMI_EINVAL, "NULL argument");
    if (dtype != BMI_C64 && dtype != BMI_C32) return fail(BMI_EINVAL, "bad dtype");
    bmi_context *ctx = m->ctx;
    BMI_HIP(hipSetDevice(ctx->device));
    WorkStream ws(ctx->stream);
    hipStream_t s = ctx->stream;
    {
        const bmi_status st = ensure_staging(ctx);
        if (st != BMI_OK) return st;
    }
    const int src32 = m->dtype == BMI_C32;
    const size_t es = src32 ? 8 : 16;
    const size_t n = (size_t)m->rows * (size_t)m->cols;
    const size_t ce = kStage / es;
    const size_t nk = (n + ce - 1) / ce;
    auto issue = [&](size_t k) -> hipError_t {
        const size_t o = k * ce, c = std::min(ce, n - o);
        hipError_t e = hipMemcpyAsync(ctx->stage[k & 1], (const unsigned char *)m->d + o * es, c * es,
                                      hipMemcpyDeviceToHost, s);
        if (e != hipSuccess) return e;
        return hipEventRecord(ctx->stage_ev[k & 1], s);
    };
    // two-deep pipeline: chunk k+1 is in flight while chunk k is copied out
    BMI_HIP(issue(0));
    for (size_t k = 0; k < nk; ++k) {
        if (k + 1 < nk) BMI_HIP(issue(k + 1));
        BMI_HIP(hipEventSynchronize(ctx->stage_ev[k & 1]));
        const size_t o = k * ce, c = std::min(ce, n - o);
        consume_chunk(ctx->stage[k & 1], src32, dst, dtype == BMI_C32, o, c);
    }
    return BMI_OK;
} catch (...) {
    return fail_exception();
}

bmi_status bmi_matrix_apply(const bmi_matrix *m, const double *x, double *y, int32_t nrhs) try {
    if (!m) return fail(BMI_EINVAL, "NULL argument");
    return matrix_apply_on(m->ctx, m, x, y, nrhs);
} catch (...) {
    return fail_exception();
}

bmi_status bmi_matrix_compare(const bmi_matrix *a, const bmi_matrix *ref, double out[3]) try {
    if (!a || !ref || !out) return fail(BMI_EINVAL, "NULL argument");
    if (a->rows != ref->rows || a->cols != ref->cols) return fail(BMI_EINVAL, "matrix shapes differ");
    if (a->ctx != ref->ctx) return fail(BMI_EINVAL, "matrices belong to different contexts");
    BMI_HIP(hipSetDevice(a->ctx->device));
    WorkStream ws(a->ctx->stream);
    hipStream_t s = a->ctx->stream;
    const int groups = 4096;
    DevBuf<double> part;
    hipError_t e = part.alloc(3 * groups);
    if (e != hipSuccess) return fail(BMI_ENOMEM, "compare buffer");
    BMI_HIP(launch_compare(a->d, a->dtype == BMI_C32, ref->d, ref->dtype == BMI_C32, a->rows * a->cols, groups,
                           part.p, s));
    std::vector<double> h(3 * groups);
    BMI_HIP(hipMemcpyAsync(h.data(), part.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost, s));
    BMI_HIP(hipStreamSynchronize(s));
    double d2 = 0.0, r2 = 0.0, dm = 0.0;
    for (int g = 0; g < groups; ++g) {
        d2 += h[3 * g];
        r2 += h[3 * g + 1];
        dm = std::max(dm, h[3 * g + 2]);
    }
    out[0] = std::sqrt(d2);
    out[1] = std::sqrt(r2);
    out[2] = std::sqrt(dm);
    return BMI_OK;
} catch (...) {
    return fail_exception();
}

}  // extern "C"

// ------------------------------------------------------------------ batches

namespace {

// Problems of a batch that share one far-pair kernel.
struct BatchBucket {
    int op_kind = 0, mixed = 0, fast = 0, c32 = 0;
    std::vector<int> problems;
    int ncm = 1;
    int nitems = 0;
    std::vector<int> colour_off;    // items of colour c: [colour_off[c], colour_off[c + 1])
    DevBuf<PanelArgs> d_args;
    DevBuf<int4> d_items;           // (problem in bucket, panel, patch, -), colour-major
    int npitems = 0;
    DevBuf<int2> d_pitems;          // (problem in bucket, panel), largest first: the persistent queue
    DevBuf<int> d_head;             // the persistent queue's head
    int mode = 0, wpc = 0;
    double tuned_ms = 0.0;
    double cand_ms[BMI_BATCH_CANDIDATES] = {};  // per-colour, persistent with 1, 2, 3 work-groups per CU, per-problem
};

// A bucket signature's tuned schedule, kept for the process.
struct TunedSchedule {
    int mode = 0, wpc = 0;
    double tuned_ms = 0.0;
    double cand_ms[BMI_BATCH_CANDIDATES] = {};
};

std::mutex g_tune_mu;
std::vector<std::pair<std::string, TunedSchedule>> g_tune_cache;

}  // namespace

// Problems of a batch with one operator kind: one launch of the special-pair
// kernels over all of them.
struct BatchOpGroup {
    int op_kind = 0, mixed = 0, ng5 = 1, fast = 1;
    std::vector<int> problems;
    int total_chunks = 0, total_entries = 0, cs = 256;
    DevBuf<SpecialArgs> d_sargs;
    DevBuf<int> d_cbase, d_ebase;
    // classification of the group's problems (one launch)
    int total_groups = 0;
    DevBuf<ClassifyArgs> d_cargs;
    DevBuf<int> d_gbase;
    DevBuf<int *> d_counters;       // each problem's classifier counter
};

struct bmi_batch {
    bmi_context *ctx = nullptr;
    int device = 0;
    std::vector<bmi_plan *> plans;
    std::vector<bmi_matrix *> mats;
    std::deque<BatchOpGroup> opgroups;
    std::vector<int> singles;
    std::deque<BatchBucket> buckets;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    // graph branches: one stream per operator group (default priority), the
    // fork event and one join event per branch
    std::vector<hipStream_t> branch;
    std::vector<hipEvent_t> join;
    hipEvent_t fork = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int cus = 0;
    double last_ms = 0.0;
};

namespace {

// The far phase of one bucket on stream s with the bucket's schedule: one
// launch per trial colour over every problem's blocks (mode 0), one
// persistent panel-major launch of wpc work-groups per CU (mode 1), or each
// problem's own per-colour launches, one problem after another (mode 2).
bmi_status enqueue_bucket(bmi_batch *Bt, BatchBucket &B, int mode, int wpc, hipStream_t s) {
    if (mode == 1) {
        BMI_HIP(launch_zero_ints(B.d_head.p, 1, s));
        BMI_HIP(launch_panels_persist(B.d_args.p, B.d_pitems.p, B.npitems, B.ncm, B.d_head.p, Bt->cus * wpc,
                                      B.op_kind, B.mixed, B.fast, B.c32, s));
    } else if (mode == 2) {
        for (int p : B.problems) {
            int launches = 0;
            const bmi_status st = enqueue_far(Bt->plans[p], Bt->mats[p], s, &launches);
            if (st != BMI_OK) return st;
        }
    } else {
        for (size_t c = 0; c + 1 < B.colour_off.size(); ++c)
            BMI_HIP(launch_panels_items(B.d_args.p, B.d_items.p + B.colour_off[c],
                                        B.colour_off[c + 1] - B.colour_off[c], B.ncm, B.op_kind, B.mixed, B.fast,
                                        B.c32, s));
    }
    return BMI_OK;
}

std::string bucket_signature(const BatchBucket &B) {
    int lg = 0;
    while ((1 << lg) < B.nitems) ++lg;
    return std::to_string(B.op_kind) + "/" + std::to_string(B.mixed) + "/" + std::to_string(B.fast) + "/" +
           std::to_string(B.c32) + "/" + std::to_string(B.problems.size()) + "/" + std::to_string(lg);
}

// Builds the colour-major item list of a bucket.  Problems with more blocks
// come first inside each colour, so the queue's tail holds small blocks.
bmi_status build_bucket(bmi_batch *Bt, BatchBucket &B) {
    std::sort(B.problems.begin(), B.problems.end(), [&](int a, int b) {
        const bmi_plan *pa = Bt->plans[a], *pb = Bt->plans[b];
        return (int64_t)pa->X->npanel * pa->npb > (int64_t)pb->X->npanel * pb->npb;
    });
    std::vector<PanelArgs> args;
    int maxcol = 0;
    for (int p : B.problems) {
        const bmi_plan *P = Bt->plans[p];
        args.push_back(panel_args(P, Bt->mats[p]));
        B.ncm = std::max(B.ncm, P->Y->ncmax);
        maxcol = std::max(maxcol, P->Y->ncolour);
    }
    std::vector<int4> items;
    B.colour_off.assign(1, 0);
    for (int c = 0; c < maxcol; ++c) {
        for (size_t k = 0; k < B.problems.size(); ++k) {
            const bmi_plan *P = Bt->plans[B.problems[k]];
            if (c >= P->Y->ncolour) continue;
            const std::vector<int> &qs = P->Y->by_colour[c];
            for (int pan = 0; pan < P->X->npanel; ++pan)
                for (int q : qs) items.push_back(make_int4((int)k, pan, q, 0));
        }
        B.colour_off.push_back((int)items.size());
    }
    B.nitems = (int)items.size();
    // persistent queue: one item per (problem, test panel), by decreasing
    // work (sub-batches of the panel times trial patches), ties in order
    std::vector<std::pair<int64_t, int2>> pq;
    for (size_t k = 0; k < B.problems.size(); ++k) {
        const bmi_plan *P = Bt->plans[B.problems[k]];
        for (int pan = 0; pan < P->X->npanel; ++pan) {
            const int64_t w = (int64_t)(P->X->sb_off[pan + 1] - P->X->sb_off[pan]) * P->Y->npatch;
            pq.push_back({w, make_int2((int)k, pan)});
        }
    }
    std::stable_sort(pq.begin(), pq.end(), [](const std::pair<int64_t, int2> &a, const std::pair<int64_t, int2> &b) {
        return a.first > b.first;
    });
    std::vector<int2> pitems;
    for (auto &v : pq) pitems.push_back(v.second);
    B.npitems = (int)pitems.size();
    hipError_t e;
    if ((e = B.d_args.upload(args)) != hipSuccess || (e = B.d_items.upload(items)) != hipSuccess ||
        (e = B.d_pitems.upload(pitems)) != hipSuccess || (e = B.d_head.alloc(1)) != hipSuccess)
        return fail(BMI_ENOMEM, std::string("batch buffers: ") + hipGetErrorString(e));
    return BMI_OK;
}

// Measures the candidate schedules of a bucket (the prologues have run) and
// keeps the fastest; a bucket signature seen before reuses its choice.
// Work-groups per CU that the far kernels' launch bounds let reside at once.
int max_wgs_per_cu(const BatchBucket &B) { return (B.op_kind == OPK_DL && !B.mixed) ? 2 : 3; }

bmi_status tune_bucket(bmi_batch *Bt, BatchBucket &B) {
    const std::string sig = bucket_signature(B);
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        for (auto &kv : g_tune_cache)
            if (kv.first == sig) {
                const TunedSchedule &t = kv.second;
                B.mode = t.mode;
                B.wpc = t.wpc;
                B.tuned_ms = t.tuned_ms;
                std::copy(t.cand_ms, t.cand_ms + BMI_BATCH_CANDIDATES, B.cand_ms);
                return BMI_OK;
            }
    }
    hipStream_t s = Bt->ctx->stream;
    const int maxwpc = max_wgs_per_cu(B);
    double best = 1e300;
    // candidate c: 0 per-colour, 1 .. maxwpc persistent with c work-groups
    // per CU, BMI_BATCH_CANDIDATES - 1 per-problem launches
    for (int c = 0; c < BMI_BATCH_CANDIDATES; ++c) {
        if (c > maxwpc && c < BMI_BATCH_CANDIDATES - 1) continue;
        const int mode = c == 0 ? 0 : (c == BMI_BATCH_CANDIDATES - 1 ? 2 : 1), wpc = mode == 1 ? c : 0;
        float ms = 0.0f;
        for (int rep = 0; rep < 2; ++rep) {
            BMI_HIP(hipEventRecord(Bt->ev0, s));
            bmi_status st = enqueue_bucket(Bt, B, mode, wpc, s);
            if (st != BMI_OK) return st;
            BMI_HIP(hipEventRecord(Bt->ev1, s));
            BMI_HIP(hipStreamSynchronize(s));
            BMI_HIP(hipEventElapsedTime(&ms, Bt->ev0, Bt->ev1));
        }
        B.cand_ms[c] = ms;
        if (ms < best) {
            best = ms;
            B.mode = mode;
            B.wpc = wpc;
            B.tuned_ms = ms;
        }
    }
    TunedSchedule t;
    t.mode = B.mode;
    t.wpc = B.wpc;
    t.tuned_ms = B.tuned_ms;
    std::copy(B.cand_ms, B.cand_ms + BMI_BATCH_CANDIDATES, t.cand_ms);
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune_cache.push_back({sig, t});
    return BMI_OK;
}

// Classification of an operator group's problems (one launch) and zero fill
// of their uncovered rows: on stream s.
bmi_status enqueue_group_classify(bmi_batch *Bt, const BatchOpGroup &G, hipStream_t s) {
    const int n = (int)G.problems.size();
    BMI_HIP(launch_zero_each(G.d_counters.p, n, s));
    BMI_HIP(launch_classify(G.d_cargs.p, G.d_gbase.p, n, G.total_groups, s));
    for (int i : G.problems) {
        const bmi_plan *P = Bt->plans[i];
        const bmi_matrix *M = Bt->mats[i];
        P->filled = true;
        if (P->X->uncovered || P->Y->uncovered)
            BMI_HIP(launch_fill_zero(M->d, (size_t)M->rows * M->cols * (M->dtype == BMI_C64 ? 16 : 8), s));
    }
    return BMI_OK;
}

// The special pairs of one operator group, on stream s.
hipError_t launch_group_special(const BatchOpGroup &G, hipStream_t s) {
    return launch_special(G.d_sargs.p, G.d_cbase.p, G.d_ebase.p, (int)G.problems.size(), G.total_chunks,
                          G.total_entries, G.cs, G.mixed, G.op_kind, G.ng5, G.fast, s);
}

// Classification, zero fills and the special pairs of every operator group,
// all on stream s (the batch before the far phase, outside a graph).
bmi_status enqueue_batch_prologue(bmi_batch *Bt, hipStream_t s) {
    for (auto &G : Bt->opgroups) {
        const bmi_status st = enqueue_group_classify(Bt, G, s);
        if (st != BMI_OK) return st;
        BMI_HIP(launch_group_special(G, s));
    }
    return BMI_OK;
}

// Device tables of the batched prologue.
bmi_status build_prologue_tables(bmi_batch *Bt) {
    const int n = (int)Bt->plans.size();
    hipError_t e;
    for (int i = 0; i < n; ++i) {
        const bmi_plan *P = Bt->plans[i];
        BatchOpGroup *G = nullptr;
        for (auto &g : Bt->opgroups)
            if (g.op_kind == P->op_kind) G = &g;
        if (!G) {
            Bt->opgroups.emplace_back();
            G = &Bt->opgroups.back();
            G->op_kind = P->op_kind;
        }
        G->problems.push_back(i);
        G->mixed |= P->opts.precision == BMI_PREC_MIXED;
        G->ng5 &= gl5(P);
        G->fast &= fast_rule(P);
    }
    for (auto &G : Bt->opgroups) {
        std::vector<ClassifyArgs> ca;
        std::vector<int> gb = {0};
        std::vector<int *> counters;
        for (int i : G.problems) {
            ca.push_back(Bt->plans[i]->cargs);
            gb.push_back(gb.back() + Bt->plans[i]->npa);
            counters.push_back(Bt->plans[i]->d_total.p);
        }
        G.total_groups = gb.back();
        if ((e = G.d_cargs.upload(ca)) != hipSuccess || (e = G.d_gbase.upload(gb)) != hipSuccess ||
            (e = G.d_counters.upload(counters)) != hipSuccess)
            return fail(BMI_ENOMEM, std::string("batch tables: ") + hipGetErrorString(e));
        int cap = 0;
        for (int i : G.problems) cap += Bt->plans[i]->capacity;
        G.cs = special_chunk(cap);
        std::vector<SpecialArgs> sa;
        std::vector<int> cb = {0}, eb = {0};
        for (int i : G.problems) {
            const bmi_plan *P = Bt->plans[i];
            sa.push_back(P->sargs);
            cb.push_back(cb.back() + (P->capacity + G.cs - 1) / G.cs);
            eb.push_back(eb.back() + P->capacity);
        }
        G.total_chunks = cb.back();
        G.total_entries = eb.back();
        if ((e = G.d_sargs.upload(sa)) != hipSuccess || (e = G.d_cbase.upload(cb)) != hipSuccess ||
            (e = G.d_ebase.upload(eb)) != hipSuccess)
            return fail(BMI_ENOMEM, std::string("batch tables: ") + hipGetErrorString(e));
    }
    return BMI_OK;
}

// The whole batch as one graph: one branch per operator group, the
// classification (and zero fills) of its problems, its special pairs, then
// the far buckets of that operator (the groups' problems are disjoint, so the
// branches overlap each other's launch tails and the latency-bound
// classification of one group overlaps the other's Sauter-Schwab work),
// joined back before the problems outside the buckets.
bmi_status capture_batch(bmi_batch *Bt) {
    hipStream_t s = Bt->ctx->stream;
    const size_t ng = Bt->opgroups.size();
    while (Bt->branch.size() < ng) {
        hipStream_t b = nullptr;
        hipEvent_t e = nullptr;
        BMI_HIP(hipStreamCreateWithFlags(&b, hipStreamNonBlocking));
        Bt->branch.push_back(b);
        BMI_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        Bt->join.push_back(e);
    }
    if (!Bt->fork) BMI_HIP(hipEventCreateWithFlags(&Bt->fork, hipEventDisableTiming));
    BMI_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    bmi_status st = BMI_OK;
    auto body = [&]() -> bmi_status {
        BMI_HIP(hipEventRecord(Bt->fork, s));
        for (size_t gi = 0; gi < ng; ++gi) {
            const BatchOpGroup &G = Bt->opgroups[gi];
            hipStream_t b = Bt->branch[gi];
            BMI_HIP(hipStreamWaitEvent(b, Bt->fork, 0));
            bmi_status r0 = enqueue_group_classify(Bt, G, b);
            if (r0 != BMI_OK) return r0;
            BMI_HIP(launch_group_special(G, b));
            for (auto &B : Bt->buckets) {
                if (B.op_kind != G.op_kind) continue;
                bmi_status r = enqueue_bucket(Bt, B, B.mode, B.wpc, b);
                if (r != BMI_OK) return r;
            }
            BMI_HIP(hipEventRecord(Bt->join[gi], b));
            BMI_HIP(hipStreamWaitEvent(s, Bt->join[gi], 0));
        }
        for (int p : Bt->singles) {
            int launches = 0;
            bmi_status r = enqueue_far(Bt->plans[p], Bt->mats[p], s, &launches);
            if (r != BMI_OK) return r;
        }
        return BMI_OK;
    };
    st = body();
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &g);
    if (st != BMI_OK) {
        if (g) (void)hipGraphDestroy(g);
        return st;
    }
    if (e != hipSuccess) return fail(BMI_EDEVICE, std::string("graph capture: ") + hipGetErrorString(e));
    Bt->graph = g;
    BMI_HIP(hipGraphInstantiate(&Bt->exec, g, nullptr, nullptr, 0));
    return BMI_OK;
}

void batch_destroy(bmi_batch *Bt) {
    if (Bt->exec) (void)hipGraphExecDestroy(Bt->exec);
    if (Bt->graph) (void)hipGraphDestroy(Bt->graph);
    if (Bt->ev0) (void)hipEventDestroy(Bt->ev0);
    if (Bt->ev1) (void)hipEventDestroy(Bt->ev1);
    if (Bt->fork) (void)hipEventDestroy(Bt->fork);
    for (hipEvent_t e : Bt->join) (void)hipEventDestroy(e);
    for (hipStream_t b : Bt->branch) (void)hipStreamDestroy(b);
    delete Bt;
}

}  // namespace

extern "C" {

bmi_status bmi_batch_create(bmi_context *ctx, bmi_plan *const *plans, bmi_matrix *const *matrices, int32_t n,
                            bmi_batch **out) try {
    if (!ctx || !plans || !matrices || !out || n < 1) return fail(BMI_EINVAL, "bad argument");
    *out = nullptr;
    for (int i = 0; i < n; ++i) {
        if (!plans[i] || !matrices[i]) return fail(BMI_EINVAL, "NULL plan or matrix");
        if (plans[i]->ctx != ctx) return fail(BMI_EINVAL, "plans must belong to the batch's context");
        const bmi_status st = check_target(plans[i], matrices[i]);
        if (st != BMI_OK) return st;
        for (int j = 0; j < i; ++j)
            if (plans[j] == plans[i] || matrices[j] == matrices[i])
                return fail(BMI_EINVAL, "a plan or matrix appears twice in the batch");
    }
    BMI_HIP(hipSetDevice(ctx->device));
    WorkStream wsc(ctx->stream);
    std::unique_ptr<bmi_batch, void (*)(bmi_batch *)> Bt(new bmi_batch, batch_destroy);
    Bt->ctx = ctx;
    Bt->device = ctx->device;
    Bt->plans.assign(plans, plans + n);
    Bt->mats.assign(matrices, matrices + n);
    BMI_HIP(hipDeviceGetAttribute(&Bt->cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
    BMI_HIP(hipEventCreate(&Bt->ev0));
    BMI_HIP(hipEventCreate(&Bt->ev1));
    // buckets by far-pair kernel
    for (int i = 0; i < n; ++i) {
        const bmi_plan *P = plans[i];
        const PanelArgs a = panel_args(P, matrices[i]);
        if (!a.rr) {
            Bt->singles.push_back(i);
            continue;
        }
        const int mixed = P->opts.precision == BMI_PREC_MIXED, c32 = matrices[i]->dtype == BMI_C32;
        BatchBucket *B = nullptr;
        for (auto &b : Bt->buckets)
            if (b.op_kind == P->op_kind && b.mixed == mixed && b.fast == a.fast && b.c32 == c32) B = &b;
        if (!B) {
            if ((int)Bt->buckets.size() == BMI_BATCH_MAX_BUCKETS) {
                Bt->singles.push_back(i);
                continue;
            }
            Bt->buckets.emplace_back();
            B = &Bt->buckets.back();
            B->op_kind = P->op_kind;
            B->mixed = mixed;
            B->fast = a.fast;
            B->c32 = c32;
        }
        B->problems.push_back(i);
    }
    for (auto &B : Bt->buckets) {
        const bmi_status st = build_bucket(Bt.get(), B);
        if (st != BMI_OK) return st;
    }
    // prologues once (the far kernels read their special pairs), then tune
    hipStream_t s = ctx->stream;
    {
        bmi_status st = build_prologue_tables(Bt.get());
        if (st == BMI_OK) st = enqueue_batch_prologue(Bt.get(), s);
        if (st != BMI_OK) return st;
    }
    BMI_HIP(hipStreamSynchronize(s));
    for (auto &B : Bt->buckets) {
        const bmi_status st = tune_bucket(Bt.get(), B);
        if (st != BMI_OK) return st;
    }
    const bmi_status st = capture_batch(Bt.get());
    if (st != BMI_OK) return st;
    *out = Bt.release();
    return BMI_OK;
} catch (...) {
    return fail_exception();
}

bmi_status bmi_batch_assemble(bmi_batch *Bt) try {
    if (!Bt) return fail(BMI_EINVAL, "NULL argument");
    BMI_HIP(hipSetDevice(Bt->ctx->device));
    hipStream_t s = Bt->ctx->stream;
    BMI_HIP(hipEventRecord(Bt->ev0, s));
    BMI_HIP(hipGraphLaunch(Bt->exec, s));
    BMI_HIP(hipEventRecord(Bt->ev1, s));
    BMI_HIP(hipStreamSynchronize(s));
    float ms = 0.0f;
    BMI_HIP(hipEventElapsedTime(&ms, Bt->ev0, Bt->ev1));
    Bt->last_ms = ms;
    return BMI_OK;
} catch (...) {
    return fail_exception();
}

bmi_status bmi_batch_stats_get(const bmi_batch *Bt, bmi_batch_stats *st) try {
    if (!Bt || !st) return fail(BMI_EINVAL, "NULL argument");
    *st = bmi_batch_stats{};
    st->last_ms = Bt->last_ms;
    st->problems = (int32_t)Bt->plans.size();
    st->buckets = (int32_t)Bt->buckets.size();
    st->singles = (int32_t)Bt->singles.size();
    for (size_t b = 0; b < Bt->buckets.size(); ++b) {
        st->bucket_problems[b] = (int32_t)Bt->buckets[b].problems.size();
        st->bucket_mode[b] = Bt->buckets[b].mode;
        st->bucket_wgs_per_cu[b] = Bt->buckets[b].wpc;
        st->bucket_tuned_ms[b] = Bt->buckets[b].tuned_ms;
        for (int c = 0; c < BMI_BATCH_CANDIDATES; ++c) st->bucket_candidate_ms[b][c] = Bt->buckets[b].cand_ms[c];
    }
    return BMI_OK;
} catch (...) {
    return fail_exception();
}

bmi_status bmi_batch_set_schedule(bmi_batch *Bt, int32_t bucket, int32_t mode, int32_t wgs_per_cu) try {
    if (!Bt) return fail(BMI_EINVAL, "NULL argument");
    if (bucket < 0 || bucket >= (int)Bt->buckets.size()) return fail(BMI_EINVAL, "no such bucket");
    BatchBucket &B = Bt->buckets[bucket];
    if (mode == 0 || mode == 2) wgs_per_cu = 0;
    else if (mode != 1 || wgs_per_cu < 1 || wgs_per_cu > max_wgs_per_cu(B))
        return fail(BMI_EINVAL, "mode 0, mode 2, or mode 1 with 1 .. " + std::to_string(max_wgs_per_cu(B)) +
                                    " work-groups per CU for this bucket");
    BMI_HIP(hipSetDevice(Bt->device));
    BMI_HIP(hipStreamSynchronize(Bt->ctx->stream));
    if (Bt->exec) BMI_HIP(hipGraphExecDestroy(Bt->exec));
    Bt->exec = nullptr;
    if (Bt->graph) BMI_HIP(hipGraphDestroy(Bt->graph));
    Bt->graph = nullptr;
    B.mode = mode;
    B.wpc = wgs_per_cu;
    return capture_batch(Bt);
} catch (...) {
    return fail_exception();
}

bmi_status bmi_batch_free(bmi_batch *Bt) try {
    if (Bt) {
        (void)hipSetDevice(Bt->device);
        batch_destroy(Bt);
    }
    return BMI_OK;
} catch (...) {
    return fail_exception();
}

}  // extern "C"
