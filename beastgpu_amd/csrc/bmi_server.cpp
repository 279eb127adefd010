// Request loop of libbmi (bmi_server_* in include/bmi.h): continuous batching
// of assembly, assemble-and-apply and apply requests on one device.
//
// Each of the `slots` worker threads owns a context (a stream, staging and
// apply workspace) and a grow-only matrix buffer, and takes the oldest queued
// request that can start now: for an assembly, the device memory it adds (its
// plan buffers, plus the growth of the worker's matrix buffer or, for a kept
// matrix, the whole matrix) must fit into the budget beside what is resident
// (kept matrices, worker buffers, running plans).  The worker prepares the
// request on the host (mesh, space, plan), assembles on its stream, applies,
// and keeps the matrix or leaves it in its buffer for the next request; in
// steady state no large allocation happens (growing the device pool while
// other requests run stalls every thread of the process).  Work of different
// requests overlaps on the device through the workers' streams; a request's
// arithmetic does not depend on what else runs.
//
// The reference has no request loop; its only scheduling precedent is the
// producer / Channel tiling of GPU_scheduler.jl:160-206.
#include "../../include/bmi.h"
#include "bmi_internal.h"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

namespace {

using Clock = std::chrono::steady_clock;

double ms_between(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

struct Request {
    uint64_t id = 0;
    int32_t kind = 0;
    std::vector<double> xyz;
    std::vector<int32_t> tri;
    int64_t nv = 0, nt = 0;
    int32_t space_kind = 0;
    std::vector<int64_t> gidx;
    std::vector<double> coeff;
    int64_t nfun = 0;
    bmi_operator op{};
    bmi_options opts{};
    int32_t dtype = BMI_C64;
    uint64_t matrix = 0;
    int32_t nrhs = 0;
    std::vector<double> x;
    double *y = nullptr;
    int64_t mbytes = 0;     // matrix bytes
    int64_t pbytes = 0;     // plan bytes (bound)
    int64_t grow = 0;       // growth of the worker's matrix buffer granted at admission
    Clock::time_point t_submit, t_admit;
    bool done = false;
    bmi_result res{};
};

// A matrix kept resident by an assembly request (m is null while the
// request is queued or running).
struct Resident {
    bmi_matrix *m = nullptr;
    int64_t nfun = 0;
    int64_t bytes = 0;
    int users = 0;               // apply requests running on it
    bool released = false;
};

struct Worker {
    bmi_context *ctx = nullptr;
    void *buf = nullptr;  // matrix buffer of assemble-and-apply requests
    int64_t cap = 0;
};

}  // namespace

struct bmi_server {
    int device = 0;
    int64_t budget = 0;
    std::vector<std::unique_ptr<Worker>> w;
    std::vector<std::thread> workers;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<Request>> queue;
    std::unordered_map<uint64_t, std::shared_ptr<Request>> pending;  // submitted, result not yet returned
    std::unordered_map<uint64_t, Resident> resident_m;
    int64_t resident = 0, peak = 0;
    uint64_t next_id = 1;
    bool stop = false;
};

namespace {

int64_t matrix_bytes(int64_t n, int32_t dtype) { return n * n * (dtype == BMI_C32 ? 8 : 16); }

}  // namespace

// Device bytes an assembly's plan holds besides its matrix, bounded from the
// mesh: geometry and slot records per cell, the classifier's block tables,
// and 154 bytes per special-pair entry (nine complex values, two cell ids,
// class, slot) for at most 3 entries per pair (a test cell sits in at most 3
// panels).  The pairs that can be special: the touching ones (at most the sum
// over vertices of degree^2), and those whose centroids are within the reach
// of the near test (dmin < 0.8 sqrt(A) of the trial cell, so within twice the
// largest cell radius plus 0.8 sqrt of the largest area) or, for the mixed
// path, within near_factor times the longest edge -- counted on a grid of
// that spacing.
int64_t plan_bytes_bound(const double *xyz, int64_t nv, const int32_t *tri, int64_t nt, const bmi_options &o) {
    std::vector<int64_t> deg(nv, 0);
    std::vector<double> cen(3 * nt);
    double rad = 0.0, amax = 0.0, hmax = 0.0;
    for (int64_t k = 0; k < nt; ++k) {
        const double *v[3];
        for (int i = 0; i < 3; ++i) {
            const int32_t vi = tri[3 * k + i];
            if (vi < 0 || vi >= nv) return -1;
            ++deg[vi];
            v[i] = xyz + 3 * (int64_t)vi;
        }
        double e1[3], e2[3];
        for (int d = 0; d < 3; ++d) {
            cen[3 * k + d] = (v[0][d] + v[1][d] + v[2][d]) / 3.0;
            e1[d] = v[0][d] - v[2][d];
            e2[d] = v[1][d] - v[2][d];
        }
        const double n0 = e1[1] * e2[2] - e1[2] * e2[1], n1 = e1[2] * e2[0] - e1[0] * e2[2],
                     n2 = e1[0] * e2[1] - e1[1] * e2[0];
        amax = std::max(amax, 0.5 * std::sqrt(n0 * n0 + n1 * n1 + n2 * n2));
        for (int i = 0; i < 3; ++i) {
            double r2 = 0.0, h2 = 0.0;
            for (int d = 0; d < 3; ++d) {
                const double x = v[i][d] - cen[3 * k + d], y = v[i][d] - v[(i + 1) % 3][d];
                r2 += x * x;
                h2 += y * y;
            }
            rad = std::max(rad, std::sqrt(r2));
            hmax = std::max(hmax, std::sqrt(h2));
        }
    }
    double touching = 0.0;
    for (int64_t d : deg) touching += (double)d * (double)d;
    double reach = 2.0 * rad + 0.8 * std::sqrt(amax);
    if (o.precision == BMI_PREC_MIXED) reach = std::max(reach, o.near_factor * hmax);
    reach = std::max(reach * (1.0 + 1e-9), 1e-300);
    // pairs with centroids closer than reach, on a grid of spacing reach
    double lo[3] = {1e300, 1e300, 1e300};
    for (int64_t k = 0; k < nt; ++k)
        for (int d = 0; d < 3; ++d) lo[d] = std::min(lo[d], cen[3 * k + d]);
    std::vector<std::pair<uint64_t, int>> keys(nt);
    auto cell_of = [&](int64_t k, int d) { return (int64_t)std::floor((cen[3 * k + d] - lo[d]) / reach); };
    auto pack = [](int64_t a, int64_t b, int64_t c) {
        return ((uint64_t)(a & 0x1fffff) << 42) | ((uint64_t)(b & 0x1fffff) << 21) | (uint64_t)(c & 0x1fffff);
    };
    for (int64_t k = 0; k < nt; ++k) keys[k] = {pack(cell_of(k, 0), cell_of(k, 1), cell_of(k, 2)), (int)k};
    std::sort(keys.begin(), keys.end());
    double close = 0.0;
    const double r2 = reach * reach;
    for (int64_t k = 0; k < nt; ++k) {
        const int64_t c0 = cell_of(k, 0), c1 = cell_of(k, 1), c2 = cell_of(k, 2);
        for (int64_t a = c0 - 1; a <= c0 + 1; ++a)
            for (int64_t b = c1 - 1; b <= c1 + 1; ++b)
                for (int64_t c = c2 - 1; c <= c2 + 1; ++c) {
                    if (a < 0 || b < 0 || c < 0) continue;
                    const uint64_t key = pack(a, b, c);
                    auto it = std::lower_bound(keys.begin(), keys.end(), std::make_pair(key, -1));
                    for (; it != keys.end() && it->first == key; ++it) {
                        const int64_t L = it->second;
                        double d2 = 0.0;
                        for (int d = 0; d < 3; ++d) {
                            const double x = cen[3 * k + d] - cen[3 * L + d];
                            d2 += x * x;
                        }
                        close += d2 < r2;
                    }
                }
    }
    const double pairs = std::min((double)nt * (double)nt, touching + close);
    const double groups = (double)nt / 8.0 + 64.0;
    return (int64_t)((64 << 20) + (double)nt * 8192.0 + groups * groups * 8.0 + 3.0 * pairs * 154.0);
}

namespace {

void finish(bmi_server *S, const std::shared_ptr<Request> &r, int32_t status, const char *msg) {
    r->res.status = status;
    if (status != BMI_OK && msg) {
        std::strncpy(r->res.message, msg, sizeof(r->res.message) - 1);
        r->res.message[sizeof(r->res.message) - 1] = 0;
    }
    r->res.total_ms = ms_between(r->t_submit, Clock::now());
    r->done = true;
}

// Runs an admitted request on the worker's context; returns its status.
int32_t run(bmi_server *S, Worker &W, Request &r, bmi_matrix **kept) {
    bmi_context *ctx = W.ctx;
    const auto t0 = Clock::now();
    if (r.kind == BMI_REQ_APPLY) {
        bmi_matrix *m = nullptr;
        {
            std::lock_guard<std::mutex> lk(S->mu);
            auto it = S->resident_m.find(r.matrix);
            if (it != S->resident_m.end()) m = it->second.m;
        }
        if (!m) return set_error(BMI_EINVAL, "the matrix of this apply request was not assembled or was released");
        // runs on this worker's stream; the matrix's assembly has completed
        // (bmi_assemble returns after its stream drains)
        const int32_t st = matrix_apply_on(ctx, m, r.x.data(), r.y, r.nrhs);
        r.res.device_ms = ms_between(t0, Clock::now());
        return st;
    }
    bmi_mesh *mesh = nullptr;
    bmi_space *space = nullptr;
    bmi_plan *plan = nullptr;
    bmi_matrix *m = nullptr;
    int32_t st = bmi_mesh_create(ctx, r.xyz.data(), r.nv, r.tri.data(), r.nt, &mesh);
    if (st == BMI_OK) st = bmi_space_create(mesh, r.space_kind, r.gidx.data(), r.coeff.data(), r.nfun, &space);
    if (st == BMI_OK) st = bmi_plan_create(ctx, &r.op, space, space, &r.opts, &plan);
    if (st == BMI_OK) {
        if (r.kind == BMI_REQ_ASSEMBLE) {
            st = bmi_matrix_create(ctx, r.nfun, r.nfun, r.dtype, &m);
        } else {
            if (r.mbytes > W.cap) {
                WorkStream ws(context_stream(ctx));
                dev_free(W.buf);
                W.buf = nullptr;
                W.cap = 0;
                const hipError_t e = dev_alloc(&W.buf, (size_t)r.mbytes);
                if (e != hipSuccess) {
                    W.buf = nullptr;
                    st = set_error(BMI_ENOMEM, std::string("matrix buffer: ") + hipGetErrorString(e));
                } else {
                    W.cap = r.mbytes;
                }
            }
            if (st == BMI_OK) st = matrix_wrap(ctx, W.buf, r.nfun, r.nfun, r.dtype, &m);
        }
    }
    const auto t1 = Clock::now();
    r.res.prep_ms = ms_between(t0, t1);
    if (st == BMI_OK) st = bmi_assemble(plan, m);
    if (st == BMI_OK) r.res.entries = r.nfun * r.nfun;
    if (st == BMI_OK && r.kind == BMI_REQ_ASSEMBLE_APPLY) st = bmi_matrix_apply(m, r.x.data(), r.y, r.nrhs);
    r.res.device_ms = ms_between(t1, Clock::now());
    if (plan) bmi_plan_free(plan);
    if (space) bmi_space_free(space);
    if (mesh) bmi_mesh_free(mesh);
    if (st == BMI_OK && r.kind == BMI_REQ_ASSEMBLE) {
        *kept = m;
    } else if (m) {
        bmi_matrix_free(m);
    }
    return st;
}

// The oldest queued request that worker W may start now (nullptr if none):
// applies need their matrix; assemblies need room in the budget for what
// they add.  An assembly that does not fit blocks the assemblies behind it
// (first come, first served), not the applies.
std::shared_ptr<Request> take(bmi_server *S, Worker &W, bool &assembly_blocked) {
    assembly_blocked = false;
    for (auto it = S->queue.begin(); it != S->queue.end(); ++it) {
        Request &r = **it;
        int64_t need = 0;
        if (r.kind == BMI_REQ_APPLY) {
            auto m = S->resident_m.find(r.matrix);
            if (m != S->resident_m.end()) {
                if (m->second.m == nullptr) continue;  // still being assembled
                m->second.users++;
            }  // else: gone (failed or released); the request runs and fails
        } else {
            if (assembly_blocked) continue;
            r.grow = r.kind == BMI_REQ_ASSEMBLE ? 0 : std::max<int64_t>(0, r.mbytes - W.cap);
            need = r.pbytes + (r.kind == BMI_REQ_ASSEMBLE ? r.mbytes : r.grow);
            if (S->resident + need > S->budget) {
                assembly_blocked = true;
                continue;
            }
        }
        auto p = *it;
        S->queue.erase(it);
        S->resident += need;
        S->peak = std::max(S->peak, S->resident);
        return p;
    }
    return nullptr;
}

void worker(bmi_server *S, Worker *W) {
    (void)hipSetDevice(S->device);
    std::unique_lock<std::mutex> lk(S->mu);
    for (;;) {
        std::shared_ptr<Request> r;
        S->cv_work.wait(lk, [&] {
            bool blocked = false;
            r = take(S, *W, blocked);
            if (!r && blocked && W->cap > 0) {
                // an assembly waits for memory: an idle worker gives its buffer back
                dev_free(W->buf);
                S->resident -= W->cap;
                W->buf = nullptr;
                W->cap = 0;
                S->cv_work.notify_all();
                r = take(S, *W, blocked);
            }
            return r != nullptr || (S->stop && S->queue.empty());
        });
        if (!r) return;
        r->t_admit = Clock::now();
        r->res.queue_ms = ms_between(r->t_submit, r->t_admit);
        const int64_t cap0 = W->cap;
        lk.unlock();
        bmi_matrix *kept = nullptr;
        int32_t st;
        try {
            st = run(S, *W, *r, &kept);
        } catch (...) {
            st = fail_exception();
        }
        const std::string msg = st == BMI_OK ? std::string() : std::string(bmi_last_error());
        lk.lock();
        if (r->kind == BMI_REQ_APPLY) {
            auto m = S->resident_m.find(r->matrix);
            if (m != S->resident_m.end()) {
                m->second.users--;
                if (m->second.released && m->second.users == 0) {
                    bmi_matrix_free(m->second.m);
                    S->resident -= m->second.bytes;
                    S->resident_m.erase(m);
                }
            }
        } else {
            S->resident -= r->pbytes;
            // the worker buffer as it is now, against the growth granted
            S->resident += (W->cap - cap0) - r->grow;
            if (r->kind == BMI_REQ_ASSEMBLE) {
                Resident &R = S->resident_m[r->id];
                if (kept) {
                    R.m = kept;
                    R.bytes = r->mbytes;
                } else {
                    S->resident -= r->mbytes;
                    S->resident_m.erase(r->id);
                }
            }
        }
        finish(S, r, st, msg.c_str());
        S->cv_done.notify_all();
        S->cv_work.notify_all();
    }
}

}  // namespace

extern "C" {

bmi_status bmi_server_create(int32_t device, const bmi_server_options *opts, bmi_server **out) try {
    if (!out) return set_error(BMI_EINVAL, "out is NULL");
    *out = nullptr;
    const int slots = opts ? opts->slots : 4;
    if (slots < 1 || slots > 16) return set_error(BMI_EINVAL, "slots must be 1..16");
    auto S = std::make_unique<bmi_server>();
    S->device = device;
    for (int i = 0; i < slots; ++i) {
        auto W = std::make_unique<Worker>();
        const int32_t st = bmi_context_create(device, &W->ctx);
        if (st != BMI_OK) {
            for (auto &o : S->w) bmi_context_destroy(o->ctx);
            return st;
        }
        S->w.push_back(std::move(W));
    }
    int64_t fr = 0, tot = 0;
    int32_t st = bmi_device_memory(S->w[0]->ctx, &fr, &tot);
    if (st != BMI_OK) {
        for (auto &o : S->w) bmi_context_destroy(o->ctx);
        return st;
    }
    S->budget = (opts && opts->memory_budget > 0) ? opts->memory_budget : std::max<int64_t>(fr - (4LL << 30), 1LL << 30);
    for (auto &W : S->w) S->workers.emplace_back(worker, S.get(), W.get());
    *out = S.release();
    return BMI_OK;
} catch (...) {
    return fail_exception();
}

bmi_status bmi_server_destroy(bmi_server *S) try {
    if (!S) return BMI_OK;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
    }
    S->cv_work.notify_all();
    for (auto &t : S->workers) t.join();
    for (auto &kv : S->resident_m)
        if (kv.second.m) bmi_matrix_free(kv.second.m);
    for (auto &W : S->w) {
        (void)hipSetDevice(S->device);
        dev_free(W->buf);
        bmi_context_destroy(W->ctx);
    }
    delete S;
    return BMI_OK;
} catch (...) {
    return fail_exception();
}

bmi_status bmi_server_submit(bmi_server *S, const bmi_request *q, double *y, uint64_t *id) try {
    if (!S || !q || !id) return set_error(BMI_EINVAL, "NULL argument");
    auto r = std::make_shared<Request>();
    r->kind = q->kind;
    r->nrhs = q->nrhs;
    r->y = y;
    if (q->kind != BMI_REQ_ASSEMBLE && q->kind != BMI_REQ_ASSEMBLE_APPLY && q->kind != BMI_REQ_APPLY)
        return set_error(BMI_EINVAL, "unknown request kind");
    if (q->kind != BMI_REQ_ASSEMBLE) {
        if (q->nrhs < 1 || q->nrhs > 8) return set_error(BMI_EINVAL, "nrhs must be 1..8");
        if (!q->x || !y) return set_error(BMI_EINVAL, "apply requests need x and y");
    }
    int64_t nfun = q->nfun;
    if (q->kind == BMI_REQ_APPLY) {
        std::lock_guard<std::mutex> lk(S->mu);
        auto m = S->resident_m.find(q->matrix);
        if (m == S->resident_m.end() || m->second.released)
            return set_error(BMI_EINVAL, "no resident matrix with that id");
        r->matrix = q->matrix;
        nfun = m->second.nfun;
    } else {
        if (!q->xyz || !q->tri || !q->gidx || !q->coeff || q->nv < 3 || q->nt < 1 || q->nfun < 1)
            return set_error(BMI_EINVAL, "assembly requests need a mesh and a space");
        if (q->dtype != BMI_C64 && q->dtype != BMI_C32) return set_error(BMI_EINVAL, "bad dtype");
        r->xyz.assign(q->xyz, q->xyz + 3 * q->nv);
        r->tri.assign(q->tri, q->tri + 3 * q->nt);
        r->nv = q->nv;
        r->nt = q->nt;
        r->space_kind = q->space_kind;
        r->gidx.assign(q->gidx, q->gidx + 3 * q->nt);
        r->coeff.assign(q->coeff, q->coeff + 3 * q->nt);
        r->op = q->op;
        r->opts = q->opts;
        r->dtype = q->dtype;
        r->mbytes = matrix_bytes(q->nfun, q->dtype);
        r->pbytes = plan_bytes_bound(q->xyz, q->nv, q->tri, q->nt, q->opts);
        if (r->pbytes < 0) return set_error(BMI_EINVAL, "vertex index out of range");
        if (r->mbytes + r->pbytes > S->budget)
            return set_error(BMI_ENOMEM, "request needs more device memory than the budget");
    }
    r->nfun = nfun;
    if (q->kind != BMI_REQ_ASSEMBLE) r->x.assign(q->x, q->x + 2 * nfun * q->nrhs);
    r->t_submit = Clock::now();
    {
        std::lock_guard<std::mutex> lk(S->mu);
        if (S->stop) return set_error(BMI_ECANCELED, "server is shutting down");
        r->id = S->next_id++;
        if (r->kind == BMI_REQ_ASSEMBLE) S->resident_m[r->id].nfun = r->nfun;
        S->pending[r->id] = r;
        S->queue.push_back(r);
        *id = r->id;
    }
    S->cv_work.notify_all();
    return BMI_OK;
} catch (...) {
    return fail_exception();
}

bmi_status bmi_server_wait(bmi_server *S, uint64_t id, int32_t timeout_ms, bmi_result *res) try {
    if (!S || !res) return set_error(BMI_EINVAL, "NULL argument");
    std::unique_lock<std::mutex> lk(S->mu);
    auto it = S->pending.find(id);
    if (it == S->pending.end()) return set_error(BMI_EINVAL, "unknown request id");
    std::shared_ptr<Request> r = it->second;
    auto ready = [&] { return r->done; };
    if (timeout_ms < 0) S->cv_done.wait(lk, ready);
    else if (!S->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready))
        return set_error(BMI_EBUSY, "request not complete");
    *res = r->res;
    S->pending.erase(id);
    return BMI_OK;
} catch (...) {
    return fail_exception();
}

bmi_status bmi_server_release(bmi_server *S, uint64_t id) try {
    if (!S) return set_error(BMI_EINVAL, "NULL argument");
    std::lock_guard<std::mutex> lk(S->mu);
    auto m = S->resident_m.find(id);
    if (m == S->resident_m.end() || m->second.released) return set_error(BMI_EINVAL, "no resident matrix with that id");
    if (!m->second.m) return set_error(BMI_EBUSY, "matrix is still being assembled");
    m->second.released = true;
    if (m->second.users == 0) {
        bmi_matrix_free(m->second.m);
        S->resident -= m->second.bytes;
        S->resident_m.erase(m);
        S->cv_work.notify_all();
    }
    return BMI_OK;
} catch (...) {
    return fail_exception();
}

bmi_status bmi_request_bytes(const bmi_request *q, int64_t *matrix_bytes_out, int64_t *plan_bytes_out) try {
    if (!q || !matrix_bytes_out || !plan_bytes_out) return set_error(BMI_EINVAL, "NULL argument");
    if (q->kind == BMI_REQ_APPLY) {
        *matrix_bytes_out = *plan_bytes_out = 0;
        return BMI_OK;
    }
    if (!q->xyz || !q->tri || q->nv < 3 || q->nt < 1 || q->nfun < 1)
        return set_error(BMI_EINVAL, "assembly requests need a mesh and a space");
    if (q->dtype != BMI_C64 && q->dtype != BMI_C32) return set_error(BMI_EINVAL, "bad dtype");
    const int64_t pb = plan_bytes_bound(q->xyz, q->nv, q->tri, q->nt, q->opts);
    if (pb < 0) return set_error(BMI_EINVAL, "vertex index out of range");
    *matrix_bytes_out = matrix_bytes(q->nfun, q->dtype);
    *plan_bytes_out = pb;
    return BMI_OK;
} catch (...) {
    return fail_exception();
}

bmi_status bmi_server_memory(const bmi_server *S, int64_t *resident, int64_t *peak) try {
    if (!S || !resident || !peak) return set_error(BMI_EINVAL, "NULL argument");
    std::lock_guard<std::mutex> lk(const_cast<bmi_server *>(S)->mu);
    *resident = S->resident;
    *peak = S->peak;
    return BMI_OK;
} catch (...) {
    return fail_exception();
}

}  // extern "C"
