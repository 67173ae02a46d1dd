// server.hip -- continuous-batching request loop over H-matrices cached in HBM
// (host code; no counterpart in the reference, which is a synchronous library).
//
// Requests arrive from any thread.  Worker threads ("slots") take jobs:
//   - a product pass on one cached H-matrix: every apply request waiting on it,
//     in arrival order, up to max_rhs_per_pass right-hand sides, gathered into
//     one column block, multiplied with one hmg_apply_h_matrix_mvp_multi call on
//     the H-matrix's own streams and scattered back;
//   - a build: init, tree (setup), admission against the memory budget (evicting
//     least recently used idle H-matrices), ACA factors.
// At most one job runs on an H-matrix at a time (its engine holds per-product
// scratch); jobs on different H-matrices run concurrently on the device.  The
// oldest waiting request decides which job a free slot takes next, except that
// builds never occupy the last product_slots slots.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "hmg_internal.h"
#include "../../include/hmglib_server.h"

namespace hmg {
namespace {

using Clock = std::chrono::steady_clock;

struct Request {
    int64_t id = 0;
    bool build = false;
    int64_t matrix = 0;
    // build
    hmg_build_request bp{};
    std::vector<double> coords[2];
    // apply
    const double* X = nullptr;
    double* Y = nullptr;
    int nrhs = 0;
    hmg_request_info info{-1, 0.0, -1.0, -1.0, 0, 0, {0}};
    bool done = false;
};

struct Matrix {
    int64_t id = 0;
    h_matrix_data data{};
    bool inited = false;
    system_assembler* kern = nullptr;
    bool ready = false;          // built and usable
    bool failed = false;
    bool busy = false;           // a job runs on it
    bool evicted = false;        // freed to make room (or released)
    bool release_pending = false;
    int64_t bytes = 0;           // device bytes counted against the budget
    double last_use = 0.0;
    std::deque<Request*> pending;  // apply requests waiting for a pass
    Request* build_req = nullptr;  // build not yet started
    DevBuf<double> Xb, Yb;         // merged pass buffers
};

}  // namespace
}  // namespace hmg

struct hmg_server {
    hmg_server_config cfg{};
    hmg::Clock::time_point t0;
    std::mutex mu;
    std::condition_variable cv_work;  // a job may have become available
    std::condition_variable cv_done;  // a request completed or memory was freed
    std::unordered_map<int64_t, std::unique_ptr<hmg::Request>> reqs;
    std::unordered_map<int64_t, std::unique_ptr<hmg::Matrix>> mats;
    std::deque<int64_t> builds;       // matrices whose build has not started
    std::vector<std::thread> workers;
    // one stream per slot, created together with the first build: with one
    // hardware queue per stream (HIP hands queues out round-robin) a long build
    // on one slot never holds back the kernels of another slot's job
    std::vector<hipStream_t> slot_streams;
    bool stop = false;
    int running = 0;                  // jobs on slots
    int running_builds = 0;           // builds among them
    int waiting = 0;                  // builds waiting for memory to free up
    int64_t used = 0;                 // bytes of cached and building H-matrices
    int64_t next_req = 1, next_mat = 1;
    int64_t n_evict = 0, n_done = 0, n_failed = 0, n_passes = 0, n_shared = 0, n_builds = 0;
    int64_t outstanding = 0;          // submitted, not completed
    // kernel descriptors of evicted H-matrices, reused by later builds (hipFree
    // would synchronize the whole device while other requests run)
    std::vector<system_assembler*> free_kernels;

    double now() const { return std::chrono::duration<double>(hmg::Clock::now() - t0).count(); }
};

namespace hmg {
namespace {

void finish(hmg_server* s, Request* r, int status, double t_end)
{
    r->info.status = status;
    r->info.t_end = t_end;
    if (status != HMG_OK) {
        // the failing call's message (set on this worker thread)
        std::snprintf(r->info.message, sizeof(r->info.message), "%s", hmg_last_error());
    }
    r->done = true;
    if (status == HMG_OK)
        s->n_done++;
    else
        s->n_failed++;
    s->outstanding--;
}

// free a matrix's device memory (stream-ordered on its engine's stream); the
// caller holds no lock and the matrix is no longer reachable for new work; its
// kernel descriptor goes back to the server's free list
void destroy_matrix(hmg_server* s, Matrix* m)
{
    m->Xb.release();
    m->Yb.release();
    if (m->inited) (void)hmg_destroy_h_matrix_data(&m->data);
    m->inited = false;
    if (m->kern) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->free_kernels.push_back(m->kern);
    }
    m->kern = nullptr;
}

// a kernel descriptor holding (kind, regularization, length scale), reusing a
// freed one when available; written on `st`
system_assembler* kernel_for(hmg_server* s, const hmg_build_request& b, hipStream_t st)
{
    system_assembler* k = nullptr;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (!s->free_kernels.empty()) {
            k = s->free_kernels.back();
            s->free_kernels.pop_back();
        }
    }
    if (!k) return hmg_kernel_new(b.kernel_kind, b.regularization, b.length_scale);
    if (b.kernel_kind < HMG_KERNEL_GAUSSIAN_REF || b.kernel_kind > HMG_KERNEL_EXPONENTIAL) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->free_kernels.push_back(k);
        set_error("unknown kernel kind");
        return nullptr;
    }
    system_assembler h{};
    h.kind = b.kernel_kind;
    h.regularization = b.regularization;
    h.length_scale = b.length_scale;
    if (upload_bytes(k, &h, sizeof(h), st) != HMG_OK) {
        set_error("kernel descriptor upload failed");
        return nullptr;
    }
    return k;
}

// the job a free slot takes: the matrix whose oldest waiting request is oldest
// (apply passes on ready, idle matrices; builds in arrival order)
Matrix* pick_job(hmg_server* s, bool* is_build)
{
    Matrix* best = nullptr;
    double best_t = 0.0;
    for (auto& kv : s->mats) {
        Matrix* m = kv.second.get();
        if (m->busy || !m->ready || m->pending.empty()) continue;
        const double t = m->pending.front()->info.t_submit;
        if (!best || t < best_t) {
            best = m;
            best_t = t;
            *is_build = false;
        }
    }
    // builds leave product_slots slots to product passes (a build runs for up to a
    // second, a product for milliseconds)
    const int build_slots = s->cfg.slots - s->cfg.product_slots;
    if (!s->builds.empty() && s->running_builds < build_slots) {
        Matrix* m = s->mats.at(s->builds.front()).get();
        const double t = m->build_req->info.t_submit;
        if (!best || t < best_t) {
            best = m;
            *is_build = true;
        }
    }
    return best;
}

// evict least recently used idle matrices without pending work until `need`
// more bytes fit; returns the matrices to destroy (outside the lock)
std::vector<std::unique_ptr<Matrix>> make_room(hmg_server* s, int64_t need)
{
    std::vector<std::unique_ptr<Matrix>> out;
    while (s->used + need > s->cfg.hbm_budget_bytes) {
        Matrix* victim = nullptr;
        for (auto& kv : s->mats) {
            Matrix* m = kv.second.get();
            if (!m->ready || m->busy || !m->pending.empty()) continue;
            if (!victim || m->last_use < victim->last_use) victim = m;
        }
        if (!victim) break;
        s->used -= victim->bytes;
        s->n_evict++;
        auto it = s->mats.find(victim->id);
        out.push_back(std::move(it->second));
        s->mats.erase(it);
    }
    return out;
}

hmg_status run_build(hmg_server* s, Matrix* m, Request* r, hipStream_t slot, std::unique_lock<std::mutex>& lk)
{
    const hmg_build_request& b = r->bp;
    lk.unlock();
    h_matrix_data& d = m->data;
    int pc[2] = {b.n[0], b.n[1]};
    hmg_status st = engine_init(&d, pc, b.dim, b.bits, slot);
    if (st == HMG_OK) {
        m->inited = true;
        Engine* e = engine_of(&d);
        for (int i = 0; i < 2 && st == HMG_OK; i++) {
            const std::vector<double>& c = r->coords[r->coords[1].empty() ? 0 : i];
            for (int k = 0; k < b.dim && st == HMG_OK; k++)
                st = upload_bytes(d.coords_d[i][k], c.data() + (size_t)k * b.n[i], sizeof(double) * b.n[i],
                                  e->stream);  // complete when it returns, error or not
        }
        r->coords[0].clear();
        r->coords[0].shrink_to_fit();
        r->coords[1].clear();
        r->coords[1].shrink_to_fit();
    }
    if (st == HMG_OK) {
        d.eta = b.eta;
        d.max_level = b.max_level;
        d.c_leaf = b.c_leaf;
        d.k = b.k;
        d.epsilon = b.epsilon;
        d.max_batched_dense_size = 1 << 20;
        d.max_batched_aca_size = 1 << 30;
        st = hmg_set_option(&d, HMG_OPT_PRECISION, s->cfg.precision);
        if (st == HMG_OK) st = hmg_set_option(&d, HMG_OPT_ACA_TOLERANCE, b.aca_tolerance);
        if (st == HMG_OK) st = hmg_set_option(&d, HMG_OPT_EXACT_BOXES, b.exact_boxes);
        // every FP64 pass through the block product: a request's result is then
        // bitwise the same whatever other requests share its pass
        if (st == HMG_OK) st = hmg_set_option(&d, HMG_OPT_FP64_BLOCK_RHS, 1);
    }
    if (st == HMG_OK) st = hmg_setup_h_matrix(&d);
    if (st == HMG_OK) {
        m->kern = kernel_for(s, b, engine_of(&d)->stream);
        if (!m->kern) st = HMG_ERR_INVALID;
        d.assem = m->kern;
    }
    int64_t mem[2] = {0, 0};
    if (st == HMG_OK) st = hmg_memory(&d, mem);
    lk.lock();
    if (st != HMG_OK) return st;
    // admission: the tree is built; the factors and product buffers must fit
    const int64_t need = mem[0] + mem[1];
    for (;;) {
        std::vector<std::unique_ptr<Matrix>> victims = make_room(s, need);
        if (!victims.empty()) {
            lk.unlock();
            for (auto& v : victims) destroy_matrix(s, v.get());
            victims.clear();
            lk.lock();
        }
        if (s->used + need <= s->cfg.hbm_budget_bytes) break;
        // fail unless some other job that is not itself waiting can still free memory
        if (s->running - s->waiting <= 1) {
            set_error("H-matrix needs " + std::to_string(need) + " bytes; the budget of " +
                      std::to_string(s->cfg.hbm_budget_bytes) + " bytes cannot hold it");
            return HMG_ERR_OOM;
        }
        s->waiting++;
        s->cv_done.wait(lk);  // other jobs will finish or leave evictable matrices
        s->waiting--;
    }
    s->used += need;
    m->bytes = need;
    lk.unlock();
    st = hmg_precompute_aca(&d);
    lk.lock();
    return st;
}

hmg_status run_pass(hmg_server* s, Matrix* m, std::vector<Request*>& batch, int total, hipStream_t slot,
                    std::unique_lock<std::mutex>& lk)
{
    lk.unlock();
    hmg_status st = HMG_OK;
    Engine* e = engine_of(&m->data);
    engine_bind_stream(e, slot, &m->data);
    AllocScope scope(e->stream);  // merged pass buffers are stream-ordered
    const int64_t n1 = m->data.point_count[0], n2 = m->data.point_count[1];
    const double* X = batch[0]->X;
    double* Y = batch[0]->Y;
    if (batch.size() > 1) {
        // gather the requests' right-hand sides into one column block
        st = m->Xb.alloc((size_t)n2 * total);
        if (st == HMG_OK) st = m->Yb.alloc((size_t)n1 * total);
        int64_t c = 0;
        for (Request* r : batch) {
            if (st != HMG_OK) break;
            if (hipMemcpyAsync(m->Xb.p + c * n2, r->X, sizeof(double) * n2 * r->nrhs, hipMemcpyDeviceToDevice,
                               e->stream) != hipSuccess)
                st = HMG_ERR_HIP;
            c += r->nrhs;
        }
        X = m->Xb.p;
        Y = m->Yb.p;
    }
    if (st == HMG_OK) st = hmg_apply_h_matrix_mvp_multi(X, Y, total, &m->data);
    if (st == HMG_OK && batch.size() > 1) {
        int64_t c = 0;
        for (Request* r : batch) {
            if (hipMemcpyAsync(r->Y, m->Yb.p + c * n1, sizeof(double) * n1 * r->nrhs, hipMemcpyDeviceToDevice,
                               e->stream) != hipSuccess) {
                st = HMG_ERR_HIP;
                break;
            }
            c += r->nrhs;
        }
        if (st == HMG_OK && hipStreamSynchronize(e->stream) != hipSuccess) st = HMG_ERR_HIP;
    }
    lk.lock();
    return st;
}

void worker(hmg_server* s, int slot_index)
{
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        bool is_build = false;
        Matrix* m = nullptr;
        s->cv_work.wait(lk, [&] { return s->stop || (m = pick_job(s, &is_build)) != nullptr; });
        if (!m) return;  // stopping with nothing to do
        const hipStream_t slot = s->slot_streams[slot_index];
        m->busy = true;
        s->running++;
        const double t_start = s->now();
        if (is_build) {
            s->builds.pop_front();
            Request* r = m->build_req;
            m->build_req = nullptr;
            r->info.t_start = t_start;
            s->running_builds++;
            const hmg_status st = run_build(s, m, r, slot, lk);
            s->running_builds--;
            const double t_end = s->now();
            if (st == HMG_OK) {
                m->ready = true;
                // admitted at the build's peak (the mixed-precision ACA works in FP64
                // temporaries); held from now on: what it owns plus its first product's buffers
                int64_t mem[2] = {0, 0};
                if (hmg_memory(&m->data, mem) == HMG_OK) {
                    s->used += mem[0] + mem[1] - m->bytes;
                    m->bytes = mem[0] + mem[1];
                }
                s->n_builds++;
                finish(s, r, HMG_OK, t_end);
            } else {
                m->failed = true;
                finish(s, r, st, t_end);
                for (Request* a : m->pending) finish(s, a, st, t_end);
                m->pending.clear();
            }
            m->last_use = t_end;
        } else {
            // merge waiting requests in arrival order up to max_rhs_per_pass
            std::vector<Request*> batch;
            int total = 0;
            while (!m->pending.empty()) {
                Request* r = m->pending.front();
                if (!batch.empty() && total + r->nrhs > s->cfg.max_rhs_per_pass) break;
                batch.push_back(r);
                total += r->nrhs;
                m->pending.pop_front();
            }
            for (Request* r : batch) {
                r->info.t_start = t_start;
                r->info.merged = (int)batch.size();
            }
            const hmg_status st = run_pass(s, m, batch, total, slot, lk);
            const double t_end = s->now();
            s->n_passes++;
            if (batch.size() > 1) s->n_shared += (int64_t)batch.size();
            for (Request* r : batch) finish(s, r, st, t_end);
            m->last_use = t_end;
        }
        m->busy = false;
        s->running--;
        std::unique_ptr<Matrix> drop;
        if ((m->failed || m->release_pending) && m->pending.empty()) {
            s->used -= m->bytes;
            m->evicted = true;
            auto it = s->mats.find(m->id);
            drop = std::move(it->second);
            s->mats.erase(it);
        }
        s->cv_done.notify_all();
        s->cv_work.notify_all();
        if (drop) {
            lk.unlock();
            destroy_matrix(s, drop.get());
            lk.lock();
        }
    }
}

}  // namespace
}  // namespace hmg

using namespace hmg;

extern "C" {

hmg_status hmg_server_create(const hmg_server_config* cfg, hmg_server** out)
{
    if (!out) {
        set_error("null output");
        return HMG_ERR_INVALID;
    }
    hmg_server_config c{(int64_t)200e9, 4, kMxRhs, HMG_PREC_MIXED};
    if (cfg) c = *cfg;
    if (c.hbm_budget_bytes <= 0 || c.slots < 1 || c.slots > 64 || c.max_rhs_per_pass < 1 ||
        (c.precision != HMG_PREC_FP64 && c.precision != HMG_PREC_MIXED) || c.product_slots < 0 ||
        c.product_slots >= c.slots) {
        set_error("invalid server configuration");
        return HMG_ERR_INVALID;
    }
    hmg_server* s = new hmg_server();
    s->cfg = c;
    s->t0 = Clock::now();
    for (int i = 0; i < c.slots; i++) s->workers.emplace_back(worker, s, i);
    *out = s;
    return HMG_OK;
}

hmg_status hmg_server_drain(hmg_server* s)
{
    if (!s) return HMG_ERR_INVALID;
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [&] { return s->outstanding == 0; });
    return HMG_OK;
}

hmg_status hmg_server_destroy(hmg_server* s)
{
    if (!s) return HMG_OK;
    hmg_server_drain(s);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    for (auto& t : s->workers) t.join();
    for (auto& kv : s->mats) destroy_matrix(s, kv.second.get());
    for (system_assembler* k : s->free_kernels) hmg_kernel_free(k);
    for (hipStream_t st : s->slot_streams) {
        (void)hipStreamSynchronize(st);
        release_scratch(st);
        (void)hipStreamSynchronize(st);
        (void)hipStreamDestroy(st);
    }
    delete s;
    return HMG_OK;
}

hmg_status hmg_server_submit_build(hmg_server* s, const hmg_build_request* b, int64_t* request, int64_t* matrix)
{
    if (!s || !b || b->dim < 1 || b->dim > MAX_POINT_DIMENSION || b->n[0] < 1 || b->n[1] < 1 ||
        !b->coords[0] || !b->coords[1] || b->k < 0 || b->c_leaf < 0 || b->max_level < 0) {
        set_error("invalid build request");
        return HMG_ERR_INVALID;
    }
    auto r = std::make_unique<Request>();
    r->build = true;
    r->bp = *b;
    const bool same = b->coords[1] == b->coords[0] && b->n[1] == b->n[0];
    r->coords[0].assign(b->coords[0], b->coords[0] + (size_t)b->dim * b->n[0]);
    if (!same) r->coords[1].assign(b->coords[1], b->coords[1] + (size_t)b->dim * b->n[1]);
    auto m = std::make_unique<Matrix>();
    std::lock_guard<std::mutex> lk(s->mu);
    // the slot streams, all at once with the first build (every job follows one)
    while ((int)s->slot_streams.size() < s->cfg.slots) {
        hipStream_t st = nullptr;
        if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
            set_error("stream creation failed");
            return HMG_ERR_HIP;
        }
        s->slot_streams.push_back(st);
    }
    r->id = s->next_req++;
    m->id = s->next_mat++;
    r->matrix = m->id;
    r->info.matrix = m->id;
    r->info.t_submit = s->now();
    m->build_req = r.get();
    if (request) *request = r->id;
    if (matrix) *matrix = m->id;
    s->builds.push_back(m->id);
    s->mats.emplace(m->id, std::move(m));
    s->reqs.emplace(r->id, std::move(r));
    s->outstanding++;
    s->cv_work.notify_one();
    return HMG_OK;
}

hmg_status hmg_server_submit_apply(hmg_server* s, int64_t matrix, const double* X, double* Y, int nrhs,
                                   int64_t* request)
{
    if (!s || !X || !Y || nrhs < 1) {
        set_error("invalid apply request");
        return HMG_ERR_INVALID;
    }
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->mats.find(matrix);
    if (it == s->mats.end() || it->second->failed || it->second->release_pending) {
        set_error("H-matrix " + std::to_string(matrix) + " is not cached (unknown, failed, evicted or released)");
        return HMG_ERR_NOT_READY;
    }
    auto r = std::make_unique<Request>();
    r->id = s->next_req++;
    r->matrix = matrix;
    r->X = X;
    r->Y = Y;
    r->nrhs = nrhs;
    r->info.matrix = matrix;
    r->info.t_submit = s->now();
    it->second->pending.push_back(r.get());
    if (request) *request = r->id;
    s->reqs.emplace(r->id, std::move(r));
    s->outstanding++;
    s->cv_work.notify_one();
    return HMG_OK;
}

hmg_status hmg_server_query(hmg_server* s, int64_t request, hmg_request_info* info)
{
    if (!s) return HMG_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(request);
    if (it == s->reqs.end()) {
        set_error("unknown request");
        return HMG_ERR_INVALID;
    }
    if (info) *info = it->second->info;
    return HMG_OK;
}

hmg_status hmg_server_wait(hmg_server* s, int64_t request, hmg_request_info* info)
{
    if (!s) return HMG_ERR_INVALID;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(request);
    if (it == s->reqs.end()) {
        set_error("unknown request");
        return HMG_ERR_INVALID;
    }
    Request* r = it->second.get();
    s->cv_done.wait(lk, [&] { return r->done; });
    if (info) *info = r->info;
    return HMG_OK;
}

hmg_status hmg_server_release(hmg_server* s, int64_t matrix)
{
    if (!s) return HMG_ERR_INVALID;
    std::unique_ptr<Matrix> drop;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->mats.find(matrix);
        if (it == s->mats.end()) return HMG_OK;
        Matrix* m = it->second.get();
        m->release_pending = true;
        if (m->ready && !m->busy && m->pending.empty()) {
            s->used -= m->bytes;
            drop = std::move(it->second);
            s->mats.erase(it);
        }
    }
    if (drop) destroy_matrix(s, drop.get());
    s->cv_done.notify_all();
    return HMG_OK;
}

hmg_status hmg_server_stats(hmg_server* s, int64_t out[8])
{
    if (!s || !out) return HMG_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    int64_t cached = 0;
    for (auto& kv : s->mats) cached += kv.second->ready ? 1 : 0;
    out[0] = s->used;
    out[1] = cached;
    out[2] = s->n_evict;
    out[3] = s->n_done;
    out[4] = s->n_failed;
    out[5] = s->n_passes;
    out[6] = s->n_shared;
    out[7] = s->n_builds;
    return HMG_OK;
}

int hmg_server_resident(hmg_server* s, int64_t* ids, int cap)
{
    if (!s) return 0;
    std::lock_guard<std::mutex> lk(s->mu);
    std::vector<int64_t> v;
    for (auto& kv : s->mats)
        if (kv.second->ready && !kv.second->release_pending) v.push_back(kv.first);
    std::sort(v.begin(), v.end());
    const int n = (int)v.size();
    for (int i = 0; i < n && i < cap && ids; i++) ids[i] = v[i];
    return n;
}

}  // extern "C"
