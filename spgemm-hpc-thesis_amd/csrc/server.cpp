// server.cpp -- served products: continuous batching over resident batches.
//
// Requests (A, B) arrive at any time from any host thread (spg_server_submit).
// A builder thread admits queued requests in arrival order while their
// memory estimate fits the HBM budget, packs them into one batch arena and
// uploads it on its own stream; a runner thread executes the built batches on
// a second stream, so the upload of the next batch overlaps the compute of the
// current one, and requests that arrive while a batch runs join the next one.
// Results stay in device memory until the client copies them out
// (spg_server_result) and releases them (spg_server_release); their memory
// stays charged to the budget until then.
//
// Operands either come with the request (host views, copied) or are registered
// once (spg_server_register: resident in device memory, charged to the budget
// until unregistered) and referenced by handle; batches of resident requests
// are packed on the device.
//
// Admission estimate per request: the operands' arena copy (12 B per entry +
// 8 B per row; twice for host operands, which also pass through a staging
// copy) plus 12 B per scalar product of A.B (the symbolic phase's upper bound
// on C's entries, summed on the host and cached per resident pair) plus 32 B
// per row of workspace.  A request larger than the whole budget is rejected
// with SPG_ENOMEM at submission; one whose A has a column outside B's rows
// with SPG_ERANGE.
//
// Every result is bit-identical to computing the same product alone: a batch
// computes each product's rows independently of its neighbours.
#include "spgemm_amd.h"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

namespace {

struct HostCsr {
    int32_t rows = 0, cols = 0;
    std::vector<int64_t> rp;
    std::vector<int32_t> col;
    std::vector<double> val;
    spg_csr_view view() const {
        spg_csr_view v{};
        v.rows = rows;
        v.cols = cols;
        v.nnz = rp.empty() ? 0 : rp.back();
        v.row_ptr = rp.data();
        v.col_idx = col.data();
        v.values = val.data();
        v.location = SPG_HOST;
        return v;
    }
};

// A matrix registered with the server: resident in device memory, with host
// copies of its row offsets and columns for admission estimates.
struct Resident {
    spg_device_matrix *m = nullptr;
    int32_t rows = 0, cols = 0;
    std::vector<int64_t> rp;
    std::vector<int32_t> col;
    uint64_t bytes = 0;
};

struct Request {
    int64_t id = 0;
    HostCsr a, b;                       // host operands (copied at submission)
    std::shared_ptr<Resident> ra, rb;   // or registered operands
    uint64_t estimate = 0;
    double t_submit = 0, t_admit = 0, t_done = 0;
    bool done = false;
    spg_status status = SPG_OK;
    std::string error;
    std::shared_ptr<spg_batch> batch;  // holds the result until released
    int32_t index = 0;                 // product index inside the batch
    int64_t nnz = 0;
};

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

struct spg_server {
    spg_context *build_ctx = nullptr;  // arenas are built and uploaded on its stream
    spg_context *run_ctx = nullptr;    // batches run on its stream
    spg_context *reg_ctx = nullptr;    // registrations (client threads, under reg_mu)
    std::mutex reg_mu;
    uint64_t budget = 0;
    int max_batch = 0;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<std::shared_ptr<Request>> queue;  // submitted, not admitted
    std::deque<std::pair<std::shared_ptr<spg_batch>, std::vector<std::shared_ptr<Request>>>> built;
    std::map<int64_t, std::shared_ptr<Request>> live;  // not yet released
    std::map<int64_t, std::shared_ptr<Resident>> resident;
    std::map<std::pair<int64_t, int64_t>, uint64_t> pair_products;  // cached work bounds of resident pairs
    int64_t next_handle = 1;
    uint64_t charged = 0;
    int64_t next_id = 1;
    bool stop = false;
    std::thread builder, runner;
    std::vector<spg_batch *> free_batches;  // built batches whose results were all released
    std::vector<spg_batch *> all_batches;
    // statistics
    int64_t batches = 0, completed = 0;
    double busy_s = 0;
};

namespace {

// Scalar products of A.B (the symbolic phase's bound on C's entries); false
// when an A column lies outside B's rows.
bool work_bound(const std::vector<int64_t> &arp, const std::vector<int32_t> &acol, const std::vector<int64_t> &brp,
                int32_t brows, uint64_t &p) {
    p = 0;
    const int64_t n = arp.empty() ? 0 : arp.back();
    for (int64_t e = 0; e < n; ++e) {
        const int32_t k = acol[(size_t)e];
        if (k < 0 || k >= brows) return false;
        p += (uint64_t)(brp[(size_t)k + 1] - brp[(size_t)k]);
    }
    return true;
}

uint64_t operand_bytes(int64_t rows, int64_t nnz) { return 12 * (uint64_t)nnz + 8 * (uint64_t)(rows + 1); }

void builder_loop(spg_server *s) {
    spg_context_bind(s->build_ctx);  // this thread's current device
    for (;;) {
        std::vector<std::shared_ptr<Request>> take;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv.wait(lk, [&] {
                return s->stop || (!s->queue.empty() && s->charged + s->queue.front()->estimate <= s->budget &&
                                   s->built.size() < 2);
            });
            if (s->stop) return;
            // admit in arrival order while the budget allows (one kind of
            // operands per batch: all host or all resident)
            const bool resident = (bool)s->queue.front()->ra;
            while (!s->queue.empty() && (int)take.size() < s->max_batch &&
                   (bool)s->queue.front()->ra == resident &&
                   s->charged + s->queue.front()->estimate <= s->budget) {
                auto r = s->queue.front();
                s->queue.pop_front();
                s->charged += r->estimate;
                r->t_admit = now_s();
                take.push_back(r);
            }
        }
        std::vector<spg_csr_view> av, bv;
        for (auto &r : take) {
            if (r->ra) {
                spg_csr_view x{}, y{};
                spg_device_matrix_view(r->ra->m, &x);
                spg_device_matrix_view(r->rb->m, &y);
                av.push_back(x);
                bv.push_back(y);
            } else {
                av.push_back(r->a.view());
                bv.push_back(r->b.view());
            }
        }
        spg_batch *raw = nullptr;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            if (!s->free_batches.empty()) {
                raw = s->free_batches.back();
                s->free_batches.pop_back();
            }
        }
        spg_status st;
        std::string msg;
        if (raw) {
            st = spg_batch_rebuild(raw, (int32_t)take.size(), av.data(), bv.data());
            if (st != SPG_OK) msg = spg_last_error(s->build_ctx);
        } else {
            st = spg_batch_create(s->build_ctx, (int32_t)take.size(), av.data(), bv.data(), &raw);
            if (st != SPG_OK) msg = spg_last_error(s->build_ctx);
        }
        std::lock_guard<std::mutex> lk(s->mu);
        if (raw && std::find(s->all_batches.begin(), s->all_batches.end(), raw) == s->all_batches.end())
            s->all_batches.push_back(raw);
        if (st != SPG_OK) {
            if (raw) s->free_batches.push_back(raw);
            for (auto &r : take) {
                r->status = st;
                r->error = msg;
                r->done = true;
                r->t_done = now_s();
            }
        } else {
            // back to the free pool when its last request is released
            std::shared_ptr<spg_batch> b(raw, [s](spg_batch *p) {
                std::lock_guard<std::mutex> g(s->mu);
                s->free_batches.push_back(p);
                s->cv.notify_all();
            });
            for (size_t i = 0; i < take.size(); ++i) {
                take[i]->batch = b;
                take[i]->index = (int32_t)i;
            }
            s->built.emplace_back(b, take);
        }
        s->cv.notify_all();
    }
}

void runner_loop(spg_server *s) {
    void *run_stream = nullptr;
    spg_context_bind(s->run_ctx);
    spg_context_stream(s->run_ctx, &run_stream);
    for (;;) {
        std::shared_ptr<spg_batch> b;
        std::vector<std::shared_ptr<Request>> reqs;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv.wait(lk, [&] { return s->stop || !s->built.empty(); });
            if (s->stop) return;
            b = s->built.front().first;
            reqs = std::move(s->built.front().second);
            s->built.pop_front();
            s->cv.notify_all();  // the builder may prepare the next batch
        }
        spg_batch_set_stream(b.get(), run_stream);
        const double t0 = now_s();
        int64_t total = 0;
        spg_status st = spg_batch_run(b.get(), &total);
        std::string msg = st == SPG_OK ? "" : spg_batch_last_error(b.get());
        std::vector<int64_t> nnz(reqs.size(), 0);
        for (size_t i = 0; st == SPG_OK && i < reqs.size(); ++i) {
            st = spg_batch_product_nnz(b.get(), (int32_t)i, &nnz[i]);
            if (st != SPG_OK) msg = spg_batch_last_error(b.get());
        }
        const double t1 = now_s();
        std::lock_guard<std::mutex> lk(s->mu);
        s->busy_s += t1 - t0;
        s->batches += 1;
        for (size_t i = 0; i < reqs.size(); ++i) {
            reqs[i]->status = st;
            reqs[i]->error = msg;
            reqs[i]->nnz = nnz[i];
            reqs[i]->done = true;
            reqs[i]->t_done = t1;
            reqs[i]->a = HostCsr{};  // operands are no longer needed
            reqs[i]->b = HostCsr{};
            reqs[i]->ra.reset();
            reqs[i]->rb.reset();
            s->completed += 1;
        }
        s->cv.notify_all();
    }
}

spg_status copy_view(const spg_csr_view *v, HostCsr &out) {
    if (!v || v->location != SPG_HOST || !v->row_ptr || v->rows < 0 || v->cols < 0 || v->nnz < 0) return SPG_EINVAL;
    if (v->row_ptr[0] != 0 || v->row_ptr[v->rows] != v->nnz) return SPG_EINVAL;
    if (v->nnz > 0 && (!v->col_idx || !v->values)) return SPG_EINVAL;
    out.rows = v->rows;
    out.cols = v->cols;
    out.rp.assign(v->row_ptr, v->row_ptr + v->rows + 1);
    out.col.assign(v->col_idx, v->col_idx + v->nnz);
    out.val.assign(v->values, v->values + v->nnz);
    return SPG_OK;
}

}  // namespace

extern "C" {

spg_status spg_server_create(int device, uint64_t budget_bytes, int max_batch, spg_server **out) {
    if (!out || budget_bytes == 0) return SPG_EINVAL;
    *out = nullptr;
    spg_server *s = new (std::nothrow) spg_server();
    if (!s) return SPG_ENOMEM;
    spg_status st = spg_create(&s->build_ctx, device);
    if (st == SPG_OK) st = spg_create(&s->run_ctx, device);
    if (st == SPG_OK) st = spg_create(&s->reg_ctx, device);
    if (st != SPG_OK) {
        if (s->run_ctx) spg_destroy(s->run_ctx);
        if (s->build_ctx) spg_destroy(s->build_ctx);
        delete s;
        return st;
    }
    s->budget = budget_bytes;
    s->max_batch = max_batch > 0 ? max_batch : 4096;
    s->builder = std::thread(builder_loop, s);
    s->runner = std::thread(runner_loop, s);
    *out = s;
    return SPG_OK;
}

spg_status spg_server_submit(spg_server *s, const spg_csr_view *A, const spg_csr_view *B, int64_t *request_id) {
    if (!s || !request_id) return SPG_EINVAL;
    auto r = std::make_shared<Request>();
    if (copy_view(A, r->a) != SPG_OK || copy_view(B, r->b) != SPG_OK) return SPG_EINVAL;
    if (r->a.cols != r->b.rows) return SPG_EINVAL;
    uint64_t p = 0;
    if (!work_bound(r->a.rp, r->a.col, r->b.rp, r->b.rows, p)) return SPG_ERANGE;  // fail() in the reference
    r->estimate = 2 * (operand_bytes(r->a.rows, r->a.rp.back()) + operand_bytes(r->b.rows, r->b.rp.back())) +
                  12 * p + 32 * (uint64_t)r->a.rows;
    if (r->estimate > s->budget) return SPG_ENOMEM;
    r->t_submit = now_s();
    std::lock_guard<std::mutex> lk(s->mu);
    r->id = s->next_id++;
    s->live[r->id] = r;
    s->queue.push_back(r);
    *request_id = r->id;
    s->cv.notify_all();
    return SPG_OK;
}

spg_status spg_server_register(spg_server *s, const spg_csr_view *host, int64_t *handle) {
    if (!s || !handle) return SPG_EINVAL;
    auto m = std::make_shared<Resident>();
    HostCsr h;
    if (copy_view(host, h) != SPG_OK) return SPG_EINVAL;
    m->rows = h.rows;
    m->cols = h.cols;
    m->bytes = operand_bytes(h.rows, h.rp.back());
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->charged + m->bytes > s->budget) return SPG_ENOMEM;
        s->charged += m->bytes;
    }
    spg_status st;
    {
        std::lock_guard<std::mutex> rl(s->reg_mu);
        st = spg_context_bind(s->reg_ctx);
        if (st == SPG_OK) st = spg_matrix_upload(s->reg_ctx, host, &m->m);
    }
    std::lock_guard<std::mutex> lk(s->mu);
    if (st != SPG_OK) {
        s->charged -= m->bytes;
        return st;
    }
    m->rp = std::move(h.rp);
    m->col = std::move(h.col);
    *handle = s->next_handle++;
    s->resident[*handle] = m;
    return SPG_OK;
}

spg_status spg_server_unregister(spg_server *s, int64_t handle) {
    if (!s) return SPG_EINVAL;
    std::shared_ptr<Resident> m;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->resident.find(handle);
        if (it == s->resident.end()) return SPG_EINVAL;
        if (it->second.use_count() > 1) return SPG_EBUSY;  // referenced by a queued or running request
        m = it->second;
        s->charged -= m->bytes;
        s->resident.erase(it);
    }
    spg_device_matrix_free(m->m);
    return SPG_OK;
}

spg_status spg_server_submit_resident(spg_server *s, int64_t a_handle, int64_t b_handle, int64_t *request_id) {
    if (!s || !request_id) return SPG_EINVAL;
    auto r = std::make_shared<Request>();
    std::lock_guard<std::mutex> lk(s->mu);
    auto ia = s->resident.find(a_handle), ib = s->resident.find(b_handle);
    if (ia == s->resident.end() || ib == s->resident.end()) return SPG_EINVAL;
    r->ra = ia->second;
    r->rb = ib->second;
    if (r->ra->cols != r->rb->rows) return SPG_EINVAL;
    const auto key = std::make_pair(a_handle, b_handle);
    auto ip = s->pair_products.find(key);
    uint64_t p = 0;
    if (ip == s->pair_products.end()) {
        if (!work_bound(r->ra->rp, r->ra->col, r->rb->rp, r->rb->rows, p)) return SPG_ERANGE;
        s->pair_products[key] = p;
    } else {
        p = ip->second;
    }
    // the arena copy of both operands, C's bound and the row workspace
    r->estimate = operand_bytes(r->ra->rows, r->ra->rp.back()) + operand_bytes(r->rb->rows, r->rb->rp.back()) +
                  12 * p + 32 * (uint64_t)r->ra->rows;
    if (r->estimate > s->budget) return SPG_ENOMEM;
    r->t_submit = now_s();
    r->id = s->next_id++;
    s->live[r->id] = r;
    s->queue.push_back(r);
    *request_id = r->id;
    s->cv.notify_all();
    return SPG_OK;
}

spg_status spg_server_wait(spg_server *s, int64_t id, int64_t *nnz, double *latency_s) {
    if (!s) return SPG_EINVAL;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->live.find(id);
    if (it == s->live.end()) return SPG_EINVAL;
    auto r = it->second;
    s->cv.wait(lk, [&] { return r->done; });
    if (nnz) *nnz = r->nnz;
    if (latency_s) *latency_s = r->t_done - r->t_submit;
    return r->status;
}

spg_status spg_server_result(spg_server *s, int64_t id, int64_t *row_ptr, int32_t *col_idx, double *values) {
    if (!s) return SPG_EINVAL;
    std::shared_ptr<Request> r;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->live.find(id);
        if (it == s->live.end() || !it->second->done) return SPG_ESTATE;
        r = it->second;
    }
    if (r->status != SPG_OK) return r->status;
    return spg_batch_copy(r->batch.get(), r->index, row_ptr, col_idx, values);
}

spg_status spg_server_release(spg_server *s, int64_t id) {
    if (!s) return SPG_EINVAL;
    std::shared_ptr<Request> r;  // the batch is freed with its last request, outside the lock
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->live.find(id);
        if (it == s->live.end()) return SPG_EINVAL;
        if (!it->second->done) return SPG_EBUSY;
        r = it->second;
        s->charged -= r->estimate;
        s->live.erase(it);
        s->cv.notify_all();
    }
    return SPG_OK;
}

spg_status spg_server_stats(spg_server *s, int64_t *batches, int64_t *completed, double *busy_s,
                            uint64_t *charged_bytes) {
    if (!s) return SPG_EINVAL;
    std::lock_guard<std::mutex> lk(s->mu);
    if (batches) *batches = s->batches;
    if (completed) *completed = s->completed;
    if (busy_s) *busy_s = s->busy_s;
    if (charged_bytes) *charged_bytes = s->charged;
    return SPG_OK;
}

spg_status spg_server_destroy(spg_server *s) {
    if (!s) return SPG_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        s->cv.notify_all();
    }
    if (s->builder.joinable()) s->builder.join();
    if (s->runner.joinable()) s->runner.join();
    s->built.clear();
    s->live.clear();
    s->queue.clear();
    for (spg_batch *b : s->all_batches) spg_batch_destroy(b);
    s->all_batches.clear();
    s->free_batches.clear();
    for (auto &kv : s->resident) spg_device_matrix_free(kv.second->m);
    s->resident.clear();
    spg_destroy(s->reg_ctx);
    spg_destroy(s->run_ctx);
    spg_destroy(s->build_ctx);
    delete s;
    return SPG_OK;
}

}  // extern "C"
