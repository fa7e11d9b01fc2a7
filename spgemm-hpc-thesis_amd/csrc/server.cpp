// server.cpp -- served products: continuous batching over resident batches.
//
// Requests (A, B) arrive at any time from any host thread (spg_server_submit).
// A builder thread admits queued requests in arrival order while their memory
// fits the HBM budget, packs them into one batch arena, uploads it and runs
// its count phase (analyze and symbolic kernels) on its own stream; a runner
// thread computes the values of the built batches on a second stream, so the
// upload and count of the next batch overlap the values of the current one,
// and requests that arrive while a batch runs join the next one.  Results stay
// in device memory until the client copies them out (spg_server_result) and
// releases them (spg_server_release).
//
// Operands either come with the request (host views, copied) or are registered
// once (spg_server_register: resident in device memory, charged to the budget
// until unregistered) and referenced by handle; batches of resident requests
// are packed on the device.
//
// Memory.  The budget caps the device bytes the server holds: registered
// matrices, and every batch object's arena, workspace and result buffers as
// measured from its allocations (spg_batch_device_bytes), including batches
// kept idle in the pool for reuse (their buffers only grow).  Bytes not yet
// allocated are reserved, before they are, by what the allocations will add:
//   - at admission, spg_batch_estimate: the arena and count workspace of the
//     admitted products and C at its capacity bound, less what the reused
//     pooled batch already holds, plus 64 KiB for small buffers.  The
//     capacity bound is what the count phase can size C to: a row of at most
//     256 scalar products takes the single pass, which sizes it by its
//     products; a longer row is counted exactly, so at most min(products, C's
//     columns).  It is a host walk over A's rows on the builder thread,
//     cached per pair of registered matrices, never on the submitting thread;
//   - after the count phase, the batch is measured and the value phase's
//     growth is reserved as spg_batch_count gives it: C at its counted
//     capacity (exact for the rows the symbolic kernels count, the work bound
//     for single-pass rows) and the large rows' accumulator windows;
//   - after the values, the batch is measured again.
// Requests are admitted in arrival order while the estimate fits and the
// batch, past its first request, needs at most 2/5 of the bytes that the
// registrations leave (two batches fit side by side, so one is built and
// counted while the other runs), reusing the pooled batch with the most
// bytes; other idle batches are trimmed when their bytes are needed.  When
// the budget cuts a batch short while the device still has a batch to run,
// admission waits for a batch to return to the pool instead of starting a
// small one that would trim it.  When the queue's first request does not fit by its
// capacity bound and nothing held can be released (no batch built or running, no
// finished request awaiting its client), it is admitted alone by its arena
// and workspace, and its count decides: C is reserved by the counted size,
// waiting for releases if needed.  A batch whose counted C cannot fit when
// nothing else can be released is split in half (the second half goes back
// to the front of the queue); a single request that cannot fit fails with
// SPG_ENOMEM.  A request whose arena and workspace alone exceed the budget is
// rejected with SPG_ENOMEM at submission; one whose A has a column outside
// B's rows with SPG_ERANGE.
//
// Every result is bit-identical to computing the same product alone: a batch
// computes each product's rows independently of its neighbours.
#include "spgemm_amd.h"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
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
// copies of its row offsets and columns for the capacity bound, and its column
// range for the submission check.
struct Resident {
    spg_device_matrix *m = nullptr;
    int32_t rows = 0, cols = 0;
    int64_t nnz = 0;
    int32_t min_col = 0, max_col = -1;  // over its entries (meaningless when it has none)
    std::vector<int64_t> rp;
    std::vector<int32_t> col;
    uint64_t bytes = 0;
};

struct Request {
    int64_t id = 0;
    HostCsr a, b;                       // host operands (copied at submission)
    std::shared_ptr<Resident> ra, rb;   // or registered operands
    int64_t ha = 0, hb = 0;             // their handles
    int64_t rows = 0, a_nnz = 0, b_rows = 0, b_nnz = 0;  // operand sizes
    int32_t c_cols = 0;
    uint64_t base = 0;                  // arena and workspace of a batch of its own
    uint64_t cap_bound = 0;             // C's capacity bound, entries (builder thread)
    bool bounded = false;
    double t_submit = 0, t_admit = 0, t_done = 0;
    bool done = false;
    spg_status status = SPG_OK;
    std::string error;
    std::shared_ptr<spg_batch> batch;  // holds the result until released
    int32_t index = 0;                 // product index inside the batch
    int64_t nnz = 0;
};

// Device bytes of a batch object as last measured, and the value-phase bytes
// reserved for it until its values are computed.
struct Slot {
    uint64_t footprint = 0;
    uint64_t value_reserve = 0;
};

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

struct spg_server {
    spg_context *build_ctx = nullptr;  // arenas are built, uploaded and counted on its stream
    spg_context *run_ctx = nullptr;    // values run on its stream
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
    std::map<std::pair<int64_t, int64_t>, uint64_t> pair_bounds;  // builder thread only
    int64_t next_handle = 1;
    // memory (see the top of this file): used + reserved is what stats report
    uint64_t used = 0;      // registrations + batches not in the pool (measured)
    uint64_t idle = 0;      // batches in the pool (measured)
    uint64_t reserved = 0;  // admitted, not yet allocated (estimates)
    uint64_t registered = 0;  // registered matrices (part of used)
    std::map<spg_batch *, Slot> slots;
    int64_t done_unreleased = 0;  // finished requests awaiting spg_server_release
    int64_t next_id = 1;
    bool stop = false;
    bool running = false;  // the runner holds a batch
    std::thread builder, runner;
    std::vector<spg_batch *> free_batches;  // built batches whose results were all released
    std::vector<spg_batch *> all_batches;
    // statistics
    int64_t batches = 0, completed = 0;
    double busy_s = 0;
    bool trace = false;  // SPG_SERVER_TRACE=1: admission, build and run events on stderr
    double t0 = 0;
};

namespace {

void trace(const spg_server *s, const char *fmt, ...) {
    if (!s->trace) return;
    char line[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(line, sizeof line, fmt, ap);
    va_end(ap);
    std::fprintf(stderr, "[spg_server %9.3f ms] %s\n", (now_s() - s->t0) * 1e3, line);
}

double gib(uint64_t b) { return (double)b / (double)(uint64_t(1) << 30); }

// Whether every column index lies in [0, brows).
bool columns_within(const int32_t *col, int64_t n, int32_t brows) {
    if (n == 0) return true;
    const auto mm = std::minmax_element(col, col + n);
    return *mm.first >= 0 && *mm.second < brows;
}

uint64_t operand_bytes(int64_t rows, int64_t nnz) { return 12 * (uint64_t)nnz + 8 * (uint64_t)(rows + 1); }

constexpr uint64_t kSmallBuffers = uint64_t(64) << 10;  // tile state, descriptors, rounding

// Total sizes of a batch's products.
struct Shape {
    int64_t rows = 0, a_nnz = 0, b_rows = 0, b_nnz = 0, c_entries = 0;
    int32_t cols = 0;
    Shape plus(const Request &r) const {
        Shape t = *this;
        t.rows += r.rows;
        t.a_nnz += r.a_nnz;
        t.b_rows += r.b_rows;
        t.b_nnz += r.b_nnz;
        t.c_entries += (int64_t)r.cap_bound;
        t.cols = std::max(cols, r.c_cols);
        return t;
    }
};

// Device bytes building `sh` into `reuse` (a new batch when null) adds, with
// C at c_entries.
uint64_t estimate(spg_server *s, const spg_batch *reuse, const Shape &sh, int64_t c_entries) {
    uint64_t n = 0;
    if (spg_batch_estimate(s->build_ctx, reuse, sh.rows, sh.a_nnz, sh.b_rows, sh.b_nnz, sh.cols, c_entries, &n) !=
        SPG_OK)
        return UINT64_MAX / 4;
    return n + kSmallBuffers;
}

// Rows of at most this many scalar products take the single pass, whose C
// capacity for the row is its product count.
constexpr int64_t kSinglePassProducts = 256;

// C's capacity bound of A.B in entries (columns already checked against B's
// rows): per row, its scalar products when they take the single pass, else
// min(products, C's columns).
uint64_t capacity_bound(const std::vector<int64_t> &arp, const std::vector<int32_t> &acol,
                        const std::vector<int64_t> &brp, int32_t c_cols) {
    uint64_t cap = 0;
    for (size_t i = 0; i + 1 < arp.size(); ++i) {
        int64_t p = 0;
        for (int64_t e = arp[i]; e < arp[i + 1]; ++e) {
            const int32_t k = acol[(size_t)e];
            p += brp[(size_t)k + 1] - brp[(size_t)k];
        }
        cap += (uint64_t)(p <= kSinglePassProducts ? p : std::min<int64_t>(p, c_cols));
    }
    return cap;
}

// C's capacity bound of a request (builder thread, without the server lock).
void bound_request(spg_server *s, Request &r) {
    if (r.bounded) return;
    if (r.ra) {
        const auto key = std::make_pair(r.ha, r.hb);
        auto it = s->pair_bounds.find(key);
        if (it == s->pair_bounds.end())
            it = s->pair_bounds.emplace(key, capacity_bound(r.ra->rp, r.ra->col, r.rb->rp, r.rb->cols)).first;
        r.cap_bound = it->second;
    } else {
        r.cap_bound = capacity_bound(r.a.rp, r.a.col, r.b.rp, r.b.cols);
    }
    r.bounded = true;
}

uint64_t free_bytes(const spg_server *s) {
    const uint64_t held = s->used + s->idle + s->reserved;
    return held >= s->budget ? 0 : s->budget - held;
}

// Whether bytes may still be released without the builder: a built batch
// waits for the runner or is running, or a finished request awaits its
// client's release.
bool releases_pending(const spg_server *s) {
    return !s->built.empty() || s->running || s->done_unreleased > 0;
}

uint64_t measure(spg_batch *b) {
    uint64_t n = 0;
    spg_batch_device_bytes(b, &n);
    return n;
}

// A batch (not in the pool, lock held) goes to the pool.
void to_pool(spg_server *s, spg_batch *p) {
    Slot &sl = s->slots[p];
    s->used -= sl.footprint;
    s->idle += sl.footprint;
    s->free_batches.push_back(p);
}

// Free the buffers of every pooled batch but `keep` (lock held on entry and
// exit, released meanwhile); whether idle bytes went down.
bool trim_idle(spg_server *s, std::unique_lock<std::mutex> &lk, spg_batch *keep = nullptr) {
    std::vector<spg_batch *> pool;
    for (spg_batch *p : s->free_batches)
        if (p != keep && s->slots[p].footprint > 0) pool.push_back(p);
    if (pool.empty()) return false;
    // out of reach of the builder's pick while trimmed; still counted idle
    s->free_batches.erase(std::remove_if(s->free_batches.begin(), s->free_batches.end(),
                                         [&](spg_batch *p) { return std::find(pool.begin(), pool.end(), p) != pool.end(); }),
                          s->free_batches.end());
    lk.unlock();
    std::vector<uint64_t> after(pool.size());
    for (size_t i = 0; i < pool.size(); ++i) {
        spg_batch_set_stream(pool[i], nullptr);  // no work pending: its last run was synchronised
        spg_batch_trim(pool[i]);
        after[i] = measure(pool[i]);
    }
    lk.lock();
    const uint64_t before = s->idle;
    trace(s, "trimmed %zu idle batches", pool.size());
    for (size_t i = 0; i < pool.size(); ++i) {
        Slot &sl = s->slots[pool[i]];
        s->idle = s->idle - sl.footprint + after[i];
        sl.footprint = after[i];
        s->free_batches.push_back(pool[i]);
    }
    s->cv.notify_all();
    return s->idle < before;
}

void finish_failed(spg_server *s, std::vector<std::shared_ptr<Request>> &take, spg_status st,
                   const std::string &msg) {
    for (auto &r : take) {
        r->status = st;
        r->error = msg;
        r->done = true;
        r->t_done = now_s();
        s->done_unreleased += 1;
    }
    take.clear();
    s->cv.notify_all();
}

// The pooled batch with the most bytes (null when the pool is empty).
spg_batch *largest_idle(spg_server *s) {
    spg_batch *best = nullptr;
    for (spg_batch *p : s->free_batches)
        if (!best || s->slots[p].footprint > s->slots[best].footprint) best = p;
    return best;
}

// Admit the next batch (lock held; released while idle batches are trimmed).
// Empty when the builder must wait for a release, or when the first request
// failed.  `reserve` receives the bytes reserved for the batch and `reuse`
// the pooled batch it is built into (null: a new batch), taken out of the pool.
std::vector<std::shared_ptr<Request>> admit(spg_server *s, std::unique_lock<std::mutex> &lk, uint64_t &reserve,
                                            spg_batch *&reuse) {
    std::vector<std::shared_ptr<Request>> take;
    reserve = 0;
    reuse = nullptr;
    for (;;) {
        if (s->queue.empty()) return take;
        spg_batch *cand = largest_idle(s);
        const bool resident = (bool)s->queue.front()->ra;
        const uint64_t held = s->used + s->idle + s->reserved;
        // Past its first request, a batch needs at most 2/5 of the bytes the
        // registrations leave: two batches then fit beside each other even
        // after their buffers regrow (by 1/8), so the next batch is built and
        // counted while one runs.  A batch as large as the whole budget
        // serialises them, and the few requests that fit beside it form
        // batches of their own.
        const uint64_t share = (s->budget > s->registered ? s->budget - s->registered : 0) / 5 * 2;
        Shape sh;
        uint64_t est = 0;
        bool budget_cut = false;  // the budget, not the queue or the share, ended the batch
        while (!s->queue.empty() && (int)take.size() < s->max_batch) {
            const auto &r = s->queue.front();
            if (!r->bounded || (bool)r->ra != resident) break;
            const Shape s2 = sh.plus(*r);
            const uint64_t e2 = estimate(s, cand, s2, s2.c_entries);
            if (held + e2 > s->budget) {
                budget_cut = true;
                break;
            }
            if (!take.empty() && estimate(s, nullptr, s2, s2.c_entries) > share) break;
            sh = s2;
            est = e2;
            take.push_back(r);
            s->queue.pop_front();
        }
        if (!take.empty() && budget_cut && (s->running || !s->built.empty())) {
            // Only part of the next requests fits beside the batches in
            // flight.  While the device still has a batch to run, wait for a
            // batch to come back to the pool and be reused: a batch started in
            // the bytes left would take a pooled batch's bytes for its values
            // (trimming it), and those buffers would then be allocated again
            // (fresh allocations of tens of GB stall the device for seconds).
            // With the device idle, the part that fits is admitted.
            for (auto it = take.rbegin(); it != take.rend(); ++it) s->queue.push_front(*it);
            take.clear();
            return take;
        }
        if (take.empty()) {
            // other idle batches' bytes first; then wait for releases
            if (trim_idle(s, lk, cand)) continue;
            if (releases_pending(s)) return take;
            // Nothing can be released: the first request is admitted alone
            // by its arena and workspace, and its count decides.
            const auto r = s->queue.front();
            const uint64_t e1 = estimate(s, cand, Shape{}.plus(*r), 0);
            if (s->used + s->idle + s->reserved + e1 > s->budget) {
                if (trim_idle(s, lk)) continue;  // the reuse candidate too
                s->queue.pop_front();
                std::vector<std::shared_ptr<Request>> one{r};
                finish_failed(s, one, SPG_ENOMEM,
                              "the product's operands and count workspace exceed the server budget");
                return take;
            }
            s->queue.pop_front();
            take.push_back(r);
            est = e1;
        }
        reserve = est;
        reuse = cand;
        break;
    }
    if (reuse) {
        s->free_batches.erase(std::find(s->free_batches.begin(), s->free_batches.end(), reuse));
        Slot &sl = s->slots[reuse];
        s->idle -= sl.footprint;
        s->used += sl.footprint;
    }
    const double t = now_s();
    for (auto &r : take) r->t_admit = t;
    s->reserved += reserve;
    trace(s, "admit %zu requests (%zu queued), reserve %.3f GiB, reuse %.3f GiB; used %.3f idle %.3f reserved %.3f",
          take.size(), s->queue.size(), gib(reserve), reuse ? gib(s->slots[reuse].footprint) : 0.0, gib(s->used),
          gib(s->idle), gib(s->reserved));
    return take;
}

void builder_loop(spg_server *s) {
    spg_context_bind(s->build_ctx);  // this thread's current device
    for (;;) {
        // Work bounds of the requests that may be admitted next, off the
        // submitting threads and outside the lock.
        std::vector<std::shared_ptr<Request>> pending;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv.wait(lk, [&] { return s->stop || (!s->queue.empty() && s->built.size() < 2); });
            if (s->stop) return;
            for (size_t i = 0; i < s->queue.size() && (int)i < s->max_batch; ++i)
                if (!s->queue[i]->bounded) pending.push_back(s->queue[i]);
        }
        for (auto &r : pending) bound_request(s, *r);
        std::vector<std::shared_ptr<Request>> take;
        uint64_t reserve = 0;
        spg_batch *raw = nullptr;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            if (s->stop) return;
            if (s->queue.empty() || s->built.size() >= 2) continue;
            take = admit(s, lk, reserve, raw);
            if (take.empty()) {
                // wait for a release before trying again
                if (!s->stop && releases_pending(s)) s->cv.wait(lk);
                continue;
            }
        }
        // Build and count; split the batch while its C cannot fit.
        while (!take.empty()) {
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
            spg_status st;
            bool created = false;
            const double tb = now_s();
            if (raw) {
                // a pooled batch last ran on the runner's stream and has no
                // work pending; pack and count on the builder's stream
                spg_batch_set_stream(raw, nullptr);
                st = spg_batch_rebuild(raw, (int32_t)take.size(), av.data(), bv.data());
            } else {
                st = spg_batch_create(s->build_ctx, (int32_t)take.size(), av.data(), bv.data(), &raw);
                created = raw != nullptr;
            }
            std::string msg = st == SPG_OK ? "" : spg_last_error(s->build_ctx);
            uint64_t grow = 0;
            const double tc = now_s();
            if (st == SPG_OK) {
                st = spg_batch_count(raw, &grow);
                if (st != SPG_OK) msg = spg_batch_last_error(raw);
            }
            const uint64_t fp = raw ? measure(raw) : 0;
            trace(s, "built %zu requests in %.3f ms, counted in %.3f ms: %.3f GiB, value growth %.3f GiB", take.size(),
                  (tc - tb) * 1e3, (now_s() - tc) * 1e3, gib(fp), gib(grow));
            std::unique_lock<std::mutex> lk(s->mu);
            if (created) {
                s->all_batches.push_back(raw);
                s->slots[raw] = Slot{};
            }
            s->reserved -= reserve;  // the arena and workspace are allocated and measured now
            reserve = 0;
            if (raw) {
                Slot &sl = s->slots[raw];
                s->used = s->used - sl.footprint + fp;
                sl.footprint = fp;
            }
            if (st != SPG_OK) {
                if (raw) to_pool(s, raw);
                finish_failed(s, take, st, msg);
                break;
            }
            bool fits = false;
            const double tw = now_s();
            for (;;) {
                if (s->stop) return;
                if (grow <= free_bytes(s)) {
                    fits = true;
                    break;
                }
                if (s->idle > 0 && trim_idle(s, lk)) continue;
                if (!releases_pending(s)) break;
                s->cv.wait(lk);
            }
            if (fits) {
                trace(s, "values fit after %.3f ms of waiting", (now_s() - tw) * 1e3);
                s->reserved += grow;
                s->slots[raw].value_reserve = grow;
                // back to the pool when its last request is released
                std::shared_ptr<spg_batch> b(raw, [s](spg_batch *p) {
                    std::lock_guard<std::mutex> g(s->mu);
                    to_pool(s, p);
                    s->cv.notify_all();
                });
                for (size_t i = 0; i < take.size(); ++i) {
                    take[i]->batch = b;
                    take[i]->index = (int32_t)i;
                }
                s->built.emplace_back(b, take);
                take.clear();  // only the built queue and the live map hold them now
                s->cv.notify_all();
                break;
            }
            // C cannot fit: free this batch's buffers before deciding
            trace(s, "values do not fit (%zu requests): split or fail", take.size());
            lk.unlock();
            spg_batch_trim(raw);
            const uint64_t left = measure(raw);
            lk.lock();
            {
                Slot &sl = s->slots[raw];
                s->used = s->used - sl.footprint + left;
                sl.footprint = left;
            }
            if (take.size() == 1) {
                to_pool(s, raw);
                finish_failed(s, take, SPG_ENOMEM,
                              "the product's counted result (" + std::to_string(grow) +
                                  " B) does not fit the server budget beside its operands");
                break;
            }
            // the second half goes back to the front of the queue, in order
            const size_t keep = take.size() / 2;
            for (size_t i = take.size(); i-- > keep;) {
                take[i]->t_admit = 0;
                s->queue.push_front(take[i]);
            }
            take.resize(keep);
            Shape sh;
            for (auto &r : take) sh = sh.plus(*r);
            reserve = estimate(s, raw, sh, 0);
            s->reserved += reserve;
            s->cv.notify_all();
        }
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
            s->running = true;
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
        const uint64_t fp = measure(b.get());
        trace(s, "ran %zu requests in %.3f ms: %.3f GiB", reqs.size(), (t1 - t0) * 1e3, gib(fp));
        std::lock_guard<std::mutex> lk(s->mu);
        {
            Slot &sl = s->slots[b.get()];
            s->used = s->used - sl.footprint + fp;  // C and the value buffers are allocated now
            sl.footprint = fp;
            s->reserved -= sl.value_reserve;
            sl.value_reserve = 0;
        }
        s->running = false;
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
            s->done_unreleased += 1;
        }
        // Drop this thread's references before a client can see the requests
        // done: the batch then returns to the pool exactly when its last
        // request is released.  Not the last references here: the live map
        // holds every request until release.
        reqs.clear();
        b.reset();
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

spg_status enqueue(spg_server *s, const std::shared_ptr<Request> &r, int64_t *request_id) {
    r->base = estimate(s, nullptr, Shape{}.plus(*r), 0);
    if (r->base > s->budget) return SPG_ENOMEM;
    r->t_submit = now_s();
    r->id = s->next_id++;
    s->live[r->id] = r;
    s->queue.push_back(r);
    *request_id = r->id;
    s->cv.notify_all();
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
    const char *tr = std::getenv("SPG_SERVER_TRACE");
    s->trace = tr && tr[0] == '1';
    s->t0 = now_s();
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
    if (!columns_within(r->a.col.data(), (int64_t)r->a.col.size(), r->b.rows)) return SPG_ERANGE;  // fail() in the reference
    r->rows = r->a.rows;
    r->a_nnz = r->a.rp.back();
    r->b_rows = r->b.rows;
    r->b_nnz = r->b.rp.back();
    r->c_cols = r->b.cols;
    std::lock_guard<std::mutex> lk(s->mu);
    return enqueue(s, r, request_id);
}

spg_status spg_server_register(spg_server *s, const spg_csr_view *host, int64_t *handle) {
    if (!s || !handle) return SPG_EINVAL;
    auto m = std::make_shared<Resident>();
    HostCsr h;
    if (copy_view(host, h) != SPG_OK) return SPG_EINVAL;
    m->rows = h.rows;
    m->cols = h.cols;
    m->nnz = h.rp.back();
    if (!h.col.empty()) {
        const auto mm = std::minmax_element(h.col.begin(), h.col.end());
        m->min_col = *mm.first;
        m->max_col = *mm.second;
    }
    m->bytes = operand_bytes(h.rows, h.rp.back());
    {
        std::unique_lock<std::mutex> lk(s->mu);
        if (m->bytes > free_bytes(s) && s->idle > 0) trim_idle(s, lk);
        if (m->bytes > free_bytes(s)) return SPG_ENOMEM;
        s->used += m->bytes;
        s->registered += m->bytes;
    }
    spg_status st;
    {
        std::lock_guard<std::mutex> rl(s->reg_mu);
        st = spg_context_bind(s->reg_ctx);
        if (st == SPG_OK) st = spg_matrix_upload(s->reg_ctx, host, &m->m);
    }
    std::lock_guard<std::mutex> lk(s->mu);
    if (st != SPG_OK) {
        s->used -= m->bytes;
        s->registered -= m->bytes;
        s->cv.notify_all();
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
        s->used -= m->bytes;
        s->registered -= m->bytes;
        s->resident.erase(it);
        s->cv.notify_all();
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
    r->ha = a_handle;
    r->hb = b_handle;
    if (r->ra->cols != r->rb->rows) return SPG_EINVAL;
    if (r->ra->nnz > 0 && (r->ra->min_col < 0 || r->ra->max_col >= r->rb->rows)) return SPG_ERANGE;
    r->rows = r->ra->rows;
    r->a_nnz = r->ra->nnz;
    r->b_rows = r->rb->rows;
    r->b_nnz = r->rb->nnz;
    r->c_cols = r->rb->cols;
    return enqueue(s, r, request_id);
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
    std::shared_ptr<Request> r;  // the batch goes to the pool with its last request, outside the lock
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->live.find(id);
        if (it == s->live.end()) return SPG_EINVAL;
        if (!it->second->done) return SPG_EBUSY;
        r = it->second;
        s->live.erase(it);
        s->done_unreleased -= 1;
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
    if (charged_bytes) *charged_bytes = s->used + s->reserved;
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
