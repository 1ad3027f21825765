// Continuous-batching request server for pair comparisons.
//
// A worker thread owns a private engine context.  Submitted requests wait
// in a FIFO; whenever an in-flight slot is free, the worker admits the
// oldest requests whose device footprint (strings, strands, ring) fits in
// the free part of a device arena bounded by the HBM budget, uploads them,
// and launches one persistent chain per tile height on the slot's stream.
// Up to `slots` micro-batches are on the device at once, so requests that
// arrive while others run start as soon as memory and a slot free up, and
// their kernels overlap with the running ones.  When a slot's kernels are
// done the worker copies the strands back, frees the arena space and marks
// the requests done.  A request's result never depends on what it was
// batched with: every pair has its own ring and the comb is deterministic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "lcs_hip.h"
#include "lcs_internal.h"

using namespace lcs;

namespace {

using Clock = std::chrono::steady_clock;

// launch classes: tile height x variant (0..5 register chain, 6..11 LDS groups)
constexpr int kNumTiles = 6;
constexpr int kClasses = 2 * kNumTiles;
constexpr int kTiles[kNumTiles] = {1, 2, 4, 8, 16, 32};
constexpr size_t kAlign = 256;

int class_index(int tm, int group_w)
{
    for (int i = 0; i < kNumTiles; ++i)
        if (kTiles[i] == tm) return group_w == 1 ? i : group_w == 4 ? kNumTiles + i : -1;
    return -1;
}

// First-fit allocator over one device arena (offsets only).
struct Arena {
    size_t size = 0;
    std::map<size_t, size_t> free_;  // offset -> length
    void init(size_t s)
    {
        size = s;
        free_.clear();
        if (s) free_[0] = s;
    }
    bool alloc(size_t len, size_t &off)
    {
        len = (len + kAlign - 1) & ~(kAlign - 1);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= len) {
                off = it->first;
                const size_t rest = it->second - len;
                free_.erase(it);
                if (rest) free_[off + len] = rest;
                return true;
            }
        }
        return false;
    }
    void release(size_t off, size_t len)
    {
        len = (len + kAlign - 1) & ~(kAlign - 1);
        auto it = free_.emplace(off, len).first;
        auto next = std::next(it);
        if (next != free_.end() && it->first + it->second == next->first) {
            it->second += next->second;
            free_.erase(next);
        }
        if (it != free_.begin()) {
            auto prev = std::prev(it);
            if (prev->first + prev->second == it->first) {
                prev->second += it->second;
                free_.erase(it);
            }
        }
    }
    size_t free_bytes() const
    {
        size_t s = 0;
        for (auto &kv : free_) s += kv.second;
        return s;
    }
};

struct Request {
    int64_t id = 0;
    std::vector<int32_t> a, b, h, v;
    lcs_schedule sched{};
    float feat[4] = {0, 0, 0, 0};
    bool has_feat = false;
    int status = LCS_EAGAIN;   // EAGAIN while pending/running
    bool cancelled = false;
    size_t off = 0, bytes = 0;
    Clock::time_point t_submit, t_start, t_done;
};

// Bytes of a request's arena block: a, b, h, v, the ring and, for a pair
// whose register chain takes byte symbols (use_sym8), the byte copies of b
// and their range flag, each part kAlign-aligned.
size_t sym8_bytes(int32_t n) { return (size_t)16 * (size_t)lcsk_sym8_stride(n) + kAlign; }
size_t footprint(int32_t m, int32_t n, int variant)
{
    return (size_t)4 * (2 * (size_t)m + 2 * (size_t)n) + 8 * lcsk_ring_words(n) + 4 * kAlign +
           (use_sym8(variant, m, n) ? sym8_bytes(n) + kAlign : 0);
}

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    DevBuf descs;                  // every class's descriptors, one upload
    Staging staging;               // this slot's own pinned upload buffer
    std::vector<std::shared_ptr<Request>> reqs;
    bool busy = false;
};

// A cost model handed over by lcs_server_set_cost_model: the parameters are
// copied on the caller's thread and the worker builds its own model from
// them on the server's private context, so the caller's context and model
// are never touched by the worker.
struct ModelUpdate {
    bool pending = false;
    int n_in = 0, hidden = 0, n_out = 0;
    std::vector<float> params;     // empty: detach the model
    std::vector<lcs_schedule> candidates;
};

}  // namespace

struct lcs_server {
    lcs_ctx *ctx = nullptr;         // private context of the worker
    int device = 0;
    size_t budget = 0;
    void *arena_base = nullptr;
    Arena arena;
    std::vector<Slot> slots;
    unsigned *ctrl = nullptr;       // 2 words per (slot, class)
    // the worker's own copy of the cost model (on ctx) and its candidates;
    // read and written by the worker thread only
    lcs_mlp *model = nullptr;
    std::vector<lcs_schedule> candidates;
    ModelUpdate update;             // guarded by mu

    std::mutex mu;
    std::condition_variable cv;
    std::deque<std::shared_ptr<Request>> pending;
    std::map<int64_t, std::shared_ptr<Request>> all;
    int64_t next_id = 1;
    bool stop = false;
    std::thread worker;
    int max_inflight_seen = 0;
    long long completed = 0, batches = 0;
    int worker_error = LCS_OK;
};

namespace {

// Launch the admitted requests of one slot (worker thread).
int launch_slot(lcs_server *S, Slot &slot)
{
    lcs_ctx *ctx = S->ctx;
    // learned schedules for requests without one
    if (S->model && !S->candidates.empty()) {
        std::vector<float> X;
        std::vector<size_t> idx;
        for (size_t i = 0; i < slot.reqs.size(); ++i)
            if (!slot.reqs[i]->sched.tile_m && slot.reqs[i]->has_feat) {
                X.insert(X.end(), slot.reqs[i]->feat, slot.reqs[i]->feat + 4);
                idx.push_back(i);
            }
        if (!idx.empty()) {
            std::vector<int32_t> choice(idx.size());
            if (lcs_mlp_choose(S->model, X.data(), (int32_t)idx.size(), choice.data()) == LCS_OK)
                for (size_t j = 0; j < idx.size(); ++j)
                    if (choice[j] >= 0 && (size_t)choice[j] < S->candidates.size())
                        slot.reqs[idx[j]]->sched = S->candidates[(size_t)choice[j]];
        }
    }
    std::vector<std::vector<PairDesc>> groups(kClasses);
    char *base = (char *)S->arena_base;
    for (auto &R : slot.reqs) {
        const int32_t m = (int32_t)R->a.size(), n = (int32_t)R->b.size();
        int32_t *da = (int32_t *)(base + R->off);
        int32_t *db = da + m, *dh = db + n, *dv = dh + m;
        unsigned long long *E = (unsigned long long *)(((uintptr_t)(dv + n) + 255) & ~(uintptr_t)255);
        if (m && hipMemcpyAsync(da, R->a.data(), (size_t)m * 4, hipMemcpyHostToDevice, slot.stream) != hipSuccess)
            return LCS_EIO;
        if (n && hipMemcpyAsync(db, R->b.data(), (size_t)n * 4, hipMemcpyHostToDevice, slot.stream) != hipSuccess)
            return LCS_EIO;
        const int tm = R->sched.tile_m ? R->sched.tile_m : default_batch_tile_m(m);
        const int c = class_index(tm, group_width(R->sched.variant));
        if (c < 0) return LCS_EINVAL;
        PairDesc P{};
        P.a = da;
        P.b = db;
        P.h_out = dh;
        P.v_out = dv;
        P.m = m;
        P.n = n;
        P.E = E;  // prepare_pairs assigns rings from `ring`; each pair's own lies right here
        if (c < kNumTiles && use_sym8(R->sched.variant, m, n) && R->bytes >= footprint(m, n, R->sched.variant)) {
            // byte symbols (register chain): the copies follow the ring in the block
            const int32_t stride = lcsk_sym8_stride(n);
            uint32_t *copies = (uint32_t *)(((uintptr_t)(E + lcsk_ring_words(n)) + kAlign - 1) & ~(uintptr_t)(kAlign - 1));
            int32_t *bad = (int32_t *)(copies + (size_t)4 * stride);
            if (lcsk_launch_sym8(db, n, copies, stride, bad, slot.stream)) return LCS_EIO;
            P.sym8 = copies;
            P.sym8_bad = bad;
            P.sym8_col0 = 0;
            P.sym8_stride = stride;
        }
        groups[(size_t)c].push_back(P);
    }
    // every class's descriptors go up in one copy before any kernel is
    // queued, so no upload waits behind a chain kernel of this slot
    const int slot_index = (int)(&slot - S->slots.data());
    std::vector<unsigned char> blob;
    PreparedLaunch launches[kClasses];
    for (int c = 0; c < kClasses; ++c) {
        auto &g = groups[(size_t)c];
        if (g.empty()) continue;
        // each pair's ring lies inside its request's arena block (ring = NULL)
        const int rc = build_launch(ctx, g, kTiles[c % kNumTiles], 1, nullptr,
                                    S->ctrl + 2 * (slot_index * kClasses + c), c < kNumTiles ? 1 : 4, blob,
                                    launches[c]);
        if (rc) return rc;
    }
    if (blob.empty()) return hipEventRecord(slot.done, slot.stream) == hipSuccess ? LCS_OK : LCS_EIO;
    int rc = slot.descs.reserve(ctx, blob.size() + 256);
    if (rc) return rc;
    if ((rc = upload_via(slot.staging, slot.descs.ptr, blob.data(), blob.size(), slot.stream))) return rc;
    for (int c = 0; c < kClasses; ++c) {
        if (groups[(size_t)c].empty()) continue;
        rebase_launch(launches[c], slot.descs.ptr);
        launches[c].args.poll_sleep = 1;   // slots and classes overlap on the SIMDs
        if ((rc = launch_prepared(launches[c], slot.stream))) return rc;
    }
    return hipEventRecord(slot.done, slot.stream) == hipSuccess ? LCS_OK : LCS_EIO;
}

// Installs a cost model handed over by lcs_server_set_cost_model (worker
// thread, the only user of the server's context).
int apply_model_update(lcs_server *S, ModelUpdate &u)
{
    if (S->model) {
        lcs_mlp_destroy(S->model);
        S->model = nullptr;
    }
    S->candidates.clear();
    if (u.params.empty()) return LCS_OK;
    lcs_mlp *M = nullptr;
    int rc = lcs_mlp_create(S->ctx, u.n_in, u.hidden, u.n_out, 0, &M);
    if (rc) return rc;
    if ((rc = lcs_mlp_set_params(M, u.params.data()))) {
        lcs_mlp_destroy(M);
        return rc;
    }
    S->model = M;
    S->candidates = std::move(u.candidates);
    return LCS_OK;
}

void finish_slot(lcs_server *S, Slot &slot)
{
    char *base = (char *)S->arena_base;
    unsigned flags[2 * kClasses] = {};
    const int slot_index = (int)(&slot - S->slots.data());
    const bool flags_ok = hipMemcpy(flags, S->ctrl + 2 * slot_index * kClasses, sizeof(flags),
                                    hipMemcpyDeviceToHost) == hipSuccess;
    bool aborted = !flags_ok;
    for (int c = 0; c < kClasses; ++c) aborted = aborted || flags[2 * c + 1];
    for (auto &R : slot.reqs) {
        const int32_t m = (int32_t)R->a.size(), n = (int32_t)R->b.size();
        int32_t *dh = (int32_t *)(base + R->off) + m + n, *dv = dh + m;
        int st = aborted ? LCS_ETIMEDOUT : LCS_OK;
        R->h.resize((size_t)m);
        R->v.resize((size_t)n);
        if (!st && m && hipMemcpy(R->h.data(), dh, (size_t)m * 4, hipMemcpyDeviceToHost) != hipSuccess) st = LCS_EIO;
        if (!st && n && hipMemcpy(R->v.data(), dv, (size_t)n * 4, hipMemcpyDeviceToHost) != hipSuccess) st = LCS_EIO;
        std::lock_guard<std::mutex> lk(S->mu);
        S->arena.release(R->off, R->bytes);
        R->t_done = Clock::now();
        R->status = R->cancelled ? LCS_ECANCELED : st;
        ++S->completed;
    }
    slot.reqs.clear();
    slot.busy = false;
    S->cv.notify_all();
}

void worker_main(lcs_server *S)
{
    (void)hipSetDevice(S->device);
    for (;;) {
        // retire finished slots
        int busy = 0;
        for (auto &slot : S->slots) {
            if (!slot.busy) continue;
            if (hipEventQuery(slot.done) == hipSuccess) finish_slot(S, slot);
            else ++busy;
        }
        // admit into a free slot
        Slot *free_slot = nullptr;
        for (auto &slot : S->slots)
            if (!slot.busy) { free_slot = &slot; break; }
        bool launched = false;
        if (free_slot) {
            std::unique_lock<std::mutex> lk(S->mu);
            if (S->stop && S->pending.empty() && busy == 0) break;
            // a model set before a request was submitted decides that
            // request's schedule: take the update with the admission
            ModelUpdate u;
            if (S->update.pending) {
                u = std::move(S->update);
                S->update = ModelUpdate{};
            }
            while (!S->pending.empty()) {
                auto R = S->pending.front();
                if (R->cancelled) {
                    S->pending.pop_front();
                    R->status = LCS_ECANCELED;
                    R->t_done = Clock::now();
                    continue;
                }
                size_t off;
                if (!S->arena.alloc(R->bytes, off)) break;  // wait for memory (FIFO order)
                R->off = off;
                R->t_start = Clock::now();
                free_slot->reqs.push_back(R);
                S->pending.pop_front();
            }
            lk.unlock();
            if (u.pending) {
                const int urc = apply_model_update(S, u);
                if (urc) {
                    std::lock_guard<std::mutex> g(S->mu);
                    S->worker_error = urc;
                }
            }
            if (!free_slot->reqs.empty()) {
                const int rc = launch_slot(S, *free_slot);
                free_slot->busy = true;
                launched = true;
                if (rc) {
                    // the launch failed: finish the slot with the error
                    (void)hipStreamSynchronize(free_slot->stream);
                    for (auto &R : free_slot->reqs) R->cancelled = false;
                    finish_slot(S, *free_slot);
                    std::lock_guard<std::mutex> g(S->mu);
                    S->worker_error = rc;
                } else {
                    std::lock_guard<std::mutex> g(S->mu);
                    ++S->batches;
                    int inflight = 0;
                    for (auto &sl : S->slots) inflight += sl.busy;
                    S->max_inflight_seen = std::max(S->max_inflight_seen, inflight);
                }
            }
        }
        if (!launched) {
            std::unique_lock<std::mutex> lk(S->mu);
            if (S->stop && S->pending.empty()) {
                bool any = false;
                for (auto &slot : S->slots) any = any || slot.busy;
                if (!any) break;
            }
            S->cv.wait_for(lk, std::chrono::microseconds(100));
        }
    }
}

}  // namespace

extern "C" {

int lcs_server_create(lcs_ctx *ctx, size_t hbm_budget, int32_t max_inflight, lcs_server **out)
{
    if (!ctx || !out || max_inflight < 1 || max_inflight > 8) return LCS_EINVAL;
    *out = nullptr;
    lcs_server *S = new (std::nothrow) lcs_server();
    if (!S) return LCS_ENOMEM;
    S->device = ctx->device;
    int rc = lcs_ctx_create(ctx->device, 0, &S->ctx);
    if (rc) { delete S; return rc; }
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) { lcs_ctx_destroy(S->ctx); delete S; return LCS_EIO; }
    S->budget = hbm_budget ? hbm_budget : free_b / 2;
    if (S->budget > free_b / 4 * 3) S->budget = free_b / 4 * 3;
    if (hipMalloc(&S->arena_base, S->budget) != hipSuccess) { lcs_ctx_destroy(S->ctx); delete S; return LCS_ENOMEM; }
    S->arena.init(S->budget);
    void *p = nullptr;
    if (hipMalloc(&p, sizeof(unsigned) * 2 * kClasses * (size_t)max_inflight) != hipSuccess) {
        (void)hipFree(S->arena_base);
        lcs_ctx_destroy(S->ctx);
        delete S;
        return LCS_ENOMEM;
    }
    S->ctrl = (unsigned *)p;
    (void)hipMemset(S->ctrl, 0, sizeof(unsigned) * 2 * kClasses * (size_t)max_inflight);
    S->slots.resize((size_t)max_inflight);
    for (auto &slot : S->slots)
        if (hipStreamCreateWithFlags(&slot.stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&slot.done, hipEventDisableTiming) != hipSuccess || slot.staging.init() != LCS_OK) {
            lcs_server_destroy(S);
            return LCS_EIO;
        }
    S->worker = std::thread(worker_main, S);
    *out = S;
    return LCS_OK;
}

int lcs_server_destroy(lcs_server *S)
{
    if (!S) return LCS_EINVAL;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
        for (auto &R : S->pending) R->cancelled = true;
    }
    S->cv.notify_all();
    if (S->worker.joinable()) S->worker.join();
    (void)hipSetDevice(S->device);
    for (auto &slot : S->slots) {
        if (slot.stream) { (void)hipStreamSynchronize(slot.stream); (void)hipStreamDestroy(slot.stream); }
        if (slot.done) (void)hipEventDestroy(slot.done);
        slot.descs.release(S->ctx);
        slot.staging.destroy();
    }
    if (S->model) lcs_mlp_destroy(S->model);
    if (S->ctrl) (void)hipFree(S->ctrl);
    if (S->arena_base) (void)hipFree(S->arena_base);
    if (S->ctx) lcs_ctx_destroy(S->ctx);
    delete S;
    return LCS_OK;
}

int lcs_server_set_cost_model(lcs_server *S, lcs_mlp *model, const lcs_schedule *candidates, int32_t ncand)
{
    if (!S || ncand < 0 || (ncand && !candidates)) return LCS_EINVAL;
    for (int32_t i = 0; i < ncand; ++i) {
        const int gw = group_width(candidates[i].variant);
        if (!gw || !candidates[i].tile_m || class_index(candidates[i].tile_m, gw) < 0) return LCS_EINVAL;
    }
    ModelUpdate u;
    u.pending = true;
    if (model && ncand > 0) {
        // copy the parameters here, on the caller's thread (the model's
        // context belongs to the caller); the worker rebuilds the model
        int32_t count = 0;
        int rc = lcs_mlp_param_count(model, &count);
        if (!rc) rc = mlp_shape(model, &u.n_in, &u.hidden, &u.n_out);
        if (rc) return rc;
        u.params.resize((size_t)count);
        if ((rc = lcs_mlp_get_params(model, u.params.data()))) return rc;
        u.candidates.assign(candidates, candidates + ncand);
    }
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->update = std::move(u);
    }
    S->cv.notify_all();
    return LCS_OK;
}

int lcs_server_submit(lcs_server *S, const int32_t *a, int32_t m, const int32_t *b, int32_t n,
                      const lcs_schedule *sched, const float *features, int64_t *request_id)
{
    if (!S || !request_id || m < 0 || n < 0 || (m && !a) || (n && !b)) return LCS_EINVAL;
    if ((long long)m + n > 0x7fffffffLL) return LCS_ERANGE;
    if (sched && (group_width(sched->variant) == 0 ||
                  (sched->tile_m && class_index(sched->tile_m, group_width(sched->variant)) < 0)))
        return LCS_EINVAL;
    auto R = std::make_shared<Request>();
    R->a.assign(a, a + m);
    R->b.assign(b, b + n);
    if (sched) R->sched = *sched;
    R->sched.strips = 1;
    if (features) {
        std::memcpy(R->feat, features, sizeof(R->feat));
        R->has_feat = true;
    }
    R->bytes = footprint(m, n, R->sched.variant);
    if (R->bytes > S->budget) return LCS_ENOMEM;  // could never be admitted
    R->t_submit = Clock::now();
    {
        std::lock_guard<std::mutex> lk(S->mu);
        if (S->stop) return LCS_ECANCELED;
        R->id = S->next_id++;
        S->pending.push_back(R);
        S->all[R->id] = R;
        *request_id = R->id;
    }
    S->cv.notify_all();
    return LCS_OK;
}

int lcs_server_poll(lcs_server *S, int64_t id, int32_t *h_out, int32_t *v_out, int32_t *done)
{
    if (!S || !done) return LCS_EINVAL;
    std::shared_ptr<Request> R;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        auto it = S->all.find(id);
        if (it == S->all.end()) return LCS_ENOENT;
        R = it->second;
        *done = R->status != LCS_EAGAIN;
        if (!*done) return LCS_OK;
    }
    if (R->status != LCS_OK) return R->status;
    if (h_out && !R->h.empty()) std::memcpy(h_out, R->h.data(), R->h.size() * 4);
    if (v_out && !R->v.empty()) std::memcpy(v_out, R->v.data(), R->v.size() * 4);
    return LCS_OK;
}

int lcs_server_wait(lcs_server *S, int64_t id, int32_t *h_out, int32_t *v_out, int32_t timeout_ms)
{
    if (!S) return LCS_EINVAL;
    const auto deadline = Clock::now() + std::chrono::milliseconds(timeout_ms < 0 ? 0 : timeout_ms);
    {
        std::unique_lock<std::mutex> lk(S->mu);
        auto it = S->all.find(id);
        if (it == S->all.end()) return LCS_ENOENT;
        auto R = it->second;
        while (R->status == LCS_EAGAIN) {
            if (timeout_ms >= 0 && Clock::now() >= deadline) return LCS_EAGAIN;
            S->cv.wait_for(lk, std::chrono::milliseconds(1));
        }
    }
    int32_t done = 0;
    return lcs_server_poll(S, id, h_out, v_out, &done);
}

int lcs_server_release(lcs_server *S, int64_t id)
{
    if (!S) return LCS_EINVAL;
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->all.find(id);
    if (it == S->all.end()) return LCS_ENOENT;
    if (it->second->status == LCS_EAGAIN) return LCS_EBUSY;
    S->all.erase(it);
    return LCS_OK;
}

int lcs_server_cancel(lcs_server *S, int64_t id)
{
    if (!S) return LCS_EINVAL;
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->all.find(id);
    if (it == S->all.end()) return LCS_ENOENT;
    it->second->cancelled = true;  // dropped if still queued, discarded if running
    return LCS_OK;
}

int lcs_server_request_times(lcs_server *S, int64_t id, double *queued_us, double *total_us)
{
    if (!S) return LCS_EINVAL;
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->all.find(id);
    if (it == S->all.end()) return LCS_ENOENT;
    const auto &R = *it->second;
    if (R.status == LCS_EAGAIN) return LCS_EAGAIN;
    if (queued_us) *queued_us = std::chrono::duration<double, std::micro>(R.t_start - R.t_submit).count();
    if (total_us) *total_us = std::chrono::duration<double, std::micro>(R.t_done - R.t_submit).count();
    return LCS_OK;
}

int lcs_server_stats(lcs_server *S, int64_t *pending, int64_t *completed, int64_t *batches, int32_t *max_inflight,
                     size_t *free_bytes)
{
    if (!S) return LCS_EINVAL;
    std::lock_guard<std::mutex> lk(S->mu);
    if (pending) *pending = (int64_t)S->pending.size();
    if (completed) *completed = S->completed;
    if (batches) *batches = S->batches;
    if (max_inflight) *max_inflight = S->max_inflight_seen;
    if (free_bytes) *free_bytes = S->arena.free_bytes();
    return S->worker_error;
}

}  // extern "C"
