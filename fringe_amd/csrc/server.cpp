// server.cpp -- tile server: many phase-linking requests resident in HBM at
// once, admitted under a memory budget and run concurrently on a pool of
// streams (continuous batching: a request that arrives while others run is
// admitted as soon as a stream and arena space are free, and its kernels
// overlap with theirs on the device).
//
// One dispatcher thread owns every HIP launch.  Completion is signalled by a
// host function enqueued behind each request's kernels; it only touches host
// state (arena bookkeeping, request state) under the server mutex.
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "engine_internal.hpp"
#include "fringe_amd.h"

namespace {

using Clock = std::chrono::steady_clock;

constexpr size_t kAlign = 256;
constexpr int kMaxSkips = 4;   // admissions past the oldest request before it blocks the queue

size_t align_up(size_t b) { return (b + kAlign - 1) & ~(kAlign - 1); }

// First-fit allocator over [0, size) with coalescing free list.
struct Arena {
    size_t size = 0;
    std::map<size_t, size_t> free_;   // offset -> bytes
    void init(size_t s) {
        size = s;
        free_.clear();
        if (s) free_[0] = s;
    }
    bool alloc(size_t bytes, size_t* off) {
        if (bytes == 0) {
            *off = 0;
            return true;
        }
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            *off = it->first;
            const size_t rest = it->second - bytes, at = it->first + bytes;
            free_.erase(it);
            if (rest) free_[at] = rest;
            return true;
        }
        return false;
    }
    void release(size_t off, size_t bytes) {
        if (bytes == 0) return;
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
};

// Byte layout of one request inside the arena.
struct Layout {
    bool nmap = false;
    size_t nmap_ws = 0, link_ws = 0;
    // outputs (held until release)
    size_t o_phase = 0, o_tcorr = 0, o_comp = 0, o_wts = 0, o_count = 0, out_bytes = 0;
    // workspace (held until completion)
    size_t w_nmap = 0, w_link = 0, work_bytes = 0;
};

int plan(const fringe_server_request* r, Layout* L) {
    if (!r || !r->d_slc) return FRINGE_ERR_ARGUMENT;
    const int st = fringe::link_check(&r->tile);
    if (st) return st;
    const fringe_link_tile& t = r->tile;
    const size_t npx = (size_t)t.rows * t.cols;
    L->nmap = (r->d_wts == nullptr);
    if (L->nmap) {
        if (!(r->prob > 0.0 && r->prob < 1.0)) return FRINGE_ERR_ARGUMENT;
        fringe_nmap_tile nt = {t.rows, t.cols, t.nbands, t.Nx, t.Ny, r->prob, 0, t.row_begin, t.row_end};
        L->nmap_ws = fringe_nmap_workspace_bytes(&nt);
        if (L->nmap_ws == 0) return FRINGE_ERR_ARGUMENT;
    }
    L->link_ws = fringe_link_workspace_bytes(&t);
    if (L->link_ws == 0) return FRINGE_ERR_ARGUMENT;
    size_t b = 0;
    L->o_phase = b;
    b += align_up(npx * t.nbands * 8);
    L->o_tcorr = b;
    b += align_up(npx * 4);
    L->o_comp = b;
    b += align_up(npx * 8);
    if (L->nmap) {
        L->o_wts = b;
        b += align_up(npx * (size_t)fringe_mask_words(t.Nx, t.Ny) * 4);
        L->o_count = b;
        b += align_up(npx * 4);
    }
    L->out_bytes = b;
    b = 0;
    L->w_nmap = b;
    b += align_up(L->nmap_ws);
    L->w_link = b;
    b += align_up(L->link_ws);
    L->work_bytes = b;
    return 0;
}

struct Request {
    long long id = 0;
    fringe_server_request r{};
    Layout lay;
    size_t work_off = 0, out_off = 0;
    int state = FRINGE_REQ_QUEUED;
    int code = 0;
    int stream = -1;
    int skips = 0;
    Clock::time_point t_submit, t_start, t_done;
    fringe_server* srv = nullptr;
};

}  // namespace

struct fringe_server {
    int device = 0;
    char* base = nullptr;
    Arena arena;
    std::vector<hipStream_t> streams;
    std::vector<char> busy;
    std::mutex mu;
    std::condition_variable wake;   // dispatcher: submit, completion, stop
    std::condition_variable done;   // waiters
    std::deque<Request*> queue;
    std::unordered_map<long long, std::unique_ptr<Request>> reqs;
    long long next_id = 1;
    int running = 0;
    bool stop = false;
    std::thread th;
};

namespace {

void on_done(void* p) {
    Request* q = static_cast<Request*>(p);
    fringe_server* S = q->srv;
    std::lock_guard<std::mutex> lk(S->mu);
    q->t_done = Clock::now();
    q->state = FRINGE_REQ_DONE;
    S->arena.release(q->work_off, q->lay.work_bytes);
    S->busy[q->stream] = 0;
    S->running--;
    S->wake.notify_all();
    S->done.notify_all();
}

// Launch every kernel of an admitted request on its stream (no lock held).
void launch(fringe_server* S, Request* q) {
    hipStream_t st = S->streams[q->stream];
    const fringe_link_tile& t = q->r.tile;
    const Layout& L = q->lay;
    char* out = S->base + q->out_off;
    char* work = S->base + q->work_off;
    const size_t npx = (size_t)t.rows * t.cols;
    const uint32_t* wts = q->r.d_wts;
    int rc = 0;
    if (L.nmap) {
        fringe_nmap_tile nt = {t.rows, t.cols, t.nbands, t.Nx, t.Ny, q->r.prob, 0, t.row_begin, t.row_end};
        rc = fringe_nmap_run(&nt, q->r.d_slc, q->r.d_mask, nullptr, (uint32_t*)(out + L.o_wts),
                             (int32_t*)(out + L.o_count), work + L.w_nmap, L.nmap_ws, st);
        wts = (const uint32_t*)(out + L.o_wts);
    }
    // rows outside [row_begin, row_end) are not written by the link: clear them
    if (rc == 0 && (t.row_begin > 0 || t.row_end < t.rows)) {
        if (hipMemsetAsync(out + L.o_phase, 0, npx * t.nbands * 8, st) != hipSuccess ||
            hipMemsetAsync(out + L.o_tcorr, 0, npx * 4, st) != hipSuccess ||
            hipMemsetAsync(out + L.o_comp, 0, npx * 8, st) != hipSuccess)
            rc = FRINGE_ERR_DEVICE;
    }
    if (rc == 0)
        rc = fringe_link_run(&t, q->r.d_slc, wts, out + L.o_phase, (float*)(out + L.o_tcorr), out + L.o_comp,
                             work + L.w_link, L.link_ws, st);
    {
        std::lock_guard<std::mutex> lk(S->mu);
        q->code = rc;
    }
    if (hipLaunchHostFunc(st, on_done, q) != hipSuccess) {
        hipStreamSynchronize(st);
        {
            std::lock_guard<std::mutex> lk(S->mu);
            if (q->code == 0) q->code = FRINGE_ERR_DEVICE;
        }
        on_done(q);
    }
}

int free_stream(fringe_server* S) {
    for (size_t i = 0; i < S->busy.size(); ++i)
        if (!S->busy[i]) return (int)i;
    return -1;
}

// Reserve stream + arena space for q; false if either is missing.
bool try_admit(fringe_server* S, Request* q) {
    const int si = free_stream(S);
    if (si < 0) return false;
    size_t wo = 0, oo = 0;
    if (!S->arena.alloc(q->lay.work_bytes, &wo)) return false;
    if (!S->arena.alloc(q->lay.out_bytes, &oo)) {
        S->arena.release(wo, q->lay.work_bytes);
        return false;
    }
    q->work_off = wo;
    q->out_off = oo;
    q->stream = si;
    S->busy[si] = 1;
    S->running++;
    q->state = FRINGE_REQ_RUNNING;
    q->t_start = Clock::now();
    return true;
}

void dispatcher(fringe_server* S) {
    hipSetDevice(S->device);
    std::unique_lock<std::mutex> lk(S->mu);
    for (;;) {
        if (S->stop && S->queue.empty() && S->running == 0) break;
        std::vector<Request*> go;
        // oldest first; later requests may pass a request that does not fit
        // yet, until it has been passed kMaxSkips times
        Request* head = S->queue.empty() ? nullptr : S->queue.front();
        bool head_waiting = false;
        for (auto it = S->queue.begin(); it != S->queue.end();) {
            Request* q = *it;
            if (free_stream(S) < 0) break;
            if (q != head && head_waiting && head->skips >= kMaxSkips) break;
            if (try_admit(S, q)) {
                if (q != head && head_waiting) head->skips++;
                go.push_back(q);
                it = S->queue.erase(it);
                continue;
            }
            if (q == head) head_waiting = true;
            ++it;
        }
        if (!go.empty()) {
            lk.unlock();
            for (Request* q : go) launch(S, q);
            lk.lock();
            continue;
        }
        // stopping with nothing running: a queued request that could not be
        // admitted now never will be (the arena space it waits for is held
        // by outputs nobody releases any more); fail it so the loop ends
        if (S->stop && S->running == 0 && !S->queue.empty()) {
            const auto now = Clock::now();
            for (Request* q : S->queue) {
                q->code = FRINGE_ERR_WORKSPACE;
                q->state = FRINGE_REQ_DONE;
                q->t_start = q->t_done = now;
            }
            S->queue.clear();
            S->done.notify_all();
            continue;
        }
        S->wake.wait(lk);
    }
}

Request* find(fringe_server* S, long long id) {
    auto it = S->reqs.find(id);
    return it == S->reqs.end() ? nullptr : it->second.get();
}

double ms_between(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

}  // namespace

extern "C" {

size_t fringe_server_request_bytes(const fringe_server_request* r) {
    Layout L;
    if (plan(r, &L) != 0) return 0;
    return L.out_bytes + L.work_bytes;
}

int fringe_server_create(int device, size_t budget_bytes, int streams, fringe_server** out) {
    if (!out || budget_bytes == 0 || streams <= 0 || streams > 64) return FRINGE_ERR_ARGUMENT;
    *out = nullptr;
    int prev = 0;
    if (hipGetDevice(&prev) != hipSuccess) return FRINGE_ERR_DEVICE;
    if (hipSetDevice(device) != hipSuccess) return FRINGE_ERR_DEVICE;
    std::unique_ptr<fringe_server> S(new fringe_server());
    S->device = device;
    const size_t budget = budget_bytes & ~(kAlign - 1);
    int rc = 0;
    if (hipMalloc((void**)&S->base, budget) != hipSuccess) {
        fringe::set_last_error("fringe_server_create: arena allocation failed");
        rc = FRINGE_ERR_DEVICE;
    }
    for (int i = 0; rc == 0 && i < streams; ++i) {
        hipStream_t s;
        if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) {
            rc = FRINGE_ERR_DEVICE;
            break;
        }
        S->streams.push_back(s);
    }
    if (rc) {
        for (hipStream_t s : S->streams) hipStreamDestroy(s);
        if (S->base) hipFree(S->base);
        hipSetDevice(prev);
        return rc;
    }
    S->busy.assign(S->streams.size(), 0);
    S->arena.init(budget);
    fringe_server* raw = S.get();
    S->th = std::thread(dispatcher, raw);
    hipSetDevice(prev);
    *out = S.release();
    return 0;
}

int fringe_server_submit(fringe_server* S, const fringe_server_request* r, long long* id) {
    if (!S || !id) return FRINGE_ERR_ARGUMENT;
    std::unique_ptr<Request> q(new Request());
    const int st = plan(r, &q->lay);
    if (st) return st;
    if (q->lay.out_bytes + q->lay.work_bytes > S->arena.size) return FRINGE_ERR_WORKSPACE;
    q->r = *r;
    q->srv = S;
    q->t_submit = Clock::now();
    std::lock_guard<std::mutex> lk(S->mu);
    if (S->stop) return FRINGE_ERR_ARGUMENT;
    q->id = S->next_id++;
    *id = q->id;
    S->queue.push_back(q.get());
    S->reqs.emplace(q->id, std::move(q));
    S->wake.notify_all();
    return 0;
}

int fringe_server_poll(fringe_server* S, long long id, int* state) {
    if (!S) return FRINGE_ERR_ARGUMENT;
    std::lock_guard<std::mutex> lk(S->mu);
    Request* q = find(S, id);
    if (!q) return FRINGE_ERR_ARGUMENT;
    if (state) *state = q->state;
    return q->state == FRINGE_REQ_DONE ? q->code : 0;
}

int fringe_server_wait(fringe_server* S, long long id) {
    if (!S) return FRINGE_ERR_ARGUMENT;
    std::unique_lock<std::mutex> lk(S->mu);
    Request* q = find(S, id);
    if (!q) return FRINGE_ERR_ARGUMENT;
    S->done.wait(lk, [q] { return q->state == FRINGE_REQ_DONE; });
    return q->code;
}

int fringe_server_result(fringe_server* S, long long id, void** d_phase, float** d_tcorr, void** d_comp,
                         uint32_t** d_wts, int32_t** d_count) {
    if (!S) return FRINGE_ERR_ARGUMENT;
    std::lock_guard<std::mutex> lk(S->mu);
    Request* q = find(S, id);
    if (!q || q->state != FRINGE_REQ_DONE) return FRINGE_ERR_ARGUMENT;
    char* out = S->base + q->out_off;
    if (d_phase) *d_phase = out + q->lay.o_phase;
    if (d_tcorr) *d_tcorr = (float*)(out + q->lay.o_tcorr);
    if (d_comp) *d_comp = out + q->lay.o_comp;
    if (d_wts) *d_wts = q->lay.nmap ? (uint32_t*)(out + q->lay.o_wts) : nullptr;
    if (d_count) *d_count = q->lay.nmap ? (int32_t*)(out + q->lay.o_count) : nullptr;
    return q->code;
}

int fringe_server_times(fringe_server* S, long long id, double* queued_ms, double* total_ms) {
    if (!S) return FRINGE_ERR_ARGUMENT;
    std::lock_guard<std::mutex> lk(S->mu);
    Request* q = find(S, id);
    if (!q || q->state != FRINGE_REQ_DONE) return FRINGE_ERR_ARGUMENT;
    if (queued_ms) *queued_ms = ms_between(q->t_submit, q->t_start);
    if (total_ms) *total_ms = ms_between(q->t_submit, q->t_done);
    return 0;
}

int fringe_server_release(fringe_server* S, long long id) {
    if (!S) return FRINGE_ERR_ARGUMENT;
    std::lock_guard<std::mutex> lk(S->mu);
    Request* q = find(S, id);
    if (!q || q->state != FRINGE_REQ_DONE) return FRINGE_ERR_ARGUMENT;
    S->arena.release(q->out_off, q->lay.out_bytes);
    S->reqs.erase(id);
    S->wake.notify_all();
    return 0;
}

int fringe_server_destroy(fringe_server* S) {
    if (!S) return FRINGE_ERR_ARGUMENT;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
        S->wake.notify_all();
    }
    if (S->th.joinable()) S->th.join();
    int prev = 0;
    hipGetDevice(&prev);
    hipSetDevice(S->device);
    int rc = 0;
    for (hipStream_t s : S->streams) {
        if (hipStreamSynchronize(s) != hipSuccess) rc = FRINGE_ERR_DEVICE;
        hipStreamDestroy(s);
    }
    if (S->base && hipFree(S->base) != hipSuccess) rc = FRINGE_ERR_DEVICE;
    hipSetDevice(prev);
    delete S;
    return rc;
}

}  // extern "C"
