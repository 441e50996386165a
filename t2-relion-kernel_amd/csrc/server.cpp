// Serving engine: continuous batching of coarse-search requests on one GPU.
//
// Requests (a micrograph's particles against 1..n classes) are submitted from
// any host thread; their inputs are copied into a pinned host arena so the
// caller may reuse its buffers at once.  One worker thread owns the device:
// it admits queued requests in arrival order whenever an in-flight slot is
// free AND the request's device memory (inputs, plan buffers, diff2 output)
// fits in the HBM budget arena, launches the whole search of a request on
// the slot's own stream (so requests overlap on the device), and retires
// requests whose completion event has fired, returning their memory to the
// arena.  Projector volumes stay device-resident: requests reference
// rc_projector objects the caller keeps alive.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "plan_internal.hpp"
#include "rc_internal.hpp"

namespace rc {

static size_t align256(size_t v) { return (v + 255) / 256 * 256; }

static double now_s()
{
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// First-fit allocator over one contiguous buffer (host bookkeeping only).
class Arena {
public:
    void init(size_t size)
    {
        size_ = size;
        free_.clear();
        if (size) free_[0] = size;
        used_ = peak_ = 0;
    }
    bool alloc(size_t n, size_t *off)
    {
        n = align256(std::max<size_t>(n, 1));
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= n) {
                *off = it->first;
                const size_t rest = it->second - n, start = it->first + n;
                free_.erase(it);
                if (rest) free_[start] = rest;
                used_ += n;
                peak_ = std::max(peak_, used_);
                return true;
            }
        }
        return false;
    }
    void release(size_t off, size_t n)
    {
        n = align256(std::max<size_t>(n, 1));
        used_ -= n;
        auto it = free_.emplace(off, n).first;
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
    size_t size() const { return size_; }
    size_t used() const { return used_; }
    size_t peak() const { return peak_; }

private:
    size_t size_ = 0, used_ = 0, peak_ = 0;
    std::map<size_t, size_t> free_;
};

struct Request {
    uint64_t ticket = 0;
    int state = 0;                  // 0 queued, 1 running, 2 done
    bool keep = false;
    int P = 0, C = 0, O = 0, npix = 0, full = 0, form = 0;
    std::vector<rc_projector_kernel> pks;
    size_t total_trans = 0, n_diff2 = 0;
    // pinned host staging: inputs, plan tables, results
    size_t h_off = 0, h_bytes = 0;
    unsigned char *h = nullptr;
    size_t h_in = 0, h_tab = 0, h_best = 0, h_idx = 0;
    // device memory in the arena
    size_t d_off = 0, d_bytes = 0;
    unsigned char *d = nullptr;
    size_t d_plan = 0, d_diff2 = 0, d_best = 0, d_idx = 0;
    size_t in_bytes = 0;
    rc_plan plan;
    PlanTables tab;
    int slot = -1;
    hipEvent_t done = nullptr;
    double t_submit = 0, t_start = 0, t_done = 0;
    int status = RC_OK;
};

// Offsets of the inputs inside a request's input block (host and device
// copies share one layout).
struct InLayout {
    size_t off_trans, tx, ty, eul, re, im, corr, xy, total;
};

static InLayout in_layout(int P, int O, size_t T, int full)
{
    InLayout l;
    size_t o = 0;
    l.off_trans = o;
    o = align256(o + (size_t)(P + 1) * sizeof(int));
    l.tx = o;
    o = align256(o + T * sizeof(float));
    l.ty = o;
    o = align256(o + T * sizeof(float));
    l.eul = o;
    o = align256(o + (size_t)O * 9 * sizeof(float));
    l.re = o;
    o = align256(o + (size_t)P * full * sizeof(float));
    l.im = o;
    o = align256(o + (size_t)P * full * sizeof(float));
    l.corr = o;
    o = align256(o + (size_t)P * full * sizeof(float));
    l.xy = o;
    o = align256(o + (size_t)full * 2 * sizeof(float));
    l.total = o;
    return l;
}

}  // namespace rc

struct rc_server {
    rc_server_config cfg{};
    int device = 0;
    unsigned char *dmem = nullptr;
    unsigned char *hmem = nullptr;
    rc::Arena darena, harena;
    std::vector<hipStream_t> streams;
    std::vector<int> slot_busy;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<rc::Request *> queue;
    std::map<uint64_t, std::unique_ptr<rc::Request>> reqs;
    std::vector<rc::Request *> running;
    uint64_t next_ticket = 1;
    bool stop = false;
    std::thread worker;
    // statistics
    uint64_t completed = 0;
    int peak_inflight = 0;
    double busy_since = 0, first_start = 0, last_done = 0;
};

namespace rc {

static rc_batch_desc device_desc(Request &r)
{
    const InLayout l = in_layout(r.P, r.O, r.total_trans, r.full);
    rc_batch_desc d{};
    d.n_particles = r.P;
    d.n_classes = r.C;
    d.orientation_num = r.O;
    d.npix = r.npix;
    d.full_npix = r.full;
    d.form = r.form;
    d.projectors = r.pks.data();
    d.host_trans_offset = reinterpret_cast<const int *>(r.h + r.h_in + l.off_trans);
    d.d_trans_offset = reinterpret_cast<const int *>(r.d + l.off_trans);
    d.d_trans_x = reinterpret_cast<const float *>(r.d + l.tx);
    d.d_trans_y = reinterpret_cast<const float *>(r.d + l.ty);
    d.d_eulers = reinterpret_cast<const float *>(r.d + l.eul);
    d.d_real = reinterpret_cast<const float *>(r.d + l.re);
    d.d_imag = reinterpret_cast<const float *>(r.d + l.im);
    d.d_corr = reinterpret_cast<const float *>(r.d + l.corr);
    d.d_coor_xy = reinterpret_cast<const float *>(r.d + l.xy);
    d.d_diff2_init = nullptr;
    d.d_diff2 = reinterpret_cast<float *>(r.d + r.d_diff2);
    d.d_best_diff2 = reinterpret_cast<double *>(r.d + r.d_best);
    d.d_best_index = reinterpret_cast<int64_t *>(r.d + r.d_idx);
    return d;
}

// Launch the whole search of an admitted request on its slot's stream.
static int launch_request(rc_server *s, Request &r)
{
    hipStream_t st = s->streams[r.slot];
    RC_HIP(hipMemcpyAsync(r.d, r.h + r.h_in, r.in_bytes, hipMemcpyHostToDevice, st));
    plan_bind(&r.plan, r.d + r.d_plan);
    if (int e = plan_upload_async(&r.plan, r.tab, r.h + r.h_tab, st)) return e;
    rc_batch_desc d = device_desc(r);
    if (int e = plan_run(&r.plan, &d, st)) return e;
    RC_HIP(hipMemcpyAsync(r.h + r.h_best, r.d + r.d_best, (size_t)r.P * sizeof(double),
                          hipMemcpyDeviceToHost, st));
    RC_HIP(hipMemcpyAsync(r.h + r.h_idx, r.d + r.d_idx, (size_t)r.P * sizeof(int64_t),
                          hipMemcpyDeviceToHost, st));
    RC_HIP(hipEventRecord(r.done, st));
    return RC_OK;
}

static void retire(rc_server *s, Request *r)
{
    r->t_done = now_s();
    r->state = 2;
    s->slot_busy[r->slot] = 0;
    if (!r->keep) {
        s->darena.release(r->d_off, r->d_bytes);
        r->d = nullptr;
    }
    ++s->completed;
    s->last_done = r->t_done;
}

static void worker_loop(rc_server *s)
{
    (void)hipSetDevice(s->device);
    std::unique_lock<std::mutex> lk(s->mu);
    while (true) {
        // retire finished requests
        bool changed = false;
        for (size_t k = 0; k < s->running.size();) {
            Request *r = s->running[k];
            const hipError_t e = hipEventQuery(r->done);
            if (e == hipErrorNotReady) {
                ++k;
                continue;
            }
            if (e != hipSuccess) r->status = RC_ERR_DEVICE;
            retire(s, r);
            s->running.erase(s->running.begin() + k);
            changed = true;
        }
        if (changed) s->cv_done.notify_all();
        // admit in arrival order while a slot and the memory are free
        while (!s->queue.empty()) {
            int slot = -1;
            for (int k = 0; k < (int)s->slot_busy.size(); ++k)
                if (!s->slot_busy[k]) {
                    slot = k;
                    break;
                }
            if (slot < 0) break;
            Request *r = s->queue.front();
            size_t off = 0;
            if (!s->darena.alloc(r->d_bytes, &off)) break;
            s->queue.pop_front();
            r->d_off = off;
            r->d = s->dmem + off;
            r->slot = slot;
            r->state = 1;
            r->t_start = now_s();
            if (s->first_start == 0) s->first_start = r->t_start;
            s->slot_busy[slot] = 1;
            const int st = launch_request(s, *r);
            if (st != RC_OK) {
                r->status = st;
                (void)hipStreamSynchronize(s->streams[slot]);
                retire(s, r);
                s->cv_done.notify_all();
                continue;
            }
            s->running.push_back(r);
            s->peak_inflight = std::max(s->peak_inflight, (int)s->running.size());
        }
        if (s->stop && s->queue.empty() && s->running.empty()) break;
        if (s->running.empty() && s->queue.empty())
            s->cv_work.wait(lk);
        else
            s->cv_work.wait_for(lk, std::chrono::microseconds(50));
    }
}

}  // namespace rc

using namespace rc;

extern "C" {

int rc_server_create(const rc_server_config *cfg, rc_server **out)
{
    RC_REQUIRE(cfg && out, RC_ERR_INVALID, "NULL argument");
    *out = nullptr;
    RC_REQUIRE(cfg->max_inflight >= 1 && cfg->max_inflight <= 64, RC_ERR_INVALID,
               "max_inflight must be 1..64");
    RC_REQUIRE(cfg->hbm_budget_bytes >= (1u << 20) && cfg->pinned_budget_bytes >= (1u << 20),
               RC_ERR_INVALID, "budgets must be at least 1 MiB");
    RC_REQUIRE(cfg->form == RC_FORM_FP32 || cfg->form == RC_FORM_BF16X3 ||
                   cfg->form == RC_FORM_FP16,
               RC_ERR_INVALID, "unknown form %d", cfg->form);
    auto s = std::make_unique<rc_server>();
    s->cfg = *cfg;
    RC_HIP(hipGetDevice(&s->device));
    if (hipMalloc(&s->dmem, cfg->hbm_budget_bytes) != hipSuccess) {
        set_error("hipMalloc of the %zu-byte HBM budget failed", cfg->hbm_budget_bytes);
        return RC_ERR_NO_MEMORY;
    }
    if (hipHostMalloc(&s->hmem, cfg->pinned_budget_bytes, hipHostMallocDefault) != hipSuccess) {
        (void)hipFree(s->dmem);
        set_error("hipHostMalloc of %zu pinned bytes failed", cfg->pinned_budget_bytes);
        return RC_ERR_NO_MEMORY;
    }
    s->darena.init(cfg->hbm_budget_bytes);
    s->harena.init(cfg->pinned_budget_bytes);
    s->streams.resize(cfg->max_inflight);
    s->slot_busy.assign(cfg->max_inflight, 0);
    for (auto &st : s->streams) RC_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    rc_server *raw = s.release();
    raw->worker = std::thread(worker_loop, raw);
    *out = raw;
    return RC_OK;
}

int rc_server_submit(rc_server *s, const rc_request *q, uint64_t *ticket)
{
    RC_REQUIRE(s && q && ticket, RC_ERR_INVALID, "NULL argument");
    RC_REQUIRE(q->n_particles > 0 && q->n_classes > 0 && q->orientation_num > 0 && q->npix > 0 &&
                   q->full_npix >= q->npix && q->projectors && q->trans_offset && q->trans_x &&
                   q->trans_y && q->eulers && q->real && q->imag && q->corr && q->coor_xy,
               RC_ERR_INVALID, "incomplete request");
    auto r = std::make_unique<Request>();
    r->P = q->n_particles;
    r->C = q->n_classes;
    r->O = q->orientation_num;
    r->npix = q->npix;
    r->full = q->full_npix;
    r->form = s->cfg.form;
    r->keep = q->keep_diff2 != 0;
    r->pks.assign(q->projectors, q->projectors + q->n_classes);
    r->total_trans = (size_t)q->trans_offset[q->n_particles];
    const InLayout l = in_layout(r->P, r->O, r->total_trans, r->full);
    r->in_bytes = l.total;
    // plan (host side) against a descriptor with the caller's host offsets
    rc_batch_desc hd{};
    hd.n_particles = r->P;
    hd.n_classes = r->C;
    hd.orientation_num = r->O;
    hd.npix = r->npix;
    hd.full_npix = r->full;
    hd.form = r->form;
    hd.projectors = r->pks.data();
    hd.host_trans_offset = q->trans_offset;
    if (int st = plan_setup(&r->plan, &hd, r->tab)) return st;
    const size_t plan_bytes = plan_bind(&r->plan, nullptr);
    r->n_diff2 = r->total_trans * r->C * r->O;
    // device block: inputs | plan | diff2 | best | index
    size_t o = align256(l.total);
    r->d_plan = o;
    o = align256(o + plan_bytes);
    r->d_diff2 = o;
    o = align256(o + r->n_diff2 * sizeof(float));
    r->d_best = o;
    o = align256(o + (size_t)r->P * sizeof(double));
    r->d_idx = o;
    o = align256(o + (size_t)r->P * sizeof(int64_t));
    r->d_bytes = o;
    RC_REQUIRE(r->d_bytes <= s->darena.size(), RC_ERR_LIMIT,
               "request needs %zu device bytes, more than the %zu-byte budget", r->d_bytes,
               s->darena.size());
    // pinned block: inputs | tables | best | index
    size_t h = 0;
    r->h_in = h;
    h = align256(h + l.total);
    r->h_tab = h;
    h = align256(h + plan_table_bytes(r->tab));
    r->h_best = h;
    h = align256(h + (size_t)r->P * sizeof(double));
    r->h_idx = h;
    h = align256(h + (size_t)r->P * sizeof(int64_t));
    r->h_bytes = h;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        RC_REQUIRE(s->harena.alloc(r->h_bytes, &r->h_off), RC_ERR_NO_MEMORY,
                   "pinned staging full (%zu bytes needed, %zu in use of %zu)", r->h_bytes,
                   s->harena.used(), s->harena.size());
    }
    r->h = s->hmem + r->h_off;
    unsigned char *in = r->h + r->h_in;
    const size_t pf = (size_t)r->P * r->full * sizeof(float);
    memcpy(in + l.off_trans, q->trans_offset, (size_t)(r->P + 1) * sizeof(int));
    memcpy(in + l.tx, q->trans_x, r->total_trans * sizeof(float));
    memcpy(in + l.ty, q->trans_y, r->total_trans * sizeof(float));
    memcpy(in + l.eul, q->eulers, (size_t)r->O * 9 * sizeof(float));
    memcpy(in + l.re, q->real, pf);
    memcpy(in + l.im, q->imag, pf);
    memcpy(in + l.corr, q->corr, pf);
    memcpy(in + l.xy, q->coor_xy, (size_t)r->full * 2 * sizeof(float));
    if (hipEventCreateWithFlags(&r->done, hipEventDisableTiming) != hipSuccess) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->harena.release(r->h_off, r->h_bytes);
        set_error("hipEventCreate failed");
        return RC_ERR_DEVICE;
    }
    std::lock_guard<std::mutex> lk(s->mu);
    r->ticket = s->next_ticket++;
    r->t_submit = now_s();
    *ticket = r->ticket;
    s->queue.push_back(r.get());
    s->reqs[r->ticket] = std::move(r);
    s->cv_work.notify_one();
    return RC_OK;
}

int rc_server_wait(rc_server *s, uint64_t ticket, double *best, int64_t *best_index,
                   double *latency_ms)
{
    RC_REQUIRE(s, RC_ERR_INVALID, "NULL server");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    RC_REQUIRE(it != s->reqs.end(), RC_ERR_INVALID, "unknown ticket %llu",
               (unsigned long long)ticket);
    Request *r = it->second.get();
    s->cv_done.wait(lk, [&] { return r->state == 2; });
    if (best) memcpy(best, r->h + r->h_best, (size_t)r->P * sizeof(double));
    if (best_index) memcpy(best_index, r->h + r->h_idx, (size_t)r->P * sizeof(int64_t));
    if (latency_ms) *latency_ms = (r->t_done - r->t_submit) * 1e3;
    const int st = r->status;
    if (!r->keep) {
        s->harena.release(r->h_off, r->h_bytes);
        (void)hipEventDestroy(r->done);
        s->reqs.erase(it);
    }
    RC_REQUIRE(st == RC_OK, st, "request %llu failed", (unsigned long long)ticket);
    return RC_OK;
}

int rc_server_diff2(rc_server *s, uint64_t ticket, float *host_out, size_t count)
{
    RC_REQUIRE(s && host_out, RC_ERR_INVALID, "NULL argument");
    unsigned char *src = nullptr;
    size_t n = 0;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        auto it = s->reqs.find(ticket);
        RC_REQUIRE(it != s->reqs.end(), RC_ERR_INVALID, "unknown ticket");
        Request *r = it->second.get();
        s->cv_done.wait(lk, [&] { return r->state == 2; });
        RC_REQUIRE(r->keep && r->d, RC_ERR_INVALID, "request was not submitted with keep_diff2");
        src = r->d + r->d_diff2;
        n = r->n_diff2;
    }
    RC_REQUIRE(count >= n, RC_ERR_INVALID, "output holds %zu values, %zu needed", count, n);
    RC_HIP(hipMemcpy(host_out, src, n * sizeof(float), hipMemcpyDeviceToHost));
    return RC_OK;
}

int rc_server_release(rc_server *s, uint64_t ticket)
{
    RC_REQUIRE(s, RC_ERR_INVALID, "NULL server");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    RC_REQUIRE(it != s->reqs.end(), RC_ERR_INVALID, "unknown ticket");
    Request *r = it->second.get();
    s->cv_done.wait(lk, [&] { return r->state == 2; });
    if (r->d) s->darena.release(r->d_off, r->d_bytes);
    s->harena.release(r->h_off, r->h_bytes);
    (void)hipEventDestroy(r->done);
    s->reqs.erase(it);
    s->cv_work.notify_one();
    return RC_OK;
}

int rc_server_stats(rc_server *s, rc_server_counters *out)
{
    RC_REQUIRE(s && out, RC_ERR_INVALID, "NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    out->completed = s->completed;
    out->queued = (int)s->queue.size();
    out->inflight = (int)s->running.size();
    out->peak_inflight = s->peak_inflight;
    out->hbm_used = s->darena.used();
    out->hbm_peak = s->darena.peak();
    out->hbm_budget = s->darena.size();
    out->busy_s = s->last_done > s->first_start ? s->last_done - s->first_start : 0.0;
    return RC_OK;
}

int rc_server_destroy(rc_server *s)
{
    if (!s) return RC_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        s->cv_work.notify_all();
    }
    if (s->worker.joinable()) s->worker.join();
    for (auto &kv : s->reqs)
        if (kv.second->done) (void)hipEventDestroy(kv.second->done);
    for (auto st : s->streams) (void)hipStreamDestroy(st);
    (void)hipFree(s->dmem);
    (void)hipHostFree(s->hmem);
    delete s;
    return RC_OK;
}

}  // extern "C"
