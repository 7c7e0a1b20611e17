// alto_serve.hip -- device-resident request loop with continuous batching.
//
// The reference serves requests with reader/writer threads around one engine thread that
// drains the whole queue into one synchronous alto_answer call per batch
// (src/bin/altopir.rs:372-496).  Here every resident database keeps a pool of request
// slots (pinned host staging + a device arena) and several engine contexts:
//   submit   copies the frame into a free pinned slot and starts its upload at once on the
//            database's upload stream (an event per slot marks its arrival in HBM);
//   dispatch whenever a context is free, takes every waiting request of that database (up
//            to max_batch, never waiting to fill a batch), makes the context's stream wait
//            for their uploads and answers them with alto_answer_gather; responses are
//            written by the tail kernel straight into mapped pinned host memory;
//   complete a completion thread waits for each batch, stamps the completion time and
//            hands the tickets to poll().
// Batches of one database overlap on the device as far as its contexts allow; the number
// of contexts is bounded by the device memory free at start (memory admission).  Frames
// whose length is wrong get status 1 at once, and status 9 when every slot is taken
// (src/bin/altopir.rs:417-436); the device validates the rest (statuses 2 and 3).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/altopir_engine.h"

namespace {

int64_t now_ns() {
    return std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

size_t round_up(size_t x, size_t m) { return (x + m - 1) / m * m; }

// Upload modes (alto_server_set_upload).
constexpr int kCopyAtSubmit = 0, kZeroCopy = 1, kCopyAtDispatch = 2, kCopyBoth = 3;
// By default frames up to this size are copied per batch at dispatch and the batch's
// responses written to HBM and copied back per batch: a copy call per frame costs more than
// the transfer of a small frame, and kernels reading or writing host memory over PCIe hold
// the CUs (1 GiB databases served at saturation: 61-67k QPS with p99 11-14 ms, against
// 54-68k with p99 10-22 ms copying only the frames, 55-60k reading the slots and 46.5k
// copying at submit).  Larger frames are copied at submit, so that their transfer overlaps
// earlier batches' work (8 GiB: 22.7k against 16-18k at dispatch and 14.8k reading the
// slots).
constexpr size_t kSmallFrame = 512 << 10;
int default_upload(size_t frame_len) { return frame_len <= kSmallFrame ? kCopyBoth : kCopyAtSubmit; }

struct Server;

struct Ctx {
    void *ctx = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    int *d_idx = nullptr;   // device: slot index of each batch position
    int *h_idx = nullptr;   // pinned staging of d_idx
    int *h_status = nullptr;  // mapped pinned: statuses written by the device
    int *d_status = nullptr;  // device view of h_status
    std::vector<int> batch;   // slots of the batch in flight
    std::vector<int> runs;    // the batch's slots, sorted (copy at dispatch)
    bool busy = false;
    bool failed = false;      // a copy of the batch failed: every request reports -1
};

struct Db {
    void *engine = nullptr;
    size_t L = 0, RL = 0, S = 0;
    int nslots = 0;
    uint8_t *h_req = nullptr, *d_req = nullptr;      // nslots x S (pinned, mapped / device arena)
    uint8_t *h_req_dev = nullptr;                    // device view of h_req
    // how frames reach the device (alto_server_set_upload): kCopyAtSubmit each frame is
    // copied into d_req at submit; kZeroCopy the batch's kernels read the pinned slots over
    // PCIe; kCopyAtDispatch the batch's slots are copied into d_req on the batch's stream, a
    // few large copies per batch (slots go out in release order, so a batch's slots form a
    // few contiguous runs); kCopyBoth that, and the responses go to d_resp_arena and are
    // copied back the same way.  Otherwise the tail writes the responses over PCIe.
    std::atomic<int> upload_mode{0};
    uint8_t *h_resp = nullptr, *d_resp = nullptr;    // nslots x RL, mapped pinned
    uint8_t *d_resp_arena = nullptr;                 // nslots x RL in HBM (kCopyBoth)
    std::vector<hipEvent_t> slot_ev;
    std::deque<int> free_slots;  // handed out from the front, returned at the back
    std::deque<int> pending;
    std::vector<uint64_t> tag;
    std::vector<int64_t> t_submit;
    // one upload stream per database: with the contexts' streams it stays within the
    // device's hardware queues, so an upload never queues behind a batch's kernels
    hipStream_t upload = nullptr;
    std::mutex up_mu;  // serialises uploads on this database's stream
    std::vector<Ctx> ctx;
};

struct Server {
    std::vector<std::unique_ptr<Db>> dbs;
    int max_batch = 128;
    std::mutex mu;
    std::condition_variable cv_dispatch, cv_done, cv_complete;
    std::deque<alto_completion> completed;
    std::deque<std::pair<int, int>> inflight;  // (db, ctx) in launch order
    std::atomic<bool> stop{false};
    std::vector<std::thread> threads;
    std::string err;
    alto_server_stats stats{};
    int device = 0;
};

bool ok(hipError_t e, Server *s, const char *what) {
    if (e == hipSuccess) return true;
    s->err = std::string(what) + ": " + hipGetErrorString(e);
    return false;
}

void dispatcher(Server *s, int di) {
    (void)hipSetDevice(s->device);
    Db &d = *s->dbs[di];
    std::unique_lock<std::mutex> lk(s->mu);
    while (true) {
        int ci = -1;
        s->cv_dispatch.wait(lk, [&] {
            if (s->stop) return true;
            if (d.pending.empty()) return false;
            for (size_t i = 0; i < d.ctx.size(); i++)
                if (!d.ctx[i].busy) {
                    ci = (int)i;
                    return true;
                }
            return false;
        });
        if (s->stop) return;
        Ctx &c = d.ctx[ci];
        c.batch.clear();
        while (!d.pending.empty() && (int)c.batch.size() < s->max_batch) {
            c.batch.push_back(d.pending.front());
            d.pending.pop_front();
        }
        c.busy = true;
        c.failed = false;
        const int B = (int)c.batch.size();
        lk.unlock();
        const int mode = d.upload_mode;
        const bool zc = mode == kZeroCopy;
        for (int i = 0; i < B; i++) {
            c.h_idx[i] = c.batch[i];
            c.h_status[i] = 0;
            if (mode == kCopyAtSubmit) (void)hipStreamWaitEvent(c.stream, d.slot_ev[c.batch[i]], 0);
        }
        bool good = true;
        if (mode == kCopyAtDispatch || mode == kCopyBoth) {
            // one copy per run of consecutive slots
            c.runs.assign(c.batch.begin(), c.batch.end());
            std::sort(c.runs.begin(), c.runs.end());
            for (size_t i = 0; good && i < c.runs.size();) {
                size_t j = i + 1;
                while (j < c.runs.size() && c.runs[j] == c.runs[j - 1] + 1) j++;
                const size_t off = (size_t)c.runs[i] * d.S, bytes = (size_t)(c.runs[j - 1] - c.runs[i] + 1) * d.S;
                good = ok(hipMemcpyAsync(d.d_req + off, d.h_req + off, bytes, hipMemcpyHostToDevice, c.stream), s,
                          "batch upload");
                i = j;
            }
        }
        good = good && ok(hipMemcpyAsync(c.d_idx, c.h_idx, B * sizeof(int), hipMemcpyHostToDevice, c.stream), s, "index");
        uint8_t *resp = mode == kCopyBoth ? d.d_resp_arena : d.d_resp;
        const int rc = good ? alto_answer_gather(c.ctx, zc ? d.h_req_dev : d.d_req, d.S, c.d_idx, B, resp, d.RL,
                                                 c.d_idx, c.d_status, c.stream)
                            : -1;
        // kCopyBoth: the responses were written to HBM; one copy back per run of slots
        for (size_t i = 0; rc == 0 && mode == kCopyBoth && i < c.runs.size();) {
            size_t j = i + 1;
            while (j < c.runs.size() && c.runs[j] == c.runs[j - 1] + 1) j++;
            const size_t off = (size_t)c.runs[i] * d.RL, bytes = (size_t)(c.runs[j - 1] - c.runs[i] + 1) * d.RL;
            if (!ok(hipMemcpyAsync(d.h_resp + off, d.d_resp_arena + off, bytes, hipMemcpyDeviceToHost, c.stream), s,
                    "response download"))
                c.failed = true;  // applied at completion (the device still writes the statuses)
            i = j;
        }
        if (rc != 0) s->err = std::string("alto_answer_gather: ") + alto_last_error();
        (void)hipEventRecord(c.done, c.stream);
        lk.lock();
        if (rc != 0)
            for (int i = 0; i < B; i++) c.h_status[i] = -1;  // reported as an engine failure
        s->inflight.emplace_back(di, ci);
        s->stats.batches++;
        s->stats.answered_in_batches += B;
        s->stats.max_batch_seen = std::max<long long>(s->stats.max_batch_seen, B);
        s->cv_done.notify_one();
    }
}

void completer(Server *s) {
    (void)hipSetDevice(s->device);
    std::unique_lock<std::mutex> lk(s->mu);
    while (true) {
        s->cv_done.wait(lk, [&] { return s->stop || !s->inflight.empty(); });
        if (s->inflight.empty() && s->stop) return;
        auto [di, ci] = s->inflight.front();
        s->inflight.pop_front();
        Db &d = *s->dbs[di];
        Ctx &c = d.ctx[ci];
        lk.unlock();
        (void)hipEventSynchronize(c.done);
        const int64_t t = now_ns();
        lk.lock();
        const int B = (int)c.batch.size();
        for (int i = 0; i < B; i++) {
            const int slot = c.batch[i];
            alto_completion x{};
            x.tag = d.tag[slot];
            x.db = di;
            x.status = c.failed ? -1 : c.h_status[i];
            x.slot = x.status == 0 ? slot : -1;
            x.response = x.status == 0 ? d.h_resp + (size_t)slot * d.RL : nullptr;
            x.response_len = x.status == 0 ? d.RL : 0;
            x.t_submit_ns = d.t_submit[slot];
            x.t_done_ns = t;
            x.batch = B;
            if (x.status != 0) d.free_slots.push_back(slot);  // nothing to read back
            s->completed.push_back(x);
        }
        c.busy = false;
        s->cv_complete.notify_all();
        s->cv_dispatch.notify_all();
    }
}

}  // namespace

extern "C" {

void *alto_server_new(void *const *engines, int n_engines, int max_batch, int slots_per_db, int contexts_per_db) {
    if (!engines || n_engines <= 0 || max_batch <= 0 || slots_per_db <= 0 || contexts_per_db <= 0) return nullptr;
    auto *s = new Server;
    s->max_batch = max_batch;
    (void)hipGetDevice(&s->device);
    for (int i = 0; i < n_engines; i++) s->dbs.emplace_back(new Db);
    bool good = true;
    for (int i = 0; good && i < n_engines; i++) {
        Db &d = *s->dbs[i];
        d.engine = engines[i];
        d.L = alto_request_len(d.engine);
        d.RL = alto_response_len(d.engine);
        d.S = round_up(ALTO_FRAME_OFFSET + d.L, 256);
        d.nslots = slots_per_db;
        d.upload_mode = default_upload(d.L);
        good = ok(hipHostMalloc((void **)&d.h_req, (size_t)d.nslots * d.S, hipHostMallocMapped), s, "pinned requests") &&
               ok(hipHostGetDevicePointer((void **)&d.h_req_dev, d.h_req, 0), s, "request mapping") &&
               ok(hipMalloc(&d.d_req, (size_t)d.nslots * d.S), s, "request arena") &&
               ok(hipHostMalloc((void **)&d.h_resp, (size_t)d.nslots * d.RL, hipHostMallocMapped), s, "responses") &&
               ok(hipHostGetDevicePointer((void **)&d.d_resp, d.h_resp, 0), s, "response mapping") &&
               ok(hipMalloc(&d.d_resp_arena, (size_t)d.nslots * d.RL), s, "response arena");
        good = good && ok(hipStreamCreateWithFlags(&d.upload, hipStreamNonBlocking), s, "upload stream");
        d.slot_ev.resize(d.nslots);
        for (auto &ev : d.slot_ev) good = good && ok(hipEventCreateWithFlags(&ev, hipEventDisableTiming), s, "event");
        d.tag.resize(d.nslots);
        d.t_submit.resize(d.nslots);
        for (int k = 0; k < d.nslots; k++) d.free_slots.push_back(k);
        // memory admission: as many contexts (each reserved for max_batch queries) as fit in
        // 90% of the device memory free now, at least one, at most contexts_per_db
        for (int k = 0; good && k < contexts_per_db; k++) {
            size_t free_b = 0, total_b = 0;
            (void)hipMemGetInfo(&free_b, &total_b);
            Ctx c;
            c.ctx = alto_context_new(d.engine);
            if (!c.ctx) {
                good = k > 0;
                break;
            }
            if (alto_context_reserve(c.ctx, max_batch) != 0) {
                alto_context_free(c.ctx);
                good = k > 0;
                if (!good) s->err = "not enough device memory for one batch of max_batch queries";
                break;
            }
            size_t free_after = 0;
            (void)hipMemGetInfo(&free_after, &total_b);
            good = ok(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking), s, "stream") &&
                   ok(hipEventCreateWithFlags(&c.done, hipEventDisableTiming), s, "event") &&
                   ok(hipMalloc(&c.d_idx, max_batch * sizeof(int)), s, "index") &&
                   ok(hipHostMalloc((void **)&c.h_idx, max_batch * sizeof(int), hipHostMallocDefault), s, "index") &&
                   ok(hipHostMalloc((void **)&c.h_status, max_batch * sizeof(int), hipHostMallocMapped), s, "status") &&
                   ok(hipHostGetDevicePointer((void **)&c.d_status, c.h_status, 0), s, "status mapping");
            d.ctx.push_back(c);
            const size_t used = free_b > free_after ? free_b - free_after : 0;
            if (free_after < used + free_b / 10) break;
        }
    }
    if (!good) {
        std::string e = s->err;
        alto_server_free(s);
        fprintf(stderr, "alto_server_new: %s\n", e.c_str());
        return nullptr;
    }
    for (int i = 0; i < n_engines; i++) s->threads.emplace_back(dispatcher, s, i);
    s->threads.emplace_back(completer, s);
    return s;
}

int alto_server_contexts(void *p, int db) {
    auto *s = (Server *)p;
    if (!s || db < 0 || db >= (int)s->dbs.size()) return -1;
    return (int)s->dbs[db]->ctx.size();
}

int alto_server_submit(void *p, int db, const uint8_t *frame, size_t len, uint64_t tag) {
    auto *s = (Server *)p;
    if (!s || db < 0 || db >= (int)s->dbs.size()) return -1;
    Db &d = *s->dbs[db];
    const int64_t t0 = now_ns();
    int slot = -1;
    int reject = 0;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stats.submitted++;
        if (!frame || len != d.L) reject = 1;
        else if (d.free_slots.empty()) reject = 9;
        else {
            slot = d.free_slots.front();
            d.free_slots.pop_front();
        }
        if (reject) {
            alto_completion x{};
            x.tag = tag;
            x.db = db;
            x.status = reject;
            x.slot = -1;
            x.t_submit_ns = t0;
            x.t_done_ns = now_ns();
            s->completed.push_back(x);
            if (reject == 9) s->stats.queue_full++;
            s->cv_complete.notify_all();
            return reject;
        }
    }
    uint8_t *h = d.h_req + (size_t)slot * d.S;
    memcpy(h + ALTO_FRAME_OFFSET, frame, len);
    hipError_t ue = hipSuccess;
    if (d.upload_mode == kCopyAtSubmit) {
        std::lock_guard<std::mutex> ul(d.up_mu);
        ue = hipMemcpyAsync(d.d_req + (size_t)slot * d.S, h, ALTO_FRAME_OFFSET + len, hipMemcpyHostToDevice,
                            d.upload);
        if (ue == hipSuccess) ue = hipEventRecord(d.slot_ev[slot], d.upload);
    }
    if (ue != hipSuccess) {
        // the frame never reached the device: report an engine failure for it at once
        std::lock_guard<std::mutex> lk(s->mu);
        s->err = std::string("request upload: ") + hipGetErrorString(ue);
        alto_completion x{};
        x.tag = tag;
        x.db = db;
        x.status = -1;
        x.slot = -1;
        x.t_submit_ns = t0;
        x.t_done_ns = now_ns();
        s->completed.push_back(x);
        d.free_slots.push_back(slot);
        s->cv_complete.notify_all();
        return -1;
    }
    bool wake = false;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        d.tag[slot] = tag;
        d.t_submit[slot] = t0;
        d.pending.push_back(slot);
        // the dispatcher waits for a free context; when all are busy, the completion of one
        // wakes it and it then takes everything pending
        for (const Ctx &c : d.ctx) wake = wake || !c.busy;
    }
    if (wake) s->cv_dispatch.notify_all();
    return 0;
}

int alto_server_poll(void *p, alto_completion *out, int max, int timeout_ms) {
    auto *s = (Server *)p;
    if (!s || !out || max <= 0) return -1;
    std::unique_lock<std::mutex> lk(s->mu);
    if (s->completed.empty() && timeout_ms > 0)
        s->cv_complete.wait_for(lk, std::chrono::milliseconds(timeout_ms), [&] { return !s->completed.empty(); });
    int k = 0;
    while (k < max && !s->completed.empty()) {
        out[k++] = s->completed.front();
        s->completed.pop_front();
    }
    return k;
}

int alto_server_release(void *p, int db, int slot) {
    auto *s = (Server *)p;
    if (!s || db < 0 || db >= (int)s->dbs.size() || slot < 0 || slot >= s->dbs[db]->nslots) return -1;
    std::lock_guard<std::mutex> lk(s->mu);
    s->dbs[db]->free_slots.push_back(slot);  // no dispatcher waits on slots
    return 0;
}

int alto_server_get_stats(void *p, alto_server_stats *out) {
    auto *s = (Server *)p;
    if (!s || !out) return -1;
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->stats;
    return 0;
}

const char *alto_server_error(void *p) { return p ? ((Server *)p)->err.c_str() : ""; }

void alto_server_free(void *p) {
    auto *s = (Server *)p;
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_dispatch.notify_all();
    s->cv_done.notify_all();
    s->cv_complete.notify_all();
    for (auto &t : s->threads) t.join();
    for (auto &dp : s->dbs) {
        Db &d = *dp;
        for (auto &c : d.ctx) {
            if (c.stream) (void)hipStreamSynchronize(c.stream);
            if (c.ctx) alto_context_free(c.ctx);
            if (c.stream) (void)hipStreamDestroy(c.stream);
            if (c.done) (void)hipEventDestroy(c.done);
            if (c.d_idx) (void)hipFree(c.d_idx);
            if (c.h_idx) (void)hipHostFree(c.h_idx);
            if (c.h_status) (void)hipHostFree(c.h_status);
        }
        if (d.upload) (void)hipStreamSynchronize(d.upload);
        for (auto &ev : d.slot_ev)
            if (ev) (void)hipEventDestroy(ev);
        if (d.upload) (void)hipStreamDestroy(d.upload);
        if (d.h_req) (void)hipHostFree(d.h_req);
        if (d.d_req) (void)hipFree(d.d_req);
        if (d.h_resp) (void)hipHostFree(d.h_resp);
        if (d.d_resp_arena) (void)hipFree(d.d_resp_arena);
    }
    delete s;
}

int alto_server_set_upload(void *p, int mode) {
    auto *s = (Server *)p;
    if (!s || mode < -1 || mode > 3) return -1;
    std::lock_guard<std::mutex> lk(s->mu);
    for (auto &dp : s->dbs) {
        if (!dp->pending.empty()) return -1;
        for (auto &c : dp->ctx)
            if (c.busy) return -1;
        if (dp->upload) (void)hipStreamSynchronize(dp->upload);
    }
    for (auto &dp : s->dbs) dp->upload_mode = mode < 0 ? default_upload(dp->L) : mode;
    return 0;
}

int alto_server_upload_mode(void *p, int db) {
    auto *s = (Server *)p;
    if (!s || db < 0 || db >= (int)s->dbs.size()) return -1;
    return s->dbs[db]->upload_mode;
}

int alto_server_set_graphs(void *p, int mode) {
    auto *s = (Server *)p;
    if (!s || mode < -1 || mode > 1) return -1;
    std::lock_guard<std::mutex> lk(s->mu);
    for (auto &dp : s->dbs)
        for (auto &c : dp->ctx) {
            if (c.busy) (void)hipStreamSynchronize(c.stream);
            if (alto_context_set_graphs(c.ctx, mode) != 0) return -1;
        }
    return 0;
}

int alto_server_replay(void *p, int db, const uint8_t *const *frames, const size_t *lens, int nframes,
                       long long total, int threads, int window, alto_replay_result *out) {
    auto *s = (Server *)p;
    if (!s || db < 0 || db >= (int)s->dbs.size() || !frames || !lens || nframes <= 0 || total <= 0 ||
        threads <= 0 || window <= 0 || !out)
        return -1;
    alto_server_stats st0{}, st1{};
    (void)alto_server_get_stats(s, &st0);
    std::mutex wmu;
    std::condition_variable wcv;
    long long outstanding = 0;
    std::vector<double> lat_ms((size_t)total, -1.0);
    long long got = 0, answered = 0, rejected = 0;
    const int64_t t0 = now_ns();
    std::thread collector([&] {
        std::vector<alto_completion> buf(512);
        while (got < total) {
            const int k = alto_server_poll(s, buf.data(), (int)buf.size(), 20);
            for (int i = 0; i < k; i++) {
                const alto_completion &c = buf[i];
                if (c.status == 0) {
                    answered++;
                    (void)alto_server_release(s, c.db, c.slot);
                } else {
                    rejected++;
                }
                if (c.tag < (uint64_t)total) lat_ms[c.tag] = (c.t_done_ns - c.t_submit_ns) / 1e6;
                got++;
            }
            if (k > 0) {
                std::lock_guard<std::mutex> lk(wmu);
                outstanding -= k;
                wcv.notify_all();
            }
        }
    });
    std::vector<std::thread> subs;
    for (int t = 0; t < threads; t++)
        subs.emplace_back([&, t] {
            for (long long i = t; i < total; i += threads) {
                {
                    std::unique_lock<std::mutex> lk(wmu);
                    wcv.wait(lk, [&] { return outstanding < window; });
                    outstanding++;
                }
                const int f = (int)(i % nframes);
                (void)alto_server_submit(s, db, frames[f], lens[f], (uint64_t)i);
            }
        });
    for (auto &t : subs) t.join();
    collector.join();
    const double wall = (now_ns() - t0) / 1e9;
    (void)alto_server_get_stats(s, &st1);
    std::vector<double> v;
    for (double x : lat_ms)
        if (x >= 0) v.push_back(x);
    std::sort(v.begin(), v.end());
    auto pct = [&](double q) { return v.empty() ? 0.0 : v[std::min(v.size() - 1, (size_t)(q / 100.0 * v.size()))]; };
    out->answered = answered;
    out->rejected = rejected;
    out->batches = st1.batches - st0.batches;
    out->wall_s = wall;
    out->qps = wall > 0 ? answered / wall : 0.0;
    out->mean_batch = out->batches ? (double)(st1.answered_in_batches - st0.answered_in_batches) / out->batches : 0.0;
    out->p50_ms = pct(50);
    out->p90_ms = pct(90);
    out->p99_ms = pct(99);
    out->max_ms = v.empty() ? 0.0 : v.back();
    return 0;
}

}  // extern "C"
