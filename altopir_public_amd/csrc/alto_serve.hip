// alto_serve.hip -- device-resident request loop with continuous batching.
//
// The reference serves requests with reader/writer threads around one engine thread that
// drains the whole queue into one synchronous alto_answer call per batch
// (src/bin/altopir.rs:372-496).  Here every resident database keeps a pool of request
// slots (pinned host staging + a device arena) and several engine contexts:
//   submit   copies the frame into a free pinned slot and hands it to the database's
//            uploader thread (by default; other upload modes hand it straight to the
//            dispatcher and copy a batch's frames when it is dispatched, or let the kernels
//            read the pinned slots);
//   upload   the uploader copies every frame handed to it since its last round into HBM,
//            one DMA copy per run of consecutive slots, waits for them on the host and
//            passes the frames on;
//   dispatch whenever a context is free, takes every waiting request of that database (up
//            to max_batch, never waiting to fill a batch) and answers them with
//            alto_answer_gather on the context's stream; responses are written by the tail
//            kernel straight into mapped pinned host memory (or into HBM and copied back
//            per batch);
//   complete a completion thread per context waits for that context's batch, stamps the
//            completion time and hands the tickets to poll(), so batches complete in the
//            order they end, whatever database they belong to.
// Batches of one database overlap on the device as far as its contexts allow.  A database
// starts with one context; the dispatcher admits another (up to contexts_per_db) when
// requests wait, every context is busy and the device has the memory for one more context's
// workspace with as much again to spare (memory admission at run time).  Frames
// whose length is wrong get status 1 at once, and status 9 when every slot is taken
// (src/bin/altopir.rs:417-436); the device validates the rest (statuses 2 and 3).
#include <hip/hip_runtime.h>
#include <hsa/hsa.h>
#include <hsa/hsa_ext_amd.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdio>
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

// Writes every page of a fresh pinned allocation once, with a few threads.  Its pages are
// mapped into the process on the first CPU access, one fault per page: without this, the
// first frame written into each slot takes those faults inside submit (about 2,000 per
// 8.45 MB frame), which slows the first pass over the slot pool several times over.
void prefault(uint8_t *p, size_t bytes) {
    constexpr size_t kChunk = 64 << 20;
    const size_t chunks = (bytes + kChunk - 1) / kChunk;
    const unsigned nt = (unsigned)std::min<size_t>(chunks, 8);
    std::atomic<size_t> next{0};
    std::vector<std::thread> ts;
    for (unsigned t = 0; t < nt; t++)
        ts.emplace_back([&] {
            for (size_t c; (c = next.fetch_add(1)) < chunks;)
                memset(p + c * kChunk, 0, std::min(kChunk, bytes - c * kChunk));
        });
    for (auto &t : ts) t.join();
}

// Upload modes (alto_server_set_upload).
constexpr int kCopyOnArrival = 0, kZeroCopy = 1, kCopyAtDispatch = 2, kCopyBoth = 3;
// The default: frames copied by the uploader as they arrive, so that their transfer overlaps
// the batches in flight, and responses written by the tail into the pinned slots.  Served at
// saturation (batch 128, 2 contexts) this beats copying each batch's frames at dispatch with
// or without copying its responses back: 92k against 81-88k QPS (1 GiB, p99 8.6 against
// 8.7-14.5 ms), 24.4k against 23.7k (8 GiB), 3.6k against 3.5k (64 GiB); copying the
// responses back as well (uploader frames) gave 79-87k, 22.5-23.4k and 3.55-3.6k.
int default_upload(size_t) { return kCopyOnArrival; }

struct Server;

constexpr int kNoGraphChange = 2;  // Ctx::graph_req: no graph-mode change pending

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
    bool launched = false;    // the batch is on the device; the context's completer waits for it
    bool failed = false;      // a copy of the batch failed: every request reports -1
    // graph mode asked for by alto_server_set_graphs, applied by the dispatcher before the
    // context's next batch (only that thread drives the context): kNoGraphChange = none
    int graph_req = kNoGraphChange;
};

struct Db {
    void *engine = nullptr;
    size_t L = 0, RL = 0, S = 0;
    int nslots = 0;
    uint8_t *h_req = nullptr, *d_req = nullptr;      // nslots x S (pinned, mapped / device arena)
    uint8_t *h_req_dev = nullptr;                    // device view of h_req
    // how frames reach the device (alto_server_set_upload): kCopyOnArrival the uploader
    // thread copies frames into d_req as they arrive; kZeroCopy the batch's kernels read the pinned slots over
    // PCIe; kCopyAtDispatch the batch's slots are copied into d_req on the batch's stream, a
    // few large copies per batch (slots go out in release order, so a batch's slots form a
    // few contiguous runs); kCopyBoth that, and the responses go to d_resp_arena and are
    // copied back the same way.  Otherwise the tail writes the responses over PCIe.
    std::atomic<int> upload_mode{0};
    uint8_t *h_resp = nullptr, *d_resp = nullptr;    // nslots x RL, mapped pinned
    uint8_t *d_resp_arena = nullptr;                 // nslots x RL in HBM (kCopyBoth)
    std::deque<int> free_slots;  // handed out from the front, returned at the back
    std::vector<uint8_t> leased; // per slot: its response is handed out and not yet released
    std::deque<int> to_upload;   // submitted, not yet handed to the uploader
    int uploading = 0;           // frames of the uploader's current round
    std::condition_variable cv_upload;
    std::deque<int> pending;     // in HBM (or readable by the kernels), waiting for a batch
    std::vector<uint64_t> tag;
    std::vector<int64_t> t_submit;
    // the uploader's copies go straight to the device's DMA engines, one of these signals
    // each, and are waited for on them: nothing of them enters a stream, whose completion
    // markers can sit in a hardware queue shared with a context's stream behind that batch's
    // kernels (a waited-for upload stream returned only when the batch in flight ended)
    std::vector<hsa_signal_t> up_sigs;
    // contexts (a deque: the completer of each keeps a reference to its element while more
    // are admitted); at most max_ctx, each needing about ctx_bytes of device memory
    std::deque<Ctx> ctx;
    int max_ctx = 1;
    size_t ctx_bytes = 0;
    bool admitting = false;            // the dispatcher is creating a context
    long long refused_at_batch = -1;   // admission refused when this many batches had run
};

struct Server {
    std::vector<std::unique_ptr<Db>> dbs;
    int max_batch = 128;
    std::mutex mu;
    std::condition_variable cv_dispatch, cv_done, cv_complete;
    std::deque<alto_completion> completed;
    std::atomic<bool> stop{false};
    std::vector<std::thread> threads;
    std::string err;
    alto_server_stats stats{};
    int device = 0;
    int max_batch_ctx = 0;  // queries each context is reserved for
    int graph_mode = kNoGraphChange;  // the last alto_server_set_graphs mode, for new contexts
    hsa_agent_t gpu_agent{}, cpu_agent{};  // the device and a host agent, for the DMA copies
};

struct AgentSearch {
    uint32_t bdf = 0, domain = 0;
    hsa_agent_t gpu{}, cpu{};
    bool gpu_found = false, cpu_found = false;
};

hsa_status_t match_agent(hsa_agent_t a, void *p) {
    auto *f = (AgentSearch *)p;
    hsa_device_type_t t;
    if (hsa_agent_get_info(a, HSA_AGENT_INFO_DEVICE, &t) != HSA_STATUS_SUCCESS) return HSA_STATUS_SUCCESS;
    if (t == HSA_DEVICE_TYPE_CPU && !f->cpu_found) {
        f->cpu = a;
        f->cpu_found = true;
    } else if (t == HSA_DEVICE_TYPE_GPU) {
        uint32_t bdf = 0, dom = 0;
        if (hsa_agent_get_info(a, (hsa_agent_info_t)HSA_AMD_AGENT_INFO_BDFID, &bdf) == HSA_STATUS_SUCCESS &&
            hsa_agent_get_info(a, (hsa_agent_info_t)HSA_AMD_AGENT_INFO_DOMAIN, &dom) == HSA_STATUS_SUCCESS &&
            bdf == f->bdf && dom == f->domain) {
            f->gpu = a;
            f->gpu_found = true;
        }
    }
    return HSA_STATUS_SUCCESS;
}

// The HSA agents of the server's device (matched by PCI location) and of the host, for the
// uploader's DMA copies.
bool find_agents(Server *s) {
    // one reference on the HSA runtime for the life of the process, as the HIP runtime holds
    // its own: releasing it with hsa_shut_down when a server ends would look like the end of
    // the process to tools that intercept that call (the profiler's tracing did)
    static const hsa_status_t init = hsa_init();
    if (init != HSA_STATUS_SUCCESS) {
        s->err = "hsa_init failed";
        return false;
    }
    char bus[64] = {0};
    unsigned dom = 0, b = 0, dv = 0, fn = 0;
    if (hipDeviceGetPCIBusId(bus, sizeof bus, s->device) != hipSuccess ||
        sscanf(bus, "%x:%x:%x.%x", &dom, &b, &dv, &fn) != 4) {
        s->err = "no PCI location for the device";
        return false;
    }
    AgentSearch f;
    f.bdf = (b << 8) | (dv << 3) | fn;
    f.domain = dom;
    (void)hsa_iterate_agents(match_agent, &f);
    if (!f.gpu_found || !f.cpu_found) {
        s->err = std::string("no HSA agent for device ") + bus;
        return false;
    }
    s->gpu_agent = f.gpu;
    s->cpu_agent = f.cpu;
    return true;
}

bool ok(hipError_t e, Server *s, const char *what) {
    if (e == hipSuccess) return true;
    s->err = std::string(what) + ": " + hipGetErrorString(e);
    return false;
}

// Sorts `runs` (slots) and calls f(first slot, slot count) for each run of consecutive
// slots; stops at the first call that returns false.
template <class F>
bool for_runs(std::vector<int> &runs, F f) {
    std::sort(runs.begin(), runs.end());
    for (size_t i = 0; i < runs.size();) {
        size_t j = i + 1;
        while (j < runs.size() && runs[j] == runs[j - 1] + 1) j++;
        if (!f((size_t)runs[i], j - i)) return false;
        i = j;
    }
    return true;
}

// One host-to-device copy per run of consecutive slots, on `st`.
hipError_t copy_runs(const Db &d, std::vector<int> &runs, hipStream_t st) {
    hipError_t e = hipSuccess;
    for_runs(runs, [&](size_t first, size_t n) {
        e = hipMemcpyAsync(d.d_req + first * d.S, d.h_req + first * d.S, n * d.S, hipMemcpyHostToDevice, st);
        return e == hipSuccess;
    });
    return e;
}

// Signals per database for the uploader's copies: runs beyond this many in one round are
// issued after the earlier ones have ended.
constexpr size_t kUploadSignals = 16;

// One DMA copy per run of consecutive slots, each completing its own signal (set to 1, counted
// down by the copy); returns once every issued copy has ended, true if all were issued and
// none failed.
bool dma_runs(Server *s, Db &d, std::vector<int> &runs) {
    bool good = true;
    size_t k = 0;  // copies in flight: up_sigs[0, k)
    auto drain = [&] {
        for (size_t i = 0; i < k; i++) {
            hsa_signal_value_t v;
            while ((v = hsa_signal_wait_scacquire(d.up_sigs[i], HSA_SIGNAL_CONDITION_LT, 1, UINT64_MAX,
                                                  HSA_WAIT_STATE_BLOCKED)) > 0) {
            }
            good = good && v == 0;  // a failed copy leaves its signal negative
        }
        k = 0;
    };
    for_runs(runs, [&](size_t first, size_t n) {
        if (k == d.up_sigs.size()) drain();
        hsa_signal_store_screlease(d.up_sigs[k], 1);
        if (hsa_amd_memory_async_copy(d.d_req + first * d.S, s->gpu_agent, d.h_req + first * d.S, s->cpu_agent,
                                      n * d.S, 0, nullptr, d.up_sigs[k]) != HSA_STATUS_SUCCESS) {
            good = false;
            return false;
        }
        k++;
        return true;
    });
    drain();
    return good;
}

// One context of database d, reserved for max_batch queries: its engine context, stream,
// completion event and index/status buffers.  On failure frees what it made.
bool make_context(Server *s, Db &d, Ctx &c) {
    const int mb = s->max_batch_ctx;
    c = Ctx{};
    c.ctx = alto_context_new(d.engine);
    bool good = c.ctx && alto_context_reserve(c.ctx, (size_t)mb) == 0;
    good = good && hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking) == hipSuccess &&
           hipEventCreateWithFlags(&c.done, hipEventDisableTiming) == hipSuccess &&
           hipMalloc(&c.d_idx, mb * sizeof(int)) == hipSuccess &&
           hipHostMalloc((void **)&c.h_idx, mb * sizeof(int), hipHostMallocDefault) == hipSuccess &&
           hipHostMalloc((void **)&c.h_status, mb * sizeof(int), hipHostMallocMapped) == hipSuccess &&
           hipHostGetDevicePointer((void **)&c.d_status, c.h_status, 0) == hipSuccess;
    if (!good) {
        if (c.ctx) alto_context_free(c.ctx);
        if (c.stream) (void)hipStreamDestroy(c.stream);
        if (c.done) (void)hipEventDestroy(c.done);
        if (c.d_idx) (void)hipFree(c.d_idx);
        if (c.h_idx) (void)hipHostFree(c.h_idx);
        if (c.h_status) (void)hipHostFree(c.h_status);
        c = Ctx{};
    }
    return good;
}

void completer(Server *s, int di, int ci);

void dispatcher(Server *s, int di) {
    (void)hipSetDevice(s->device);
    Db &d = *s->dbs[di];
    std::unique_lock<std::mutex> lk(s->mu);
    while (true) {
        int ci = -1;
        bool admit = false;
        s->cv_dispatch.wait(lk, [&] {
            if (s->stop) return true;
            if (d.pending.empty()) return false;
            for (size_t i = 0; i < d.ctx.size(); i++)
                if (!d.ctx[i].busy) {
                    ci = (int)i;
                    return true;
                }
            // every context busy: try one more, unless the last attempt was refused and no
            // batch has completed since
            if ((int)d.ctx.size() < d.max_ctx && !d.admitting && d.refused_at_batch != s->stats.batches) {
                admit = true;
                return true;
            }
            return false;
        });
        if (s->stop) return;
        if (admit) {
            // memory admission: a context only if the device then still has its size free
            d.admitting = true;
            const size_t need = d.ctx_bytes;
            lk.unlock();
            size_t free_b = 0, total_b = 0;
            (void)hipMemGetInfo(&free_b, &total_b);
            Ctx c;
            const bool made = free_b >= 2 * need && make_context(s, d, c);
            lk.lock();
            d.admitting = false;
            if (!made) {
                d.refused_at_batch = s->stats.batches;
                s->stats.contexts_refused++;
                continue;
            }
            c.graph_req = s->graph_mode;
            d.ctx.push_back(c);
            s->stats.contexts_admitted++;
            if (!s->stop) s->threads.emplace_back(completer, s, di, (int)d.ctx.size() - 1);
            continue;  // the new context is free: it takes the waiting requests next
        }
        Ctx &c = d.ctx[ci];
        c.batch.clear();
        while (!d.pending.empty() && (int)c.batch.size() < s->max_batch) {
            c.batch.push_back(d.pending.front());
            d.pending.pop_front();
        }
        c.busy = true;
        c.failed = false;
        const int B = (int)c.batch.size();
        const int graph_req = c.graph_req;
        c.graph_req = kNoGraphChange;
        lk.unlock();
        if (graph_req != kNoGraphChange) (void)alto_context_set_graphs(c.ctx, graph_req);
        const int mode = d.upload_mode;
        const bool zc = mode == kZeroCopy;
        for (int i = 0; i < B; i++) {
            c.h_idx[i] = c.batch[i];
            c.h_status[i] = 0;
        }
        bool good = true;
        if (mode == kCopyAtDispatch || mode == kCopyBoth) {
            // one copy per run of consecutive slots (c.runs: the batch's slots, sorted)
            c.runs.assign(c.batch.begin(), c.batch.end());
            good = ok(copy_runs(d, c.runs, c.stream), s, "batch upload");
        }
        good = good && ok(hipMemcpyAsync(c.d_idx, c.h_idx, B * sizeof(int), hipMemcpyHostToDevice, c.stream), s, "index");
        const bool back = mode == kCopyBoth;
        uint8_t *resp = back ? d.d_resp_arena : d.d_resp;
        const int rc = good ? alto_answer_gather(c.ctx, zc ? d.h_req_dev : d.d_req, d.S, c.d_idx, B, resp, d.RL,
                                                 c.d_idx, c.d_status, c.stream)
                            : -1;
        // kCopyBoth: the responses were written to HBM; one copy back per run of slots
        // (c.runs: the batch's slots, sorted by the upload above)
        for (size_t i = 0; rc == 0 && back && i < c.runs.size();) {
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
        c.launched = true;
        s->stats.batches++;
        s->stats.answered_in_batches += B;
        s->stats.max_batch_seen = std::max<long long>(s->stats.max_batch_seen, B);
        s->cv_done.notify_all();
    }
}

// Frames copied on arrival: moves the frames handed over since the last round into HBM (at most
// max_batch per round, so that a batch can start on the first ones while later ones are
// still on their way) and then passes them, in submission order, to the dispatcher.
void uploader(Server *s, int di) {
    (void)hipSetDevice(s->device);
    Db &d = *s->dbs[di];
    std::vector<int> group, runs;
    std::unique_lock<std::mutex> lk(s->mu);
    while (true) {
        d.cv_upload.wait(lk, [&] { return s->stop || !d.to_upload.empty(); });
        if (s->stop) return;
        group.clear();
        while (!d.to_upload.empty() && (int)group.size() < s->max_batch) {
            group.push_back(d.to_upload.front());
            d.to_upload.pop_front();
        }
        d.uploading = (int)group.size();
        lk.unlock();
        runs.assign(group.begin(), group.end());
        const bool good = dma_runs(s, d, runs);
        lk.lock();
        d.uploading = 0;
        if (good) {
            for (int slot : group) d.pending.push_back(slot);
            bool wake = false;
            for (const Ctx &c : d.ctx) wake = wake || !c.busy;
            if (wake) s->cv_dispatch.notify_all();
        } else {
            // these frames never reached the device: an engine failure for each
            s->err = "request upload: a DMA copy failed";
            const int64_t t = now_ns();
            for (int slot : group) {
                alto_completion x{};
                x.tag = d.tag[slot];
                x.db = di;
                x.status = -1;
                x.slot = -1;
                x.t_submit_ns = d.t_submit[slot];
                x.t_done_ns = t;
                s->completed.push_back(x);
                d.free_slots.push_back(slot);
            }
            s->cv_complete.notify_all();
        }
    }
}

// One completer per context: it waits for that context's batch alone, so a batch of a
// small database is completed as soon as it ends, not behind an earlier batch of a larger
// one on another context.
void completer(Server *s, int di, int ci) {
    (void)hipSetDevice(s->device);
    Db &d = *s->dbs[di];
    Ctx &c = d.ctx[ci];
    std::unique_lock<std::mutex> lk(s->mu);
    while (true) {
        s->cv_done.wait(lk, [&] { return s->stop || c.launched; });
        if (!c.launched && s->stop) return;
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
            else d.leased[slot] = 1;                          // until alto_server_release
            s->completed.push_back(x);
        }
        c.launched = false;
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
    s->max_batch_ctx = max_batch;
    (void)hipGetDevice(&s->device);
    for (int i = 0; i < n_engines; i++) s->dbs.emplace_back(new Db);
    bool good = find_agents(s);
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
        if (good) {
            prefault(d.h_req, (size_t)d.nslots * d.S);
            prefault(d.h_resp, (size_t)d.nslots * d.RL);
        }
        for (size_t k = 0; good && k < kUploadSignals; k++) {
            hsa_signal_t sig;
            if (hsa_signal_create(0, 0, nullptr, &sig) == HSA_STATUS_SUCCESS) d.up_sigs.push_back(sig);
            else {
                s->err = "upload signal";
                good = false;
            }
        }
        d.tag.resize(d.nslots);
        d.leased.assign(d.nslots, 0);
        d.t_submit.resize(d.nslots);
        for (int k = 0; k < d.nslots; k++) d.free_slots.push_back(k);
        // one context to start with (more are admitted at run time, up to contexts_per_db,
        // each by its device footprint)
        d.max_ctx = contexts_per_db;
        d.ctx_bytes = alto_context_bytes(d.engine, (size_t)max_batch) + 2 * (size_t)max_batch * sizeof(int);
        if (good) {
            Ctx c;
            good = make_context(s, d, c);
            if (!good) s->err = "not enough device memory for one batch of max_batch queries";
            else d.ctx.push_back(c);
        }
    }
    if (!good) {
        std::string e = s->err;
        alto_server_free(s);
        fprintf(stderr, "alto_server_new: %s\n", e.c_str());
        return nullptr;
    }
    {
        std::lock_guard<std::mutex> lk(s->mu);  // dispatchers add completers under the lock
        for (int i = 0; i < n_engines; i++) {
            s->threads.emplace_back(uploader, s, i);
            for (int k = 0; k < (int)s->dbs[i]->ctx.size(); k++) s->threads.emplace_back(completer, s, i, k);
            s->threads.emplace_back(dispatcher, s, i);
        }
    }
    return s;
}

int alto_server_contexts(void *p, int db) {
    auto *s = (Server *)p;
    if (!s || db < 0 || db >= (int)s->dbs.size()) return -1;
    std::lock_guard<std::mutex> lk(s->mu);
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
    memcpy(d.h_req + (size_t)slot * d.S + ALTO_FRAME_OFFSET, frame, len);
    bool wake = false, upload = false;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        d.tag[slot] = tag;
        d.t_submit[slot] = t0;
        if (d.upload_mode == kCopyOnArrival) {
            d.to_upload.push_back(slot);
            upload = true;
        } else {
            d.pending.push_back(slot);
            // the dispatcher waits for a free context; when all are busy, the completion of
            // one wakes it and it then takes everything pending
            for (const Ctx &c : d.ctx) wake = wake || !c.busy;
        }
    }
    if (upload) d.cv_upload.notify_one();
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
    Db &d = *s->dbs[db];
    if (!d.leased[slot]) return -1;  // not handed out with a response, or released already
    d.leased[slot] = 0;
    d.free_slots.push_back(slot);  // no dispatcher waits on slots
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
    for (auto &dp : s->dbs) dp->cv_upload.notify_all();
    // no thread is added once stop is set (the dispatchers add completers under the lock
    // only while it is clear), so the list is fixed while it is joined
    for (size_t i = 0; i < s->threads.size(); i++) s->threads[i].join();
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
        for (hsa_signal_t sig : d.up_sigs) (void)hsa_signal_destroy(sig);
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
        if (!dp->pending.empty() || !dp->to_upload.empty() || dp->uploading) return -1;
        for (auto &c : dp->ctx)
            if (c.busy) return -1;
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
    // recorded here and applied by each context's dispatcher before its next batch: a busy
    // context may be inside alto_answer_gather (capturing or launching a graph) right now
    std::lock_guard<std::mutex> lk(s->mu);
    s->graph_mode = mode;  // for contexts admitted later
    for (auto &dp : s->dbs)
        for (auto &c : dp->ctx) c.graph_req = mode;
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
