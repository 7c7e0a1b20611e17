// TCP front end of the serving loop: the reference server's wire protocol
// (src/bin/altopir.rs:282-308 frames, :403-442 connection readers, :383-400 writers) in front
// of alto_server_* instead of the reference's engine thread.
//
//   request  [u32 len][u64 id][len bytes]            little-endian
//   reply    [u32 len][u64 id][i32 status][len bytes] status 0: the response; else len = 0
//   stop     len = 0, id = 2^64 - 1
//
// One listening socket per database.  Each accepted connection gets a reader thread (frames
// are read whole and handed to alto_server_submit, which copies them into a pinned slot; a
// frame longer than any valid request is discarded without buffering and answered with
// status 1, as the reference does) and a writer thread (replies in completion order, written
// straight from the leased pinned response slot, which is released once written).  One
// completion thread polls the server and routes each completion to its connection's writer.
// A full slot pool answers status 9 at once (the reference's busy status for a full queue),
// a frame of the wrong length status 1, and the device's validation the rest.
#include <arpa/inet.h>
#include <netinet/in.h>
#include <netinet/tcp.h>
#include <sys/socket.h>
#include <sys/uio.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cerrno>
#include <chrono>
#include <climits>
#include <cmath>
#include <condition_variable>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/altopir_engine.h"

namespace {

constexpr int kStatusBadLength = 1;  // src/bin/altopir.rs:53
constexpr uint64_t kStopId = ~0ull;

struct Reply {
    uint64_t id;
    int status;
    int db, slot;  // leased response slot (status 0), released after the write
    const uint8_t *body;
    size_t len;
};

// A connection's socket is closed by its writer once the reader has stopped (end of stream,
// a stop frame or an error) and every reply owed has been written, or by the teardown.
struct Conn {
    int fd = -1;  // guarded by mu once the threads run
    int db = 0;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<Reply> q;
    bool closing = false;      // the front end is being freed
    bool reader_done = false;
    std::atomic<long long> owed{0};   // replies not yet written (frames read, not answered)
    std::atomic<bool> broken{false};  // a write failed: replies are dropped (slots released)
    std::thread reader, writer;
};

struct Net {
    void *srv = nullptr;
    std::vector<size_t> req_len;
    std::vector<int> lfd, port;
    std::vector<std::thread> acceptors;
    std::thread completer;
    std::atomic<bool> stopping{false};

    std::mutex mu;  // conns, tags
    std::vector<std::shared_ptr<Conn>> conns;
    std::unordered_map<uint64_t, std::pair<std::shared_ptr<Conn>, uint64_t>> tags;  // tag -> (conn, id)
    uint64_t next_tag = 1;

    std::mutex smu;
    std::condition_variable scv;
    bool stop_frame = false;

    std::atomic<long long> connections{0}, frames{0}, answered{0}, rejected{0}, busy{0}, discarded{0},
        stop_frames{0}, bytes_in{0}, bytes_out{0};
};

bool read_full(int fd, void *p, size_t n) {
    auto *b = (uint8_t *)p;
    while (n) {
        const ssize_t r = recv(fd, b, n, 0);
        if (r > 0) {
            b += r;
            n -= (size_t)r;
        } else if (r < 0 && errno == EINTR) {
            continue;
        } else {
            return false;
        }
    }
    return true;
}

bool write_all(int fd, iovec *iov, int cnt) {
    while (cnt) {
        msghdr m{};
        m.msg_iov = iov;
        m.msg_iovlen = (size_t)cnt;
        const ssize_t w = sendmsg(fd, &m, MSG_NOSIGNAL);
        if (w < 0) {
            if (errno == EINTR) continue;
            return false;
        }
        size_t left = (size_t)w;
        while (cnt && left >= iov->iov_len) {
            left -= iov->iov_len;
            ++iov;
            --cnt;
        }
        if (cnt) {
            iov->iov_base = (uint8_t *)iov->iov_base + left;
            iov->iov_len -= left;
        }
    }
    return true;
}

void enqueue(Conn &c, const Reply &r) {
    {
        std::lock_guard<std::mutex> lk(c.mu);
        c.q.push_back(r);
    }
    c.cv.notify_one();
}

void writer_loop(Net *n, Conn *c) {
    for (;;) {
        Reply r;
        {
            std::unique_lock<std::mutex> lk(c->mu);
            c->cv.wait(lk, [&] { return !c->q.empty() || c->closing || (c->reader_done && c->owed == 0); });
            if (c->q.empty()) {
                if (!c->closing && c->fd >= 0) {  // reader stopped, nothing owed: close
                    shutdown(c->fd, SHUT_RDWR);
                    close(c->fd);
                    c->fd = -1;
                }
                return;
            }
            r = c->q.front();
            c->q.pop_front();
        }
        if (!c->broken.load()) {
            uint8_t h[16];
            const uint32_t len = r.status == 0 ? (uint32_t)r.len : 0u;
            memcpy(h, &len, 4);
            memcpy(h + 4, &r.id, 8);
            memcpy(h + 12, &r.status, 4);
            iovec iov[2] = {{h, 16}, {(void *)r.body, len}};
            if (write_all(c->fd, iov, len ? 2 : 1)) {
                n->bytes_out += 16 + len;
                if (r.status == 0) n->answered++;
                else n->rejected++;
            } else {
                c->broken = true;
                shutdown(c->fd, SHUT_RDWR);
            }
        }
        if (r.status == 0) alto_server_release(n->srv, r.db, r.slot);
        c->owed--;
    }
}

void reader_loop(Net *n, std::shared_ptr<Conn> cp) {
    Conn &c = *cp;
    const size_t ql = n->req_len[c.db];
    struct Done {  // on any exit: the writer may close once the replies owed are written
        Conn &c;
        ~Done() {
            {
                std::lock_guard<std::mutex> lk(c.mu);
                c.reader_done = true;
            }
            c.cv.notify_all();
        }
    } done{c};
    std::vector<uint8_t> buf(ql ? ql : 1);
    std::vector<uint8_t> sink(64 << 10);
    for (;;) {
        uint8_t h[12];
        if (!read_full(c.fd, h, 12)) break;
        uint32_t len;
        uint64_t id;
        memcpy(&len, h, 4);
        memcpy(&id, h + 4, 8);
        n->bytes_in += 12;
        if (len == 0 && id == kStopId) {
            n->stop_frames++;
            {
                std::lock_guard<std::mutex> lk(n->smu);
                n->stop_frame = true;
            }
            n->scv.notify_all();
            break;
        }
        n->frames++;
        if (len > ql) {
            // longer than any valid request: discarded without buffering (src/bin/altopir.rs:417-425)
            size_t left = len;
            bool ok = true;
            while (left && ok) {
                const size_t k = left < sink.size() ? left : sink.size();
                ok = read_full(c.fd, sink.data(), k);
                left -= k;
            }
            if (!ok) break;
            n->bytes_in += len;
            n->discarded++;
            c.owed++;
            enqueue(c, Reply{id, kStatusBadLength, c.db, -1, nullptr, 0});
            continue;
        }
        if (len && !read_full(c.fd, buf.data(), len)) break;
        n->bytes_in += len;
        uint64_t tag;
        c.owed++;
        {
            std::lock_guard<std::mutex> lk(n->mu);
            tag = n->next_tag++;
            n->tags.emplace(tag, std::make_pair(cp, id));
        }
        // copies the frame into a pinned slot before it returns; 1 and 9 complete at once
        const int rc = alto_server_submit(n->srv, c.db, len ? buf.data() : nullptr, len, tag);
        if (rc == 9) n->busy++;
        if (rc < 0) {
            {
                std::lock_guard<std::mutex> lk(n->mu);
                n->tags.erase(tag);
            }
            enqueue(c, Reply{id, -1, c.db, -1, nullptr, 0});
        }
    }
}

void completer_loop(Net *n) {
    std::vector<alto_completion> done(256);
    while (!n->stopping.load()) {
        const int k = alto_server_poll(n->srv, done.data(), (int)done.size(), 50);
        for (int i = 0; i < k; i++) {
            const alto_completion &x = done[i];
            std::shared_ptr<Conn> c;
            uint64_t id = 0;
            {
                std::lock_guard<std::mutex> lk(n->mu);
                auto it = n->tags.find(x.tag);
                if (it != n->tags.end()) {
                    c = it->second.first;
                    id = it->second.second;
                    n->tags.erase(it);
                }
            }
            if (!c) {  // not submitted by this front end
                if (x.status == 0) alto_server_release(n->srv, x.db, x.slot);
                continue;
            }
            enqueue(*c, Reply{id, x.status, x.db, x.status == 0 ? x.slot : -1, x.response, x.response_len});
        }
    }
}

void accept_loop(Net *n, int db) {
    for (;;) {
        const int fd = accept(n->lfd[db], nullptr, nullptr);
        if (fd < 0) {
            if (errno == EINTR || errno == ECONNABORTED) continue;
            return;  // the listening socket was shut down
        }
        if (n->stopping.load()) {
            close(fd);
            return;
        }
        const int one = 1;
        setsockopt(fd, IPPROTO_TCP, TCP_NODELAY, &one, sizeof one);
        auto c = std::make_shared<Conn>();
        c->fd = fd;
        c->db = db;
        {
            std::lock_guard<std::mutex> lk(n->mu);
            n->conns.push_back(c);
        }
        n->connections++;
        c->writer = std::thread(writer_loop, n, c.get());
        c->reader = std::thread(reader_loop, n, c);
    }
}

void teardown(Net *n) {
    n->stopping = true;
    for (int fd : n->lfd)
        if (fd >= 0) shutdown(fd, SHUT_RDWR);
    for (auto &t : n->acceptors)
        if (t.joinable()) t.join();
    std::vector<std::shared_ptr<Conn>> conns;
    {
        std::lock_guard<std::mutex> lk(n->mu);
        conns = n->conns;
    }
    for (auto &c : conns) {
        std::lock_guard<std::mutex> lk(c->mu);
        if (c->fd >= 0) shutdown(c->fd, SHUT_RDWR);  // wakes its reader and writer
    }
    for (auto &c : conns)
        if (c->reader.joinable()) c->reader.join();
    if (n->completer.joinable()) n->completer.join();
    for (auto &c : conns) {
        {
            std::lock_guard<std::mutex> lk(c->mu);
            c->closing = true;
        }
        c->cv.notify_all();
        if (c->writer.joinable()) c->writer.join();  // drains its queue, releasing the slots
        if (c->fd >= 0) close(c->fd);
        c->fd = -1;
    }
    for (int fd : n->lfd)
        if (fd >= 0) close(fd);
    {
        std::lock_guard<std::mutex> lk(n->mu);
        n->tags.clear();
        n->conns.clear();
    }
}

}  // namespace

extern "C" {

void *alto_net_new(void *srv, void *const *engines, int n_engines, const char *host, const int *ports) {
    if (!srv || !engines || n_engines <= 0 || !ports) return nullptr;
    auto *n = new Net;
    n->srv = srv;
    in_addr addr{};
    if (inet_pton(AF_INET, host ? host : "127.0.0.1", &addr) != 1) {
        fprintf(stderr, "alto_net_new: %s is not an IPv4 address\n", host);
        delete n;
        return nullptr;
    }
    bool ok = true;
    for (int i = 0; i < n_engines && ok; i++) {
        if (!engines[i]) {
            ok = false;
            break;
        }
        n->req_len.push_back(alto_request_len(engines[i]));
        const int fd = socket(AF_INET, SOCK_STREAM | SOCK_CLOEXEC, 0);
        n->lfd.push_back(fd);
        n->port.push_back(-1);
        if (fd < 0) {
            ok = false;
            break;
        }
        const int one = 1;
        setsockopt(fd, SOL_SOCKET, SO_REUSEADDR, &one, sizeof one);
        sockaddr_in sa{};
        sa.sin_family = AF_INET;
        sa.sin_port = htons((uint16_t)ports[i]);
        sa.sin_addr = addr;
        socklen_t sl = sizeof sa;
        if (bind(fd, (sockaddr *)&sa, sizeof sa) != 0 || listen(fd, 128) != 0 ||
            getsockname(fd, (sockaddr *)&sa, &sl) != 0) {
            fprintf(stderr, "alto_net_new: cannot listen on port %d: %s\n", ports[i], strerror(errno));
            ok = false;
            break;
        }
        n->port.back() = ntohs(sa.sin_port);
    }
    if (!ok) {
        for (int fd : n->lfd)
            if (fd >= 0) close(fd);
        delete n;
        return nullptr;
    }
    n->completer = std::thread(completer_loop, n);
    for (int i = 0; i < n_engines; i++) n->acceptors.emplace_back(accept_loop, n, i);
    return n;
}

int alto_net_port(const void *p, int db) {
    auto *n = (const Net *)p;
    if (!n || db < 0 || db >= (int)n->port.size()) return -1;
    return n->port[db];
}

int alto_net_wait(void *p, int timeout_ms) {
    auto *n = (Net *)p;
    if (!n) return -1;
    std::unique_lock<std::mutex> lk(n->smu);
    const auto pred = [&] { return n->stop_frame; };
    if (timeout_ms < 0) n->scv.wait(lk, pred);
    else n->scv.wait_for(lk, std::chrono::milliseconds(timeout_ms), pred);
    return n->stop_frame ? 1 : 0;
}

int alto_net_get_stats(const void *p, alto_net_stats *out) {
    auto *n = (const Net *)p;
    if (!n || !out) return -1;
    out->connections = n->connections;
    out->frames = n->frames;
    out->answered = n->answered;
    out->rejected = n->rejected;
    out->busy = n->busy;
    out->discarded = n->discarded;
    out->stop_frames = n->stop_frames;
    out->bytes_in = n->bytes_in;
    out->bytes_out = n->bytes_out;
    return 0;
}

void alto_net_free(void *p) {
    auto *n = (Net *)p;
    if (!n) return;
    teardown(n);
    delete n;
}

}  // extern "C"

// ------------------------------------------------------------ native load client --

namespace {

// xoshiro256** seeded through splitmix64 (the schedule's own generator)
struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed) {
        for (auto &w : s) {
            seed += 0x9e3779b97f4a7c15ull;
            uint64_t z = seed;
            z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
            z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
            w = z ^ (z >> 31);
        }
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        const uint64_t r = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }  // [0, 1)
};

int64_t mono_ns() {
    return std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

int dial(const char *host, int port, int timeout_s) {
    in_addr addr{};
    if (inet_pton(AF_INET, host, &addr) != 1) return -1;
    const int fd = socket(AF_INET, SOCK_STREAM | SOCK_CLOEXEC, 0);
    if (fd < 0) return -1;
    sockaddr_in sa{};
    sa.sin_family = AF_INET;
    sa.sin_port = htons((uint16_t)port);
    sa.sin_addr = addr;
    if (connect(fd, (sockaddr *)&sa, sizeof sa) != 0) {
        close(fd);
        return -1;
    }
    const int one = 1;
    setsockopt(fd, IPPROTO_TCP, TCP_NODELAY, &one, sizeof one);
    timeval tv{timeout_s, 0};
    setsockopt(fd, SOL_SOCKET, SO_RCVTIMEO, &tv, sizeof tv);
    return fd;
}

}  // namespace

extern "C" int alto_net_load(const char *host, int port, const uint8_t *const *frames, const size_t *lens, int nframes,
                             const int *expect_status, const uint8_t *const *expect, size_t expect_len, double rate,
                             double seconds, int conns, uint64_t seed, int drain_ms, alto_net_load_result *out) {
    if (!host || !frames || !lens || nframes <= 0 || !(rate > 0) || !(seconds > 0) || conns <= 0 || !out) return -1;
    // the schedule (src/bin/altopir.rs:612-623): exponential gaps at `rate`, frames uniform
    Rng rng(seed);
    std::vector<int64_t> due;
    double t = 0;
    for (;;) {
        t += -std::log(1.0 - rng.uniform()) / rate;
        if (t >= seconds) break;
        due.push_back((int64_t)(t * 1e9));
    }
    const size_t n = due.size();
    std::vector<int> idx(n);
    for (auto &v : idx) v = (int)(rng.next() % (uint64_t)nframes);
    std::vector<int> fds(conns, -1);
    for (int c = 0; c < conns; c++) {
        fds[c] = dial(host, port, (int)seconds + drain_ms / 1000 + 5);
        if (fds[c] < 0) {
            for (int fd : fds)
                if (fd >= 0) close(fd);
            return -1;
        }
    }
    std::vector<int> status(n, INT32_MIN);
    std::vector<int64_t> t_recv(n, 0);
    std::vector<uint8_t> mismatch(n, 0);
    const int64_t start = mono_ns() + 200000000;  // 0.2 s to let every thread start
    std::vector<std::thread> th;
    for (int c = 0; c < conns; c++) {
        th.emplace_back([&, c] {  // sender (src/bin/altopir.rs:665-684)
            for (size_t i = (size_t)c; i < n; i += (size_t)conns) {
                const int64_t at = start + due[i];
                for (;;) {
                    const int64_t now = mono_ns();
                    if (now >= at) break;
                    if (at - now > 200000) std::this_thread::sleep_for(std::chrono::nanoseconds(at - now - 150000));
                }
                uint8_t h[12];
                const uint32_t len = (uint32_t)lens[idx[i]];
                const uint64_t id = i;
                memcpy(h, &len, 4);
                memcpy(h + 4, &id, 8);
                iovec iov[2] = {{h, 12}, {(void *)frames[idx[i]], len}};
                if (!write_all(fds[c], iov, len ? 2 : 1)) return;
            }
            std::this_thread::sleep_for(std::chrono::milliseconds(drain_ms));
            shutdown(fds[c], SHUT_WR);
        });
        th.emplace_back([&, c] {  // receiver (src/bin/altopir.rs:650-664)
            const size_t mine = n > (size_t)c ? (n - (size_t)c + conns - 1) / (size_t)conns : 0;
            std::vector<uint8_t> body(expect_len > 0 ? expect_len : 1);
            for (size_t k = 0; k < mine; k++) {
                uint8_t h[16];
                if (!read_full(fds[c], h, 16)) return;
                uint32_t len;
                uint64_t id;
                int st;
                memcpy(&len, h, 4);
                memcpy(&id, h + 4, 8);
                memcpy(&st, h + 12, 4);
                if (len > body.size()) body.resize(len);
                if (len && !read_full(fds[c], body.data(), len)) return;
                if (id >= n) continue;
                t_recv[id] = mono_ns();
                status[id] = st;
                if (st == 0 && expect && expect[idx[id]])
                    mismatch[id] = (len != expect_len || memcmp(body.data(), expect[idx[id]], len) != 0) ? 2 : 1;
            }
        });
    }
    for (auto &x : th) x.join();
    for (int fd : fds) close(fd);
    alto_net_load_result r{};
    r.requests = (long long)n;
    r.seconds = seconds;
    std::vector<double> lat;
    for (size_t i = 0; i < n; i++) {
        if (status[i] == INT32_MIN) {
            r.missing++;
            continue;
        }
        if (status[i] == 0) {
            r.answered++;
            lat.push_back((double)(t_recv[i] - start - due[i]) * 1e-6);
        } else if (status[i] == 9) {
            r.busy++;
        } else {
            r.rejected++;
        }
        if (expect_status && status[i] != 9 && status[i] != expect_status[idx[i]]) r.wrong_status++;
        if (mismatch[i]) r.checked++;
        if (mismatch[i] == 2) r.bad++;
    }
    std::sort(lat.begin(), lat.end());
    auto pct = [&](double p) { return lat.empty() ? 0.0 : lat[std::min(lat.size() - 1, (size_t)(lat.size() * p))]; };
    r.answered_qps = (double)r.answered / seconds;
    r.p50_ms = pct(0.5);
    r.p90_ms = pct(0.9);
    r.p99_ms = pct(0.99);
    r.max_ms = lat.empty() ? 0.0 : lat.back();
    *out = r;
    return 0;
}
