// ss_server.cpp -- continuous-batching request loop over the batched pipeline.
//
// Device memory: one arena allocated at creation; its size is the admission
// cap.  Each admitted request takes three blocks (image, max, min); each launch
// takes one workspace block for the tile and seam-DSU state of its whole batch.
// A scheduler thread admits pending requests in arrival order while their
// blocks fit, folds them into one ss_extrema_batch_u8 launch on a free stream
// (H2D copies, pipeline, D2H copies, event), and a completion thread retires
// launches in the order they finish, frees their blocks and wakes waiters.  With two streams the copies
// of one launch overlap the kernels of the other, and requests arriving while
// launches run are folded into the next one.
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <new>
#include <random>
#include <thread>
#include <unordered_map>
#include <vector>

#include "ss_internal.hpp"

namespace {

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// First-fit allocator over [0, cap) with 256-byte granules and coalescing.
class Arena {
public:
    void reset(size_t cap) {
        free_.clear();
        free_[0] = cap;
        used_ = peak_ = 0;
    }
    bool alloc(size_t bytes, size_t* off) {
        bytes = ssx::align256(bytes ? bytes : 1);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            *off = it->first;
            const size_t rest = it->second - bytes;
            const size_t next = it->first + bytes;
            free_.erase(it);
            if (rest) free_[next] = rest;
            used_ += bytes;
            peak_ = std::max(peak_, used_);
            return true;
        }
        return false;
    }
    void release(size_t off, size_t bytes) {
        bytes = ssx::align256(bytes ? bytes : 1);
        used_ -= bytes;
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
    size_t peak() const { return peak_; }

private:
    std::map<size_t, size_t> free_;
    size_t used_ = 0, peak_ = 0;
};

struct Request {
    ss_request r{};
    double submit = 0;
    size_t px = 0;
    size_t off_img = 0, off_max = 0, off_min = 0;
};

struct Launch {
    std::vector<uint64_t> ids;
    size_t ws_off = 0, ws_bytes = 0;
    hipEvent_t done = nullptr;
    int stream = 0;
    double launch = 0;
    ss_status status = SS_OK;
};

}  // namespace

struct ss_server {
    ss_server_config cfg{};
    ss_config kcfg{};
    int device = 0;
    char* base = nullptr;
    Arena arena;
    std::vector<hipStream_t> streams;
    std::vector<int> stream_busy;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<uint64_t> pending;
    std::unordered_map<uint64_t, Request> reqs;
    std::unordered_map<uint64_t, ss_completion> done;
    std::deque<Launch> inflight;
    uint64_t next_id = 1;
    bool stop = false;
    ss_server_stats stats{};
    std::thread sched, reaper;

    void schedule_loop();
    void reap_loop();
    void free_request(Request& q) {
        arena.release(q.off_img, q.px);
        arena.release(q.off_max, q.px);
        arena.release(q.off_min, q.px);
    }
};

void ss_server::schedule_loop() {
    (void)hipSetDevice(device);
    std::unique_lock<std::mutex> lk(mu);
    while (!stop) {
        int free_stream = -1;
        for (size_t i = 0; i < streams.size(); ++i)
            if (!stream_busy[i]) { free_stream = (int)i; break; }
        if (pending.empty() || free_stream < 0) {
            cv_work.wait(lk);
            continue;
        }
        // Admission in arrival order while the blocks fit.
        std::vector<uint64_t> ids;
        uint64_t batch_px = 0;
        while (!pending.empty() && (cfg.max_batch <= 0 || (int)ids.size() < cfg.max_batch) &&
               (cfg.max_batch_pixels == 0 || batch_px < cfg.max_batch_pixels)) {
            Request& q = reqs[pending.front()];
            size_t a, b, c;
            if (!arena.alloc(q.px, &a)) break;
            if (!arena.alloc(q.px, &b)) { arena.release(a, q.px); break; }
            if (!arena.alloc(q.px, &c)) { arena.release(a, q.px); arena.release(b, q.px); break; }
            q.off_img = a; q.off_max = b; q.off_min = c;
            ids.push_back(pending.front());
            pending.pop_front();
            batch_px += q.px;
        }
        std::vector<ss_image_desc> d;
        size_t ws_bytes = 0, ws_off = 0;
        for (;;) {
            d.clear();
            for (uint64_t id : ids) {
                const Request& q = reqs[id];
                d.push_back(ss_image_desc{reinterpret_cast<const uint8_t*>(base + q.off_img),
                                          reinterpret_cast<uint8_t*>(base + q.off_max),
                                          reinterpret_cast<uint8_t*>(base + q.off_min), (uint64_t)q.r.width,
                                          (uint64_t)q.r.width, q.r.width, q.r.height});
            }
            if (d.empty()) break;
            if (ss_batch_workspace_size(d.data(), (int)d.size(), &kcfg, &ws_bytes) == SS_OK &&
                arena.alloc(ws_bytes, &ws_off))
                break;
            // Workspace does not fit: hand the newest admitted request back.
            Request& q = reqs[ids.back()];
            free_request(q);
            pending.push_front(ids.back());
            ids.pop_back();
        }
        if (ids.empty()) {
            cv_work.wait(lk);  // wait for a launch to retire and free memory
            continue;
        }
        Launch L;
        L.ids = ids;
        L.ws_off = ws_off;
        L.ws_bytes = ws_bytes;
        L.stream = free_stream;
        L.launch = now_s();
        stream_busy[free_stream] = 1;
        std::vector<Request> snap;
        for (uint64_t id : ids) snap.push_back(reqs[id]);
        lk.unlock();
        hipStream_t st = streams[free_stream];
        ss_status status = SS_OK;
        for (const Request& q : snap)
            if (hipMemcpyAsync(base + q.off_img, q.r.host_img, q.px, hipMemcpyHostToDevice, st) != hipSuccess)
                status = SS_ERR_HIP;
        if (status == SS_OK)
            status = ss_extrema_batch_u8(d.data(), (int)d.size(), base + ws_off, ws_bytes, &kcfg, st);
        for (const Request& q : snap) {
            if (hipMemcpyAsync(q.r.host_max, base + q.off_max, q.px, hipMemcpyDeviceToHost, st) != hipSuccess ||
                hipMemcpyAsync(q.r.host_min, base + q.off_min, q.px, hipMemcpyDeviceToHost, st) != hipSuccess)
                status = SS_ERR_HIP;
        }
        if (hipEventCreateWithFlags(&L.done, hipEventDisableTiming) != hipSuccess) {
            L.done = nullptr;
            status = SS_ERR_HIP;
        } else if (hipEventRecord(L.done, st) != hipSuccess) {
            (void)hipEventDestroy(L.done);
            L.done = nullptr;
            status = SS_ERR_HIP;
        }
        // Without a recorded completion event the reaper cannot tell when the
        // launch's copies and kernels end, and it would free their arena
        // blocks at once: drain the stream here so that nothing still uses
        // them when the launch is retired.
        if (!L.done) (void)hipStreamSynchronize(st);
        L.status = status;
        lk.lock();
        stats.launches++;
        inflight.push_back(L);
        cv_done.notify_all();
    }
}

void ss_server::reap_loop() {
    (void)hipSetDevice(device);
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
        if (inflight.empty()) {
            if (stop) return;
            cv_done.wait(lk);
            continue;
        }
        // Retire whichever launch has finished first: launches on different
        // streams complete out of order, and a small batch must not wait
        // behind a large one's copies.
        size_t pick = inflight.size();
        for (size_t i = 0; i < inflight.size(); ++i) {
            const Launch& c = inflight[i];
            if (!c.done || hipEventQuery(c.done) != hipErrorNotReady) {
                pick = i;
                break;
            }
        }
        if (pick == inflight.size()) {
            lk.unlock();
            std::this_thread::sleep_for(std::chrono::microseconds(20));
            lk.lock();
            continue;
        }
        Launch L = inflight[pick];
        inflight.erase(inflight.begin() + (std::ptrdiff_t)pick);
        lk.unlock();
        ss_status status = L.status;
        if (L.done && hipEventSynchronize(L.done) != hipSuccess) status = SS_ERR_HIP;
        if (status == SS_OK) {
            ss_stats st{};
            const ss_status rs = ss_read_stats(base + L.ws_off, L.ws_bytes, &st, streams[L.stream]);
            if (rs != SS_OK) status = rs;
        }
        if (L.done) (void)hipEventDestroy(L.done);
        const double t = now_s();
        lk.lock();
        arena.release(L.ws_off, L.ws_bytes);
        stream_busy[L.stream] = 0;
        for (uint64_t id : L.ids) {
            Request& q = reqs[id];
            free_request(q);
            ss_completion c{};
            c.id = id;
            c.tag = q.r.tag;
            c.status = status;
            c.batch_size = (int)L.ids.size();
            c.submit_s = q.submit;
            c.launch_s = L.launch;
            c.done_s = t;
            done[id] = c;
            stats.completed++;
            stats.pixels += q.px;
            reqs.erase(id);
        }
        stats.peak_inflight_bytes = arena.peak();
        cv_done.notify_all();
        cv_work.notify_all();
    }
}

extern "C" {

ss_status ss_server_create(const ss_server_config* cfg, ss_server** out) {
    if (!cfg || !out || cfg->hbm_cap_bytes < 4096) return SS_ERR_ARG;
    *out = nullptr;
    ss_server* s = new (std::nothrow) ss_server();
    if (!s) return SS_ERR_CAPACITY;
    s->cfg = *cfg;
    if (s->cfg.streams <= 0) s->cfg.streams = 2;
    s->kcfg = ssx::resolve_config(nullptr);
    if (cfg->tile) s->kcfg.tile = cfg->tile;
    if (cfg->threads) s->kcfg.threads = cfg->threads;
    if (!ssx::config_valid(s->kcfg.tile, s->kcfg.threads)) {
        delete s;
        return SS_ERR_ARG;
    }
    if (hipGetDevice(&s->device) != hipSuccess || hipMalloc(&s->base, cfg->hbm_cap_bytes) != hipSuccess) {
        delete s;
        return SS_ERR_HIP;
    }
    s->arena.reset(cfg->hbm_cap_bytes);
    s->streams.resize((size_t)s->cfg.streams);
    s->stream_busy.assign((size_t)s->cfg.streams, 0);
    for (size_t i = 0; i < s->streams.size(); ++i) {
        if (hipStreamCreateWithFlags(&s->streams[i], hipStreamNonBlocking) != hipSuccess) {
            for (size_t j = 0; j < i; ++j) (void)hipStreamDestroy(s->streams[j]);
            (void)hipFree(s->base);
            delete s;
            return SS_ERR_HIP;
        }
    }
    s->sched = std::thread([s] { s->schedule_loop(); });
    s->reaper = std::thread([s] { s->reap_loop(); });
    *out = s;
    return SS_OK;
}

ss_status ss_server_submit(ss_server* s, const ss_request* req, uint64_t* id) {
    if (!s || !req || !req->host_img || !req->host_max || !req->host_min || req->width <= 0 || req->height <= 0)
        return SS_ERR_ARG;
    const size_t px = (size_t)req->width * (size_t)req->height;
    size_t ws = 0;
    if (ss_workspace_size(req->width, req->height, &s->kcfg, &ws) != SS_OK) return SS_ERR_ARG;
    // a lone request needs its three blocks plus a one-image batch workspace
    const size_t need = 3 * ssx::align256(px) + ssx::align256(ws) + 4096;
    std::lock_guard<std::mutex> lk(s->mu);
    s->stats.submitted++;
    if (need > s->cfg.hbm_cap_bytes) {
        s->stats.rejected++;
        return SS_ERR_CAPACITY;
    }
    const uint64_t rid = s->next_id++;
    Request& q = s->reqs[rid];
    q.r = *req;
    q.px = px;
    q.submit = now_s();
    s->pending.push_back(rid);
    if (id) *id = rid;
    s->cv_work.notify_all();
    return SS_OK;
}

ss_status ss_server_wait(ss_server* s, uint64_t id, ss_completion* out) {
    if (!s) return SS_ERR_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        auto it = s->done.find(id);
        if (it != s->done.end()) {
            if (out) *out = it->second;
            const ss_status st = (ss_status)it->second.status;
            s->done.erase(it);
            return st;
        }
        if (!s->reqs.count(id)) return SS_ERR_ARG;
        s->cv_done.wait(lk);
    }
}

ss_status ss_server_drain(ss_server* s) {
    if (!s) return SS_ERR_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [s] { return s->reqs.empty(); });
    return SS_OK;
}

ss_status ss_server_get_stats(ss_server* s, ss_server_stats* out) {
    if (!s || !out) return SS_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->stats;
    return SS_OK;
}

ss_status ss_server_destroy(ss_server* s) {
    if (!s) return SS_OK;
    (void)ss_server_drain(s);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    s->cv_done.notify_all();
    if (s->sched.joinable()) s->sched.join();
    if (s->reaper.joinable()) s->reaper.join();
    for (auto& st : s->streams) (void)hipStreamDestroy(st);
    if (s->base) (void)hipFree(s->base);
    delete s;
    return SS_OK;
}

ss_status ss_server_run_open_loop(ss_server* s, const ss_request* reqs, int n, double qps, uint64_t seed,
                                  ss_completion* out) {
    if (!s || !reqs || n <= 0 || !(qps > 0) || !out) return SS_ERR_ARG;
    std::mt19937_64 rng(seed);
    std::exponential_distribution<double> gap(qps);
    std::vector<uint64_t> ids((size_t)n, 0);
    std::vector<ss_status> sub((size_t)n, SS_OK);
    const auto t0 = std::chrono::steady_clock::now();
    double t = 0.0;
    for (int i = 0; i < n; ++i) {
        t += gap(rng);
        std::this_thread::sleep_until(t0 + std::chrono::duration_cast<std::chrono::steady_clock::duration>(
                                               std::chrono::duration<double>(t)));
        sub[(size_t)i] = ss_server_submit(s, &reqs[i], &ids[(size_t)i]);
    }
    ss_status result = SS_OK;
    for (int i = 0; i < n; ++i) {
        ss_completion c{};
        if (sub[(size_t)i] != SS_OK) {
            c.status = sub[(size_t)i];
            c.tag = reqs[i].tag;
            result = sub[(size_t)i];
        } else {
            const ss_status st = ss_server_wait(s, ids[(size_t)i], &c);
            if (st != SS_OK) result = st;
        }
        out[i] = c;
    }
    return result;
}

}  // extern "C"
