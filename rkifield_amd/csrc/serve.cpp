// serve.cpp — request server over resident sparse scenes.
//
// Requests (views, shadow-ray batches, point-query batches) arrive from any
// host thread. A worker thread admits them in arrival order against a
// device-memory arena allocated once (the memory budget) and a slot count,
// then launches rounds: each round takes every admitted request (up to the
// round caps), copies inputs in, runs one batched launch over all of them
// (batch.hip: each request computed exactly as alone), copies outputs out and
// records an event. Up to rounds_in_flight rounds run on separate streams, so
// the next round starts filling the device while the previous one drains, and
// requests that arrive meanwhile join the next round as slots and memory free
// up (continuous batching). The per-request policy runs the plain single-batch
// kernels, one request at a time: the baseline the batched policy is measured
// against.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <limits>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "host_internal.h"

namespace {

constexpr uint64_t kAlign = 256;
constexpr uint32_t kMaxRounds = 4;
// Rounds below this many 64-lane work units (4 M rays) take the low-latency
// schedule under the default shape: their time is set by their slowest rays
// (one 1080p view is 32,400 units); larger rounds are bound by total work.

uint64_t align_up(uint64_t v) { return (v + kAlign - 1) & ~(kAlign - 1); }

// First-fit allocator over [0, size) with a coalescing free list.
class Arena {
  public:
    void init(uint64_t size) {
        free_.clear();
        if (size) free_[0] = size;
        size_ = size;
        used_ = peak_ = 0;
    }
    bool alloc(uint64_t bytes, uint64_t &off) {
        if (bytes == 0) {
            off = 0;
            return true;
        }
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            off = it->first;
            uint64_t rest = it->second - bytes;
            free_.erase(it);
            if (rest) free_[off + bytes] = rest;
            used_ += bytes;
            peak_ = std::max(peak_, used_);
            return true;
        }
        return false;
    }
    void release(uint64_t off, uint64_t bytes) {
        if (bytes == 0) return;
        uint64_t start = off, len = bytes;
        auto next = free_.lower_bound(off);
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == off) {
                start = prev->first;
                len += prev->second;
                free_.erase(prev);
            }
        }
        if (next != free_.end() && off + bytes == next->first) {
            len += next->second;
            free_.erase(next);
        }
        free_[start] = len;
        used_ -= bytes;
    }
    uint64_t size() const { return size_; }
    uint64_t used() const { return used_; }
    uint64_t peak() const { return peak_; }

  private:
    std::map<uint64_t, uint64_t> free_;
    uint64_t size_ = 0, used_ = 0, peak_ = 0;
};

struct Req {
    uint64_t id = 0;
    rkfm_request r{};
    uint64_t work = 0;   // rays + queries (views: primary pixels)
    uint64_t bytes = 0;  // arena footprint
    uint64_t off = 0;    // arena offset of the footprint
    bool holds = false;  // footprint currently allocated
    // offsets inside the footprint
    uint64_t in_off = 0, out_off = 0;
    uint64_t plane_off[5] = {0, 0, 0, 0, 0};  // position, normal, material, motion, shadow
    rkfm_request_status st{};
};

const uint64_t kPlaneBytes[5] = {16, 8, 4, 16, 4};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t ev_start = nullptr, ev_end = nullptr;
    rkf::BatchDev *h_items = nullptr;  // pinned
    uint32_t *h_begins = nullptr;      // pinned
    DevBuf d_items, d_begins, d_counter;
    bool busy = false;
    uint64_t round = 0;
    std::vector<Req *> reqs;
    int error = 0;
};

double seconds_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

struct rkfm_server {
    rkfm_ctx *ctx = nullptr;
    int device = 0;
    std::vector<const rkfm_sparse_scene *> scenes;
    rkfm_server_config cfg{};
    rkfm_batch_tuning tuning{};
    bool adaptive = false;  // zero tuning given: the default shape
    uint32_t waves = 0;
    unsigned char *arena_base = nullptr;
    Arena arena;
    Slot slots[kMaxRounds];
    uint32_t n_slots = 0;
    std::chrono::steady_clock::time_point t0;

    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Req *> pending, admitted;
    std::unordered_map<uint64_t, Req *> reqs;
    uint32_t running = 0;
    uint64_t admitted_work = 0;  // rays + queries of the admitted queue
    uint64_t next_id = 1, round_seq = 0, completions = 0;
    bool stop = false;
    rkfm_server_stats stats{};
    std::thread worker;

    double now() const { return seconds_since(t0); }
    void run();
    bool admit_locked();
    bool reap_locked();
    void finish_locked(Slot &sl, int error);
    int launch(Slot &sl);
    int launch_batched(Slot &sl);
    int launch_plain(Slot &sl);
    int copies_in(Slot &sl);
    int copies_out(Slot &sl);
};

namespace {

// Bytes per input / output record of a query kind.
uint64_t in_record_bytes(uint32_t kind) { return kind == RKFM_BATCH_SPHERES ? sizeof(rkfm_sphere_cast) : 12u; }
uint64_t out_record_bytes(uint32_t kind) {
    return kind == RKFM_BATCH_SHADOW ? 4u : kind == RKFM_BATCH_POINTS ? sizeof(rkfm_point_result)
                                                                      : sizeof(rkfm_cast_result);
}

// Footprint layout of one request inside its arena region.
int plan_request(const rkfm_server *s, const rkfm_request &r, Req &q) {
    if (r.scene >= s->scenes.size())
        return rkf_set_error(RKFM_ERR_INVALID_ARG, "server: scene %u of %zu", r.scene, s->scenes.size());
    if (r.ext.hit_eps_mode > RKFM_EPS_RELATIVE)
        return rkf_set_error(RKFM_ERR_INVALID_ARG, "server: hit_eps_mode %u", r.ext.hit_eps_mode);
    // A round's batched launch has no per-request tile-interval buffers and a
    // round of one request takes the plain kernel without them, so a flagged
    // request is refused here, before admission, instead of failing (or being
    // silently run unculled) together with the requests of its round.
    if (r.ext.march_flags != 0)
        return rkf_set_error(RKFM_ERR_INVALID_ARG, "server: ext.march_flags %u (the server takes 0)",
                             r.ext.march_flags);
    uint64_t off = 0;
    if (r.kind == RKFM_BATCH_VIEW) {
        if (r.width == 0 || r.height == 0 || (uint64_t)r.width * r.height > (1ull << 30))
            return rkf_set_error(RKFM_ERR_INVALID_ARG, "server: view size %ux%u", r.width, r.height);
        uint64_t px = (uint64_t)r.width * r.height;
        for (int p = 0; p < 5; p++) {
            if (p == 4 && !r.ext.shadow_enable) continue;
            q.plane_off[p] = off;
            off += align_up(px * kPlaneBytes[p]);
        }
        q.work = px;
    } else if (r.kind == RKFM_BATCH_SHADOW || r.kind == RKFM_BATCH_POINTS || r.kind == RKFM_BATCH_SPHERES) {
        if (r.n && (!r.in || !r.out)) return rkf_set_error(RKFM_ERR_INVALID_ARG, "server: null host buffer");
        q.in_off = 0;
        off = align_up((uint64_t)r.n * in_record_bytes(r.kind));
        q.out_off = off;
        off += align_up((uint64_t)r.n * out_record_bytes(r.kind));
        q.work = r.n;
    } else {
        return rkf_set_error(RKFM_ERR_INVALID_ARG, "server: request kind %u", r.kind);
    }
    q.bytes = off;
    if (q.bytes > s->arena.size())
        return rkf_set_error(RKFM_ERR_OOM, "server: request needs %llu bytes, budget %llu",
                             (unsigned long long)q.bytes, (unsigned long long)s->arena.size());
    return RKFM_OK;
}

}  // namespace

// Admission in arrival order: the head of the queue waits until a slot and
// its footprint are free (no overtaking, so a large request is not starved).
bool rkfm_server::admit_locked() {
    bool any = false;
    while (!pending.empty()) {
        if (admitted.size() + running >= cfg.max_requests) break;
        Req *q = pending.front();
        uint64_t off = 0;
        if (!arena.alloc(q->bytes, off)) break;
        q->off = off;
        q->holds = true;
        q->st.state = RKFM_REQ_ADMITTED;
        q->st.t_admit = now();
        pending.pop_front();
        admitted.push_back(q);
        admitted_work += q->work;
        any = true;
    }
    stats.arena_in_use = arena.used();
    stats.arena_peak = arena.peak();
    return any;
}

void rkfm_server::finish_locked(Slot &sl, int error) {
    double t = now();
    float ms = 0.0f;
    if (!error && hipEventElapsedTime(&ms, sl.ev_start, sl.ev_end) == hipSuccess) stats.gpu_seconds += ms * 1e-3;
    for (Req *q : sl.reqs) {
        q->st.t_done = t;
        if (error) {
            q->st.state = RKFM_REQ_FAILED;
            q->st.error = error;
            stats.failed++;
        } else {
            q->st.state = RKFM_REQ_DONE;
            stats.completed++;
            stats.work_items += q->work;
        }
        bool keep = !error && q->r.kind == RKFM_BATCH_VIEW && (q->r.flags & RKFM_REQ_KEEP_ON_DEVICE);
        if (!keep && q->holds) {
            arena.release(q->off, q->bytes);
            q->holds = false;
        }
    }
    running -= (uint32_t)sl.reqs.size();
    sl.reqs.clear();
    sl.busy = false;
    completions++;
    stats.arena_in_use = arena.used();
    cv_done.notify_all();
}

bool rkfm_server::reap_locked() {
    bool any = false;
    for (uint32_t i = 0; i < n_slots; i++) {
        Slot &sl = slots[i];
        if (!sl.busy) continue;
        hipError_t e = hipEventQuery(sl.ev_end);
        if (e == hipErrorNotReady) continue;
        int err = 0;
        if (e != hipSuccess) err = rkf_hip_check(e, "server round");
        finish_locked(sl, err);
        any = true;
    }
    return any;
}

int rkfm_server::copies_in(Slot &sl) {
    for (Req *q : sl.reqs) {
        if (q->r.kind == RKFM_BATCH_VIEW || q->r.n == 0) continue;
        hipError_t e = hipMemcpyAsync(arena_base + q->off + q->in_off, q->r.in,
                                      (uint64_t)q->r.n * in_record_bytes(q->r.kind), hipMemcpyHostToDevice, sl.stream);
        if (e != hipSuccess) return rkf_hip_check(e, "server input copy");
    }
    return RKFM_OK;
}

int rkfm_server::copies_out(Slot &sl) {
    for (Req *q : sl.reqs) {
        const rkfm_request &r = q->r;
        unsigned char *base = arena_base + q->off;
        if (r.kind == RKFM_BATCH_VIEW) {
            if (r.flags & RKFM_REQ_KEEP_ON_DEVICE) continue;
            void *dst[5] = {r.view_out.position, r.view_out.normal, r.view_out.material, r.view_out.motion,
                            r.view_out.shadow};
            uint64_t px = (uint64_t)r.width * r.height;
            for (int p = 0; p < 5; p++) {
                if (!dst[p] || (p == 4 && !r.ext.shadow_enable)) continue;
                hipError_t e = hipMemcpyAsync(dst[p], base + q->plane_off[p], px * kPlaneBytes[p],
                                              hipMemcpyDeviceToHost, sl.stream);
                if (e != hipSuccess) return rkf_hip_check(e, "server view copy");
            }
        } else if (r.n) {
            uint64_t bytes = (uint64_t)r.n * out_record_bytes(r.kind);
            hipError_t e = hipMemcpyAsync(r.out, base + q->out_off, bytes, hipMemcpyDeviceToHost, sl.stream);
            if (e != hipSuccess) return rkf_hip_check(e, "server output copy");
        }
    }
    return RKFM_OK;
}

namespace {

rkfm_batch_item item_of(const rkfm_server &s, const Req &q) {
    rkfm_batch_item it;
    memset(&it, 0, sizeof(it));
    const rkfm_request &r = q.r;
    unsigned char *base = s.arena_base + q.off;
    it.kind = r.kind;
    it.n = r.n;
    it.scene = s.scenes[r.scene];
    it.camera = r.camera;
    it.su = r.su;
    it.ext = r.ext;
    if (r.kind == RKFM_BATCH_VIEW) {
        it.out.position = base + q.plane_off[0];
        it.out.normal = base + q.plane_off[1];
        it.out.material = base + q.plane_off[2];
        it.out.motion = base + q.plane_off[3];
        it.out.shadow = r.ext.shadow_enable ? base + q.plane_off[4] : nullptr;
        it.out.width = r.width;
        it.out.height = r.height;
    } else if (r.kind == RKFM_BATCH_SHADOW) {
        it.origins = (const float *)(base + q.in_off);
        it.visibility = (float *)(base + q.out_off);
    } else if (r.kind == RKFM_BATCH_POINTS) {
        it.xyz = (const float *)(base + q.in_off);
        it.results = (rkfm_point_result *)(base + q.out_off);
    } else {
        it.casts = (const rkfm_sphere_cast *)(base + q.in_off);
        it.cast_results = (rkfm_cast_result *)(base + q.out_off);
    }
    return it;
}

}  // namespace

int rkfm_server::launch_batched(Slot &sl) {
    std::vector<rkfm_batch_item> items;
    items.reserve(sl.reqs.size());
    for (Req *q : sl.reqs) items.push_back(item_of(*this, *q));
    std::vector<rkf::BatchDev> dev;
    std::vector<uint32_t> begins;
    uint32_t total = 0;
    int rc = rkf_batch_prepare(nullptr, items.data(), (uint32_t)items.size(), tuning, dev, begins, total);
    if (rc) return rc;
    memcpy(sl.h_items, dev.data(), dev.size() * sizeof(rkf::BatchDev));
    memcpy(sl.h_begins, begins.data(), begins.size() * 4);
    rc = rkf_hip_check(hipMemcpyAsync(sl.d_items.ptr, sl.h_items, dev.size() * sizeof(rkf::BatchDev),
                                      hipMemcpyHostToDevice, sl.stream),
                       "server descriptors");
    if (!rc)
        rc = rkf_hip_check(
            hipMemcpyAsync(sl.d_begins.ptr, sl.h_begins, begins.size() * 4, hipMemcpyHostToDevice, sl.stream),
            "server begins");
    if (rc) return rc;
    // Default shape (set at creation): the 8x8 grid with the low-latency
    // per-ray code, the faster schedule for single views and mixed sets alike.
    const rkfm_batch_tuning &t = tuning;
    return rkf_batch_enqueue(t, (const rkf::BatchDev *)sl.d_items.ptr, (const uint32_t *)sl.d_begins.ptr,
                             (uint32_t)items.size(), total, (uint32_t *)sl.d_counter.ptr, waves, sl.stream);
}

// One request with its plain single-batch kernel (views: the 8x8 per-pixel
// march; points: the query kernel; shadow batches have no single-batch
// kernel and use the batched launch with one item).
int rkfm_server::launch_plain(Slot &sl) {
    Req &q = *sl.reqs[0];
    rkfm_batch_item it = item_of(*this, q);
    if (q.r.kind == RKFM_BATCH_VIEW) {
        rkf::SparseMarchParams P;
        memset(&P, 0, sizeof(P));
        int rc = rkf_fill_gbuf(&it.out, P.g);
        if (rc) return rc;
        P.S = it.scene->dev();
        P.cam = it.camera;
        P.su = it.su;
        P.ext = it.ext;
        return rkf_hip_check(rkf_launch_sparse_march(&P, 0, 0, sl.stream), "server plain march");
    }
    if (q.r.kind == RKFM_BATCH_POINTS) {
        rkf::SparseDev D = it.scene->dev();
        return rkf_hip_check(rkf_launch_sparse_query(&D, it.xyz, it.n, it.results, sl.stream), "server plain query");
    }
    if (q.r.kind == RKFM_BATCH_SPHERES) {
        rkf::SparseDev D = it.scene->dev();
        return rkf_hip_check(rkf_launch_sparse_cast(&D, it.casts, it.n, it.cast_results, sl.stream),
                             "server plain sphere cast");
    }
    return launch_batched(sl);
}

int rkfm_server::launch(Slot &sl) {
    int rc = rkf_hip_check(hipEventRecord(sl.ev_start, sl.stream), "server event");
    if (!rc) rc = copies_in(sl);
    // with the default shape, a round of one request takes the plain
    // kernel: no descriptor copies, the single-batch kernels' launch shapes
    const bool plain = cfg.policy == RKFM_SERVE_PER_REQUEST || (adaptive && sl.reqs.size() == 1);
    if (!rc) rc = plain ? launch_plain(sl) : launch_batched(sl);
    if (!rc) rc = copies_out(sl);
    hipError_t e = hipEventRecord(sl.ev_end, sl.stream);
    if (!rc && e != hipSuccess) rc = rkf_hip_check(e, "server event");
    return rc;
}

void rkfm_server::run() {
    (void)hipSetDevice(device);
    std::unique_lock<std::mutex> lk(mu);
    while (true) {
        bool progressed = reap_locked();
        if (stop) {
            // fail everything that has not been launched
            for (std::deque<Req *> *qq : {&pending, &admitted}) {
                for (Req *q : *qq) {
                    if (q->holds) {
                        arena.release(q->off, q->bytes);
                        q->holds = false;
                    }
                    q->st.state = RKFM_REQ_FAILED;
                    q->st.error = RKFM_ERR_BUSY;
                    q->st.t_done = now();
                    stats.failed++;
                }
                qq->clear();
            }
            admitted_work = 0;
            cv_done.notify_all();
            bool busy = false;
            for (uint32_t i = 0; i < n_slots; i++) busy |= slots[i].busy;
            if (!busy) break;
            lk.unlock();
            std::this_thread::yield();
            lk.lock();
            continue;
        }
        progressed |= admit_locked();
        Slot *free_slot = nullptr;
        for (uint32_t i = 0; i < n_slots && !free_slot; i++)
            if (!slots[i].busy) free_slot = &slots[i];
        if (free_slot && !admitted.empty()) {
            Slot &sl = *free_slot;
            sl.reqs.clear();
            uint64_t work = 0;
            uint32_t cap = cfg.policy == RKFM_SERVE_PER_REQUEST ? 1u : cfg.max_round_requests;
            // a round takes at least its share of the queued work, so a deep
            // queue (a burst) runs in fewer, larger launches
            const uint64_t cap_work = std::max<uint64_t>(cfg.max_round_work, admitted_work / n_slots);
            while (!admitted.empty() && sl.reqs.size() < cap) {
                Req *q = admitted.front();
                if (!sl.reqs.empty() && cap_work && work + q->work > cap_work) break;
                admitted.pop_front();
                admitted_work -= q->work;
                work += q->work;
                sl.reqs.push_back(q);
            }
            sl.busy = true;
            sl.round = ++round_seq;
            double t = now();
            for (Req *q : sl.reqs) {
                q->st.state = RKFM_REQ_RUNNING;
                q->st.t_launch = t;
                q->st.round = sl.round;
                q->st.round_requests = (uint32_t)sl.reqs.size();
            }
            running += (uint32_t)sl.reqs.size();
            stats.rounds++;
            stats.max_round_requests_seen = std::max<uint32_t>(stats.max_round_requests_seen, (uint32_t)sl.reqs.size());
            lk.unlock();
            int rc = launch(sl);
            lk.lock();
            if (rc) {
                // the stream may still hold part of the round: wait for it
                lk.unlock();
                (void)hipStreamSynchronize(sl.stream);
                lk.lock();
                finish_locked(sl, rc);
            }
            continue;
        }
        if (!progressed) {
            bool busy = running > 0;
            if (!busy) {
                cv_work.wait(lk);
            } else {
                lk.unlock();
                std::this_thread::yield();
                lk.lock();
            }
        }
    }
}

extern "C" {

int rkfm_host_alloc(uint64_t bytes, void **p) {
    if (!p) return rkf_set_error(RKFM_ERR_INVALID_ARG, "host_alloc: null argument");
    *p = nullptr;
    return rkf_hip_check(hipHostMalloc(p, bytes ? bytes : 1, hipHostMallocDefault), "hipHostMalloc");
}

int rkfm_host_free(void *p) {
    if (!p) return RKFM_OK;
    return rkf_hip_check(hipHostFree(p), "hipHostFree");
}

int rkfm_server_create(rkfm_ctx *c, const rkfm_sparse_scene *const *scenes, uint32_t n_scenes,
                       const rkfm_server_config *cfg, rkfm_server **out) {
    if (!c || !out || !scenes || n_scenes == 0) return rkf_set_error(RKFM_ERR_INVALID_ARG, "server_create: null argument");
    *out = nullptr;
    for (uint32_t i = 0; i < n_scenes; i++)
        if (!scenes[i] || std::find(c->scenes.begin(), c->scenes.end(), scenes[i]) == c->scenes.end())
            return rkf_set_error(RKFM_ERR_INVALID_ARG, "server_create: scene %u not resident in this context", i);
    rkfm_server_config k{};
    if (cfg) k = *cfg;
    if (k.policy > RKFM_SERVE_PER_REQUEST) return rkf_set_error(RKFM_ERR_INVALID_ARG, "server_create: policy %u", k.policy);
    if (k.memory_budget == 0) k.memory_budget = c->memory_budget ? c->memory_budget : (1ull << 30);
    if (k.max_requests == 0) k.max_requests = 1024;
    if (k.rounds_in_flight == 0) k.rounds_in_flight = 4;
    if (k.max_round_work == 0) k.max_round_work = 1ull << 20;  // a 1080p view runs in a round of its own
    if (k.rounds_in_flight > kMaxRounds)
        return rkf_set_error(RKFM_ERR_INVALID_ARG, "server_create: rounds_in_flight %u > %u", k.rounds_in_flight,
                             kMaxRounds);
    if (k.policy == RKFM_SERVE_PER_REQUEST) k.rounds_in_flight = 1;
    if (k.max_round_requests == 0) k.max_round_requests = 512;
    if (k.max_pending == 0) k.max_pending = 65536;
    rkfm_batch_tuning t = k.tuning;
    const bool adaptive = t.tile_w == 0 && t.tile_h == 0;
    if (adaptive) t = rkfm_batch_tuning{RKFM_SCHEDULE_GRID_LOWLAT, 8u, 8u, 0u};
    if (!rkf_batch_valid_shape(t)) return rkf_set_error(RKFM_ERR_INVALID_ARG, "server_create: invalid launch shape");
    k.tuning = t;
    int rc = rkf_hip_check(hipSetDevice(c->device), "hipSetDevice");
    if (rc) return rc;
    rkfm_server *s = new rkfm_server();
    s->ctx = c;
    s->device = c->device;
    s->scenes.assign(scenes, scenes + n_scenes);
    s->cfg = k;
    s->tuning = t;
    s->adaptive = adaptive;
    s->n_slots = k.rounds_in_flight;
    s->t0 = std::chrono::steady_clock::now();
    if (t.schedule == RKFM_SCHEDULE_PERSISTENT) rc = rkf_device_waves(t.waves_per_cu, s->waves);
    void *base = nullptr;
    if (!rc) rc = rkf_hip_check(hipMalloc(&base, k.memory_budget), "server arena");
    s->arena_base = (unsigned char *)base;
    s->arena.init(base ? k.memory_budget : 0);
    for (uint32_t i = 0; !rc && i < s->n_slots; i++) {
        Slot &sl = s->slots[i];
        rc = rkf_hip_check(hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking), "server stream");
        if (!rc) rc = rkf_hip_check(hipEventCreate(&sl.ev_start), "server event");
        if (!rc) rc = rkf_hip_check(hipEventCreate(&sl.ev_end), "server event");
        uint64_t ib = (uint64_t)k.max_round_requests * sizeof(rkf::BatchDev);
        if (!rc) rc = rkf_hip_check(hipHostMalloc((void **)&sl.h_items, ib, hipHostMallocDefault), "server staging");
        if (!rc)
            rc = rkf_hip_check(hipHostMalloc((void **)&sl.h_begins, (uint64_t)k.max_round_requests * 4, hipHostMallocDefault),
                               "server staging");
        if (!rc) rc = sl.d_items.reserve(ib, false);
        if (!rc) rc = sl.d_begins.reserve((uint64_t)k.max_round_requests * 4, false);
        if (!rc) rc = sl.d_counter.reserve(256, false);
    }
    if (rc) {
        rkfm_server_destroy(s);
        return rc;
    }
    s->stats.arena_bytes = k.memory_budget;
    s->worker = std::thread([s] { s->run(); });
    *out = s;
    return RKFM_OK;
}

int rkfm_server_submit(rkfm_server *s, const rkfm_request *req, uint64_t *id) {
    if (!s || !req) return rkf_set_error(RKFM_ERR_INVALID_ARG, "server_submit: null argument");
    Req *q = new Req();
    q->r = *req;
    int rc = plan_request(s, *req, *q);
    if (rc) {
        delete q;
        return rc;
    }
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stop) {
            delete q;
            return rkf_set_error(RKFM_ERR_BUSY, "server_submit: server is stopping");
        }
        if (s->pending.size() >= s->cfg.max_pending) {
            delete q;
            return rkf_set_error(RKFM_ERR_BUSY, "server_submit: %u requests already queued", s->cfg.max_pending);
        }
        q->id = s->next_id++;
        q->st.state = RKFM_REQ_QUEUED;
        q->st.tag = req->tag;
        q->st.t_submit = s->now();
        s->reqs[q->id] = q;
        s->pending.push_back(q);
        s->stats.submitted++;
    }
    s->cv_work.notify_one();
    if (id) *id = q->id;
    return RKFM_OK;
}

int rkfm_server_status(rkfm_server *s, uint64_t id, rkfm_request_status *st) {
    if (!s || !st) return rkf_set_error(RKFM_ERR_INVALID_ARG, "server_status: null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return rkf_set_error(RKFM_ERR_INVALID_ARG, "server_status: unknown request %llu", (unsigned long long)id);
    *st = it->second->st;
    return RKFM_OK;
}

int rkfm_server_wait(rkfm_server *s, uint64_t id, double timeout_s, rkfm_request_status *st) {
    if (!s) return rkf_set_error(RKFM_ERR_INVALID_ARG, "server_wait: null server");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return rkf_set_error(RKFM_ERR_INVALID_ARG, "server_wait: unknown request %llu", (unsigned long long)id);
    Req *q = it->second;
    auto done = [q] { return q->st.state == RKFM_REQ_DONE || q->st.state == RKFM_REQ_FAILED; };
    if (timeout_s < 0) {
        s->cv_done.wait(lk, done);
    } else if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), done)) {
        if (st) *st = q->st;
        return rkf_set_error(RKFM_ERR_NOT_READY, "server_wait: request %llu not done", (unsigned long long)id);
    }
    if (st) *st = q->st;
    return RKFM_OK;
}

int rkfm_server_view_planes(rkfm_server *s, uint64_t id, rkfm_gbuffer *planes) {
    if (!s || !planes) return rkf_set_error(RKFM_ERR_INVALID_ARG, "server_view_planes: null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return rkf_set_error(RKFM_ERR_INVALID_ARG, "server_view_planes: unknown request");
    Req *q = it->second;
    if (q->r.kind != RKFM_BATCH_VIEW || !(q->r.flags & RKFM_REQ_KEEP_ON_DEVICE) || !q->holds ||
        q->st.state != RKFM_REQ_DONE)
        return rkf_set_error(RKFM_ERR_NOT_READY, "server_view_planes: not a finished view kept on the device");
    unsigned char *base = s->arena_base + q->off;
    memset(planes, 0, sizeof(*planes));
    planes->position = base + q->plane_off[0];
    planes->normal = base + q->plane_off[1];
    planes->material = base + q->plane_off[2];
    planes->motion = base + q->plane_off[3];
    planes->shadow = q->r.ext.shadow_enable ? base + q->plane_off[4] : nullptr;
    planes->width = q->r.width;
    planes->height = q->r.height;
    return RKFM_OK;
}

int rkfm_server_release(rkfm_server *s, uint64_t id) {
    if (!s) return rkf_set_error(RKFM_ERR_INVALID_ARG, "server_release: null server");
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->reqs.find(id);
        if (it == s->reqs.end()) return rkf_set_error(RKFM_ERR_INVALID_ARG, "server_release: unknown request");
        Req *q = it->second;
        if (q->st.state != RKFM_REQ_DONE && q->st.state != RKFM_REQ_FAILED)
            return rkf_set_error(RKFM_ERR_NOT_READY, "server_release: request %llu still in flight", (unsigned long long)id);
        if (q->holds) {
            s->arena.release(q->off, q->bytes);
            q->holds = false;
            s->stats.arena_in_use = s->arena.used();
        }
        s->reqs.erase(it);
        delete q;
    }
    s->cv_work.notify_one();  // freed memory may admit the queue head
    return RKFM_OK;
}

int rkfm_server_stats_get(rkfm_server *s, rkfm_server_stats *st) {
    if (!s || !st) return rkf_set_error(RKFM_ERR_INVALID_ARG, "server_stats: null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    *st = s->stats;
    st->queued = (uint32_t)s->pending.size();
    st->admitted = (uint32_t)s->admitted.size();
    st->running = s->running;
    return RKFM_OK;
}

double rkfm_server_now(rkfm_server *s) { return s ? s->now() : 0.0; }

int rkfm_server_destroy(rkfm_server *s) {
    if (!s) return RKFM_OK;
    if (s->worker.joinable()) {
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->stop = true;
        }
        s->cv_work.notify_all();
        s->worker.join();
    }
    for (uint32_t i = 0; i < s->n_slots; i++) {
        Slot &sl = s->slots[i];
        if (sl.stream) {
            (void)hipStreamSynchronize(sl.stream);
            (void)hipStreamDestroy(sl.stream);
        }
        if (sl.ev_start) (void)hipEventDestroy(sl.ev_start);
        if (sl.ev_end) (void)hipEventDestroy(sl.ev_end);
        if (sl.h_items) (void)hipHostFree(sl.h_items);
        if (sl.h_begins) (void)hipHostFree(sl.h_begins);
        sl.d_items.release();
        sl.d_begins.release();
        sl.d_counter.release();
    }
    if (s->arena_base) (void)hipFree(s->arena_base);
    for (auto &kv : s->reqs) delete kv.second;
    delete s;
    return RKFM_OK;
}

int rkfm_server_replay(rkfm_server *s, const rkfm_request *reqs, const double *arrival, uint32_t n,
                       double *latency, double *wall) {
    if (!s || (n && (!reqs || !arrival || !latency)))
        return rkf_set_error(RKFM_ERR_INVALID_ARG, "server_replay: null argument");
    for (uint32_t i = 1; i < n; i++)
        if (arrival[i] < arrival[i - 1]) return rkf_set_error(RKFM_ERR_INVALID_ARG, "server_replay: arrivals not ascending");
    const double nan = std::numeric_limits<double>::quiet_NaN();
    const double t0 = s->now();
    std::vector<std::pair<uint64_t, uint32_t>> outstanding;  // (id, index)
    uint32_t next = 0, done = 0;
    uint64_t seen = ~0ull;
    double last = t0;
    while (done < n) {
        // submit what is due
        double t = s->now();
        while (next < n && t - t0 >= arrival[next]) {
            uint64_t id = 0;
            int rc = rkfm_server_submit(s, &reqs[next], &id);
            if (rc) {
                latency[next] = nan;
                done++;
            } else {
                outstanding.emplace_back(id, next);
            }
            next++;
        }
        // collect completions
        std::unique_lock<std::mutex> lk(s->mu);
        if (s->completions != seen) {
            seen = s->completions;
            size_t w = 0;
            std::vector<uint64_t> finished;
            for (size_t k = 0; k < outstanding.size(); k++) {
                Req *q = s->reqs[outstanding[k].first];
                if (q->st.state == RKFM_REQ_DONE || q->st.state == RKFM_REQ_FAILED) {
                    uint32_t i = outstanding[k].second;
                    latency[i] = q->st.state == RKFM_REQ_DONE ? q->st.t_done - (t0 + arrival[i]) : nan;
                    last = std::max(last, q->st.t_done);
                    finished.push_back(outstanding[k].first);
                    done++;
                } else {
                    outstanding[w++] = outstanding[k];
                }
            }
            outstanding.resize(w);
            lk.unlock();
            for (uint64_t id : finished) rkfm_server_release(s, id);
            continue;
        }
        if (done >= n) break;
        // wait for the next arrival or completion
        double wait = next < n ? t0 + arrival[next] - s->now() : 1e-3;
        if (wait > 2e-4) {
            s->cv_done.wait_for(lk, std::chrono::duration<double>(wait - 1e-4));
        } else {
            lk.unlock();
            std::this_thread::yield();
        }
    }
    if (wall) *wall = last - t0;
    return RKFM_OK;
}

}  // extern "C"
