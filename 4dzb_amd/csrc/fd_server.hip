// Serving form: a continuous-batching loop over resident scenes.
//
// Clients submit shadow requests (receivers + light + outputs, uneven sizes)
// at any time.  A worker thread admits pending requests in arrival order as
// soon as (a) one of `streams` batch slots is free and (b) the request's
// workspace fits in the HBM pool, packs everything admissible into one batch
// (one launch per stage over all of its instances, the same kernels as the
// batched form) and enqueues it on that slot's stream.  Batches on different
// streams run concurrently; completed batches release their workspace, which
// admits the next requests.  Each result is bit-identical to running the
// request alone through fd_launch (same arithmetic, per-request workspace).
//
// Scene moves (FD::moveModel on a resident scene) travel in the same
// arrival-ordered queue: a move waits until every running batch that holds
// a request of its scene has finished, transforms the scene's vertices in
// place, and only then are later requests admitted, so each request sees the
// geometry of its submission time.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/fd4d.h"
#include "fd_host.h"
#include "fd_kernels.h"
#include "fd_types.h"

using namespace fdh;

namespace {

// First-fit allocator over one device arena (offsets, 256-byte granules).
class Pool {
public:
    void init(size_t bytes) { free_.clear(); free_[0] = bytes; size_ = bytes; used_ = 0; }
    bool alloc(size_t bytes, size_t& off) {
        bytes = (bytes + 255) & ~(size_t)255;
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= bytes) {
                off = it->first;
                const size_t rest = it->second - bytes;
                free_.erase(it);
                if (rest) free_[off + bytes] = rest;
                used_ += bytes;
                return true;
            }
        }
        return false;
    }
    void release(size_t off, size_t bytes) {
        bytes = (bytes + 255) & ~(size_t)255;
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
    size_t size() const { return size_; }
    size_t used() const { return used_; }

private:
    std::map<size_t, size_t> free_;
    size_t size_ = 0, used_ = 0;
};

struct Req {
    long long id;
    int kind = 0;          // 0: shadow request, 1: scene move
    int moveScene = -1;
    float m[16];
    fd_request r;
    FdArgs args;
    size_t bytes = 0, off = 0;
    double tSubmit = 0, tStart = 0;
    int batchSize = 0;
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    DevBuf table;
    char* staging = nullptr;      // pinned: plan image
    FdState* states = nullptr;    // pinned: per-request state copies
    std::vector<Req> reqs;
    bool busy = false;
};

}  // namespace

struct fd_server {
    fd_server_config cfg{};
    std::vector<fd_tune_entry> tune;   // per-class layouts (fd_autotune)
    std::vector<Model> scenes;
    DevBuf poolMem;
    Pool pool;
    std::vector<Slot> slots;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<Req> incoming, pending;
    std::vector<fd_completion> done;
    long long nextId = 0, submitted = 0, completed = 0;
    std::atomic<bool> stop{false};
    std::thread worker;
    std::chrono::steady_clock::time_point t0;
    std::atomic<int> error{0};
    int device = 0;
    hipStream_t moveStream = nullptr;

    double now() const {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    void run();
    bool launch(Slot& s, std::vector<Req>& batch);
    bool try_move();
};

// The scene move at the head of the queue, once no running batch holds a
// request of its scene.  Returns true when it ran.
bool fd_server::try_move() {
    const Req& q = pending.front();
    for (const Slot& sl : slots)
        if (sl.busy)
            for (const Req& r : sl.reqs)
                if (r.r.scene == q.moveScene) return false;
    Model md;
    {
        std::lock_guard<std::mutex> lk(mu);   // fd_server_add_scene may grow scenes concurrently
        md = scenes[q.moveScene];
    }
    hipError_t e = move_vertices((float*)md.verts, md.vstride, md.nverts, q.m, moveStream);
    if (e == hipSuccess) e = hipStreamSynchronize(moveStream);
    if (e != hipSuccess) { error = FD_ERR_HIP; stop = true; cv.notify_all(); return false; }
    fd_completion c{};
    c.id = q.id;
    c.user_tag = q.r.user_tag;
    c.status = FD_OK;
    c.t_submit_ms = q.tSubmit;
    c.t_start_ms = c.t_done_ms = now();
    pending.pop_front();
    std::lock_guard<std::mutex> lk(mu);
    done.push_back(c);
    ++completed;
    cv.notify_all();
    return true;
}

bool fd_server::launch(Slot& s, std::vector<Req>& batch) {
    std::vector<FdArgs> args;
    args.reserve(batch.size());
    for (Req& q : batch) args.push_back(q.args);
    std::vector<char> image;
    FdBatchPlan P;
    bool anyTris = false;
    for (const Req& q : batch) anyTris |= q.args.ntris > 0;
    build_plan(args, anyTris, (char*)s.table.p, image, P);
    std::memcpy(s.staging, image.data(), image.size());
    hipError_t e = hipMemcpyAsync(s.table.p, s.staging, image.size(), hipMemcpyHostToDevice, s.stream);
    if (e == hipSuccess) e = fdk::enqueue_batch(P, s.stream);
    for (size_t i = 0; i < batch.size() && e == hipSuccess; ++i)
        e = hipMemcpyAsync(&s.states[i], batch[i].args.st, sizeof(FdState), hipMemcpyDeviceToHost, s.stream);
    if (e == hipSuccess) e = hipEventRecord(s.done, s.stream);
    if (e != hipSuccess) return false;
    const double t = now();
    for (Req& q : batch) { q.tStart = t; q.batchSize = (int)batch.size(); }
    s.reqs = std::move(batch);
    s.busy = true;
    return true;
}

void fd_server::run() {
    (void)hipSetDevice(device);
    while (!stop.load()) {
        bool progress = false;
        {
            std::unique_lock<std::mutex> lk(mu);
            while (!incoming.empty()) { pending.push_back(std::move(incoming.front())); incoming.pop_front(); }
        }
        // retire finished batches
        for (Slot& s : slots) {
            if (!s.busy || hipEventQuery(s.done) != hipSuccess) continue;
            const double t = now();
            std::vector<fd_completion> out;
            for (size_t i = 0; i < s.reqs.size(); ++i) {
                const Req& q = s.reqs[i];
                fd_completion c{};
                c.id = q.id;
                c.user_tag = q.r.user_tag;
                c.exit_code = s.states[i].status;
                c.status = c.exit_code == FD_EXIT_CAPACITY ? FD_ERR_CAPACITY : FD_OK;
                c.t_submit_ms = q.tSubmit;
                c.t_start_ms = q.tStart;
                c.t_done_ms = t;
                c.batch_size = q.batchSize;
                c.pairs = s.states[i].pairs;
                out.push_back(c);
                pool.release(q.off, q.bytes);
            }
            s.reqs.clear();
            s.busy = false;
            progress = true;
            std::lock_guard<std::mutex> lk(mu);
            for (auto& c : out) done.push_back(c);
            completed += (long long)out.size();
            cv.notify_all();
        }
        // a scene move at the head of the queue runs once its scene is idle
        if (!pending.empty() && pending.front().kind == 1 && try_move()) progress = true;
        // admit and launch, in arrival order (a move ends the batch)
        for (Slot& s : slots) {
            if (s.busy || pending.empty() || pending.front().kind == 1) continue;
            std::vector<Req> batch;
            while (!pending.empty() && pending.front().kind == 0 && (int)batch.size() < cfg.max_batch) {
                Req& q = pending.front();
                size_t off;
                if (!pool.alloc(q.bytes, off)) break;   // wait for memory: keep arrival order
                q.off = off;
                layout(q.args, (char*)poolMem.p + off, false);
                batch.push_back(std::move(q));
                pending.pop_front();
            }
            if (batch.empty()) break;
            if (!launch(s, batch)) { error = FD_ERR_HIP; stop = true; cv.notify_all(); break; }
            progress = true;
        }
        if (!progress) std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
}

extern "C" {

int fd_server_create(const fd_server_config* cfg, fd_server** out) {
    if (!cfg || !out || cfg->hbm_budget_bytes < (64u << 20)) return FD_ERR_ARG;
    int nd = 0;
    if (hipGetDeviceCount(&nd) != hipSuccess || nd <= 0) return FD_ERR_NO_DEVICE;
    fd_server* s = new fd_server();
    s->cfg = *cfg;
    (void)hipGetDevice(&s->device);
    if (s->cfg.streams <= 0) s->cfg.streams = 4;
    if (s->cfg.max_batch <= 0) s->cfg.max_batch = 16;
    if (s->cfg.bins_per_side == 0) s->cfg.bins_per_side = 16;
    if (s->cfg.tile_width == 0) s->cfg.tile_width = 8;
    if (s->cfg.ntune < 0 || (s->cfg.ntune > 0 && !s->cfg.tune)) { delete s; return FD_ERR_ARG; }
    s->tune.assign(s->cfg.tune, s->cfg.tune + s->cfg.ntune);
    s->cfg.tune = nullptr;
    if (!s->cfg.max_pairs) s->cfg.max_pairs = 16ull << 20;
    hipError_t e = s->poolMem.ensure(cfg->hbm_budget_bytes);
    if (e != hipSuccess) { delete s; return hip_status(e); }
    s->pool.init(cfg->hbm_budget_bytes);
    s->slots.resize(s->cfg.streams);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->moveStream, hipStreamNonBlocking);
    for (Slot& sl : s->slots) {
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&sl.done, hipEventDisableTiming);
        if (e == hipSuccess) e = sl.table.ensure(plan_bytes(s->cfg.max_batch));
        if (e == hipSuccess) e = hipHostMalloc((void**)&sl.staging, plan_bytes(s->cfg.max_batch), 0);
        if (e == hipSuccess) e = hipHostMalloc((void**)&sl.states, sizeof(FdState) * s->cfg.max_batch, 0);
    }
    if (e != hipSuccess) { fd_server_destroy(s); return hip_status(e); }
    s->t0 = std::chrono::steady_clock::now();
    s->worker = std::thread([s] { s->run(); });
    *out = s;
    return FD_OK;
}

int fd_server_add_scene(fd_server* s, const float* v, size_t vsz, int nv, const int* idx, int nt) {
    if (!s || !v || !idx || nv <= 0 || nt < 0 || vsz < 12 || vsz % 4) return FD_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    s->scenes.push_back(Model{v, (int)(vsz / 4), nv, idx, nt});
    return (int)s->scenes.size() - 1;
}

long long fd_server_submit(fd_server* s, const fd_request* r) {
    if (!s || !r || !r->d_position4 || !r->d_normal4 || !r->d_shadow || r->width <= 0 || r->height <= 0)
        return FD_ERR_ARG;
    if (r->light.res_w != r->light.res_h || r->light.res_w < 1 || r->light.res_w > 64) return FD_ERR_UNSUPPORTED_LIGHT;
    Req q;
    Model md;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (r->scene < 0 || r->scene >= (int)s->scenes.size()) return FD_ERR_ARG;
        md = s->scenes[r->scene];   // pointers and counts only: moves transform the buffers in place
    }
    q.r = *r;
    q.tSubmit = s->now();
    FdLight L;
    light_setup(r->light, L);
    // layout of this request's class: the tuning table, else the defaults
    int bins = s->cfg.bins_per_side, tw = s->cfg.tile_width, waves = 0;   // 0: the default, once the frame is laid out
    int order = FD_TRI_ORDER_AUTO, split = 0;
    for (const fd_tune_entry& t : s->tune)
        if ((t.scene < 0 || t.scene == r->scene) && t.width == r->width && t.height == r->height &&
            t.light_res == r->light.res_w && t.ms >= 0.0f) {
            bins = t.bins_per_side;
            tw = t.tile_width;
            if (fd_valid_waves(t.occ_waves)) waves = t.occ_waves;
            if (t.tri_order >= FD_TRI_ORDER_LIST && t.tri_order <= FD_TRI_ORDER_MESH) order = t.tri_order;
            if (t.list_split > 0 && t.list_split <= FD_LIST_SPLIT_MAX) split = t.list_split;
            break;
        }
    int g = 0, tl = 0;
    while ((1 << g) < bins) ++g;
    while ((1 << tl) < tw) ++tl;
    if ((1 << g) != bins || g < 3 || g > 6 || (1 << tl) != tw || tl < 3 || tl > 6) return FD_ERR_ARG;
    fill_args(q.args, md, (const float4*)r->d_position4, (const float4*)r->d_normal4, r->width,
              r->height, L, s->cfg.eps, s->cfg.backface_cull, g, tl, r->d_shadow, r->d_masks, s->cfg.max_pairs);
    q.args.occWaves = waves ? waves : fd_default_waves(L.wpr, q.args.chunkSlots);
    q.args.triOrder = order;
    apply_split(q.args, split);
    FdArgs probe = q.args;
    q.bytes = layout(probe, nullptr, false);
    if (q.bytes > s->pool.size()) return FD_ERR_CAPACITY;   // can never be admitted
    // the id and the queue position are taken in one critical section, so
    // that a concurrent move is ordered either before or after this request
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->error.load()) return s->error.load();
    q.id = s->nextId++;
    const long long id = q.id;
    s->incoming.push_back(std::move(q));
    ++s->submitted;
    return id;
}

int fd_server_poll(fd_server* s, fd_completion* out, int max) {
    if (!s || (!out && max > 0)) return FD_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    int n = std::min<int>(max, (int)s->done.size());
    for (int i = 0; i < n; ++i) out[i] = s->done[i];
    s->done.erase(s->done.begin(), s->done.begin() + n);
    return n;
}

int fd_server_drain(fd_server* s, int timeout_ms) {
    if (!s) return FD_ERR_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    const bool ok = s->cv.wait_for(lk, std::chrono::milliseconds(timeout_ms),
                                   [s] { return s->completed >= s->submitted || s->error.load(); });
    if (s->error.load()) return s->error.load();
    return ok ? FD_OK : FD_ERR_STATE;
}

long long fd_server_move_scene(fd_server* s, int scene, const float m[16], unsigned long long user_tag) {
    if (!s || !m) return FD_ERR_ARG;
    Req q;
    q.kind = 1;
    std::memcpy(q.m, m, sizeof(q.m));
    q.r = fd_request{};
    q.r.user_tag = user_tag;
    q.r.scene = -1;
    std::lock_guard<std::mutex> lk(s->mu);
    if (scene < 0 || scene >= (int)s->scenes.size()) return FD_ERR_ARG;
    if (s->error.load()) return s->error.load();
    q.moveScene = scene;
    q.id = s->nextId++;
    q.tSubmit = s->now();
    const long long id = q.id;
    s->incoming.push_back(std::move(q));
    ++s->submitted;
    return id;
}

double fd_server_now_ms(fd_server* s) { return s ? s->now() : 0.0; }

int fd_server_destroy(fd_server* s) {
    if (!s) return FD_ERR_ARG;
    s->stop = true;
    if (s->worker.joinable()) s->worker.join();
    (void)hipDeviceSynchronize();
    for (Slot& sl : s->slots) {
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
        if (sl.done) (void)hipEventDestroy(sl.done);
        sl.table.release();
        if (sl.staging) (void)hipHostFree(sl.staging);
        if (sl.states) (void)hipHostFree(sl.states);
    }
    if (s->moveStream) (void)hipStreamDestroy(s->moveStream);
    s->poolMem.release();
    delete s;
    return FD_OK;
}

}  // extern "C"
