// Serving loop: continuous batching of tessellation requests on one GPU.
//
// Requests (one path + options, copied at submit) wait in a FIFO.  A dispatcher thread
// admits them into batches of one kind and one size class (factor 4 in flattened points)
// while their worst-case output fits the unreserved part of the
// device pool, and hands each batch to an idle worker.  Every worker owns a context (its
// own HIP stream), so up to `max_inflight` batches are on the device at once: requests
// that arrive while others run are batched as soon as a worker and pool room free up.
// A batch runs the batched kernels (stroke: staged emit, scan, compaction; fill: sweep
// rounds, scan, copy); once its totals are known, its output is allocated from the pool at its exact
// size and the reservation is returned.  Results stay in the pool (device pointers,
// request-local vertex ids) until the client releases them.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "ctx.h"
#include "internal.h"
#include "lyg_math.h"

using Clock = std::chrono::steady_clock;

namespace {

// First-fit allocator over one device allocation, with coalescing on free.
class Pool {
  public:
    bool init(size_t bytes) {
        if (hipMalloc(&base_, bytes) != hipSuccess) return false;
        size_ = bytes;
        free_[0] = bytes;
        return true;
    }
    void destroy() {
        if (base_) (void)hipFree(base_);
        base_ = nullptr;
    }
    // Returns an offset or SIZE_MAX.
    size_t alloc(size_t bytes) {
        bytes = (bytes + 255) & ~(size_t)255;
        if (bytes == 0) bytes = 256;
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= bytes) {
                size_t off = it->first, len = it->second;
                free_.erase(it);
                if (len > bytes) free_[off + bytes] = len - bytes;
                used_ += bytes;
                high_ = std::max(high_, used_);
                live_[off] = bytes;
                return off;
            }
        }
        return SIZE_MAX;
    }
    void release(size_t off) {
        auto l = live_.find(off);
        if (l == live_.end()) return;
        size_t len = l->second;
        live_.erase(l);
        used_ -= len;
        auto it = free_.emplace(off, len).first;
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
    uint8_t* ptr(size_t off) const { return static_cast<uint8_t*>(base_) + off; }
    size_t size() const { return size_; }
    size_t used() const { return used_; }
    size_t high_water() const { return high_; }

  private:
    void* base_ = nullptr;
    size_t size_ = 0, used_ = 0, high_ = 0;
    std::map<size_t, size_t> free_, live_;
};

enum Kind : uint32_t { K_FILL = 0, K_STROKE = 1 };
enum State : uint32_t { S_QUEUED = 0, S_RUNNING = 1, S_DONE = 2 };

struct Request {
    uint64_t ticket = 0;
    Kind kind = K_FILL;
    std::vector<uint8_t> verbs;
    std::vector<float> points;
    lyg_fill_options fopts{};
    lyg_stroke_options sopts{};
    uint64_t reserve = 0;        // worst-case output bytes (admission)
    uint32_t size_class = 0;     // log4 of the flattened-point bound: batches mix one class
    bool banded = false;         // a fill large enough for the banded fill
    double crossings = 0;        // fill: estimated crossings of its edges (admission)
    Clock::time_point t_submit, t_done;
    State state = S_QUEUED;
    int32_t status = LYG_OK;
    size_t out_off = SIZE_MAX;   // pool allocation holding the batch output
    std::shared_ptr<size_t> batch_refs;  // requests of the batch not yet released
    size_t v_off = 0, i_off = 0;  // this request's vertices / indices in the pool (bytes)
    uint32_t nv = 0, ni = 0;
};

// Upper bound of the flattened points of a path at `tol` (Wang's counts, +1 per curve
// for the split quadratics of the stroker) plus one per line / sub-path end.
uint64_t flattened_bound(const std::vector<uint8_t>& verbs, const std::vector<float>& pts,
                         float tol) {
    uint64_t f = 0;
    size_t p = 0;
    lyg::vec2 cur = lyg::mk(0, 0);
    auto at = [&](size_t i) {
        return i < pts.size() / 2 ? lyg::mk(pts[2 * i], pts[2 * i + 1]) : lyg::mk(0, 0);
    };
    for (uint8_t v : verbs) {
        switch (v) {
            case LYG_VERB_BEGIN: cur = at(p++); f += 1; break;
            case LYG_VERB_LINE: cur = at(p++); f += 1; break;
            case LYG_VERB_QUADRATIC: {
                lyg::vec2 c = at(p), t = at(p + 1);
                p += 2;
                f += 2 * (uint64_t)lyg::seg_count(lyg::quad_wang(cur, c, t, tol)) + 2;
                cur = t;
                break;
            }
            case LYG_VERB_CUBIC: {
                lyg::vec2 c1 = at(p), c2 = at(p + 1), t = at(p + 2);
                p += 3;
                f += (uint64_t)lyg::seg_count(lyg::cubic_wang(cur, c1, c2, t, tol)) + 1;
                cur = t;
                break;
            }
            default:
                if (v == LYG_VERB_CLOSE) p++;
                f += 2;
                break;
        }
    }
    return f;
}

// Estimated crossings between the path's segments: over the pairs of segments whose
// control-point boxes overlap (the only pairs that can cross), Bezout's count (curves of
// degrees m and n cross at most m n times), plus one per cubic (a cubic crosses itself at
// most once).  Pairs are counted exactly up to 4,096 of them; beyond, 4,096 pairs drawn
// with a fixed seed stand for all.  The flattened polylines follow the curves within the
// tolerance, so this is the crossing count the fill's admission reserves room for.
double crossing_estimate(const std::vector<uint8_t>& verbs, const std::vector<float>& pts) {
    struct Seg {
        float x0, y0, x1, y1;
        float deg;
    };
    std::vector<Seg> seg;
    seg.reserve(verbs.size());
    size_t p = 0;
    float cx = 0, cy = 0, fx = 0, fy = 0;
    double self = 0;
    auto at = [&](size_t i, float* x, float* y) {
        *x = i < pts.size() / 2 ? pts[2 * i] : 0.0f;
        *y = i < pts.size() / 2 ? pts[2 * i + 1] : 0.0f;
    };
    for (uint8_t v : verbs) {
        int np = v == LYG_VERB_LINE ? 1 : v == LYG_VERB_QUADRATIC ? 2 : v == LYG_VERB_CUBIC ? 3 : 0;
        if (v == LYG_VERB_BEGIN) {
            at(p++, &cx, &cy);
            fx = cx;
            fy = cy;
            continue;
        }
        Seg s{cx, cy, cx, cy, (float)np};
        if (v == LYG_VERB_CLOSE || v == LYG_VERB_END) {
            // the closing edge back to the sub-path's first point
            s = Seg{std::min(cx, fx), std::min(cy, fy), std::max(cx, fx), std::max(cy, fy), 1.0f};
            if (v == LYG_VERB_CLOSE) p++;
            seg.push_back(s);
            cx = fx;
            cy = fy;
            continue;
        }
        for (int k = 0; k < np; ++k) {
            float x, y;
            at(p++, &x, &y);
            s.x0 = std::min(s.x0, x);
            s.y0 = std::min(s.y0, y);
            s.x1 = std::max(s.x1, x);
            s.y1 = std::max(s.y1, y);
            cx = x;
            cy = y;
        }
        if (v == LYG_VERB_CUBIC) self += 1.0;
        seg.push_back(s);
    }
    const size_t n = seg.size();
    if (n < 2) return self;
    auto term = [&](size_t i, size_t j) {
        const Seg& a = seg[i];
        const Seg& b = seg[j];
        const bool hit = a.x0 <= b.x1 && b.x0 <= a.x1 && a.y0 <= b.y1 && b.y0 <= a.y1;
        return hit ? (double)a.deg * (double)b.deg : 0.0;
    };
    const double pairs = 0.5 * (double)n * (double)(n - 1);
    double sum = 0;
    if (pairs <= 4096.0) {
        for (size_t i = 0; i < n; ++i)
            for (size_t j = i + 1; j < n; ++j) sum += term(i, j);
        return sum + self;
    }
    uint64_t st = 0x9e3779b97f4a7c15ull ^ (uint64_t)n;
    auto next = [&]() {
        st = st * 6364136223846793005ull + 1442695040888963407ull;
        return (uint32_t)(st >> 33);
    };
    for (int k = 0; k < 4096; ++k) {
        size_t i = next() % n, j = next() % (n - 1);
        if (j >= i) ++j;
        sum += term(i, j);
    }
    return sum / 4096.0 * pairs + self;
}

// Worst-case output bytes of a request, for admission.  Stroke: per flattened point at
// most 4 join vertices plus a round join's arc (2^k vertices for an arc of up to pi at the
// circle flattening step of the half width) and as many triangles plus 2.  Fill: per event
// of the path (edge ends and the estimated crossings) two vertices and four triangles of
// the sweep, or the banded fill's three trapezoids of four vertices and two triangles.
uint64_t worst_case_bytes(Request& r) {
    if (r.kind == K_STROKE) {
        const float tol = r.sopts.tolerance > 0 ? r.sopts.tolerance : 0.1f;
        const uint64_t f = flattened_bound(r.verbs, r.points, tol);
        double hw = 0.5 * (double)r.sopts.line_width;
        double arc = 2.0;
        if (r.sopts.line_join == LYG_JOIN_ROUND || r.sopts.start_cap == LYG_CAP_ROUND ||
            r.sopts.end_cap == LYG_CAP_ROUND) {
            double step = hw > tol ? 2.0 * std::acos(1.0 - (double)tol / hw) : 3.14159;
            double n = std::ceil(3.14159265 / std::max(step, 1e-6));
            arc = std::pow(2.0, std::ceil(std::log2(std::max(n, 1.0)))) + 2.0;
        }
        const double nv = (double)f * (4.0 + arc) + 64.0;
        const double nt = nv + 2.0 * (double)f + 64.0;
        return (uint64_t)(8.0 * nv + 12.0 * nt);
    }
    const float tol = r.fopts.tolerance > 0 ? r.fopts.tolerance : 0.1f;
    const double e = 2.0 * (double)flattened_bound(r.verbs, r.points, tol) + 16.0 +
                     2.0 * r.crossings;
    const double per_event = r.banded ? 3.0 * (4.0 * 8.0 + 2.0 * 12.0) : 2.0 * 8.0 + 4.0 * 12.0;
    const double b = per_event * e + 8.0 * 256 + 12.0 * 512;
    return b < 1.8e19 ? (uint64_t)b : UINT64_MAX;
}

}  // namespace

struct lyg_server {
    lyg_server_config cfg{};
    Pool pool;
    std::mutex mu;
    std::condition_variable cv_dispatch, cv_done;
    std::deque<Request*> pending;
    std::unordered_map<uint64_t, std::unique_ptr<Request>> reqs;
    uint64_t next_ticket = 1;
    uint64_t reserved = 0;
    bool stop = false;
    // stats: the latencies of the last LAT_WINDOW completed requests, as a ring
    static constexpr uint64_t LAT_WINDOW = 1ull << 20;
    std::vector<double> lat_ring;
    uint64_t lat_count = 0;
    void record_latency(double us) {
        if (lat_ring.size() < LAT_WINDOW) lat_ring.push_back(us);
        else lat_ring[lat_count % LAT_WINDOW] = us;
        ++lat_count;
    }
    // The latencies recorded after the first `since` ones (at most the last LAT_WINDOW).
    std::vector<double> latencies_since(uint64_t since) const {
        const uint64_t n = std::min<uint64_t>(lat_count - std::min(since, lat_count), lat_ring.size());
        std::vector<double> l(n);
        for (uint64_t k = 0; k < n; ++k) l[k] = lat_ring[(lat_count - n + k) % LAT_WINDOW];
        return l;
    }
    uint64_t completed = 0, triangles = 0, vertices = 0, batches = 0, batched_requests = 0;
    uint64_t rejected = 0;
    Clock::time_point t_first_submit{}, t_last_done{};
    bool any_submit = false;

    struct Worker {
        lyg_ctx* ctx = nullptr;
        std::thread th;
        std::vector<Request*> batch;
        bool busy = false;
        std::condition_variable cv;
        // per-worker device inputs (grow-only) and pinned staging
        DevBuf in_verbs, in_points, in_offs, in_opts, offs;
        std::vector<uint8_t> h_verbs;
        std::vector<float> h_points;
        std::vector<uint32_t> h_offs;
    };
    std::vector<std::unique_ptr<Worker>> workers;
    std::thread dispatcher;

    void run_dispatcher();
    void run_worker(Worker* w);
    void run_batch(Worker* w);
};

void lyg_server::run_dispatcher() {
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
        cv_dispatch.wait(lk, [&] {
            if (stop) return true;
            if (pending.empty()) return false;
            for (auto& w : workers)
                if (!w->busy) return true;
            return false;
        });
        if (stop) return;
        Worker* idle = nullptr;
        for (auto& w : workers)
            if (!w->busy) {
                idle = w.get();
                break;
            }
        // let a batch grow for up to linger_us while the queue is short
        if (cfg.linger_us && pending.size() < cfg.max_batch_paths) {
            lk.unlock();
            std::this_thread::sleep_for(std::chrono::microseconds(cfg.linger_us));
            lk.lock();
            if (stop) return;
        }
        // admit FIFO requests of the oldest request's kind and size class while they fit
        // the pool (a batch completes with its longest path: mixing sizes would make the
        // short requests wait for the long ones)
        const Kind kind = pending.front()->kind;
        const uint32_t cls = pending.front()->size_class;
        std::vector<Request*> batch;
        for (auto it = pending.begin(); it != pending.end() && batch.size() < cfg.max_batch_paths;) {
            Request* r = *it;
            if (r->kind != kind || r->size_class != cls || r->banded != pending.front()->banded) {
                ++it;
                continue;
            }
            // results not yet released hold pool space too; a request larger than the whole
            // pool runs alone once the pool is empty
            const bool empty = batch.empty() && reserved == 0 && pool.used() == 0;
            if (pool.used() + reserved + r->reserve > pool.size() && !empty) break;
            reserved += r->reserve;
            r->state = S_RUNNING;
            batch.push_back(r);
            it = pending.erase(it);
        }
        if (batch.empty()) {
            // the oldest request waits for pool room: wake on the next completion/release
            cv_dispatch.wait(lk);
            continue;
        }
        idle->batch.swap(batch);
        idle->busy = true;
        idle->cv.notify_one();
    }
}

void lyg_server::run_worker(Worker* w) {
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
        w->cv.wait(lk, [&] { return stop || w->busy; });
        if (stop) return;
        lk.unlock();
        run_batch(w);
        lk.lock();
        w->busy = false;
        w->batch.clear();
        cv_dispatch.notify_one();
    }
}

void lyg_server::run_batch(Worker* w) {
    lyg_ctx* c = w->ctx;
    std::vector<Request*>& B = w->batch;
    const uint32_t n = (uint32_t)B.size();
    const Kind kind = B[0]->kind;
    int32_t st = LYG_OK;
    // pack inputs
    w->h_verbs.clear();
    w->h_points.clear();
    w->h_offs.assign(2 * (n + 1), 0);
    std::vector<uint8_t> h_opts;
    const size_t osz = kind == K_FILL ? sizeof(lyg_fill_options) : sizeof(lyg_stroke_options);
    h_opts.resize(osz * n);
    for (uint32_t i = 0; i < n; ++i) {
        Request* r = B[i];
        w->h_offs[i] = (uint32_t)w->h_verbs.size();
        w->h_offs[n + 1 + i] = (uint32_t)(w->h_points.size() / 2);
        w->h_verbs.insert(w->h_verbs.end(), r->verbs.begin(), r->verbs.end());
        w->h_points.insert(w->h_points.end(), r->points.begin(), r->points.end());
        std::memcpy(h_opts.data() + osz * i, kind == K_FILL ? (const void*)&r->fopts : (const void*)&r->sopts, osz);
    }
    w->h_offs[n] = (uint32_t)w->h_verbs.size();
    w->h_offs[2 * n + 1] = (uint32_t)(w->h_points.size() / 2);
    hipStream_t s = c->stream;
    if (!ensure(c, w->in_verbs, w->h_verbs.size()) || !ensure(c, w->in_points, 4 * w->h_points.size()) ||
        !ensure(c, w->in_offs, 4 * w->h_offs.size()) || !ensure(c, w->in_opts, h_opts.size()) ||
        !ensure(c, w->offs, 8ull * (n + 1)) ||
        hipMemcpyAsync(w->in_verbs.p, w->h_verbs.data(), w->h_verbs.size(), hipMemcpyHostToDevice, s) != hipSuccess ||
        (w->h_points.size() &&
         hipMemcpyAsync(w->in_points.p, w->h_points.data(), 4 * w->h_points.size(), hipMemcpyHostToDevice, s) != hipSuccess) ||
        hipMemcpyAsync(w->in_offs.p, w->h_offs.data(), 4 * w->h_offs.size(), hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(w->in_opts.p, h_opts.data(), h_opts.size(), hipMemcpyHostToDevice, s) != hipSuccess)
        st = LYG_ERR_DEVICE;
    lygi::BatchSpec b{};
    b.in.verbs = as<uint8_t>(w->in_verbs);
    b.in.points = as<float>(w->in_points);
    b.in.verb_off = as<uint32_t>(w->in_offs);
    b.in.point_off = as<uint32_t>(w->in_offs) + (n + 1);
    b.in.n_paths = n;
    b.opts = w->in_opts.p;
    b.opts_stride = 1;
    b.vertex_offset = 0;
    b.local_ids = true;
    b.vtx_off = as<uint32_t>(w->offs);
    b.idx_off = as<uint32_t>(w->offs) + (n + 1);
    size_t out_off = SIZE_MAX;
    auto alloc = [&](uint64_t nv, uint64_t ni, float** v, uint32_t** i) -> int32_t {
        const size_t vb = (8 * nv + 255) & ~(size_t)255;
        const size_t bytes = vb + 4 * ni;
        std::unique_lock<std::mutex> lk(mu);
        // wait (bounded) for room freed by releases
        for (int tries = 0; tries < 2000; ++tries) {
            out_off = pool.alloc(bytes);
            if (out_off != SIZE_MAX) break;
            cv_done.wait_for(lk, std::chrono::milliseconds(5));
        }
        if (out_off == SIZE_MAX) return LYG_ERR_NO_SPACE;
        *v = reinterpret_cast<float*>(pool.ptr(out_off));
        *i = reinterpret_cast<uint32_t*>(pool.ptr(out_off + vb));
        return LYG_OK;
    };
    uint64_t tv = 0, ti = 0;
    std::vector<uint32_t> status;
    if (st == LYG_OK) {
        if (kind == K_STROKE) st = lygi::stroke_batch(c, b, alloc, &tv, &ti);
        else if (B[0]->banded) st = lygi::fill_batch_banded(c, b, alloc, &tv, &ti, &status);
        else st = lygi::fill_batch(c, b, alloc, &tv, &ti, &status);
    }
    std::vector<uint32_t> offs(2 * (n + 1), 0);
    if (st == LYG_OK &&
        (hipMemcpyAsync(offs.data(), w->offs.p, 8ull * (n + 1), hipMemcpyDeviceToHost, s) != hipSuccess ||
         hipStreamSynchronize(s) != hipSuccess))
        st = LYG_ERR_DEVICE;
    const Clock::time_point done = Clock::now();
    std::lock_guard<std::mutex> lk(mu);
    auto refs = std::make_shared<size_t>(n);
    const size_t vb = (8 * tv + 255) & ~(size_t)255;
    for (uint32_t i = 0; i < n; ++i) {
        Request* r = B[i];
        reserved -= r->reserve;
        r->t_done = done;
        r->state = S_DONE;
        r->out_off = out_off;
        r->batch_refs = refs;
        if (st != LYG_OK) {
            r->status = st;
        } else {
            r->status = kind == K_FILL && i < status.size() ? (int32_t)status[i] : LYG_OK;
            r->nv = offs[i + 1] - offs[i];
            r->ni = offs[n + 1 + i + 1] - offs[n + 1 + i];
            r->v_off = out_off + 8ull * offs[i];
            r->i_off = out_off + vb + 4ull * offs[n + 1 + i];
            triangles += r->ni / 3;
            vertices += r->nv;
        }
        record_latency(std::chrono::duration<double, std::micro>(done - r->t_submit).count());
        completed++;
    }
    if (st != LYG_OK && out_off != SIZE_MAX) {
        pool.release(out_off);
        for (Request* r : B) r->out_off = SIZE_MAX;
    }
    batches++;
    batched_requests += n;
    t_last_done = done;
    cv_done.notify_all();
    cv_dispatch.notify_one();
}

namespace {

int32_t submit(lyg_server* s, Kind kind, const lyg_path_view* path, const void* opts,
               uint64_t* ticket) {
    if (!s || !path || !opts || !ticket) return LYG_ERR_INVALID_ARGUMENT;
    if ((path->n_verbs && !path->verbs) || (path->n_points && !path->points))
        return LYG_ERR_INVALID_ARGUMENT;
    auto r = std::make_unique<Request>();
    r->kind = kind;
    r->verbs.assign(path->verbs, path->verbs + path->n_verbs);
    r->points.assign(path->points, path->points + 2ull * path->n_points);
    if (kind == K_FILL) r->fopts = *static_cast<const lyg_fill_options*>(opts);
    else r->sopts = *static_cast<const lyg_stroke_options*>(opts);
    {
        const float tol = kind == K_FILL ? r->fopts.tolerance : r->sopts.tolerance;
        const uint64_t f = flattened_bound(r->verbs, r->points, tol > 0 ? tol : 0.1f);
        r->size_class = (uint32_t)(f > 1 ? (63 - __builtin_clzll(f)) / 2 : 0);
        r->banded = kind == K_FILL && s->cfg.fill_banded_min_edges != 0 &&
                    f >= s->cfg.fill_banded_min_edges;
    }
    if (kind == K_FILL) r->crossings = crossing_estimate(r->verbs, r->points);
    r->reserve = worst_case_bytes(*r);
    const uint64_t limit = s->cfg.max_request_bytes ? s->cfg.max_request_bytes
                                                    : s->cfg.pool_bytes;
    const bool too_many_crossings = kind == K_FILL && s->cfg.max_fill_crossings != 0 &&
                                    r->crossings > (double)s->cfg.max_fill_crossings;
    r->t_submit = Clock::now();
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) return LYG_ERR_INVALID_ARGUMENT;
    r->ticket = s->next_ticket++;
    *ticket = r->ticket;
    if (!s->any_submit) {
        s->any_submit = true;
        s->t_first_submit = r->t_submit;
    }
    if (r->reserve > limit || too_many_crossings) {
        // refused at admission: complete at once, nothing reserved or allocated
        r->reserve = 0;
        r->status = LYG_ERR_NO_SPACE;
        r->state = S_DONE;
        r->t_done = r->t_submit;
        s->rejected++;
        s->reqs.emplace(r->ticket, std::move(r));
        s->cv_done.notify_all();
        return LYG_OK;
    }
    s->pending.push_back(r.get());
    s->reqs.emplace(r->ticket, std::move(r));
    s->cv_dispatch.notify_one();
    return LYG_OK;
}

void fill_result(lyg_server* s, const Request& r, lyg_server_result* out) {
    out->status = r.status;
    out->vertices = r.status == LYG_OK && r.nv ? reinterpret_cast<const float*>(s->pool.ptr(r.v_off)) : nullptr;
    out->indices = r.status == LYG_OK && r.ni ? reinterpret_cast<const uint32_t*>(s->pool.ptr(r.i_off)) : nullptr;
    out->n_vertices = r.status == LYG_OK ? r.nv : 0;
    out->n_indices = r.status == LYG_OK ? r.ni : 0;
    out->latency_us = std::chrono::duration<double, std::micro>(r.t_done - r.t_submit).count();
    out->banded = r.banded ? 1 : 0;
}

}  // namespace

extern "C" {

lyg_server* lyg_server_create(const lyg_server_config* cfg) {
    if (!cfg || cfg->pool_bytes == 0 || cfg->max_inflight == 0 || cfg->max_batch_paths == 0 ||
        cfg->stroke_sizing > 1)
        return nullptr;
    auto s = std::make_unique<lyg_server>();
    s->cfg = *cfg;
    if (hipSetDevice(cfg->device) != hipSuccess || !s->pool.init(cfg->pool_bytes)) return nullptr;
    for (uint32_t i = 0; i < cfg->max_inflight; ++i) {
        auto w = std::make_unique<lyg_server::Worker>();
        w->ctx = lygi::ctx_create(cfg->device, false);
        if (w->ctx && cfg->stroke_sizing == 1) lyg_ctx_set_stroke_staging(w->ctx, 0, 1.0f);
        if (w->ctx && cfg->stroke_split_windows)
            lyg_ctx_set_stroke_split(w->ctx, cfg->stroke_split_windows);
        if (!w->ctx) {
            for (auto& x : s->workers)
                if (x->ctx) lyg_ctx_destroy(x->ctx);
            s->pool.destroy();
            return nullptr;
        }
        s->workers.push_back(std::move(w));
    }
    lyg_server* sp = s.release();
    for (auto& w : sp->workers) {
        lyg_server::Worker* wp = w.get();
        wp->th = std::thread([sp, wp] {
            (void)hipSetDevice(sp->cfg.device);
            sp->run_worker(wp);
        });
    }
    sp->dispatcher = std::thread([sp] { sp->run_dispatcher(); });
    return sp;
}

void lyg_server_destroy(lyg_server* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_dispatch.notify_all();
    s->cv_done.notify_all();
    for (auto& w : s->workers) w->cv.notify_all();
    if (s->dispatcher.joinable()) s->dispatcher.join();
    for (auto& w : s->workers) {
        if (w->th.joinable()) w->th.join();
        for (DevBuf* b : {&w->in_verbs, &w->in_points, &w->in_offs, &w->in_opts, &w->offs})
            free_buf(*b);
        lyg_ctx_destroy(w->ctx);
    }
    s->pool.destroy();
    delete s;
}

int32_t lyg_server_submit_fill(lyg_server* s, const lyg_path_view* path,
                               const lyg_fill_options* opts, uint64_t* ticket) {
    if (opts && (opts->fill_rule > 1 || opts->sweep_orientation > 1)) return LYG_ERR_INVALID_ARGUMENT;
    return submit(s, K_FILL, path, opts, ticket);
}

int32_t lyg_server_submit_stroke(lyg_server* s, const lyg_path_view* path,
                                 const lyg_stroke_options* opts, uint64_t* ticket) {
    if (opts && (opts->start_cap > 2 || opts->end_cap > 2 || opts->line_join > 3))
        return LYG_ERR_INVALID_ARGUMENT;
    return submit(s, K_STROKE, path, opts, ticket);
}

int32_t lyg_server_poll(lyg_server* s, uint64_t ticket, lyg_server_result* out) {
    if (!s || !out) return LYG_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return LYG_ERR_INVALID_ARGUMENT;
    if (it->second->state != S_DONE) return LYG_PENDING;
    fill_result(s, *it->second, out);
    return LYG_OK;
}

int32_t lyg_server_wait(lyg_server* s, uint64_t ticket, lyg_server_result* out) {
    if (!s || !out) return LYG_ERR_INVALID_ARGUMENT;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return LYG_ERR_INVALID_ARGUMENT;
    Request* r = it->second.get();
    s->cv_done.wait(lk, [&] { return r->state == S_DONE || s->stop; });
    if (r->state != S_DONE) return LYG_ERR_INVALID_ARGUMENT;
    fill_result(s, *r, out);
    return LYG_OK;
}

int32_t lyg_server_copy_result(lyg_server* s, uint64_t ticket, float* vertices,
                               uint32_t* indices) {
    if (!s) return LYG_ERR_INVALID_ARGUMENT;
    lyg_server_result r;
    int32_t st = lyg_server_poll(s, ticket, &r);
    if (st != LYG_OK) return st;
    if (r.status != LYG_OK) return r.status;
    if ((r.n_vertices && !vertices) || (r.n_indices && !indices)) return LYG_ERR_INVALID_ARGUMENT;
    if (r.n_vertices && hipMemcpy(vertices, r.vertices, 8ull * r.n_vertices, hipMemcpyDeviceToHost) != hipSuccess)
        return LYG_ERR_DEVICE;
    if (r.n_indices && hipMemcpy(indices, r.indices, 4ull * r.n_indices, hipMemcpyDeviceToHost) != hipSuccess)
        return LYG_ERR_DEVICE;
    return LYG_OK;
}

int32_t lyg_server_release(lyg_server* s, uint64_t ticket) {
    if (!s) return LYG_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return LYG_ERR_INVALID_ARGUMENT;
    Request* r = it->second.get();
    if (r->state != S_DONE) return LYG_PENDING;
    if (r->batch_refs && --*r->batch_refs == 0 && r->out_off != SIZE_MAX) s->pool.release(r->out_off);
    s->reqs.erase(it);
    s->cv_done.notify_all();
    s->cv_dispatch.notify_one();
    return LYG_OK;
}

// p50 / p99 / max of `l` (reordered), nearest rank.
static void latency_summary(std::vector<double>& l, double* p50, double* p99, double* mx) {
    *p50 = *p99 = *mx = 0;
    if (l.empty()) return;
    auto pct = [&](double q) {
        size_t k = (size_t)std::min<double>((double)l.size() - 1, std::floor(q * (double)(l.size() - 1) + 0.5));
        std::nth_element(l.begin(), l.begin() + k, l.end());
        return l[k];
    };
    *p50 = pct(0.50);
    *p99 = pct(0.99);
    *mx = *std::max_element(l.begin(), l.end());
}

int32_t lyg_server_get_stats(lyg_server* s, lyg_server_stats* out) {
    if (!s || !out) return LYG_ERR_INVALID_ARGUMENT;
    std::vector<double> l;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        std::memset(out, 0, sizeof(*out));
        out->completed = s->completed;
        out->triangles = s->triangles;
        out->vertices = s->vertices;
        out->batches = s->batches;
        out->mean_batch = s->batches ? (double)s->batched_requests / (double)s->batches : 0.0;
        out->pool_high_water = s->pool.high_water();
        out->pending = s->pending.size();
        out->rejected = s->rejected;
        if (s->any_submit && s->completed)
            out->elapsed_s = std::chrono::duration<double>(s->t_last_done - s->t_first_submit).count();
        l = s->latencies_since(0);
    }
    latency_summary(l, &out->p50_us, &out->p99_us, &out->max_us);
    return LYG_OK;
}

}  // extern "C"

// ---- serving measurement ----------------------------------------------------------------
namespace lygen {
void stroke_path(uint64_t seed, uint64_t i, uint32_t min_segs, uint32_t max_segs,
                 float coord_max, bool closed, std::vector<uint8_t>& verbs,
                 std::vector<float>& points);
void fill_path(uint64_t seed, uint64_t i, std::vector<uint8_t>& verbs, std::vector<float>& pts);
void walk_path(uint64_t seed, uint64_t i, uint32_t min_segs, uint32_t max_segs, float extent,
               std::vector<uint8_t>& verbs, std::vector<float>& points);
}  // namespace lygen

extern "C" {

int32_t lyg_gen_stroke_options(uint64_t seed, uint32_t n_paths, float tolerance, float w_min,
                               float w_max, lyg_stroke_options* out);
int32_t lyg_gen_fill_options(uint64_t seed, uint32_t n_paths, float tolerance,
                             lyg_fill_options* out);

// A seeded request stream with Poisson arrivals (see include/lyg.h).  The requests draw
// from `n_distinct` pre-generated paths of each kind (generating multi-thousand-segment
// paths at the arrival rate would measure the generator, not the server).
int32_t lyg_server_run_stream(lyg_server* s, uint64_t seed, uint32_t n_requests,
                              double rate_per_s, float stroke_frac, uint32_t min_segs,
                              uint32_t max_segs, float tolerance, uint32_t n_distinct,
                              lyg_stream_stats* out) {
    return lyg_server_run_stream_mixed(s, seed, n_requests, rate_per_s, stroke_frac, min_segs,
                                       max_segs, tolerance, n_distinct, 0, out);
}

int32_t lyg_server_run_stream_mixed(lyg_server* s, uint64_t seed, uint32_t n_requests,
                                    double rate_per_s, float stroke_frac, uint32_t min_segs,
                                    uint32_t max_segs, float tolerance, uint32_t n_distinct,
                                    int fill_kind, lyg_stream_stats* out) {
    if (!s || !out || n_requests == 0 || !(rate_per_s > 0) || n_distinct == 0 || min_segs < 1 ||
        max_segs < min_segs)
        return LYG_ERR_INVALID_ARGUMENT;
    struct P {
        std::vector<uint8_t> v;
        std::vector<float> p;
    };
    std::vector<P> strokes(n_distinct), fills(n_distinct);
    std::vector<lyg_stroke_options> sopts(n_distinct);
    std::vector<lyg_fill_options> fopts(n_distinct);
    for (uint32_t i = 0; i < n_distinct; ++i) {
        lygen::stroke_path(seed ^ 0x5e7e5e7eull, i, min_segs, max_segs, 4096.0f, false,
                           strokes[i].v, strokes[i].p);
        if (fill_kind == 1)
            lygen::stroke_path(seed ^ 0xf111f111ull, i, min_segs, max_segs, 4096.0f, true,
                               fills[i].v, fills[i].p);
        else if (fill_kind == 2)
            lygen::walk_path(seed, i, min_segs, max_segs, 4096.0f, fills[i].v, fills[i].p);
        else
            lygen::fill_path(seed, i, fills[i].v, fills[i].p);
    }
    lyg_gen_stroke_options(seed, n_distinct, tolerance, 0.5f, 8.0f, sopts.data());
    lyg_gen_fill_options(seed, n_distinct, tolerance, fopts.data());
    // arrivals and kinds
    uint64_t x = seed * 0x9E3779B97F4A7C15ull + 1;
    auto next_u = [&]() {
        x ^= x << 13;
        x ^= x >> 7;
        x ^= x << 17;
        return (double)(x >> 11) * (1.0 / 9007199254740992.0);
    };
    std::vector<double> at(n_requests);
    std::vector<uint32_t> pick(n_requests);
    std::vector<uint8_t> is_stroke(n_requests);
    double t = 0;
    for (uint32_t i = 0; i < n_requests; ++i) {
        t += -std::log(1.0 - next_u()) / rate_per_s;
        at[i] = t;
        is_stroke[i] = next_u() < stroke_frac ? 1 : 0;
        pick[i] = (uint32_t)(next_u() * n_distinct) % n_distinct;
    }
    std::vector<uint64_t> tickets(n_requests, 0);
    std::atomic<uint32_t> submitted{0};
    std::atomic<uint64_t> failed{0};
    // LYG_STREAM_PROGRESS=1: a line on stderr for every twentieth of the stream completed
    const char* pe = getenv("LYG_STREAM_PROGRESS");
    const bool progress = pe && pe[0] == '1';
    const uint32_t progress_every = std::max(1u, n_requests / 20);
    std::atomic<int32_t> err{LYG_OK};
    lyg_server_stats before{};
    lyg_server_get_stats(s, &before);
    uint64_t lat0;
    size_t batched0;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        lat0 = s->lat_count;
        batched0 = s->batched_requests;
    }
    const Clock::time_point t0 = Clock::now();
    std::thread reaper([&] {
        for (uint32_t i = 0; i < n_requests; ++i) {
            while (submitted.load(std::memory_order_acquire) <= i && err.load() == LYG_OK)
                std::this_thread::yield();
            if (err.load() != LYG_OK) return;
            lyg_server_result r;
            lyg_server_wait(s, tickets[i], &r);
            if (r.status != LYG_OK) failed.fetch_add(1);
            lyg_server_release(s, tickets[i]);
            if (progress && (i + 1) % progress_every == 0) {
                fprintf(stderr, "[lyg stream] %u of %u requests done\n", i + 1, n_requests);
                fflush(stderr);
            }
        }
    });
    for (uint32_t i = 0; i < n_requests; ++i) {
        const Clock::time_point due = t0 + std::chrono::duration_cast<Clock::duration>(
                                               std::chrono::duration<double>(at[i]));
        while (Clock::now() < due) std::this_thread::yield();
        const P& p = is_stroke[i] ? strokes[pick[i]] : fills[pick[i]];
        lyg_path_view pv{p.v.data(), (uint32_t)p.v.size(), p.p.data(), (uint32_t)(p.p.size() / 2)};
        int32_t st = is_stroke[i] ? lyg_server_submit_stroke(s, &pv, &sopts[pick[i]], &tickets[i])
                                  : lyg_server_submit_fill(s, &pv, &fopts[pick[i]], &tickets[i]);
        if (st != LYG_OK) {
            err.store(st);
            break;
        }
        submitted.store(i + 1, std::memory_order_release);
    }
    const double submit_s = std::chrono::duration<double>(Clock::now() - t0).count();
    reaper.join();
    const double total_s = std::chrono::duration<double>(Clock::now() - t0).count();
    if (err.load() != LYG_OK) return err.load();
    lyg_server_stats after{};
    lyg_server_get_stats(s, &after);
    std::memset(out, 0, sizeof(*out));
    out->requests = n_requests;
    out->offered_rate = rate_per_s;
    out->submit_rate = n_requests / submit_s;
    out->completed_rate = n_requests / total_s;
    out->triangles = after.triangles - before.triangles;
    out->triangles_per_s = (double)out->triangles / total_s;
    // latencies and batch sizes of this stream's requests only (a server that ran a warm-up
    // stream before reports its steady state here)
    std::vector<double> l;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        l = s->latencies_since(lat0);
        const uint64_t nb = after.batches - before.batches;
        out->mean_batch = nb ? (double)(s->batched_requests - batched0) / (double)nb : 0.0;
    }
    latency_summary(l, &out->p50_us, &out->p99_us, &out->max_us);
    out->batches = after.batches - before.batches;
    out->seconds = total_s;
    out->pool_high_water = after.pool_high_water;
    out->failed = failed.load();
    out->rejected = after.rejected - before.rejected;
    return LYG_OK;
}

}  // extern "C"
