// Host side of the fill: sizing, launch rounds and compaction.
//
// A fill call runs in three steps on the context's stream:
//   1. fill_bound_kernel writes an upper bound of every path's sweep events;
//   2. rounds of fill_sweep_kernel: each path is swept by one wave into a staging area
//      sized from its bound.  Paths are grouped into size classes so that the per-wave
//      working memory of a launch fits its largest path; a path that outgrows its
//      staging or working memory ends with a retry status and is swept again in the next
//      round with four times the room;
//   3. an exclusive scan over the per-path vertex/index counts, then fill_copy_kernel
//      moves every round's staged output to its final place, rebasing the indices.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <numeric>
#include <vector>

#include "ctx.h"
#include "internal.h"
#include "lyg_types.h"

extern "C" {
size_t lyg_fill_work_bytes(const lyg::fillk::FillCaps* caps, size_t hot_bytes);
int lyg_fill_occupancy(uint32_t waves_per_block, uint32_t hot_bytes);
int lyg_fill_occupancy_form(uint32_t waves_per_block, uint32_t hot_bytes, int form);
size_t lyg_fill_lds_bytes(const lyg::fillk::FillCaps* lc);
hipError_t lyg_launch_fill_bound(const lyg::FillLaunch* L, uint32_t n, uint32_t* bound,
                                 hipStream_t s);
hipError_t lyg_launch_fill_sweep(const lyg::FillLaunch* L, int grid, hipStream_t s);
size_t lyg_fill_box_bytes();
hipError_t lyg_launch_fill_inter(const lyg::FillLaunch* L, uint32_t n, void* boxes, uint32_t* inter,
                                 hipStream_t s);
hipError_t lyg_launch_fill_copy(const lyg::FillLaunch* L, hipStream_t s);
hipError_t lyg_launch_fill_status_out(const uint32_t* st, uint32_t* out, uint32_t n,
                                      hipStream_t s);
hipError_t lyg_launch_fill_list_u32(uint32_t* a, const uint32_t* list, uint32_t m, uint32_t v,
                                    hipStream_t s);
hipError_t lyg_launch_scan(const uint32_t* counts, uint32_t n, uint32_t* offsets, uint64_t* total,
                           uint32_t* overflow, hipStream_t s);
hipError_t lyg_launch_fill_u32(uint32_t* p, uint32_t n, uint32_t v, hipStream_t s);
}

namespace {

constexpr uint32_t ST_RETRY = 100;
constexpr uint32_t ST_UNSWEPT = 0xffffffffu;  // written before a round, replaced per path
constexpr int MAX_ROUNDS = 5;
constexpr uint64_t MAX_PATH_EVENTS = 1ull << 27;     // larger paths: LYG_ERR_NO_SPACE
constexpr size_t WORK_BUDGET = 16ull << 30;          // working memory of one round
constexpr size_t STAGE_BUDGET = 64ull << 30;         // staging memory of one round
constexpr int MAX_CLASSES = 8;
constexpr int N_FORMS = 3;  // placement forms of the sweep's hot arrays (FillLaunch::form)
constexpr int FORM_MIGRATE = 0, FORM_LDS = 1, FORM_PART = 2;
constexpr int MAX_LAUNCHES = N_FORMS * MAX_CLASSES;
constexpr size_t CLASS_STREAMS = 3;
// The first round sweeps a path with all its hot arrays in LDS when their estimate fits
// this many bytes per wave, else with the per-event ones in LDS when those fit, else in
// the migrating form: three waves per CU (a CU has 160 KiB; the kernel's own LDS takes
// 4 KiB), enough to keep every large path of a launch resident at once.
constexpr size_t LDS_WAVE_MAX = 52u * 1024u;
constexpr size_t LDS_CU_BYTES = 156u * 1024u;
constexpr size_t SWEEP_MAX_WAVES = 4;  // the sweep kernel's launch bound (waves per block)

using lyg::fillk::FillCaps;

// LDS-resident capacities for a path of at most `bound` events (two per edge): active
// edges (and pending edges below one point) are bounded by the edges that cross one
// sweep line, a quarter of the events plus a margin; spans by the active edges (dead
// spans wait for the next cleanup); the pending intersections by 2.5 active edges; the
// monotone node pool by the events.  The 513-point stars of the fill workload peak at
// 280 active edges, 265 spans, 678 heap entries and 1,114 nodes (bound 1,044).  A path
// that outgrows one of them is swept again in the migrating form.
// `shrink` > 1 divides them (at least 8 each), which makes paths outgrow the LDS form on
// purpose: the tests' way to run its fallback.
FillCaps lds_caps(uint64_t bound, uint32_t shrink) {
    const uint64_t d = shrink > 1 ? shrink : 1;
    auto cap = [&](uint64_t v) {
        return (uint32_t)std::min<uint64_t>(std::max<uint64_t>(v / d, 8), 0x7fffffffu);
    };
    const uint64_t act = bound / 4 + bound / 32 + 16;
    FillCaps k;
    k.ev = 0;
    k.act = cap(act);
    k.span = cap(act + 16);
    k.heap = cap(act * 5 / 2 + 64);
    k.node = cap(bound + bound / 4 + 64);
    k.below = cap(act / 4 + 16);
    return k;
}
// The form a path starts in, and its LDS capacities (spans and nodes 0 in FORM_PART).
int path_form(uint64_t bound, uint32_t mode, FillCaps* lc) {
    *lc = lds_caps(bound, mode);
    if (mode && lyg_fill_lds_bytes(lc) <= LDS_WAVE_MAX) return FORM_LDS;
    lc->span = lc->node = 0;
    if (mode && lyg_fill_lds_bytes(lc) <= LDS_WAVE_MAX) return FORM_PART;
    return FORM_MIGRATE;
}

// Working-memory capacities of one wave for paths of at most `bound` events and `inter`
// estimated self-intersections at `scale` (1, then growing with each retry round).
// Active edges, spans and pending edges are bounded by the edges; the event heap and the
// monotone node pool grow with the intersections (each one adds a vertex and can leave
// two truncated edge parts pending).
FillCaps class_caps(uint64_t bound, uint64_t inter, uint64_t scale) {
    auto cap = [](uint64_t v) { return (uint32_t)std::min<uint64_t>(v, 0x7fffffffu); };
    FillCaps k;
    k.ev = cap(bound);
    k.heap = cap(scale * (2 * bound + 2 * inter + 256));
    k.act = cap(scale * (bound + 64));
    k.span = k.act;
    k.below = k.act;
    k.node = cap(4ull * k.act + 256 + scale * 2 * inter);
    return k;
}

// Staging capacity (vertices, triangles) of one path: its events plus its intersections.
void stage_caps(uint64_t bound, uint64_t inter, uint64_t scale, uint64_t* v, uint64_t* t) {
    *v = std::min<uint64_t>(scale * (2 * bound + 2 * inter + 256), 0x7fffffffu);
    *t = std::min<uint64_t>(2 * *v, 0x7fffffffu);
}

int size_class(uint64_t bound) {
    int k = 0;
    uint64_t lim = 1u << 12;
    while (bound > lim && k < MAX_CLASSES - 1) {
        lim <<= 2;
        ++k;
    }
    return k;
}

// Streams (default priority) for running a round's size classes at once, and one event
// per launch plus the start event.
bool class_streams(lyg_ctx* c, size_t n, size_t n_launches) {
    if (n > (size_t)MAX_CLASSES || n_launches > (size_t)MAX_LAUNCHES) return false;
    for (size_t i = 0; i < n; ++i)
        if (!c->class_stream[i] &&
            !check(c, hipStreamCreateWithFlags(&c->class_stream[i], hipStreamNonBlocking),
                   "hipStreamCreate"))
            return false;
    for (size_t i = 0; i <= n_launches; ++i)
        if (!c->class_ev[i] &&
            !check(c, hipEventCreateWithFlags(&c->class_ev[i], hipEventDisableTiming),
                   "hipEventCreate"))
            return false;
    return true;
}

bool ensure_at(lyg_ctx* c, std::vector<DevBuf>& v, size_t i, size_t bytes) {
    if (v.size() <= i) v.resize(i + 1);
    return ensure(c, v[i], bytes);
}

// One sweep launch of a round: the paths list[k0, k1) of one size class and placement
// form, `blocks` blocks of `wpb` waves, each wave with `stride` bytes of working memory at
// `work_off` + its slot and `hot` bytes of LDS.
struct Launch {
    size_t k0, k1, blocks, stride, work_off;
    uint64_t est;   // the class's work: its largest path's events + 2 x intersections
    FillCaps cap;   // working-memory capacities
    int form;       // placement form of the hot arrays
    FillCaps lc;    // their LDS capacities (forms 1 and 2)
    uint32_t wpb, hot;
};

// The paths of one round in launch order with their staging offsets, and its launches.
struct RoundPlan {
    std::vector<uint32_t> list;
    std::vector<uint64_t> offs;  // staging offsets: vertices [0, m], triangles [m + 1, 2m + 1]
    std::vector<Launch> launches;
    size_t work_bytes = 0;       // working memory of all launches
};

// Plans round `round` over the paths `todo`, with `scale` times the first round's room.
int32_t plan_round(lyg_ctx* c, const std::vector<uint32_t>& todo,
                   const std::vector<uint32_t>& bound, const std::vector<uint32_t>& inter,
                   const std::vector<uint64_t>& need, int round, uint64_t scale, RoundPlan* P) {
    // order by launch: size class, then the placement form (the first round picks it
    // per path; retries migrate); then by path
    std::vector<int> cls(todo.size());
    for (size_t k = 0; k < todo.size(); ++k) {
        const uint32_t p = todo[k];
        FillCaps lc;
        const int form = round == 0 ? path_form(bound[p], c->fill_lds_resident, &lc)
                                    : FORM_MIGRATE;
        cls[k] = N_FORMS * size_class(need[p]) + form;
    }
    std::vector<size_t> order(todo.size());
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return cls[a] < cls[b]; });
    std::vector<uint32_t>& list = P->list;
    std::vector<uint64_t>& offs = P->offs;
    list.assign(todo.size(), 0);
    offs.assign(2 * (todo.size() + 1), 0);
    uint64_t* sv = offs.data();
    uint64_t* st = offs.data() + todo.size() + 1;
    sv[0] = st[0] = 0;
    for (size_t k = 0; k < order.size(); ++k) {
        uint32_t p = todo[order[k]];
        list[k] = p;
        uint64_t v, t;
        stage_caps(bound[p], inter[p], scale, &v, &t);
        sv[k + 1] = sv[k] + v;
        st[k + 1] = st[k] + t;
    }
    const size_t m = list.size();
    if (8 * sv[m] + 12 * st[m] > STAGE_BUDGET) {
        c->last_error = "fill staging exceeds the per-round budget";
        return LYG_ERR_NO_SPACE;
    }
    // One launch per size class and form, each with its own part of the working memory.
    const uint32_t wpb = c->fill_waves_per_block;
    const uint32_t hot = c->fill_hot_bytes;
    const int occ = std::max(1, lyg_fill_occupancy(wpb, hot));
    std::vector<Launch>& launches = P->launches;
    size_t want_bytes = 0;
    for (size_t k0 = 0; k0 < m;) {
        int kc = cls[order[k0]];
        size_t k1 = k0;
        uint64_t maxb = 0, maxi = 0;
        while (k1 < m && cls[order[k1]] == kc) {
            maxb = std::max<uint64_t>(maxb, bound[list[k1]]);
            maxi = std::max<uint64_t>(maxi, inter[list[k1]]);
            ++k1;
        }
        Launch ln;
        ln.k0 = k0;
        ln.k1 = k1;
        ln.est = 0;
        for (size_t k = k0; k < k1; ++k) ln.est = std::max<uint64_t>(ln.est, need[list[k]]);
        ln.cap = class_caps(std::max<uint64_t>(maxb, 1), maxi, scale);
        ln.form = kc % N_FORMS;
        ln.lc = FillCaps{};
        ln.wpb = wpb;
        ln.hot = hot;
        int ln_occ = occ;
        if (ln.form != FORM_MIGRATE) {
            // every capacity grows with the bound: the largest path's fit them all
            (void)path_form(maxb, c->fill_lds_resident, &ln.lc);
            if (ln.form == FORM_LDS) ln.lc = lds_caps(maxb, c->fill_lds_resident);
            ln.hot = (uint32_t)((lyg_fill_lds_bytes(&ln.lc) + 255) & ~(size_t)255);
            ln.wpb = (uint32_t)std::max<size_t>(
                1, std::min<size_t>(SWEEP_MAX_WAVES, LDS_CU_BYTES / ln.hot));
            ln_occ = lyg_fill_occupancy_form(ln.wpb, ln.hot, ln.form);
            while (ln_occ < 1 && ln.wpb > 1)
                ln_occ = lyg_fill_occupancy_form(--ln.wpb, ln.hot, ln.form);
            if (ln_occ < 1) {  // cannot launch: the migrating form
                ln.form = FORM_MIGRATE;
                ln.wpb = wpb;
                ln.hot = hot;
                ln_occ = occ;
            }
        }
        if (ln.form != FORM_MIGRATE) {
            // arrays that live in LDS need no global region (their overflow retries):
            // the heap is most of a wave's worst-case working memory
            ln.cap.heap = 0;
            ln.cap.below = 0;
            if (ln.form == FORM_LDS) ln.cap.span = 0;
        }
        ln.stride = lyg_fill_work_bytes(&ln.cap, ln.form != FORM_MIGRATE ? 0 : ln.hot);
        size_t blocks = (k1 - k0 + ln.wpb - 1) / ln.wpb;
        ln.blocks = std::max<size_t>(1, std::min<size_t>(blocks, (size_t)c->cu_count * ln_occ));
        ln.work_off = 0;
        want_bytes += ln.blocks * ln.stride * ln.wpb;
        launches.push_back(ln);
        k0 = k1;
    }
    // over the budget: every class gives up blocks in proportion
    size_t& work_bytes = P->work_bytes;
    work_bytes = 0;
    for (Launch& ln : launches) {
        if (want_bytes > WORK_BUDGET)
            ln.blocks = std::max<size_t>(
                1, (size_t)((double)ln.blocks * (double)WORK_BUDGET / (double)want_bytes));
        ln.work_off = work_bytes;
        work_bytes += (ln.blocks * ln.stride * ln.wpb + 255) & ~(size_t)255;
    }
    return LYG_OK;
}

// The stream of every launch, for CLASS_STREAMS streams that run at once (a process gets
// four hardware queues; the context's own stream holds the fourth): classes go largest
// first (the last launch first) to the stream with the least work queued.
std::vector<size_t> assign_streams(const RoundPlan& P, size_t n_streams) {
    std::vector<size_t> stream_of(P.launches.size(), 0);
    uint64_t load[CLASS_STREAMS] = {};
    for (size_t r = 0; r < P.launches.size(); ++r) {
        const size_t li = P.launches.size() - 1 - r;
        size_t si = 0;
        for (size_t q = 1; q < n_streams; ++q)
            if (load[q] < load[si]) si = q;
        load[si] += P.launches[li].est;
        stream_of[li] = si;
    }
    return stream_of;
}

// The fields of L that launch li of plan P sets: its paths (from the round's list and
// staging offsets on the device), working memory, capacities, form and work counter.
void set_launch(lyg::FillLaunch& L, const RoundPlan& P, size_t li, uint8_t* work,
                const uint32_t* d_list, const uint64_t* d_off, uint32_t* counters) {
    const Launch& ln = P.launches[li];
    const size_t m = P.list.size();
    L.work = work + ln.work_off;
    L.work_stride = ln.stride;
    L.caps = ln.cap;
    L.hot_bytes = ln.hot;
    L.waves_per_block = ln.wpb;
    L.form = (uint32_t)ln.form;
    L.lds_caps = ln.lc;
    L.path_list = d_list + ln.k0;
    L.n_list = (uint32_t)(ln.k1 - ln.k0);
    L.stage_v_off = d_off + ln.k0;
    L.stage_t_off = d_off + (m + 1) + ln.k0;
    L.work_counter = counters + li;  // one per launch
}

struct FillJob {
    lyg::PathBatchDev in;
    const lyg_fill_options* opts;
    uint32_t opts_stride;
    uint32_t vertex_offset;
    bool local_ids = false;
};

struct FillResult {
    std::vector<uint32_t> status;  // per path, final
    uint64_t total_v = 0, total_i = 0;
    int rounds = 0;
    std::vector<uint32_t> round_len;  // paths swept in each round
};

// Step 1 and the intersection estimate, read back to the host: every path's event bound
// and its estimated self-intersections.
bool measure(lyg_ctx* c, const lyg::FillLaunch& L, hipStream_t s, std::vector<uint32_t>* bound_out,
             std::vector<uint32_t>* inter_out) {
    const uint32_t n = L.in.n_paths;
    if (!ensure(c, c->f_bound, 4ull * n)) return false;
    std::vector<uint32_t>& bound = *bound_out;
    std::vector<uint32_t>& inter = *inter_out;
    bound.assign(n, 0);
    inter.assign(n, 0);
    uint32_t n_verbs = 0;
    if (!check(c, lyg_launch_fill_bound(&L, n, as<uint32_t>(c->f_bound), s), "fill bound") ||
        !check(c, hipMemcpyAsync(bound.data(), c->f_bound.p, 4ull * n, hipMemcpyDeviceToHost, s),
               "D2H") ||
        !check(c, hipMemcpyAsync(&n_verbs, L.in.verb_off + n, 4, hipMemcpyDeviceToHost, s), "D2H") ||
        !check(c, hipStreamSynchronize(s), "sync"))
        return false;
    // intersection estimate: one box record per verb, then a pair count per path
    if (!ensure(c, c->f_inter, 4ull * n) ||
        !ensure(c, c->f_boxes, lyg_fill_box_bytes() * ((size_t)n_verbs + 1)))
        return false;
    if (!check(c, lyg_launch_fill_inter(&L, n, c->f_boxes.p, as<uint32_t>(c->f_inter), s),
               "fill inter") ||
        !check(c, hipMemcpyAsync(inter.data(), c->f_inter.p, 4ull * n, hipMemcpyDeviceToHost, s),
               "D2H") ||
        !check(c, hipStreamSynchronize(s), "sync"))
        return false;
    return true;
}

// Steps 1-2 and the scan of step 3; per-path offsets go to vtx_off/idx_off (n+1 each).
int32_t fill_sweep(lyg_ctx* c, const FillJob& J, uint32_t* vtx_off, uint32_t* idx_off,
                   FillResult* R) {
    const uint32_t n = J.in.n_paths;
    hipStream_t s = c->stream;
    if (!ensure(c, c->f_bound, 4ull * n) || !ensure(c, c->f_status, 4ull * n) ||
        !ensure(c, c->f_round, 4ull * n) || !ensure(c, c->f_hint, 4ull * n) ||
        !ensure(c, c->counts_v, 4ull * n) || !ensure(c, c->counts_i, 4ull * n))
        return LYG_ERR_DEVICE;
    lyg::FillLaunch L{};
    L.in = J.in;
    L.opts = J.opts;
    L.opts_stride = J.opts_stride;
    L.path_nv = as<uint32_t>(c->counts_v);
    L.path_ni = as<uint32_t>(c->counts_i);
    L.path_status = as<uint32_t>(c->f_status);
    L.path_round = as<uint32_t>(c->f_round);
    L.path_hint = as<uint32_t>(c->f_hint);
    L.vertex_offset = J.vertex_offset;
    std::vector<uint32_t> bound, inter;
    if (!measure(c, L, s, &bound, &inter)) return LYG_ERR_DEVICE;
    // size classes go by the larger of the two room drivers
    std::vector<uint64_t> need(n);
    for (uint32_t i = 0; i < n; ++i) need[i] = (uint64_t)bound[i] + 2ull * inter[i];
    R->status.assign(n, 0);
    // paths still to sweep, with their scale (4^round)
    std::vector<uint32_t> todo;
    todo.reserve(n);
    for (uint32_t i = 0; i < n; ++i) {
        if (bound[i] > MAX_PATH_EVENTS) R->status[i] = LYG_ERR_NO_SPACE;
        else todo.push_back(i);
    }
    // paths rejected up front produce nothing
    if (todo.size() != n) {
        if (!check(c, lyg_launch_fill_u32(L.path_nv, n, 0, s), "zero") ||
            !check(c, lyg_launch_fill_u32(L.path_ni, n, 0, s), "zero") ||
            !check(c, lyg_launch_fill_u32(L.path_round, n, 0xffffffffu, s), "zero"))
            return LYG_ERR_DEVICE;
    }
    uint64_t scale = 1, next_scale = 4;
    int round = 0;
    for (; round < MAX_ROUNDS && !todo.empty(); ++round, scale = next_scale) {
        RoundPlan P;
        const int32_t planned = plan_round(c, todo, bound, inter, need, round, scale, &P);
        if (planned != LYG_OK) return planned;
        const std::vector<uint32_t>& list = P.list;
        const std::vector<uint64_t>& offs = P.offs;
        const size_t m = list.size();
        const uint64_t* sv = offs.data();
        const uint64_t* st = offs.data() + m + 1;
        if (!ensure_at(c, c->f_list, round, 4 * m) ||
            !ensure_at(c, c->f_stage_off, round, 8 * offs.size()) ||
            !ensure_at(c, c->f_stage_v, round, 8 * sv[m]) ||
            !ensure_at(c, c->f_stage_t, round, 12 * st[m]))
            return LYG_ERR_DEVICE;
        uint32_t* d_list = as<uint32_t>(c->f_list[round]);
        uint64_t* d_off = as<uint64_t>(c->f_stage_off[round]);
        if (!check(c, hipMemcpyAsync(d_list, list.data(), 4 * m, hipMemcpyHostToDevice, s), "H2D") ||
            !check(c, hipMemcpyAsync(d_off, offs.data(), 8 * offs.size(), hipMemcpyHostToDevice, s),
                   "H2D") ||
            !ensure(c, c->f_counters, 4ull * MAX_LAUNCHES) ||
            !check(c, lyg_launch_fill_u32(as<uint32_t>(c->f_counters), MAX_LAUNCHES, 0, s),
                   "reset") ||
            // a path of the round that no launch reaches keeps this status, never a success
            !check(c, lyg_launch_fill_list_u32(L.path_status, d_list, (uint32_t)m, ST_UNSWEPT, s),
                   "status"))
            return LYG_ERR_DEVICE;
        L.round = (uint32_t)round;
        L.stage_v = as<float>(c->f_stage_v[round]);
        L.stage_t = as<uint32_t>(c->f_stage_t[round]);
        const std::vector<Launch>& launches = P.launches;
        if (!ensure(c, c->f_work, P.work_bytes)) return LYG_ERR_NO_SPACE;
        const auto t_round = std::chrono::steady_clock::now();
        if (c->timing && round == 0) (void)hipEventRecord(c->ev[2], s);
        const size_t n_streams = std::min<size_t>(launches.size(), CLASS_STREAMS);
        if (!class_streams(c, n_streams, launches.size()) ||
            !check(c, hipEventRecord(c->class_ev[0], s), "event"))
            return LYG_ERR_DEVICE;
        const std::vector<size_t> stream_of = assign_streams(P, n_streams);
        for (size_t r = 0; r < launches.size(); ++r) {
            const size_t li = launches.size() - 1 - r;  // largest class first
            const Launch& ln = launches[li];
            hipStream_t cs = c->class_stream[stream_of[li]];
            set_launch(L, P, li, as<uint8_t>(c->f_work), d_list, d_off,
                       as<uint32_t>(c->f_counters));
            const auto t_launch = std::chrono::steady_clock::now();
            if (!check(c, hipStreamWaitEvent(cs, c->class_ev[0], 0), "wait") ||
                !check(c, lyg_launch_fill_sweep(&L, (int)ln.blocks, cs), "fill sweep") ||
                !check(c, hipEventRecord(c->class_ev[1 + li], cs), "event") ||
                !check(c, hipStreamWaitEvent(s, c->class_ev[1 + li], 0), "wait"))
                return LYG_ERR_DEVICE;
            if (trace_on()) {  // serialises the launches: for diagnosis only
                (void)hipStreamSynchronize(cs);
                fprintf(stderr,
                        "[lyg]   class launch %zu: %zu paths, %zu blocks, %.1f KiB per wave, "
                        "heap %u node %u act %u, %s %u waves x %u B LDS, %.2f ms\n",
                        li, ln.k1 - ln.k0, ln.blocks, ln.stride / 1024.0, ln.cap.heap, ln.cap.node,
                        ln.cap.act,
                        ln.form == FORM_LDS ? "LDS" : ln.form == FORM_PART ? "LDS+global" : "migrating",
                        ln.wpb, ln.hot,
                        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() -
                                                                  t_launch).count());
                fflush(stderr);
            }
        }
        if (c->timing && round == 0) (void)hipEventRecord(c->ev[3], s);
        R->round_len.push_back((uint32_t)m);
        std::vector<uint32_t> st_h(n), hint_h(n);
        if (!check(c, hipMemcpyAsync(st_h.data(), L.path_status, 4ull * n, hipMemcpyDeviceToHost, s),
                   "D2H") ||
            !check(c, hipMemcpyAsync(hint_h.data(), L.path_hint, 4ull * n, hipMemcpyDeviceToHost, s),
                   "D2H") ||
            !check(c, hipStreamSynchronize(s), "sync"))
            return LYG_ERR_DEVICE;
        std::vector<uint32_t> next;
        uint64_t grow = 4;  // next round's room: at least 4x, more if a path's hint asks
        for (uint32_t p : list) {
            if (st_h[p] == ST_UNSWEPT) {  // no wave took the path: a launch error
                c->last_error = "fill: a path was not swept";
                return LYG_ERR_INTERNAL;
            }
            if (st_h[p] == ST_RETRY) {
                next.push_back(p);
                uint64_t v, t;
                stage_caps(bound[p], inter[p], scale, &v, &t);
                const uint64_t want = (uint64_t)hint_h[p] + hint_h[p] / 4;
                while (v * grow < want && grow < 1024) grow *= 2;
            } else {
                R->status[p] = st_h[p];
            }
        }
        if (trace_on()) {
            const double ms = std::chrono::duration<double, std::milli>(
                                  std::chrono::steady_clock::now() - t_round).count();
            fprintf(stderr, "[lyg] fill round %d: %zu paths in %zu launches, %.2f ms, %zu to retry\n",
                    round, m, launches.size(), ms, next.size());
            fflush(stderr);
        }
        todo.swap(next);
        next_scale = scale * grow;
    }
    R->rounds = round;
    for (uint32_t p : todo) R->status[p] = LYG_ERR_NO_SPACE;  // still out of room
    // step 3: offsets
    if (!check(c, lyg_launch_fill_u32(c->d_counters + 4, 1, 0, s), "reset") ||
        !check(c, ctx_scan(c, L.path_nv, n, vtx_off, c->d_totals + 0, c->d_counters + 4, s), "scan") ||
        !check(c, ctx_scan(c, L.path_ni, n, idx_off, c->d_totals + 1, c->d_counters + 4, s), "scan"))
        return LYG_ERR_DEVICE;
    uint64_t* h = reinterpret_cast<uint64_t*>(c->h_pinned);
    if (!check(c, hipMemcpyAsync(h, c->d_totals, 16, hipMemcpyDeviceToHost, s), "D2H") ||
        !check(c, hipMemcpyAsync(h + 2, c->d_counters + 4, 4, hipMemcpyDeviceToHost, s), "D2H") ||
        !check(c, hipStreamSynchronize(s), "sync"))
        return LYG_ERR_DEVICE;
    R->total_v = h[0];
    R->total_i = h[1];
    if ((uint32_t)(h[2] & 0xffffffffu)) return LYG_ERR_TOO_MANY_VERTICES;
    // BuffersBuilder::add_vertex_impl: ids up to u32::MAX - 1 (geometry_builder.rs:338-350)
    if (!J.local_ids && R->total_v > 0 && (uint64_t)J.vertex_offset + R->total_v - 1 > 0xfffffffeull)
        return LYG_ERR_TOO_MANY_VERTICES;
    return LYG_OK;
}

// Step 3: every round's staged output to the final arrays.
int32_t fill_copy(lyg_ctx* c, const FillJob& J, const FillResult& R, const uint32_t* vtx_off,
                  const uint32_t* idx_off, float* out_v, uint32_t* out_i) {
    const uint32_t n = J.in.n_paths;
    // round r swept the paths of f_list[r] into its own staging; a path's output is in the
    // round recorded in f_round.
    lyg::FillLaunch L{};
    L.in = J.in;
    L.path_status = as<uint32_t>(c->f_status);
    L.path_round = as<uint32_t>(c->f_round);
    L.vtx_off = vtx_off;
    L.idx_off = idx_off;
    L.out_vertices = out_v;
    L.out_indices = out_i;
    L.vertex_offset = J.vertex_offset;
    L.local_ids = J.local_ids ? 1u : 0u;
    L.vertex_capacity = ~0ull;  // the totals were checked against the output on the host
    L.index_capacity = ~0ull;
    (void)n;
    if (c->timing) (void)hipEventRecord(c->ev[0], c->stream);
    for (int r = 0; r < R.rounds; ++r) {
        L.round = (uint32_t)r;
        L.path_list = as<uint32_t>(c->f_list[r]);
        L.stage_v = as<float>(c->f_stage_v[r]);
        L.stage_t = as<uint32_t>(c->f_stage_t[r]);
        L.n_list = R.round_len[r];
        L.stage_v_off = as<uint64_t>(c->f_stage_off[r]);
        L.stage_t_off = as<uint64_t>(c->f_stage_off[r]) + (R.round_len[r] + 1);
        if (!check(c, lyg_launch_fill_copy(&L, c->stream), "fill copy")) return LYG_ERR_DEVICE;
    }
    if (c->timing) (void)hipEventRecord(c->ev[1], c->stream);
    return LYG_OK;
}

int32_t first_error(const FillResult& R) {
    for (uint32_t s : R.status)
        if (s != 0) return (int32_t)s;
    return LYG_OK;
}

}  // namespace

int32_t lygi::fill_batch(lyg_ctx* c, const lygi::BatchSpec& b, const lygi::OutAlloc& alloc,
                         uint64_t* tv, uint64_t* ti, std::vector<uint32_t>* status) {
    FillJob J;
    J.in = b.in;
    J.in.first_path = 0;
    J.opts = reinterpret_cast<const lyg_fill_options*>(b.opts);
    J.opts_stride = b.opts_stride;
    J.vertex_offset = b.vertex_offset;
    J.local_ids = b.local_ids;
    FillResult R;
    int32_t st = fill_sweep(c, J, b.vtx_off, b.idx_off, &R);
    *tv = R.total_v;
    *ti = R.total_i;
    if (status) *status = R.status;
    if (st != LYG_OK) return st;
    float* ov = nullptr;
    uint32_t* oi = nullptr;
    st = alloc(R.total_v, R.total_i, &ov, &oi);
    if (st != LYG_OK) return st;
    st = fill_copy(c, J, R, b.vtx_off, b.idx_off, ov, oi);
    if (st != LYG_OK) return st;
    if (!check(c, hipStreamSynchronize(c->stream), "sync")) return LYG_ERR_DEVICE;
    return LYG_OK;
}

// ---- HIP graph form -------------------------------------------------------------------
// The first round of a batched fill, planned at capture time from the batch's bounds and
// intersection estimates (size classes, placement forms, staging and working memory, the
// streams of the launches), is replayed without the host: a head graph on the context's
// stream (counter reset, unswept status), one graph per class stream holding that stream's
// sweep launches, and a tail graph (offset scans, copy with the output capacity checked on
// the device, per-path statuses).  The class streams' graphs are launched on the context's
// own side streams, ordered by events: the launches of one round must run at once, and
// streams that the runtime would make for the branches of a single graph can share a
// hardware queue (a process has four), which would run them one after another.
// Every buffer the replay touches belongs to the graph: a later call can reallocate the
// context's.  A path that outgrows its planned room in a replay (its data changed since
// the capture) produces nothing and reports LYG_ERR_NO_SPACE; lyg_fill_batch sweeps it
// again with more room.
struct lygi::FillGraph {
    lyg_ctx* c = nullptr;
    uint32_t n = 0;
    uint32_t vertex_offset = 0;
    uint64_t vertex_capacity = 0, index_capacity = 0;
    DevBuf status, round, nv, ni, list, off, stage_v, stage_t, work, counters, totals, tiles;
    std::vector<hipGraph_t> graphs;
    hipGraphExec_t head = nullptr, tail = nullptr;
    std::vector<hipGraphExec_t> part;  // per class stream
    std::vector<hipStream_t> part_stream;
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> join;
};

namespace {

// Captures what `body` enqueues on s into an instantiated graph.
template <class F>
bool capture_on(lyg_ctx* c, hipStream_t s, lygi::FillGraph* g, hipGraphExec_t* exec, F&& body) {
    hipGraph_t graph = nullptr;
    if (!check(c, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "begin capture"))
        return false;
    const bool ok = body();
    const bool ended = check(c, hipStreamEndCapture(s, &graph), "end capture");
    if (graph) g->graphs.push_back(graph);
    return ok && ended && graph &&
           check(c, hipGraphInstantiate(exec, graph, nullptr, nullptr, 0), "instantiate");
}

}  // namespace

int32_t lygi::fill_graph_capture(lyg_ctx* c, const lyg_batch_in* in, const lyg_fill_options* opts,
                                 uint32_t n_opts, const lyg_batch_out* out, FillGraph** graph) {
    const uint32_t n = in->n_paths;
    hipStream_t s = c->stream;
    auto* g = new FillGraph();
    *graph = g;  // the caller destroys it on an error
    g->c = c;
    g->n = n;
    g->vertex_offset = out->vertex_offset;
    g->vertex_capacity = out->vertex_capacity;
    g->index_capacity = out->index_capacity;
    if (!ensure(c, g->status, 4ull * n) || !ensure(c, g->round, 4ull * n) ||
        !ensure(c, g->nv, 4ull * n) || !ensure(c, g->ni, 4ull * n) ||
        !ensure(c, g->counters, 4ull * (MAX_LAUNCHES + 1)) || !ensure(c, g->totals, 16) ||
        // the scans' tile buffer is allocated before the capture, not inside it
        (n > 32768 && !ensure(c, g->tiles, lyg_scan_tiles_bytes(n))))
        return LYG_ERR_DEVICE;
    uint32_t* counters = as<uint32_t>(g->counters);
    uint32_t* overflow = counters + MAX_LAUNCHES;  // the offset scans' 32-bit overflow flag
    lyg::FillLaunch L{};
    L.in.verbs = in->verbs;
    L.in.points = in->points;
    L.in.verb_off = in->verb_offsets;
    L.in.point_off = in->point_offsets;
    L.in.n_paths = n;
    L.opts = opts;
    L.opts_stride = n_opts == 1 ? 0 : 1;
    L.path_nv = as<uint32_t>(g->nv);
    L.path_ni = as<uint32_t>(g->ni);
    L.path_status = as<uint32_t>(g->status);
    L.path_round = as<uint32_t>(g->round);
    L.path_hint = nullptr;
    L.vertex_offset = out->vertex_offset;
    L.round = 0;
    // the plan: lyg_fill_batch's first round over every path
    std::vector<uint32_t> bound, inter;
    if (!measure(c, L, s, &bound, &inter)) return LYG_ERR_DEVICE;
    std::vector<uint64_t> need(n);
    std::vector<uint32_t> todo(n);
    for (uint32_t i = 0; i < n; ++i) {
        if (bound[i] > MAX_PATH_EVENTS) {
            c->last_error = "a path too large for a captured fill";
            return LYG_ERR_NO_SPACE;
        }
        need[i] = (uint64_t)bound[i] + 2ull * inter[i];
        todo[i] = i;
    }
    RoundPlan P;
    const int32_t planned = plan_round(c, todo, bound, inter, need, 0, 1, &P);
    if (planned != LYG_OK) return planned;
    const size_t m = P.list.size();
    const uint64_t* sv = P.offs.data();
    const uint64_t* st = P.offs.data() + m + 1;
    if (!ensure(c, g->list, 4 * m) || !ensure(c, g->off, 8 * P.offs.size()) ||
        !ensure(c, g->stage_v, 8 * sv[m]) || !ensure(c, g->stage_t, 12 * st[m]))
        return LYG_ERR_DEVICE;
    if (!ensure(c, g->work, P.work_bytes)) return LYG_ERR_NO_SPACE;
    uint32_t* d_list = as<uint32_t>(g->list);
    uint64_t* d_off = as<uint64_t>(g->off);
    if (!check(c, hipMemcpyAsync(d_list, P.list.data(), 4 * m, hipMemcpyHostToDevice, s), "H2D") ||
        !check(c, hipMemcpyAsync(d_off, P.offs.data(), 8 * P.offs.size(), hipMemcpyHostToDevice, s),
               "H2D") ||
        !check(c, hipStreamSynchronize(s), "sync"))
        return LYG_ERR_DEVICE;
    L.stage_v = as<float>(g->stage_v);
    L.stage_t = as<uint32_t>(g->stage_t);
    const size_t n_streams = std::min<size_t>(P.launches.size(), CLASS_STREAMS);
    if (!class_streams(c, n_streams, P.launches.size()) ||
        !check(c, hipEventCreateWithFlags(&g->fork, hipEventDisableTiming), "hipEventCreate"))
        return LYG_ERR_DEVICE;
    const std::vector<size_t> stream_of = assign_streams(P, n_streams);
    // head: counters and the status a path keeps if no launch reaches it
    if (!capture_on(c, s, g, &g->head, [&]() {
            return check(c, lyg_launch_fill_u32(counters, MAX_LAUNCHES + 1, 0, s), "reset") &&
                   check(c, lyg_launch_fill_list_u32(L.path_status, d_list, (uint32_t)m,
                                                     ST_UNSWEPT, s),
                         "status");
        }))
        return LYG_ERR_DEVICE;
    // one graph per class stream: its launches, largest class first
    for (size_t si = 0; si < n_streams; ++si) {
        hipStream_t cs = c->class_stream[si];
        hipGraphExec_t exec = nullptr;
        hipEvent_t ev = nullptr;
        if (!check(c, hipEventCreateWithFlags(&ev, hipEventDisableTiming), "hipEventCreate"))
            return LYG_ERR_DEVICE;
        g->join.push_back(ev);
        const bool ok = capture_on(c, cs, g, &exec, [&]() {
            for (size_t r = 0; r < P.launches.size(); ++r) {
                const size_t li = P.launches.size() - 1 - r;
                if (stream_of[li] != si) continue;
                set_launch(L, P, li, as<uint8_t>(g->work), d_list, d_off, counters);
                if (!check(c, lyg_launch_fill_sweep(&L, (int)P.launches[li].blocks, cs),
                           "fill sweep"))
                    return false;
            }
            return true;
        });
        if (exec) {
            g->part.push_back(exec);
            g->part_stream.push_back(cs);
        }
        if (!ok) return LYG_ERR_DEVICE;
    }
    // tail: offsets, copy (each path's range checked against the output capacity), statuses
    lyg::FillLaunch K{};
    K.in = L.in;
    K.path_status = L.path_status;
    K.path_round = L.path_round;
    K.vtx_off = out->path_vertex_offsets;
    K.idx_off = out->path_index_offsets;
    K.out_vertices = out->vertices;
    K.out_indices = out->indices;
    K.vertex_offset = out->vertex_offset;
    K.local_ids = 0;
    K.vertex_capacity = out->vertex_capacity;
    K.index_capacity = out->index_capacity;
    K.round = 0;
    K.path_list = d_list;
    K.n_list = (uint32_t)m;
    K.stage_v = L.stage_v;
    K.stage_t = L.stage_t;
    K.stage_v_off = d_off;
    K.stage_t_off = d_off + (m + 1);
    uint64_t* totals = as<uint64_t>(g->totals);
    if (!capture_on(c, s, g, &g->tail, [&]() {
            return check(c, scan_with(c, g->tiles, L.path_nv, n, out->path_vertex_offsets,
                                      totals + 0, overflow, s),
                         "scan") &&
                   check(c, scan_with(c, g->tiles, L.path_ni, n, out->path_index_offsets,
                                      totals + 1, overflow, s),
                         "scan") &&
                   check(c, lyg_launch_fill_copy(&K, s), "fill copy") &&
                   (!out->path_status ||
                    check(c, lyg_launch_fill_status_out(
                              L.path_status, reinterpret_cast<uint32_t*>(out->path_status), n, s),
                          "status"));
        }))
        return LYG_ERR_DEVICE;
    return LYG_OK;
}

int32_t lygi::fill_graph_launch(FillGraph* g) {
    lyg_ctx* c = g->c;
    hipStream_t s = c->stream;
    if (!check(c, hipGraphLaunch(g->head, s), "graph launch") ||
        !check(c, hipEventRecord(g->fork, s), "event"))
        return LYG_ERR_DEVICE;
    for (size_t i = 0; i < g->part.size(); ++i) {
        hipStream_t cs = g->part_stream[i];
        if (!check(c, hipStreamWaitEvent(cs, g->fork, 0), "wait") ||
            !check(c, hipGraphLaunch(g->part[i], cs), "graph launch") ||
            !check(c, hipEventRecord(g->join[i], cs), "event") ||
            !check(c, hipStreamWaitEvent(s, g->join[i], 0), "wait"))
            return LYG_ERR_DEVICE;
    }
    return check(c, hipGraphLaunch(g->tail, s), "graph launch") ? LYG_OK : LYG_ERR_DEVICE;
}

// Waits for the last replay; statuses in lyg_fill_batch's order: 32-bit offsets or vertex
// ids out of range, then the output capacity, then the first failing path.
int32_t lygi::fill_graph_result(FillGraph* g, uint64_t* n_vertices, uint64_t* n_indices) {
    lyg_ctx* c = g->c;
    hipStream_t s = c->stream;
    uint64_t tot[2] = {0, 0};
    uint32_t ovf = 0;
    std::vector<uint32_t> st(g->n);
    if (!check(c, hipStreamSynchronize(s), "sync") ||
        !check(c, hipMemcpy(tot, g->totals.p, 16, hipMemcpyDeviceToHost), "D2H") ||
        !check(c, hipMemcpy(&ovf, as<uint32_t>(g->counters) + MAX_LAUNCHES, 4,
                            hipMemcpyDeviceToHost),
               "D2H") ||
        !check(c, hipMemcpy(st.data(), g->status.p, 4ull * g->n, hipMemcpyDeviceToHost), "D2H"))
        return LYG_ERR_DEVICE;
    if (n_vertices) *n_vertices = tot[0];
    if (n_indices) *n_indices = tot[1];
    if (ovf) return LYG_ERR_TOO_MANY_VERTICES;
    // BuffersBuilder::add_vertex_impl: ids up to u32::MAX - 1 (geometry_builder.rs:338-350)
    if (tot[0] > 0 && (uint64_t)g->vertex_offset + tot[0] - 1 > 0xfffffffeull)
        return LYG_ERR_TOO_MANY_VERTICES;
    if (tot[0] > g->vertex_capacity || tot[1] > g->index_capacity) return LYG_ERR_CAPACITY;
    for (uint32_t v : st) {
        if (v == 0) continue;
        if (v == ST_RETRY) return LYG_ERR_NO_SPACE;
        if (v == ST_UNSWEPT) return LYG_ERR_INTERNAL;
        return (int32_t)v;
    }
    return LYG_OK;
}

void lygi::fill_graph_destroy(FillGraph* g) {
    if (!g) return;
    if (g->head) (void)hipGraphExecDestroy(g->head);
    if (g->tail) (void)hipGraphExecDestroy(g->tail);
    for (hipGraphExec_t e : g->part) (void)hipGraphExecDestroy(e);
    for (hipGraph_t gr : g->graphs) (void)hipGraphDestroy(gr);
    if (g->fork) (void)hipEventDestroy(g->fork);
    for (hipEvent_t e : g->join) (void)hipEventDestroy(e);
    for (DevBuf* b : {&g->status, &g->round, &g->nv, &g->ni, &g->list, &g->off, &g->stage_v,
                      &g->stage_t, &g->work, &g->counters, &g->totals, &g->tiles})
        free_buf(*b);
    delete g;
}

extern "C" {

// FillTessellator::tessellate_path (fill.rs:628-641) into caller-owned host arrays.
int32_t lyg_fill_path(lyg_ctx* c, const lyg_path_view* path, const lyg_fill_options* opts,
                      lyg_output* out) {
    if (!c || !path || !opts || !out) return LYG_ERR_INVALID_ARGUMENT;
    if (opts->fill_rule > 1 || opts->sweep_orientation > 1) return LYG_ERR_INVALID_ARGUMENT;
    if ((path->n_verbs && !path->verbs) || (path->n_points && !path->points))
        return LYG_ERR_INVALID_ARGUMENT;
    out->n_vertices = 0;
    out->n_indices = 0;
    // tessellate_impl checks the tolerance before anything else is observable (fill.rs:755-759)
    if (!(opts->tolerance > 0.0f)) return LYG_ERR_TOLERANCE_IS_NAN;
    if (path->n_verbs == 0) return LYG_OK;
    if (!check(c, hipSetDevice(c->device), "hipSetDevice")) return LYG_ERR_DEVICE;
    if (!ensure(c, c->in_verbs, path->n_verbs) || !ensure(c, c->in_points, 8ull * path->n_points) ||
        !ensure(c, c->in_offs, 16) || !ensure(c, c->in_opts, sizeof(lyg_fill_options)) ||
        !ensure(c, c->offs_v, 8) || !ensure(c, c->offs_i, 8))
        return LYG_ERR_DEVICE;
    uint32_t* h = reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(c->h_pinned) + 64);
    h[0] = 0;
    h[1] = path->n_verbs;
    h[2] = 0;
    h[3] = path->n_points;
    hipStream_t s = c->stream;
    if (!check(c, hipMemcpyAsync(c->in_verbs.p, path->verbs, path->n_verbs, hipMemcpyHostToDevice, s), "H2D") ||
        (path->n_points &&
         !check(c, hipMemcpyAsync(c->in_points.p, path->points, 8ull * path->n_points, hipMemcpyHostToDevice, s), "H2D")) ||
        !check(c, hipMemcpyAsync(c->in_offs.p, h, 16, hipMemcpyHostToDevice, s), "H2D") ||
        !check(c, hipMemcpyAsync(c->in_opts.p, opts, sizeof(*opts), hipMemcpyHostToDevice, s), "H2D"))
        return LYG_ERR_DEVICE;
    FillJob J;
    J.in.verbs = as<uint8_t>(c->in_verbs);
    J.in.points = as<float>(c->in_points);
    J.in.verb_off = as<uint32_t>(c->in_offs);
    J.in.point_off = as<uint32_t>(c->in_offs) + 2;
    J.in.n_paths = 1;
    J.in.first_path = 0;
    J.opts = as<lyg_fill_options>(c->in_opts);
    J.opts_stride = 0;
    J.vertex_offset = out->vertex_offset;
    FillResult R;
    int32_t st = fill_sweep(c, J, as<uint32_t>(c->offs_v), as<uint32_t>(c->offs_i), &R);
    if (st != LYG_OK) return st;
    st = first_error(R);
    if (st != LYG_OK) return st;
    const uint64_t tv = R.total_v, ti = R.total_i;
    if (tv > out->vertex_capacity || ti > out->index_capacity) {
        out->n_vertices = (uint32_t)tv;
        out->n_indices = (uint32_t)ti;
        return LYG_ERR_CAPACITY;
    }
    if (tv == 0 && ti == 0) return LYG_OK;
    if (!ensure(c, c->out_v, 8 * tv) || !ensure(c, c->out_i, 4 * ti)) return LYG_ERR_DEVICE;
    st = fill_copy(c, J, R, as<uint32_t>(c->offs_v), as<uint32_t>(c->offs_i), as<float>(c->out_v),
                   as<uint32_t>(c->out_i));
    if (st != LYG_OK) return st;
    if (!check(c, hipMemcpyAsync(out->vertices, c->out_v.p, 8 * tv, hipMemcpyDeviceToHost, s), "D2H") ||
        (ti && !check(c, hipMemcpyAsync(out->indices, c->out_i.p, 4 * ti, hipMemcpyDeviceToHost, s), "D2H")) ||
        !check(c, hipStreamSynchronize(s), "sync"))
        return LYG_ERR_DEVICE;
    out->n_vertices = (uint32_t)tv;
    out->n_indices = (uint32_t)ti;
    return LYG_OK;
}

// Batched fill: every path as one tessellate_path call into a shared VertexBuffers; a
// path that fails contributes nothing (its geometry is aborted, fill.rs:788-793) and the
// first failing path's status is returned.
int32_t lyg_fill_batch(lyg_ctx* c, const lyg_batch_in* in, const lyg_fill_options* opts,
                       uint32_t n_opts, lyg_batch_out* out) {
    if (!c || !in || !opts || !out) return LYG_ERR_INVALID_ARGUMENT;
    if (n_opts != 1 && n_opts != in->n_paths) return LYG_ERR_INVALID_ARGUMENT;
    out->n_vertices = 0;
    out->n_indices = 0;
    if (in->n_paths == 0) return LYG_OK;
    if (!check(c, hipSetDevice(c->device), "hipSetDevice")) return LYG_ERR_DEVICE;
    lygi::BatchSpec b{};
    b.in.verbs = in->verbs;
    b.in.points = in->points;
    b.in.verb_off = in->verb_offsets;
    b.in.point_off = in->point_offsets;
    b.in.n_paths = in->n_paths;
    b.opts = opts;
    b.opts_stride = n_opts == 1 ? 0 : 1;
    b.vertex_offset = out->vertex_offset;
    b.local_ids = false;
    b.vtx_off = out->path_vertex_offsets;
    b.idx_off = out->path_index_offsets;
    uint64_t tv = 0, ti = 0;
    std::vector<uint32_t> status;
    int32_t st = lygi::fill_batch(
        c, b,
        [&](uint64_t nv, uint64_t ni, float** v, uint32_t** i) -> int32_t {
            if (out->path_status &&
                !check(c, hipMemcpyAsync(out->path_status, status.data(), 4ull * in->n_paths,
                                         hipMemcpyHostToDevice, c->stream), "H2D"))
                return LYG_ERR_DEVICE;
            if (nv > out->vertex_capacity || ni > out->index_capacity) return LYG_ERR_CAPACITY;
            *v = out->vertices;
            *i = out->indices;
            return LYG_OK;
        },
        &tv, &ti, &status);
    out->n_vertices = tv;
    out->n_indices = ti;
    if (!check(c, hipStreamSynchronize(c->stream), "sync")) return LYG_ERR_DEVICE;
    if (st != LYG_OK) return st;
    for (uint32_t x : status)
        if (x != 0) return (int32_t)x;
    return LYG_OK;
}

}  // extern "C"
