// Host side of the banded fill (band.hip): buffer sizing, the chain of launches and
// scans, retries of bands that outgrew their trapezoid room, and the C ABI entry points.
#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

#include "../../include/lyg.h"
#include "ctx.h"
#include "internal.h"
#include "lyg_types.h"

#include "band_params.h"

using BandParams = lyg::band::Params;

extern "C" {
hipError_t lyg_band_verbs(const BandParams* P, hipStream_t s);
hipError_t lyg_band_begins(const BandParams* P, hipStream_t s);
hipError_t lyg_band_edge_count(const BandParams* P, hipStream_t s);
hipError_t lyg_band_edge_emit(const BandParams* P, hipStream_t s);
hipError_t lyg_band_path_bands(const BandParams* P, const uint32_t* path_edge_off, hipStream_t s);
hipError_t lyg_band_hist(const BandParams* P, hipStream_t s);
hipError_t lyg_band_bounds(const BandParams* P, hipStream_t s);
hipError_t lyg_band_count(const BandParams* P, hipStream_t s);
hipError_t lyg_band_fill(const BandParams* P, hipStream_t s);
hipError_t lyg_band_sort(const BandParams* P, int8_t* bw, hipStream_t s);
hipError_t lyg_band_cross(const BandParams* P, int emit, hipStream_t s);
hipError_t lyg_band_event_count(const BandParams* P, hipStream_t s);
hipError_t lyg_band_event_sort(const BandParams* P, void* scratch, hipStream_t s);
hipError_t lyg_band_sweep(const BandParams* P, const int8_t* bw, hipStream_t s_small,
                          hipStream_t s_mid, hipStream_t s_big);
hipError_t lyg_band_expand_count(const BandParams* P, hipStream_t s);
hipError_t lyg_band_expand_write(const BandParams* P, hipStream_t s);
hipError_t lyg_band_path_offsets(const BandParams* P, uint32_t* path_edge_off, uint32_t* p_hbins,
                                 uint32_t* vtx_off, uint32_t* idx_off, int which, hipStream_t s);
hipError_t lyg_band_path_status(const BandParams* P, const uint32_t* b_status_in, hipStream_t s);
hipError_t lyg_band_flat_error(const BandParams* P, const float* tol, uint32_t tol_stride, int rule,
                               uint32_t* path_err, hipStream_t s);
hipError_t lyg_launch_fill_u32(uint32_t* p, uint32_t n, uint32_t v, hipStream_t s);
}

struct BandBufs {
    DevBuf v_path, v_npts, v_pbase, v_isbeg, v_begord, begin_verb, v_nedge, v_ebase;
    DevBuf geo, wind, e_path;
    DevBuf p_yr, p_status, p_nband, p_band_off, p_hbins, p_hist_off, p_edge_off, hist, p_voff, p_mult;
    DevBuf b_ytop, b_ybot, b_path, b_nedge, b_edge_off, b_fill, band_edges, bgeo, bw, b_ncross,
        b_cross_off, b_status;
    DevBuf cross_y, cross_ab, b_nev, b_ev_off, ev_y, ev_x, ev_k, b_ninit, init_order, ev_scratch;
    DevBuf b_rec_cap, b_rec_off, b_nrec, rec, r_nv, r_nt, r_voff, r_toff;
};

void lyg_band_free(lyg_ctx* c) {
    if (!c->band) return;
    BandBufs* B = c->band;
    DevBuf* all[] = {&B->v_path,     &B->v_npts,     &B->v_pbase,   &B->v_isbeg,   &B->v_begord,
                     &B->begin_verb, &B->v_nedge,    &B->v_ebase,   &B->geo,       &B->wind,
                     &B->e_path,     &B->p_yr,       &B->p_status,  &B->p_nband,   &B->p_band_off,
                     &B->p_hbins,    &B->p_hist_off, &B->p_edge_off, &B->hist,     &B->p_voff,
                     &B->p_mult,
                     &B->b_ytop,     &B->b_ybot,     &B->b_path,    &B->b_nedge,   &B->b_edge_off,
                     &B->b_fill,     &B->band_edges, &B->bgeo,      &B->bw,        &B->b_ncross,
                     &B->b_cross_off, &B->b_status,  &B->cross_y,   &B->cross_ab,  &B->b_nev,
                     &B->b_ev_off,   &B->ev_y,       &B->ev_x,      &B->ev_k,      &B->b_ninit,
                     &B->init_order, &B->ev_scratch, &B->b_rec_cap, &B->b_rec_off, &B->b_nrec,
                     &B->rec,        &B->r_nv,       &B->r_nt,      &B->r_voff,    &B->r_toff};
    for (DevBuf* b : all) free_buf(*b);
    delete B;
    c->band = nullptr;
}

namespace {

constexpr uint32_t BS_DENSE = 1, BS_RECORDS = 2, BS_BLOCK = 4, BS_NAN = 8;  // band.hip
constexpr uint32_t BAND_M_MAX = 4096;  // edges one band may hold (band.hip M_MAX)

// Exclusive scan of n counts into n + 1 offsets; returns the total (host-synchronous).
bool scan_total(lyg_ctx* c, const uint32_t* counts, uint32_t n, uint32_t* offsets, uint64_t* total) {
    hipStream_t s = c->stream;
    uint64_t* h = reinterpret_cast<uint64_t*>(reinterpret_cast<char*>(c->h_pinned) + 192);
    if (!check(c, lyg_launch_fill_u32(c->d_counters + 4, 1, 0, s), "reset") ||
        !check(c, ctx_scan(c, counts, n, offsets, c->d_totals + 2, c->d_counters + 4, s), "scan") ||
        !check(c, hipMemcpyAsync(h, c->d_totals + 2, 8, hipMemcpyDeviceToHost, s), "D2H") ||
        !check(c, hipMemcpyAsync(h + 1, c->d_counters + 4, 4, hipMemcpyDeviceToHost, s), "D2H") ||
        !check(c, hipStreamSynchronize(s), "sync"))
        return false;
    if ((uint32_t)h[1] != 0) {
        c->last_error = "banded fill: a count passed 2^32";
        return false;
    }
    *total = h[0];
    return true;
}

bool zero(lyg_ctx* c, DevBuf& b, size_t n_u32) {
    return n_u32 == 0 || check(c, lyg_launch_fill_u32(as<uint32_t>(b), (uint32_t)n_u32, 0, c->stream), "zero");
}

}  // namespace

namespace lygi {

// The banded fill of a batch (every pointer of b.in device memory).  alloc is called once
// the totals are known.  status gets every path's lyg_status.
int32_t fill_batch_banded(lyg_ctx* c, const BatchSpec& b, const OutAlloc& alloc, uint64_t* tv,
                          uint64_t* ti, std::vector<uint32_t>* status) {
    *tv = *ti = 0;
    const uint32_t np = b.in.n_paths;
    if (np == 0) return LYG_OK;
    if (!c->band) c->band = new BandBufs();
    BandBufs& B = *c->band;
    hipStream_t s = c->stream;
    uint32_t* hp = reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(c->h_pinned) + 160);
    // verbs in the batch
    if (!check(c, hipMemcpyAsync(hp, b.in.verb_off + np, 4, hipMemcpyDeviceToHost, s), "D2H") ||
        !check(c, hipStreamSynchronize(s), "sync"))
        return LYG_ERR_DEVICE;
    const uint32_t nv = hp[0];
    BandParams P{};
    P.verbs = b.in.verbs;
    P.points = b.in.points;
    P.verb_off = b.in.verb_off;
    P.point_off = b.in.point_off;
    P.n_paths = np;
    P.n_verbs = nv;
    P.opts = static_cast<const lyg_fill_options*>(b.opts);
    P.opts_stride = b.opts_stride ? 1 : 0;
    P.edges_per_band = c->band_edges_per_band;
    P.max_bands = c->band_max_bands;
    const size_t n1 = (size_t)nv + 1;
    if (!ensure(c, B.v_path, 4 * n1) || !ensure(c, B.v_npts, 4 * n1) ||
        !ensure(c, B.v_pbase, 4 * n1) || !ensure(c, B.v_isbeg, 4 * n1) ||
        !ensure(c, B.v_begord, 4 * n1) || !ensure(c, B.begin_verb, 4 * n1) ||
        !ensure(c, B.v_nedge, 4 * n1) || !ensure(c, B.v_ebase, 4 * n1) ||
        !ensure(c, B.p_yr, 8ull * np) || !ensure(c, B.p_status, 4ull * np) ||
        !ensure(c, B.p_nband, 4ull * (np + 1)) || !ensure(c, B.p_band_off, 4ull * (np + 1)) ||
        !ensure(c, B.p_hbins, 4ull * (np + 1)) || !ensure(c, B.p_hist_off, 4ull * (np + 1)) ||
        !ensure(c, B.p_edge_off, 4ull * (np + 1)) || !ensure(c, B.p_voff, 4ull * (np + 1)))
        return LYG_ERR_DEVICE;
    P.v_path = as<uint32_t>(B.v_path);
    P.v_npts = as<uint32_t>(B.v_npts);
    P.v_pbase = as<uint32_t>(B.v_pbase);
    P.v_isbeg = as<uint32_t>(B.v_isbeg);
    P.v_begord = as<uint32_t>(B.v_begord);
    P.begin_verb = as<uint32_t>(B.begin_verb);
    P.v_nedge = as<uint32_t>(B.v_nedge);
    P.v_ebase = as<uint32_t>(B.v_ebase);
    P.p_yr = as<uint32_t>(B.p_yr);
    P.p_status = as<uint32_t>(B.p_status);
    P.p_nband = as<uint32_t>(B.p_nband);
    P.p_band_off = as<uint32_t>(B.p_band_off);
    P.p_hist_off = as<uint32_t>(B.p_hist_off);
    uint64_t total = 0;
    // 1. verbs: points consumed, Begin ordinals, edges
    if (!check(c, lyg_band_verbs(&P, s), "band verbs") ||
        !scan_total(c, P.v_npts, nv, P.v_pbase, &total) ||
        !scan_total(c, P.v_isbeg, nv, P.v_begord, &total) ||
        !check(c, lyg_band_begins(&P, s), "band begins") ||
        !check(c, lyg_launch_fill_u32(P.p_status, np, 0, s), "zero") ||
        !check(c, lyg_band_edge_count(&P, s), "band edge count") ||
        !scan_total(c, P.v_nedge, nv, P.v_ebase, &total))
        return LYG_ERR_DEVICE;
    const uint64_t ne = total;
    if (ne >= 0xffffffffull) return LYG_ERR_TOO_MANY_VERTICES;
    P.n_edges = (uint32_t)ne;
    if (!ensure(c, B.geo, 16 * ne) || !ensure(c, B.wind, ne) || !ensure(c, B.e_path, 4 * ne))
        return LYG_ERR_DEVICE;
    P.geo = as<float4>(B.geo);
    P.wind = as<int8_t>(B.wind);
    P.e_path = as<uint32_t>(B.e_path);
    // y range per path: min starts at all ones, max at zero
    if (!check(c, lyg_launch_fill_u32(P.p_yr, 2 * np, 0, s), "zero") ||
        !check(c, lyg_band_path_offsets(&P, as<uint32_t>(B.p_edge_off), nullptr, nullptr, nullptr, 0, s),
               "band yr init") ||
        !check(c, lyg_band_edge_emit(&P, s), "band edge emit") ||
        !check(c, lyg_band_path_offsets(&P, as<uint32_t>(B.p_edge_off), nullptr, nullptr, nullptr, 1, s),
               "band path edges"))
        return LYG_ERR_DEVICE;
    // 2. band boundaries and band membership; a path with a band of more than M_MAX edges
    // (many edges starting close together) is cut into 4x more bands, up to 3 times
    std::vector<uint32_t> mult(np, 1);
    if (!ensure(c, B.p_mult, 4ull * np)) return LYG_ERR_DEVICE;
    uint64_t nb = 0;
    for (int round = 0;; ++round) {
        P.p_mult = nullptr;
        if (round > 0) {
            if (!check(c, hipMemcpyAsync(B.p_mult.p, mult.data(), 4ull * np, hipMemcpyHostToDevice, s), "H2D"))
                return LYG_ERR_DEVICE;
            P.p_mult = as<uint32_t>(B.p_mult);
        }
        if (!check(c, lyg_band_path_bands(&P, as<uint32_t>(B.p_edge_off), s), "band path bands") ||
            !check(c, lyg_band_path_offsets(&P, nullptr, as<uint32_t>(B.p_hbins), nullptr, nullptr, 2, s),
                   "band hist bins") ||
            !scan_total(c, P.p_nband, np, P.p_band_off, &total))
            return LYG_ERR_DEVICE;
        nb = total;
        uint64_t nbins = 0;
        if (!scan_total(c, as<uint32_t>(B.p_hbins), np, P.p_hist_off, &nbins)) return LYG_ERR_DEVICE;
        P.n_bands = (uint32_t)nb;
        if (!ensure(c, B.hist, 4 * (nbins + 1)) || !ensure(c, B.b_ytop, 4 * nb) ||
            !ensure(c, B.b_ybot, 4 * nb) || !ensure(c, B.b_path, 4 * nb) ||
            !ensure(c, B.b_nedge, 4 * (nb + 1)) || !ensure(c, B.b_edge_off, 4 * (nb + 1)) ||
            !ensure(c, B.b_fill, 4 * nb) || !ensure(c, B.b_ncross, 4 * (nb + 1)) ||
            !ensure(c, B.b_cross_off, 4 * (nb + 1)) || !ensure(c, B.b_status, 4 * nb) ||
            !ensure(c, B.b_nev, 4 * (nb + 1)) || !ensure(c, B.b_ev_off, 4 * (nb + 1)) ||
            !ensure(c, B.b_ninit, 4 * nb) || !ensure(c, B.b_rec_cap, 4 * (nb + 1)) ||
            !ensure(c, B.b_rec_off, 4 * (nb + 1)) || !ensure(c, B.b_nrec, 4 * nb))
            return LYG_ERR_DEVICE;
        P.hist = as<uint32_t>(B.hist);
        P.b_ytop = as<float>(B.b_ytop);
        P.b_ybot = as<float>(B.b_ybot);
        P.b_path = as<uint32_t>(B.b_path);
        P.b_nedge = as<uint32_t>(B.b_nedge);
        P.b_edge_off = as<uint32_t>(B.b_edge_off);
        P.b_fill = as<uint32_t>(B.b_fill);
        P.b_ncross = as<uint32_t>(B.b_ncross);
        P.b_cross_off = as<uint32_t>(B.b_cross_off);
        P.b_status = as<uint32_t>(B.b_status);
        P.b_nev = as<uint32_t>(B.b_nev);
        P.b_ev_off = as<uint32_t>(B.b_ev_off);
        P.b_ninit = as<uint32_t>(B.b_ninit);
        P.b_rec_cap = as<uint32_t>(B.b_rec_cap);
        P.b_rec_off = as<uint32_t>(B.b_rec_off);
        P.b_nrec = as<uint32_t>(B.b_nrec);
        if (!zero(c, B.hist, nbins) || !check(c, lyg_band_hist(&P, s), "band hist") ||
            !check(c, lyg_band_bounds(&P, s), "band bounds") || !zero(c, B.b_nedge, nb) ||
            !zero(c, B.b_status, nb) || !check(c, lyg_band_count(&P, s), "band count") ||
            !scan_total(c, P.b_nedge, (uint32_t)nb, P.b_edge_off, &total))
            return LYG_ERR_DEVICE;
        if (round >= 3) break;
        std::vector<uint32_t> bn(nb), bp(nb), pnb(np);
        if (!check(c, hipMemcpyAsync(bn.data(), P.b_nedge, 4 * nb, hipMemcpyDeviceToHost, s), "D2H") ||
            !check(c, hipMemcpyAsync(bp.data(), P.b_path, 4 * nb, hipMemcpyDeviceToHost, s), "D2H") ||
            !check(c, hipMemcpyAsync(pnb.data(), P.p_nband, 4ull * np, hipMemcpyDeviceToHost, s), "D2H") ||
            !check(c, hipStreamSynchronize(s), "sync"))
            return LYG_ERR_DEVICE;
        bool again = false;
        std::vector<uint8_t> bump(np, 0);
        for (uint64_t k = 0; k < nb; ++k)
            if (bn[k] > BAND_M_MAX && pnb[bp[k]] < P.max_bands) bump[bp[k]] = 1;
        for (uint32_t p = 0; p < np; ++p)
            if (bump[p]) {
                mult[p] *= 4;
                again = true;
            }
        if (!again) break;
    }
    const uint64_t nbe = total;
    if (!ensure(c, B.band_edges, 4 * nbe) || !ensure(c, B.bgeo, 16 * nbe) ||
        !ensure(c, B.bw, nbe) || !ensure(c, B.init_order, 2 * nbe))
        return LYG_ERR_DEVICE;
    P.band_edges = as<uint32_t>(B.band_edges);
    P.bgeo = as<float4>(B.bgeo);
    P.init_order = as<uint16_t>(B.init_order);
    // timing (lyg_ctx_kernel_times): band preparation (edge lists, crossings, events), then
    // the sweep
    if (c->timing) (void)hipEventRecord(c->ev[0], s);
    if (!zero(c, B.b_fill, nb) || !check(c, lyg_band_fill(&P, s), "band fill") ||
        !check(c, lyg_band_sort(&P, as<int8_t>(B.bw), s), "band sort") ||
        !check(c, lyg_band_cross(&P, 0, s), "band cross count") ||
        !scan_total(c, P.b_ncross, (uint32_t)nb, P.b_cross_off, &total))
        return LYG_ERR_DEVICE;
    const uint64_t ncross = total;
    if (!ensure(c, B.cross_y, 8 * ncross) || !ensure(c, B.cross_ab, 4 * ncross)) return LYG_ERR_DEVICE;
    P.cross_y = as<double>(B.cross_y);
    P.cross_ab = as<uint32_t>(B.cross_ab);
    if (!check(c, lyg_band_cross(&P, 1, s), "band cross emit") ||
        !check(c, lyg_band_event_count(&P, s), "band event count") ||
        !scan_total(c, P.b_nev, (uint32_t)nb, P.b_ev_off, &total))
        return LYG_ERR_DEVICE;
    const uint64_t nev = total;
    if (!ensure(c, B.ev_y, 8 * nev) || !ensure(c, B.ev_x, 4 * nev) || !ensure(c, B.ev_k, 4 * nev) ||
        !ensure(c, B.ev_scratch, 32 * nev + 64))
        return LYG_ERR_DEVICE;
    P.ev_y = as<double>(B.ev_y);
    P.ev_x = as<float>(B.ev_x);
    P.ev_k = as<uint32_t>(B.ev_k);
    if (!check(c, lyg_band_event_sort(&P, B.ev_scratch.p, s), "band event sort")) return LYG_ERR_DEVICE;
    if (c->timing) (void)hipEventRecord(c->ev[1], s);
    // 3. the sweep; bands whose trapezoids outgrew their room run again with the room they
    // reported
    for (int attempt = 0;; ++attempt) {
        if (!scan_total(c, P.b_rec_cap, (uint32_t)nb, P.b_rec_off, &total)) return LYG_ERR_DEVICE;
        const uint64_t nrec = total;
        if (nrec >= 0xffffffffull) return LYG_ERR_TOO_MANY_VERTICES;
        if (!ensure(c, B.rec, 16 * nrec)) return LYG_ERR_DEVICE;
        P.rec = as<uint4>(B.rec);
        P.n_records = (uint32_t)nrec;
        if (!zero(c, B.b_nrec, nb)) return LYG_ERR_DEVICE;
        if (c->timing) (void)hipEventRecord(c->ev[2], s);
        // the sweep's three size classes on the context's stream and two of its side
        // streams (default priority, made with the context), joined by events
        for (int k = 0; k < 2; ++k)
            if (!c->class_stream[k] &&
                !check(c, hipStreamCreateWithFlags(&c->class_stream[k], hipStreamNonBlocking),
                       "hipStreamCreate"))
                return LYG_ERR_DEVICE;
        for (int k = 0; k < 3; ++k)
            if (!c->class_ev[6 + k] &&
                !check(c, hipEventCreateWithFlags(&c->class_ev[6 + k], hipEventDisableTiming),
                       "hipEventCreate"))
                return LYG_ERR_DEVICE;
        hipStream_t s1 = c->class_stream[0], s2 = c->class_stream[1];
        if (!check(c, hipEventRecord(c->class_ev[6], s), "event") ||
            !check(c, hipStreamWaitEvent(s1, c->class_ev[6], 0), "wait") ||
            !check(c, hipStreamWaitEvent(s2, c->class_ev[6], 0), "wait") ||
            !check(c, lyg_band_sweep(&P, as<int8_t>(B.bw), s, s1, s2), "band sweep") ||
            !check(c, hipEventRecord(c->class_ev[7], s1), "event") ||
            !check(c, hipEventRecord(c->class_ev[8], s2), "event") ||
            !check(c, hipStreamWaitEvent(s, c->class_ev[7], 0), "wait") ||
            !check(c, hipStreamWaitEvent(s, c->class_ev[8], 0), "wait"))
            return LYG_ERR_DEVICE;
        if (c->timing) (void)hipEventRecord(c->ev[3], s);
        // any band out of room?
        std::vector<uint32_t> bst(nb);
        if (!check(c, hipMemcpyAsync(bst.data(), P.b_status, 4 * nb, hipMemcpyDeviceToHost, s), "D2H") ||
            !check(c, hipStreamSynchronize(s), "sync"))
            return LYG_ERR_DEVICE;
        bool retry = false;
        for (uint32_t x : bst) retry |= (x & BS_RECORDS) != 0;
        if (!retry || attempt >= 3) break;
        // clear the room-exceeded bits (their paths' bits too) and sweep everything again
        for (uint32_t& x : bst) x &= ~BS_RECORDS;
        if (!check(c, hipMemcpyAsync(P.b_status, bst.data(), 4 * nb, hipMemcpyHostToDevice, s), "H2D") ||
            !check(c, lyg_band_path_status(&P, P.b_status, s), "band path status"))
            return LYG_ERR_DEVICE;
    }
    // 4. triangles
    const uint64_t nrec = P.n_records;
    if (!ensure(c, B.r_nv, 4 * (nrec + 1)) || !ensure(c, B.r_nt, 4 * (nrec + 1)) ||
        !ensure(c, B.r_voff, 4 * (nrec + 1)) || !ensure(c, B.r_toff, 4 * (nrec + 1)))
        return LYG_ERR_DEVICE;
    P.r_nv = as<uint32_t>(B.r_nv);
    P.r_nt = as<uint32_t>(B.r_nt);
    P.r_voff = as<uint32_t>(B.r_voff);
    P.r_toff = as<uint32_t>(B.r_toff);
    uint64_t total_v = 0, total_t = 0;
    if (!check(c, lyg_band_expand_count(&P, s), "band expand count") ||
        !scan_total(c, P.r_nv, (uint32_t)nrec, P.r_voff, &total_v) ||
        !scan_total(c, P.r_nt, (uint32_t)nrec, P.r_toff, &total_t))
        return LYG_ERR_DEVICE;
    if (total_v >= 0xffffffffull || 3 * total_t >= 0xffffffffull) return LYG_ERR_TOO_MANY_VERTICES;
    // per-path statuses; failed paths produce nothing (their bands are skipped)
    std::vector<uint32_t> pst(np);
    if (!check(c, hipMemcpyAsync(pst.data(), P.p_status, 4ull * np, hipMemcpyDeviceToHost, s), "D2H") ||
        !check(c, hipStreamSynchronize(s), "sync"))
        return LYG_ERR_DEVICE;
    status->assign(np, LYG_OK);
    for (uint32_t p = 0; p < np; ++p) {
        const uint32_t x = pst[p];
        if (x & BS_NAN) {
            (*status)[p] = LYG_ERR_POSITION_IS_NAN;
        } else if (x) {
            (*status)[p] = LYG_ERR_INTERNAL;
            c->last_error = (x & BS_DENSE) ? "banded fill: more than 4096 edges cross one band"
                            : (x & BS_BLOCK) ? "banded fill: an event touched more than 32 edges"
                                             : "banded fill: trapezoid room exceeded";
        }
    }
    // per-path offsets (first vertex / index of each path) and the caller's buffers
    if (!check(c, lyg_band_path_offsets(&P, nullptr, nullptr, b.vtx_off, b.idx_off, 3, s),
               "band path offsets"))
        return LYG_ERR_DEVICE;
    P.p_voff = b.vtx_off;
    P.vertex_offset = b.vertex_offset;
    P.local_ids = b.local_ids ? 1u : 0u;
    *tv = total_v;
    *ti = 3 * total_t;
    float* ov = nullptr;
    uint32_t* oi = nullptr;
    int32_t st = alloc(total_v, 3 * total_t, &ov, &oi);
    if (st != LYG_OK) return st;
    P.out_v = ov;
    P.out_i = oi;
    if (total_v && !check(c, lyg_band_expand_write(&P, s), "band expand write")) return LYG_ERR_DEVICE;
    return LYG_OK;
}

}  // namespace lygi

extern "C" {

int32_t lyg_ctx_set_band_size(lyg_ctx* c, uint32_t edges_per_band, uint32_t max_bands) {
    if (!c || edges_per_band == 0 || max_bands == 0) return LYG_ERR_INVALID_ARGUMENT;
    c->band_edges_per_band = edges_per_band;
    c->band_max_bands = max_bands;
    return LYG_OK;
}

int32_t lyg_flattening_errors(lyg_ctx* c, const lyg_batch_in* in, const float* tolerance,
                              uint32_t n_tol, int stroke_rule, float* path_error) {
    if (!c || !in || !tolerance || !path_error) return LYG_ERR_INVALID_ARGUMENT;
    if (n_tol != 1 && n_tol != in->n_paths) return LYG_ERR_INVALID_ARGUMENT;
    const uint32_t np = in->n_paths;
    if (np == 0) return LYG_OK;
    if (!check(c, hipSetDevice(c->device), "hipSetDevice")) return LYG_ERR_DEVICE;
    if (!c->band) c->band = new BandBufs();
    BandBufs& B = *c->band;
    hipStream_t s = c->stream;
    uint32_t* hp = reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(c->h_pinned) + 160);
    if (!check(c, hipMemcpyAsync(hp, in->verb_offsets + np, 4, hipMemcpyDeviceToHost, s), "D2H") ||
        !check(c, hipStreamSynchronize(s), "sync"))
        return LYG_ERR_DEVICE;
    const uint32_t nv = hp[0];
    BandParams P{};
    P.verbs = in->verbs;
    P.points = in->points;
    P.verb_off = in->verb_offsets;
    P.point_off = in->point_offsets;
    P.n_paths = np;
    P.n_verbs = nv;
    const size_t n1 = (size_t)nv + 1;
    if (!ensure(c, B.v_path, 4 * n1) || !ensure(c, B.v_npts, 4 * n1) ||
        !ensure(c, B.v_pbase, 4 * n1) || !ensure(c, B.v_isbeg, 4 * n1) ||
        !ensure(c, B.v_begord, 4 * n1) || !ensure(c, B.begin_verb, 4 * n1))
        return LYG_ERR_DEVICE;
    P.v_path = as<uint32_t>(B.v_path);
    P.v_npts = as<uint32_t>(B.v_npts);
    P.v_pbase = as<uint32_t>(B.v_pbase);
    P.v_isbeg = as<uint32_t>(B.v_isbeg);
    P.v_begord = as<uint32_t>(B.v_begord);
    P.begin_verb = as<uint32_t>(B.begin_verb);
    uint64_t total = 0;
    uint32_t* err = reinterpret_cast<uint32_t*>(path_error);
    if (!check(c, lyg_band_verbs(&P, s), "band verbs") ||
        !scan_total(c, P.v_npts, nv, P.v_pbase, &total) ||
        !scan_total(c, P.v_isbeg, nv, P.v_begord, &total) ||
        !check(c, lyg_band_begins(&P, s), "band begins") ||
        !check(c, lyg_launch_fill_u32(err, np, 0, s), "zero") ||
        !check(c, lyg_band_flat_error(&P, tolerance, n_tol == 1 ? 0 : 1, stroke_rule ? 1 : 0, err, s),
               "flattening error") ||
        !check(c, hipStreamSynchronize(s), "sync"))
        return LYG_ERR_DEVICE;
    return LYG_OK;
}

int32_t lyg_fill_batch_banded(lyg_ctx* c, const lyg_batch_in* in, const lyg_fill_options* opts,
                              uint32_t n_opts, lyg_batch_out* out) {
    if (!c || !in || !opts || !out) return LYG_ERR_INVALID_ARGUMENT;
    if (n_opts != 1 && n_opts != in->n_paths) return LYG_ERR_INVALID_ARGUMENT;
    out->n_vertices = 0;
    out->n_indices = 0;
    if (in->n_paths == 0) return LYG_OK;
    std::vector<lyg_fill_options> ho(n_opts);
    if (!check(c, hipSetDevice(c->device), "hipSetDevice") ||
        !check(c, hipMemcpy(ho.data(), opts, sizeof(lyg_fill_options) * n_opts, hipMemcpyDeviceToHost),
               "D2H"))
        return LYG_ERR_DEVICE;
    for (const auto& o : ho) {
        if (!(o.tolerance > 0.0f)) return LYG_ERR_TOLERANCE_IS_NAN;
        if (o.fill_rule > 1 || o.sweep_orientation > 1) return LYG_ERR_INVALID_ARGUMENT;
    }
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
    int32_t st = lygi::fill_batch_banded(
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

int32_t lyg_fill_path_banded(lyg_ctx* c, const lyg_path_view* path, const lyg_fill_options* opts,
                             lyg_output* out) {
    if (!c || !path || !opts || !out) return LYG_ERR_INVALID_ARGUMENT;
    if (opts->fill_rule > 1 || opts->sweep_orientation > 1) return LYG_ERR_INVALID_ARGUMENT;
    if ((path->n_verbs && !path->verbs) || (path->n_points && !path->points))
        return LYG_ERR_INVALID_ARGUMENT;
    out->n_vertices = 0;
    out->n_indices = 0;
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
    lygi::BatchSpec b{};
    b.in.verbs = as<uint8_t>(c->in_verbs);
    b.in.points = as<float>(c->in_points);
    b.in.verb_off = as<uint32_t>(c->in_offs);
    b.in.point_off = as<uint32_t>(c->in_offs) + 2;
    b.in.n_paths = 1;
    b.opts = c->in_opts.p;
    b.opts_stride = 0;
    b.vertex_offset = out->vertex_offset;
    b.local_ids = true;
    b.vtx_off = as<uint32_t>(c->offs_v);
    b.idx_off = as<uint32_t>(c->offs_i);
    uint64_t tv = 0, ti = 0;
    std::vector<uint32_t> status;
    int32_t st = lygi::fill_batch_banded(
        c, b,
        [&](uint64_t nv, uint64_t ni, float** v, uint32_t** i) -> int32_t {
            if (!status.empty() && status[0] != LYG_OK) return (int32_t)status[0];
            if (nv > out->vertex_capacity || ni > out->index_capacity) return LYG_ERR_CAPACITY;
            if (!ensure(c, c->out_v, 8 * nv) || !ensure(c, c->out_i, 4 * ni)) return LYG_ERR_DEVICE;
            *v = as<float>(c->out_v);
            *i = as<uint32_t>(c->out_i);
            return LYG_OK;
        },
        &tv, &ti, &status);
    if (st == LYG_ERR_CAPACITY) {
        out->n_vertices = (uint32_t)tv;
        out->n_indices = (uint32_t)ti;
        return st;
    }
    if (st != LYG_OK) return st;
    if (!status.empty() && status[0] != LYG_OK) return (int32_t)status[0];
    if (tv && !check(c, hipMemcpyAsync(out->vertices, c->out_v.p, 8 * tv, hipMemcpyDeviceToHost, s), "D2H"))
        return LYG_ERR_DEVICE;
    if (ti && !check(c, hipMemcpyAsync(out->indices, c->out_i.p, 4 * ti, hipMemcpyDeviceToHost, s), "D2H"))
        return LYG_ERR_DEVICE;
    if (!check(c, hipStreamSynchronize(s), "sync")) return LYG_ERR_DEVICE;
    out->n_vertices = (uint32_t)tv;
    out->n_indices = (uint32_t)ti;
    return LYG_OK;
}

}  // extern "C"
