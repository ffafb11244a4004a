// Host side of the SVG path data parser (svg.hip): lyon_extra's PathParser::parse
// (crates/extra/src/parser.rs:154-175) for many strings at once, and for one host string.
#include "ctx.h"
#include "lyg_types.h"

extern "C" {
hipError_t lyg_launch_svg_parse(const lyg::SvgLaunch* L, hipStream_t s);
hipError_t lyg_launch_svg_compact(const lyg::SvgLaunch* L, uint8_t* verbs, float* points,
                                  hipStream_t s);
hipError_t lyg_launch_fill_u32(uint32_t* p, uint32_t n, uint32_t v, hipStream_t s);
}

extern "C" {

int32_t lyg_parse_svg_paths(lyg_ctx* c, const uint8_t* text, const uint32_t* text_offsets,
                            uint32_t n, uint32_t num_attributes, int32_t stop_at,
                            uint8_t* verbs, uint64_t verb_capacity, float* points,
                            uint64_t point_capacity, uint32_t* verb_offsets,
                            uint32_t* point_offsets, int32_t* errors, uint64_t* n_verbs,
                            uint64_t* n_points, uint32_t* n_errors) {
    if (!c || (n && (!text || !text_offsets)) || !verb_offsets || !point_offsets ||
        num_attributes > 64 || stop_at < -1 || stop_at > 127)
        return LYG_ERR_INVALID_ARGUMENT;
    if (n_verbs) *n_verbs = 0;
    if (n_points) *n_points = 0;
    if (n_errors) *n_errors = 0;
    if (!check(c, hipSetDevice(c->device), "hipSetDevice")) return LYG_ERR_DEVICE;
    hipStream_t s = c->stream;
    if (n == 0) {
        return check(c, lyg_launch_fill_u32(verb_offsets, 1, 0, s), "zero") &&
                       check(c, lyg_launch_fill_u32(point_offsets, 1, 0, s), "zero") &&
                       check(c, hipStreamSynchronize(s), "sync")
                   ? LYG_OK
                   : LYG_ERR_DEVICE;
    }
    // the staging rooms follow the text: string i's verbs at text_offsets[i] + 2 i (its
    // length + 2), its points at 2 text_offsets[i] + 2 i (twice its length + 2)
    uint32_t* h32 = reinterpret_cast<uint32_t*>(c->h_pinned);
    if (!check(c, hipMemcpyAsync(h32, text_offsets + n, 4, hipMemcpyDeviceToHost, s), "D2H") ||
        !check(c, hipStreamSynchronize(s), "sync"))
        return LYG_ERR_DEVICE;
    const uint64_t text_bytes = h32[0];
    if (!ensure(c, c->counts_v, 4ull * n) || !ensure(c, c->counts_i, 4ull * n) ||
        !ensure(c, c->st_i, text_bytes + 2ull * n) ||
        !ensure(c, c->st_v, 8ull * (2 * text_bytes + 2ull * n)))
        return LYG_ERR_DEVICE;
    lyg::SvgLaunch L{};
    L.text = text;
    L.text_off = text_offsets;
    L.n = n;
    L.num_attributes = (int)num_attributes;
    L.stop_at = stop_at;
    L.counts_v = as<uint32_t>(c->counts_v);
    L.counts_p = as<uint32_t>(c->counts_i);
    L.errors = errors;
    L.n_errors = c->d_counters + 5;  // [5] paths with an error, [6] rooms overflowed
    L.verbs = as<uint8_t>(c->st_i);
    L.points = as<float>(c->st_v);
    if (!check(c, lyg_launch_fill_u32(c->d_counters, 8, 0, s), "reset") ||
        !check(c, lyg_launch_svg_parse(&L, s), "svg parse") ||
        !check(c, ctx_scan(c, L.counts_v, n, verb_offsets, c->d_totals, c->d_counters + 4, s), "scan") ||
        !check(c, ctx_scan(c, L.counts_p, n, point_offsets, c->d_totals + 1, c->d_counters + 4, s), "scan"))
        return LYG_ERR_DEVICE;
    uint64_t* h = reinterpret_cast<uint64_t*>(c->h_pinned);
    if (!check(c, hipMemcpyAsync(h, c->d_totals, 16, hipMemcpyDeviceToHost, s), "D2H") ||
        !check(c, hipMemcpyAsync(h + 2, c->d_counters + 4, 12, hipMemcpyDeviceToHost, s), "D2H") ||
        !check(c, hipStreamSynchronize(s), "sync"))
        return LYG_ERR_DEVICE;
    const uint64_t tv = h[0], tp = h[1];
    const uint32_t* hc = reinterpret_cast<const uint32_t*>(h + 2);
    const uint32_t overflow = hc[0], nerr = hc[1], room_overflow = hc[2];
    if (n_verbs) *n_verbs = tv;
    if (n_points) *n_points = tp;
    if (n_errors) *n_errors = nerr;
    if (room_overflow) {
        c->last_error = "svg parse: a path outgrew its staging room";
        return LYG_ERR_INTERNAL;
    }
    if (overflow) {
        c->last_error = "more than 2^32 verbs or points";
        return LYG_ERR_TOO_MANY_VERTICES;
    }
    if (tv > verb_capacity || tp > point_capacity || (tv && !verbs) || (tp && !points))
        return LYG_ERR_CAPACITY;
    L.off_v = verb_offsets;
    L.off_p = point_offsets;
    if (!check(c, lyg_launch_svg_compact(&L, verbs, points, s), "svg compact") ||
        !check(c, hipStreamSynchronize(s), "sync"))
        return LYG_ERR_DEVICE;
    return LYG_OK;
}

int32_t lyg_parse_svg_path(lyg_ctx* c, const char* text, uint32_t len, uint32_t num_attributes,
                           int32_t stop_at, uint8_t* verbs, uint32_t verb_capacity,
                           float* points, uint32_t point_capacity, uint32_t* n_verbs,
                           uint32_t* n_points, int32_t* error) {
    if (!c || (len && !text) || !n_verbs || !n_points) return LYG_ERR_INVALID_ARGUMENT;
    *n_verbs = *n_points = 0;
    if (!check(c, hipSetDevice(c->device), "hipSetDevice")) return LYG_ERR_DEVICE;
    hipStream_t s = c->stream;
    // in_verbs: the text; in_offs: text offsets (2), then verb / point offsets (2 + 2);
    // misc: the error record; out_v / out_i: points and verbs
    uint32_t* hp = reinterpret_cast<uint32_t*>(c->h_pinned);
    if (!ensure(c, c->in_verbs, len) || !ensure(c, c->in_offs, 32) || !ensure(c, c->misc, 16))
        return LYG_ERR_DEVICE;
    hp[0] = 0;
    hp[1] = len;
    if ((len && !check(c, hipMemcpyAsync(c->in_verbs.p, text, len, hipMemcpyHostToDevice, s), "H2D")) ||
        !check(c, hipMemcpyAsync(c->in_offs.p, hp, 8, hipMemcpyHostToDevice, s), "H2D") ||
        !check(c, hipStreamSynchronize(s), "sync"))
        return LYG_ERR_DEVICE;
    uint32_t* offs = as<uint32_t>(c->in_offs);
    uint64_t tv = 0, tp = 0;
    uint32_t nerr = 0;
    // one parse into buffers of the staging bound (every output of `len` bytes fits)
    const uint64_t cap_v = (uint64_t)len + 2, cap_p = 2ull * len + 2;
    if (!ensure(c, c->out_v, 8 * cap_p) || !ensure(c, c->out_i, cap_v)) return LYG_ERR_DEVICE;
    int32_t st = lyg_parse_svg_paths(c, as<uint8_t>(c->in_verbs), offs, 1, num_attributes,
                                     stop_at, as<uint8_t>(c->out_i), cap_v, as<float>(c->out_v),
                                     cap_p, offs + 2, offs + 4, as<int32_t>(c->misc), &tv, &tp,
                                     &nerr);
    if (st != LYG_OK) return st;
    *n_verbs = (uint32_t)tv;
    *n_points = (uint32_t)tp;
    if (error && !check(c, hipMemcpy(error, c->misc.p, 16, hipMemcpyDeviceToHost), "D2H"))
        return LYG_ERR_DEVICE;
    if (tv > verb_capacity || tp > point_capacity || (tv && !verbs) || (tp && !points))
        return LYG_ERR_CAPACITY;
    if ((tv && !check(c, hipMemcpyAsync(verbs, c->out_i.p, tv, hipMemcpyDeviceToHost, s), "D2H")) ||
        (tp && !check(c, hipMemcpyAsync(points, c->out_v.p, 8 * tp, hipMemcpyDeviceToHost, s), "D2H")) ||
        !check(c, hipStreamSynchronize(s), "sync"))
        return LYG_ERR_DEVICE;
    return LYG_OK;
}

}  // extern "C"
