// Host side of the variable-width stroke (StrokeOptions::variable_line_width): a count
// launch (one lane per path), the two offset scans, the output check, an emit launch.
#include <algorithm>
#include <vector>

#include "ctx.h"
#include "internal.h"
#include "lyg_types.h"

extern "C" {
hipError_t lyg_launch_stroke_vw(const lyg::StrokeVWLaunch* L, int emit, int grid,
                                hipStream_t s);
hipError_t lyg_launch_fill_u32(uint32_t* p, uint32_t n, uint32_t v, hipStream_t s);
}

namespace {

bool valid_opts(const lyg_stroke_options& o) {
    return o.start_cap <= 2 && o.end_cap <= 2 && o.line_join <= 3;
}

// Runs the batch `L` (inputs, options, widths set); offsets go to vtx_off / idx_off (n + 1
// each), per-path statuses to c->f_status; `alloc` sets the output arrays once the totals
// are known.
int32_t stroke_vw_run(lyg_ctx* c, lyg::StrokeVWLaunch& L, uint32_t* vtx_off, uint32_t* idx_off,
                      uint32_t vertex_offset, const lygi::OutAlloc& alloc, uint64_t* tv,
                      uint64_t* ti, std::vector<uint32_t>* status) {
    const uint32_t n = L.in.n_paths;
    hipStream_t s = c->stream;
    if (!ensure(c, c->counts_v, 4ull * n) || !ensure(c, c->counts_i, 4ull * n) ||
        !ensure(c, c->f_status, 4ull * n))
        return LYG_ERR_DEVICE;
    L.path_nv = as<uint32_t>(c->counts_v);
    L.path_ni = as<uint32_t>(c->counts_i);
    L.path_status = as<uint32_t>(c->f_status);
    L.vertex_offset = vertex_offset;
    L.err = c->d_counters + 5;
    // about six resident waves per SIMD (four SIMDs per CU), each with as few paths as
    // fill them
    const uint32_t slots = (uint32_t)std::max(1, c->cu_count) * 4u * 6u;
    L.paths_per_wave = std::min<uint32_t>(64u, std::max<uint32_t>(1u, (n + slots - 1) / slots));
    const int grid =
        (int)std::min<uint32_t>((n + L.paths_per_wave - 1) / L.paths_per_wave, 1u << 20);
    if (!check(c, lyg_launch_fill_u32(c->d_counters + 4, 2, 0, s), "reset") ||
        !check(c, lyg_launch_stroke_vw(&L, 0, grid, s), "stroke vw count") ||
        !check(c, ctx_scan(c, L.path_nv, n, vtx_off, c->d_totals + 0, c->d_counters + 4, s),
               "scan") ||
        !check(c, ctx_scan(c, L.path_ni, n, idx_off, c->d_totals + 1, c->d_counters + 4, s),
               "scan"))
        return LYG_ERR_DEVICE;
    uint64_t* h = reinterpret_cast<uint64_t*>(c->h_pinned);
    status->assign(n, 0);
    if (!check(c, hipMemcpyAsync(h, c->d_totals, 16, hipMemcpyDeviceToHost, s), "D2H") ||
        !check(c, hipMemcpyAsync(h + 2, c->d_counters + 4, 4, hipMemcpyDeviceToHost, s), "D2H") ||
        !check(c, hipMemcpyAsync(status->data(), L.path_status, 4ull * n, hipMemcpyDeviceToHost,
                                 s),
               "D2H") ||
        !check(c, hipStreamSynchronize(s), "sync"))
        return LYG_ERR_DEVICE;
    *tv = h[0];
    *ti = h[1];
    if ((uint32_t)h[2]) return LYG_ERR_TOO_MANY_VERTICES;
    // BuffersBuilder::add_vertex_impl: ids up to u32::MAX - 1 (geometry_builder.rs:338-350)
    if (*tv > 0 && (uint64_t)vertex_offset + *tv - 1 > 0xfffffffeull)
        return LYG_ERR_TOO_MANY_VERTICES;
    float* ov = nullptr;
    uint32_t* oi = nullptr;
    int32_t st = alloc(*tv, *ti, &ov, &oi);
    if (st != LYG_OK) return st;
    L.vtx_off = vtx_off;
    L.idx_off = idx_off;
    L.out_vertices = ov;
    L.out_indices = oi;
    uint32_t err = 0;
    if (!check(c, lyg_launch_stroke_vw(&L, 1, grid, s), "stroke vw emit") ||
        !check(c, hipMemcpyAsync(&err, c->d_counters + 5, 4, hipMemcpyDeviceToHost, s), "D2H") ||
        !check(c, hipStreamSynchronize(s), "sync"))
        return LYG_ERR_DEVICE;
    // the emit pass wrote past its path's counted range: the two passes disagreed
    if (err) return LYG_ERR_INTERNAL;
    return LYG_OK;
}

int32_t first_failure(const std::vector<uint32_t>& st) {
    for (uint32_t v : st)
        if (v) return (int32_t)v;
    return LYG_OK;
}

}  // namespace

extern "C" {

// StrokeTessellator::tessellate_path with StrokeOptions::variable_line_width
// (stroke.rs:653-767) for a batch in device memory.
int32_t lyg_stroke_batch_vw(lyg_ctx* c, const lyg_batch_in* in, const float* attributes,
                            uint32_t n_attributes, uint32_t width_attribute,
                            const lyg_stroke_options* opts, uint32_t n_opts, lyg_batch_out* out,
                            float* out_attributes) {
    if (!c || !in || !opts || !out || !attributes) return LYG_ERR_INVALID_ARGUMENT;
    if (n_opts != 1 && n_opts != in->n_paths) return LYG_ERR_INVALID_ARGUMENT;
    if (width_attribute >= n_attributes) return LYG_ERR_INVALID_ARGUMENT;
    out->n_vertices = 0;
    out->n_indices = 0;
    if (in->n_paths == 0) return LYG_OK;
    if (!out->path_vertex_offsets || !out->path_index_offsets) return LYG_ERR_INVALID_ARGUMENT;
    if (!check(c, hipSetDevice(c->device), "hipSetDevice")) return LYG_ERR_DEVICE;
    lyg::StrokeVWLaunch L{};
    L.in.verbs = in->verbs;
    L.in.points = in->points;
    L.in.verb_off = in->verb_offsets;
    L.in.point_off = in->point_offsets;
    L.in.n_paths = in->n_paths;
    L.opts = opts;
    L.opts_stride = n_opts == 1 ? 0 : 1;
    L.widths = attributes + width_attribute;
    L.widths_stride = n_attributes;
    L.attrs = attributes;
    L.n_attrs = n_attributes;
    L.out_attrs = out_attributes;
    uint64_t tv = 0, ti = 0;
    std::vector<uint32_t> status;
    int32_t st = stroke_vw_run(
        c, L, out->path_vertex_offsets, out->path_index_offsets, out->vertex_offset,
        [&](uint64_t nv, uint64_t ni, float** v, uint32_t** i) -> int32_t {
            if (nv > out->vertex_capacity || ni > out->index_capacity) return LYG_ERR_CAPACITY;
            *v = out->vertices;
            *i = out->indices;
            return LYG_OK;
        },
        &tv, &ti, &status);
    out->n_vertices = tv;
    out->n_indices = ti;
    if (out->path_status && !status.empty() &&
        !check(c, hipMemcpy(out->path_status, status.data(), 4ull * in->n_paths,
                            hipMemcpyHostToDevice),
               "H2D"))
        return LYG_ERR_DEVICE;
    if (st != LYG_OK) return st;
    return first_failure(status);
}

// The same for one path in host memory (tessellate_path into caller-owned arrays);
// attributes: n_points x n_attributes floats.
int32_t lyg_stroke_path_vw(lyg_ctx* c, const lyg_path_view* path, const float* attributes,
                           uint32_t n_attributes, uint32_t width_attribute,
                           const lyg_stroke_options* opts, lyg_output* out,
                           float* out_attributes) {
    if (!c || !path || !opts || !out) return LYG_ERR_INVALID_ARGUMENT;
    if (!valid_opts(*opts) || width_attribute >= n_attributes) return LYG_ERR_INVALID_ARGUMENT;
    if ((path->n_verbs && !path->verbs) || (path->n_points && !path->points) ||
        (path->n_points && !attributes))
        return LYG_ERR_INVALID_ARGUMENT;
    out->n_vertices = 0;
    out->n_indices = 0;
    if (path->n_verbs == 0) return LYG_OK;
    if (!check(c, hipSetDevice(c->device), "hipSetDevice")) return LYG_ERR_DEVICE;
    const size_t attr_bytes = 4ull * n_attributes * path->n_points;
    if (!ensure(c, c->in_verbs, path->n_verbs) || !ensure(c, c->in_points, 8ull * path->n_points) ||
        !ensure(c, c->in_offs, 16) || !ensure(c, c->in_opts, sizeof(lyg_stroke_options)) ||
        !ensure(c, c->in_attrs, attr_bytes) || !ensure(c, c->offs_v, 8) ||
        !ensure(c, c->offs_i, 8))
        return LYG_ERR_DEVICE;
    uint32_t* h = reinterpret_cast<uint32_t*>(reinterpret_cast<char*>(c->h_pinned) + 64);
    h[0] = 0;
    h[1] = path->n_verbs;
    h[2] = 0;
    h[3] = path->n_points;
    hipStream_t s = c->stream;
    if (!check(c, hipMemcpyAsync(c->in_verbs.p, path->verbs, path->n_verbs, hipMemcpyHostToDevice, s), "H2D") ||
        (path->n_points &&
         (!check(c, hipMemcpyAsync(c->in_points.p, path->points, 8ull * path->n_points, hipMemcpyHostToDevice, s), "H2D") ||
          !check(c, hipMemcpyAsync(c->in_attrs.p, attributes, attr_bytes, hipMemcpyHostToDevice, s), "H2D"))) ||
        !check(c, hipMemcpyAsync(c->in_offs.p, h, 16, hipMemcpyHostToDevice, s), "H2D") ||
        !check(c, hipMemcpyAsync(c->in_opts.p, opts, sizeof(*opts), hipMemcpyHostToDevice, s), "H2D"))
        return LYG_ERR_DEVICE;
    lyg::StrokeVWLaunch L{};
    L.in.verbs = as<uint8_t>(c->in_verbs);
    L.in.points = as<float>(c->in_points);
    L.in.verb_off = as<uint32_t>(c->in_offs);
    L.in.point_off = as<uint32_t>(c->in_offs) + 2;
    L.in.n_paths = 1;
    L.opts = as<lyg_stroke_options>(c->in_opts);
    L.opts_stride = 0;
    L.widths = as<float>(c->in_attrs) + width_attribute;
    L.widths_stride = n_attributes;
    L.attrs = as<float>(c->in_attrs);
    L.n_attrs = n_attributes;
    uint64_t tv = 0, ti = 0;
    std::vector<uint32_t> status;
    int32_t st = stroke_vw_run(
        c, L, as<uint32_t>(c->offs_v), as<uint32_t>(c->offs_i), out->vertex_offset,
        [&](uint64_t nv, uint64_t ni, float** v, uint32_t** i) -> int32_t {
            if (nv > out->vertex_capacity || ni > out->index_capacity) {
                out->n_vertices = (uint32_t)nv;
                out->n_indices = (uint32_t)ni;
                return LYG_ERR_CAPACITY;
            }
            if (!ensure(c, c->out_v, 8 * nv) || !ensure(c, c->out_i, 4 * ni) ||
                (out_attributes && !ensure(c, c->out_attrs, 4ull * n_attributes * nv)))
                return LYG_ERR_DEVICE;
            L.out_attrs = out_attributes ? as<float>(c->out_attrs) : nullptr;
            *v = as<float>(c->out_v);
            *i = as<uint32_t>(c->out_i);
            return LYG_OK;
        },
        &tv, &ti, &status);
    if (st == LYG_OK) st = first_failure(status);
    if (st != LYG_OK) return st;
    if (tv && !check(c, hipMemcpy(out->vertices, c->out_v.p, 8 * tv, hipMemcpyDeviceToHost), "D2H"))
        return LYG_ERR_DEVICE;
    if (ti && !check(c, hipMemcpy(out->indices, c->out_i.p, 4 * ti, hipMemcpyDeviceToHost), "D2H"))
        return LYG_ERR_DEVICE;
    if (out_attributes && tv &&
        !check(c, hipMemcpy(out_attributes, c->out_attrs.p, 4ull * n_attributes * tv,
                            hipMemcpyDeviceToHost),
               "D2H"))
        return LYG_ERR_DEVICE;
    out->n_vertices = (uint32_t)tv;
    out->n_indices = (uint32_t)ti;
    return LYG_OK;
}

}  // extern "C"
