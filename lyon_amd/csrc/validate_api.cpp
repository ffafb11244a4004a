// Host side of the path validators (validate.hip): lyon_algorithms' hit_test_path /
// path_winding_number_at_position (crates/algorithms/src/hit_test.rs:8-83) and
// approximate_signed_area (crates/algorithms/src/area.rs:7-19), as batched calls over
// device-resident paths and as single-path calls over host buffers.
#include <cstring>

#include "ctx.h"
#include "lyg_types.h"

extern "C" {
hipError_t lyg_launch_winding(const lyg::ValidateLaunch* L, hipStream_t s);
hipError_t lyg_launch_signed_area(const lyg::ValidateLaunch* L, hipStream_t s);
hipError_t lyg_launch_fill_u32(uint32_t* p, uint32_t n, uint32_t v, hipStream_t s);
}

namespace {

// for_each_flattened's own precondition (quadratic_bezier.rs:344, cubic_bezier.rs:564):
// tolerance >= f32::EPSILON^2.  Smaller values make the segment counts unbounded.
bool tolerance_ok(float tol) { return tol >= 1.1920929e-07f * 1.1920929e-07f; }

lyg::ValidateLaunch launch_of(const lyg_batch_in* in, float tol) {
    lyg::ValidateLaunch L{};
    L.verbs = in->verbs;
    L.points = in->points;
    L.verb_off = in->verb_offsets;
    L.point_off = in->point_offsets;
    L.n_paths = in->n_paths;
    L.tolerance = tol;
    return L;
}

// Copies one host path into the context's input buffers as a one-path batch.
bool upload_path(lyg_ctx* c, const lyg_path_view* p, lyg_batch_in* in, hipStream_t s) {
    uint32_t* h = reinterpret_cast<uint32_t*>(c->h_pinned);
    if (!ensure(c, c->in_verbs, p->n_verbs) || !ensure(c, c->in_points, 8ull * p->n_points) ||
        !ensure(c, c->in_offs, 16))
        return false;
    h[0] = 0;
    h[1] = p->n_verbs;
    h[2] = 0;
    h[3] = p->n_points;
    if ((p->n_verbs && !check(c, hipMemcpyAsync(c->in_verbs.p, p->verbs, p->n_verbs,
                                                hipMemcpyHostToDevice, s), "H2D")) ||
        (p->n_points && !check(c, hipMemcpyAsync(c->in_points.p, p->points, 8ull * p->n_points,
                                                 hipMemcpyHostToDevice, s), "H2D")) ||
        !check(c, hipMemcpyAsync(c->in_offs.p, h, 16, hipMemcpyHostToDevice, s), "H2D") ||
        !check(c, hipStreamSynchronize(s), "sync"))  // the pinned staging is reused next
        return false;
    in->verbs = as<uint8_t>(c->in_verbs);
    in->points = as<float>(c->in_points);
    in->verb_offsets = as<uint32_t>(c->in_offs);
    in->point_offsets = as<uint32_t>(c->in_offs) + 2;
    in->n_paths = 1;
    return true;
}

bool path_ok(const lyg_path_view* p) {
    return p && (p->n_verbs == 0 || p->verbs) && (p->n_points == 0 || p->points);
}

}  // namespace

extern "C" {

int32_t lyg_winding_numbers(lyg_ctx* c, const lyg_batch_in* in, const float* queries,
                            const uint32_t* query_paths, uint32_t n_queries, float tolerance,
                            uint8_t fill_rule, int32_t* winding, uint8_t* inside) {
    if (!c || !in || (n_queries && (!queries || !query_paths || !winding)) || fill_rule > 1)
        return LYG_ERR_INVALID_ARGUMENT;
    if (!tolerance_ok(tolerance)) return LYG_ERR_TOLERANCE_IS_NAN;
    if (n_queries == 0) return LYG_OK;
    if (!check(c, hipSetDevice(c->device), "hipSetDevice")) return LYG_ERR_DEVICE;
    lyg::ValidateLaunch L = launch_of(in, tolerance);
    L.queries = queries;
    L.query_path = query_paths;
    L.n_queries = n_queries;
    L.fill_rule = fill_rule;
    L.winding = winding;
    L.inside = inside;
    if (!check(c, lyg_launch_winding(&L, c->stream), "winding") ||
        !check(c, hipStreamSynchronize(c->stream), "sync"))
        return LYG_ERR_DEVICE;
    return LYG_OK;
}

int32_t lyg_signed_areas(lyg_ctx* c, const lyg_batch_in* in, float tolerance, float* areas) {
    if (!c || !in || (in->n_paths && !areas)) return LYG_ERR_INVALID_ARGUMENT;
    if (!tolerance_ok(tolerance)) return LYG_ERR_TOLERANCE_IS_NAN;
    if (in->n_paths == 0) return LYG_OK;
    if (!check(c, hipSetDevice(c->device), "hipSetDevice")) return LYG_ERR_DEVICE;
    lyg::ValidateLaunch L = launch_of(in, tolerance);
    L.areas = areas;
    if (!check(c, lyg_launch_signed_area(&L, c->stream), "signed area") ||
        !check(c, hipStreamSynchronize(c->stream), "sync"))
        return LYG_ERR_DEVICE;
    return LYG_OK;
}

int32_t lyg_path_winding_numbers(lyg_ctx* c, const lyg_path_view* path, const float* xy,
                                 uint32_t n, float tolerance, uint8_t fill_rule,
                                 int32_t* winding, uint8_t* inside) {
    if (!c || !path_ok(path) || (n && (!xy || !winding)) || fill_rule > 1)
        return LYG_ERR_INVALID_ARGUMENT;
    if (!tolerance_ok(tolerance)) return LYG_ERR_TOLERANCE_IS_NAN;
    if (n == 0) return LYG_OK;
    if (!check(c, hipSetDevice(c->device), "hipSetDevice")) return LYG_ERR_DEVICE;
    hipStream_t s = c->stream;
    lyg_batch_in in{};
    // misc holds the query points, the (all-zero) path indices and the results
    const size_t q_bytes = 8ull * n, idx_bytes = 4ull * n, w_bytes = 4ull * n;
    const size_t in_at = q_bytes, w_at = in_at + idx_bytes, b_at = w_at + w_bytes;
    if (!upload_path(c, path, &in, s) || !ensure(c, c->misc, b_at + n)) return LYG_ERR_DEVICE;
    char* m = static_cast<char*>(c->misc.p);
    if (!check(c, hipMemcpyAsync(m, xy, q_bytes, hipMemcpyHostToDevice, s), "H2D") ||
        !check(c, lyg_launch_fill_u32(reinterpret_cast<uint32_t*>(m + in_at), n, 0, s), "zero"))
        return LYG_ERR_DEVICE;
    lyg::ValidateLaunch L = launch_of(&in, tolerance);
    L.queries = reinterpret_cast<const float*>(m);
    L.query_path = reinterpret_cast<const uint32_t*>(m + in_at);
    L.n_queries = n;
    L.fill_rule = fill_rule;
    L.winding = reinterpret_cast<int32_t*>(m + w_at);
    L.inside = reinterpret_cast<uint8_t*>(m + b_at);
    if (!check(c, lyg_launch_winding(&L, s), "winding") ||
        !check(c, hipMemcpyAsync(winding, m + w_at, w_bytes, hipMemcpyDeviceToHost, s), "D2H") ||
        (inside && !check(c, hipMemcpyAsync(inside, m + b_at, n, hipMemcpyDeviceToHost, s), "D2H")) ||
        !check(c, hipStreamSynchronize(s), "sync"))
        return LYG_ERR_DEVICE;
    return LYG_OK;
}

int32_t lyg_approximate_signed_area(lyg_ctx* c, const lyg_path_view* path, float tolerance,
                                    float* area) {
    if (!c || !path_ok(path) || !area) return LYG_ERR_INVALID_ARGUMENT;
    if (!tolerance_ok(tolerance)) return LYG_ERR_TOLERANCE_IS_NAN;
    if (!check(c, hipSetDevice(c->device), "hipSetDevice")) return LYG_ERR_DEVICE;
    hipStream_t s = c->stream;
    lyg_batch_in in{};
    if (!upload_path(c, path, &in, s) || !ensure(c, c->misc, 4)) return LYG_ERR_DEVICE;
    lyg::ValidateLaunch L = launch_of(&in, tolerance);
    L.areas = as<float>(c->misc);
    float* h = reinterpret_cast<float*>(c->h_pinned);
    if (!check(c, lyg_launch_signed_area(&L, s), "signed area") ||
        !check(c, hipMemcpyAsync(h, c->misc.p, 4, hipMemcpyDeviceToHost, s), "D2H") ||
        !check(c, hipStreamSynchronize(s), "sync"))
        return LYG_ERR_DEVICE;
    *area = h[0];
    return LYG_OK;
}

}  // extern "C"
