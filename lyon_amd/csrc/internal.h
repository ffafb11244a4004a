// Internal to liblyg.so: the batched entry points behind the C ABI, with the output
// allocation left to the caller (the C ABI checks the caller's capacity; the serving
// loop allocates from its device pool once the totals are known).
#pragma once
#include <functional>
#include <vector>

#include "ctx.h"
#include "lyg_types.h"

namespace lygi {

// Called once the batch's totals are known; sets the output arrays or returns a status.
using OutAlloc = std::function<int32_t(uint64_t n_vertices, uint64_t n_indices, float** vertices,
                                       uint32_t** indices)>;

struct BatchSpec {
    lyg::PathBatchDev in;
    const void* opts;        // device: lyg_stroke_options or lyg_fill_options records
    uint32_t opts_stride;    // 0: one record for every path
    uint32_t vertex_offset;
    bool local_ids;          // ids count from vertex_offset in every path
    uint32_t* vtx_off;       // device, n + 1
    uint32_t* idx_off;       // device, n + 1
};

// Batched stroke; *tv / *ti get the totals (also on a capacity error).
int32_t stroke_batch(lyg_ctx* c, const BatchSpec& b, const OutAlloc& alloc, uint64_t* tv,
                     uint64_t* ti);
// Batched fill; `status` gets every path's status (a failed path produces nothing).
// Banded fill of a batch (band_api.cpp): lyon's fill coverage, its own triangles.
int32_t fill_batch_banded(lyg_ctx* c, const BatchSpec& b, const OutAlloc& alloc, uint64_t* tv,
                          uint64_t* ti, std::vector<uint32_t>* status);
int32_t fill_batch(lyg_ctx* c, const BatchSpec& b, const OutAlloc& alloc, uint64_t* tv,
                   uint64_t* ti, std::vector<uint32_t>* status);

// A batched fill captured as HIP graphs (lyg_fill_batch_graph): the first sweep round,
// planned from the batch at capture time, the offset scans and the copy.
struct FillGraph;
int32_t fill_graph_capture(lyg_ctx* c, const lyg_batch_in* in, const lyg_fill_options* opts,
                           uint32_t n_opts, const lyg_batch_out* out, FillGraph** graph);
int32_t fill_graph_launch(FillGraph* g);
int32_t fill_graph_result(FillGraph* g, uint64_t* n_vertices, uint64_t* n_indices);
void fill_graph_destroy(FillGraph* g);

// A context; side_streams: make the fill's side streams now (lyg_ctx_create does, so that
// they get hardware queues of their own) rather than on the first fill (the serving loop's
// workers, whose own streams should each keep a queue).
lyg_ctx* ctx_create(int device, bool side_streams);

}  // namespace lygi
