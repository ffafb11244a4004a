// Launch parameters of the banded fill (band.hip), shared with its host side.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/lyg.h"

namespace lyg {
namespace band {

struct Params {
    // input batch (lyon's flat layout, path.rs:75-79)
    const uint8_t* verbs;
    const float* points;
    const uint32_t* verb_off;
    const uint32_t* point_off;
    uint32_t n_paths;
    uint32_t n_verbs;
    const lyg_fill_options* opts;
    uint32_t opts_stride;
    // per verb
    uint32_t* v_path;
    uint32_t* v_npts;    // points consumed; scanned into v_pbase (n_verbs + 1)
    uint32_t* v_pbase;
    uint32_t* v_isbeg;   // 1 at Begin; scanned into v_begord
    uint32_t* v_begord;
    uint32_t* begin_verb;
    uint32_t* v_nedge;   // edges produced; scanned into v_ebase
    uint32_t* v_ebase;
    // edges
    float4* geo;         // (x0, y0, x1, y1): top then bottom in sweep order
    int8_t* wind;        // +1 / -1; 0 for horizontal or empty edges
    uint32_t* e_path;
    uint32_t n_edges;
    // per path
    uint32_t* p_yr;      // orderable min / max of the non-horizontal edges' y (2 per path)
    uint32_t* p_status;  // BS_* bits of the path
    uint32_t* p_nband;   // bands of the path; scanned into p_band_off (n_paths + 1)
    uint32_t* p_band_off;
    uint32_t* p_hist_off;  // histogram bins of the path (64 per band when banded)
    uint32_t* hist;
    uint32_t edges_per_band;
    uint32_t max_bands;
    const uint32_t* p_mult;  // per path: band count multiplier (null: 1)
    // per band
    float* b_ytop;
    float* b_ybot;
    uint32_t* b_path;
    uint32_t* b_nedge;   // scanned into b_edge_off
    uint32_t* b_edge_off;
    uint32_t* b_fill;    // scatter cursors
    uint32_t* band_edges;  // global edge ids, band after band
    float4* bgeo;        // the band's edges' geometry, band-local order
    uint32_t* b_ncross;  // scanned into b_cross_off
    uint32_t* b_cross_off;
    uint32_t* b_status;
    uint32_t n_bands;
    // crossings and events
    double* cross_y;     // crossing height
    uint32_t* cross_ab;  // band-local edge ids (lo 16: a, hi 16: b)
    uint32_t* b_nev;     // events of the band; scanned into b_ev_off
    uint32_t* b_ev_off;
    double* ev_y;
    float* ev_x;
    uint32_t* ev_k;      // kind in the top 2 bits (0 cross, 1 end, 2 start), ids below
    uint32_t* b_ninit;   // edges entering at the band's top
    uint16_t* init_order;  // band-local ids, left to right (at b_edge_off)
    // trapezoids
    uint32_t* b_rec_cap;  // room per band; scanned into b_rec_off (n_bands + 1)
    uint32_t* b_rec_off;
    uint32_t* b_nrec;
    uint4* rec;           // (left edge, right edge, top event, bottom event)
    // output
    uint32_t* r_nv;       // per record: vertices (0, 3 or 4), then triangles
    uint32_t* r_nt;
    uint32_t* r_voff;     // scans (records + 1)
    uint32_t* r_toff;
    uint32_t n_records;   // sum of b_rec_cap
    float* out_v;
    uint32_t* out_i;
    const uint32_t* p_voff;  // per path first vertex (local ids), n_paths + 1
    uint32_t vertex_offset;
    uint32_t local_ids;
};

}  // namespace band
}  // namespace lyg
