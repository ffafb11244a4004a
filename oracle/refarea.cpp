// ORACLE — test infrastructure only (see refgeom.h).
//
// The exact area of a path's fill, independent of any triangulation: the area of
// {p : fill_rule(winding(p))} for the path's flattened edges (flattened as lyon's event
// queue builder flattens them, reffill::flattened_edges, event_queue.rs:661-900).
//
// By the divergence theorem the area of a region is the sum over its boundary of
// integral(x dy), with the sign of each boundary piece set by which side the region lies
// on.  The region's boundary is made of pieces of the path's edges: every edge is cut at
// its crossings with other edges (found on a uniform grid with exact orientation signs in
// f64), and a piece is boundary when the fill rule gives different answers on its two
// sides.  The winding left of an edge is counted (signed crossings of a horizontal ray) at
// its first piece's midpoint and carried along it: crossing another edge changes both
// sides by that edge's winding; where another edge ends on it (a T-junction) it is counted
// again.  Each piece is straight, so its integral is its mid x times its height, exact up
// to f64 rounding.  Collinear overlapping edges (a spike, a two-point sub-path) meet the
// ray at the piece's own point: the lowest-numbered of them accounts for the boundary of the
// whole bundle, with the bundle's windings summed on its right.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>

#include "reffill.h"

namespace {

struct E {
    double x0, y0, x1, y1;  // top to bottom (y0 < y1)
    int s;                  // +1 path runs downwards, -1 upwards
};

double orient(double px, double py, double qx, double qy, double rx, double ry) {
    return (qx - px) * (ry - py) - (qy - py) * (rx - px);
}

// proper crossing (strictly inside both): t along a
bool cross_t(const E& a, const E& b, double* ta, double* tb) {
    const double d1 = orient(a.x0, a.y0, a.x1, a.y1, b.x0, b.y0);
    const double d2 = orient(a.x0, a.y0, a.x1, a.y1, b.x1, b.y1);
    if (!((d1 > 0 && d2 < 0) || (d1 < 0 && d2 > 0))) return false;
    const double d3 = orient(b.x0, b.y0, b.x1, b.y1, a.x0, a.y0);
    const double d4 = orient(b.x0, b.y0, b.x1, b.y1, a.x1, a.y1);
    if (!((d3 > 0 && d4 < 0) || (d3 < 0 && d4 > 0))) return false;
    *ta = d3 / (d3 - d4);
    *tb = d1 / (d1 - d2);
    return true;
}

bool in_rule(int rule, int w) { return rule == 0 ? (w % 2 != 0) : (w != 0); }


}  // namespace

extern "C" {

// out[0] = area of the EvenOdd fill, out[1] = of the NonZero fill, out[2] = number of edge
// crossings.
int orc_fill_area(const uint8_t* verbs, uint32_t n_verbs, const float* points, uint32_t n_points,
                  float tolerance, double* out) {
    refpath::PathView pv{verbs, n_verbs, points, n_points};
    std::vector<float> fl;
    reffill::flattened_edges(pv, tolerance, fl);
    std::vector<E> es;
    double xmin = INFINITY, xmax = -INFINITY, ymin = INFINITY, ymax = -INFINITY;
    for (size_t i = 0; i < fl.size(); i += 4) {
        double ax = fl[i], ay = fl[i + 1], bx = fl[i + 2], by = fl[i + 3];
        if (ax == bx && ay == by) continue;
        // stored top to bottom (horizontal edges: left to right); s: +1 when the path runs
        // the stored way.  Horizontal edges bound no area (dy = 0) and no horizontal ray
        // crosses them, but the winding carried along another edge changes where it
        // crosses one.
        E e;
        if (ay < by || (ay == by && ax < bx)) e = E{ax, ay, bx, by, +1};
        else e = E{bx, by, ax, ay, -1};
        es.push_back(e);
        xmin = std::min({xmin, ax, bx});
        xmax = std::max({xmax, ax, bx});
        ymin = std::min(ymin, e.y0);
        ymax = std::max(ymax, e.y1);
    }
    const size_t n = es.size();
    out[0] = out[1] = out[2] = 0;
    if (n == 0) return 0;
    // uniform grid of about n cells
    const int G = std::max(1, std::min(4096, (int)std::sqrt((double)n)));
    const double cw = std::max((xmax - xmin) / G, 1e-300), ch = std::max((ymax - ymin) / G, 1e-300);
    auto cx = [&](double x) { return std::min(G - 1, std::max(0, (int)((x - xmin) / cw))); };
    auto cy = [&](double y) { return std::min(G - 1, std::max(0, (int)((y - ymin) / ch))); };
    std::vector<uint32_t> cnt((size_t)G * G + 1, 0);
    for (const E& e : es) {
        const int x0 = cx(std::min(e.x0, e.x1)), x1 = cx(std::max(e.x0, e.x1));
        for (int r = cy(e.y0); r <= cy(e.y1); ++r)
            for (int c = x0; c <= x1; ++c) ++cnt[(size_t)r * G + c + 1];
    }
    for (size_t i = 1; i < cnt.size(); ++i) cnt[i] += cnt[i - 1];
    std::vector<uint32_t> cell(cnt.back()), cur(cnt.begin(), cnt.end() - 1);
    for (size_t i = 0; i < n; ++i) {
        const E& e = es[i];
        const int x0 = cx(std::min(e.x0, e.x1)), x1 = cx(std::max(e.x0, e.x1));
        for (int r = cy(e.y0); r <= cy(e.y1); ++r)
            for (int c = x0; c <= x1; ++c) cell[cur[(size_t)r * G + c]++] = (uint32_t)i;
    }
    // crossings per edge: (t, other); a pair is kept by the cell that holds its crossing
    // point (computed the same way for both orders)
    std::vector<std::vector<std::pair<double, uint32_t>>> xs(n);
    const size_t ncell = (size_t)G * G;
    {
        unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        std::vector<std::vector<std::pair<uint32_t, std::pair<double, uint32_t>>>> found(nt);
        std::vector<std::thread> th;
        for (unsigned t = 0; t < nt; ++t) {
            th.emplace_back([&, t] {
                for (size_t c = t; c < ncell; c += nt) {
                    const uint32_t b = cnt[c], e = cnt[c + 1];
                    for (uint32_t p = b; p < e; ++p)
                        for (uint32_t q = p + 1; q < e; ++q) {
                            uint32_t i = cell[p], j = cell[q];
                            if (i > j) std::swap(i, j);
                            double ti, tj;
                            if (!cross_t(es[i], es[j], &ti, &tj)) continue;
                            const double px = es[i].x0 + ti * (es[i].x1 - es[i].x0);
                            const double py = es[i].y0 + ti * (es[i].y1 - es[i].y0);
                            if ((size_t)cy(py) * G + cx(px) != c) continue;
                            found[t].push_back({i, {ti, j}});
                            found[t].push_back({j, {tj, i}});
                        }
                }
            });
        }
        for (auto& x : th) x.join();
        for (auto& f : found)
            for (auto& r : f) xs[r.first].push_back(r.second);
    }
    size_t ncross = 0;
    for (auto& v : xs) ncross += v.size();
    // rows for the winding queries
    const int R = G;
    std::vector<uint32_t> rc((size_t)R + 1, 0);
    for (const E& e : es)
        for (int r = cy(e.y0); r <= cy(e.y1); ++r) ++rc[r + 1];
    for (int i = 1; i <= R; ++i) rc[i] += rc[i - 1];
    std::vector<uint32_t> rows(rc[R]), rcur(rc.begin(), rc.end() - 1);
    for (size_t i = 0; i < n; ++i)
        for (int r = cy(es[i].y0); r <= cy(es[i].y1); ++r) rows[rcur[r]++] = (uint32_t)i;
    auto xat = [&](const E& e, double y) {
        return e.x0 + (y - e.y0) * ((e.x1 - e.x0) / (e.y1 - e.y0));
    };
    // T-junctions: an end point of another edge lying on an edge's interior also splits it
    // (marked with the other edge's id + n, so the walk below counts the winding again there)
    std::vector<uint8_t> has_t(n, 0);
    {
        unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
        std::vector<std::thread> th;
        for (unsigned t = 0; t < nt; ++t) {
            th.emplace_back([&, t] {
                for (size_t i = t; i < n; i += nt) {
                    const E& e = es[i];
                    const int x0 = cx(std::min(e.x0, e.x1)), x1 = cx(std::max(e.x0, e.x1));
                    const double len2 = (e.x1 - e.x0) * (e.x1 - e.x0) + (e.y1 - e.y0) * (e.y1 - e.y0);
                    for (int r = cy(e.y0); r <= cy(e.y1); ++r)
                        for (int c = x0; c <= x1; ++c)
                            for (uint32_t q = cnt[(size_t)r * G + c]; q < cnt[(size_t)r * G + c + 1]; ++q) {
                                const uint32_t j = cell[q];
                                if (j == i) continue;
                                for (int k = 0; k < 2; ++k) {
                                    const double px = k ? es[j].x1 : es[j].x0, py = k ? es[j].y1 : es[j].y0;
                                    if (orient(e.x0, e.y0, e.x1, e.y1, px, py) != 0.0) continue;
                                    const double tt = ((px - e.x0) * (e.x1 - e.x0) + (py - e.y0) * (e.y1 - e.y0)) / len2;
                                    if (tt > 0.0 && tt < 1.0) {
                                        xs[i].push_back({tt, (uint32_t)(j + n)});
                                        has_t[i] = 1;
                                    }
                                }
                            }
                    std::sort(xs[i].begin(), xs[i].end());
                }
            });
        }
        for (auto& x : th) x.join();
    }
    // signed crossings of the horizontal ray from (mx, my) towards -x (half-open in y):
    // *wl the edges strictly left, *wb the other edges through the point itself (collinear
    // overlaps, which meet the ray exactly there); *lead: no such edge comes before edge i,
    // so edge i accounts for the whole bundle's boundary
    auto winding_left = [&](size_t i, double mx, double my, int* wl, int* wb, bool* lead) {
        int w = 0, bsum = 0;
        bool ld = true;
        const double tol = 1e-12 * std::max(1.0, std::fabs(mx));
        const int r = cy(my);
        for (uint32_t k = rc[r]; k < rc[r + 1]; ++k) {
            const uint32_t j = rows[k];
            if (j == i) continue;
            const E& f = es[j];
            if (!(f.y0 <= my && my < f.y1)) continue;
            const double x = xat(f, my);
            if (x < mx - tol) {
                w += f.s;
            } else if (x <= mx + tol) {
                bsum += f.s;
                if (j < i) ld = false;
            }
        }
        *wl = w;
        *wb = bsum;
        *lead = ld;
    };
    // each piece: the winding on its left is counted at the first piece's midpoint (and
    // again after a T-junction), and carried across proper crossings, where both sides
    // cross the other edge f and change by the sign of cross(edge direction, f's path
    // direction); on its right it is that plus the edge's own
    std::vector<double> acc(32, 0.0);
    std::vector<std::thread> th;
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    for (unsigned t = 0; t < nt; ++t) {
        th.emplace_back([&, t] {
            double a_eo = 0, a_nz = 0;
            for (size_t i = t; i < n; i += nt) {
                const E& e = es[i];
                if (e.y0 == e.y1) continue;
                const auto& v = xs[i];
                const double dex = e.x1 - e.x0, dey = e.y1 - e.y0;
                double tp = 0.0;
                bool need = true, lead = true;
                int wl = 0, wb = 0;
                for (size_t q = 0; q <= v.size(); ++q) {
                    const double tn = q < v.size() ? v[q].first : 1.0;
                    if (tn > tp) {
                        const double tm = 0.5 * (tp + tn);
                        const double my = e.y0 + tm * dey, mx = e.x0 + tm * dex;
                        if (need) {
                            winding_left(i, mx, my, &wl, &wb, &lead);
                            need = false;
                        }
                        const int wr = wl + e.s + wb;
                        const double xdy = mx * ((tn - tp) * dey);
                        // region on the left of the piece: + x dy; on the right: - x dy
                        for (int rule = 0; rule < 2 && lead; ++rule) {
                            const bool il = in_rule(rule, wl), ir = in_rule(rule, wr);
                            if (il != ir) (rule ? a_nz : a_eo) += (il ? 1.0 : -1.0) * xdy;
                        }
                    }
                    if (q < v.size()) {
                        const uint32_t o = v[q].second;
                        if (o >= n) {
                            need = true;  // a T-junction: count again past it
                        } else if (!need) {
                            const E& f = es[o];
                            const double c = dex * (f.y1 - f.y0) - dey * (f.x1 - f.x0);
                            wl += (c > 0 ? 1 : -1) * f.s;
                        }
                    }
                    tp = tn;
                }
            }
            acc[2 * t] = a_eo;
            acc[2 * t + 1] = a_nz;
        });
    }
    for (auto& x : th) x.join();
    double area_eo = 0, area_nz = 0;
    for (unsigned t = 0; t < nt; ++t) {
        area_eo += acc[2 * t];
        area_nz += acc[2 * t + 1];
    }
    out[0] = area_eo;
    out[1] = area_nz;
    out[2] = (double)(ncross / 2);
    return 0;
}

}  // extern "C"
