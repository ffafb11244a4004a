// ORACLE — test infrastructure only (see refgeom.h).
//
// Region checks for fill results:
//   * orc_coverage_compare: per-sample coverage of two triangle sets on an N x N grid of
//     sample centres over a box; samples closer than `eps` to any triangle edge of either
//     set are excluded; areas are summed in f64.
//   * orc_winding_coverage_compare: the same grid, but one side is the textbook winding
//     number (crossing count, even-odd or non-zero) of the polygon formed by the path's
//     flattened edges — an independent statement of what a fill must cover, used to pin
//     the fill oracle itself.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>
#include <algorithm>

#include "reffill.h"

namespace {

struct Grid {
    double x0, y0, dx, dy;
    uint32_t n;
    double sx(uint32_t i) const { return x0 + (i + 0.5) * dx; }
    double sy(uint32_t j) const { return y0 + (j + 0.5) * dy; }
};

double seg_dist(double px, double py, double ax, double ay, double bx, double by) {
    double vx = bx - ax, vy = by - ay;
    double wx = px - ax, wy = py - ay;
    double l2 = vx * vx + vy * vy;
    double t = l2 > 0 ? (wx * vx + wy * vy) / l2 : 0.0;
    t = t < 0 ? 0 : (t > 1 ? 1 : t);
    double qx = ax + t * vx - px, qy = ay + t * vy - py;
    return std::sqrt(qx * qx + qy * qy);
}

// Rasterises triangles [t0, t1) into counts (saturating u8) and marks samples near their
// edges in `near` (rows [r0, r1) only, so threads can split the grid by rows).
void raster(const float* v, const uint32_t* idx, uint64_t t0, uint64_t t1, const Grid& g,
            double eps, uint32_t r0, uint32_t r1, uint8_t* counts, uint8_t* near, double* area) {
    double a_sum = 0;
    for (uint64_t t = t0; t < t1; ++t) {
        double P[3][2];
        for (int k = 0; k < 3; ++k) {
            P[k][0] = v[2ull * idx[3 * t + k]];
            P[k][1] = v[2ull * idx[3 * t + k] + 1];
        }
        double cr = (P[1][0] - P[0][0]) * (P[2][1] - P[0][1]) - (P[1][1] - P[0][1]) * (P[2][0] - P[0][0]);
        if (r0 == 0) a_sum += std::fabs(cr) * 0.5;
        double sgn = cr >= 0 ? 1.0 : -1.0;
        // edge k: from P[k] to P[k+1]; inside: sgn * cross(P[k+1]-P[k], S-P[k]) > 0
        double A[3], B[3], Cc[3], Ln[3];
        for (int k = 0; k < 3; ++k) {
            const double* a = P[k];
            const double* b = P[(k + 1) % 3];
            double ex = b[0] - a[0], ey = b[1] - a[1];
            // cross(e, S - a) = ex*(sy-ay) - ey*(sx-ax) = -ey*sx + ex*sy + (ey*ax - ex*ay)
            A[k] = -ey * sgn;
            B[k] = ex * sgn;
            Cc[k] = (ey * a[0] - ex * a[1]) * sgn;
            Ln[k] = std::sqrt(ex * ex + ey * ey);
        }
        double mnx = std::min({P[0][0], P[1][0], P[2][0]}) - eps;
        double mxx = std::max({P[0][0], P[1][0], P[2][0]}) + eps;
        double mny = std::min({P[0][1], P[1][1], P[2][1]}) - eps;
        double mxy = std::max({P[0][1], P[1][1], P[2][1]}) + eps;
        long j0 = (long)std::floor((mny - g.y0) / g.dy - 0.5), j1 = (long)std::ceil((mxy - g.y0) / g.dy - 0.5);
        long i0 = (long)std::floor((mnx - g.x0) / g.dx - 0.5), i1 = (long)std::ceil((mxx - g.x0) / g.dx - 0.5);
        j0 = std::max(j0, (long)r0);
        j1 = std::min(j1, (long)r1 - 1);
        i0 = std::max(i0, 0L);
        i1 = std::min(i1, (long)g.n - 1);
        for (long j = j0; j <= j1; ++j) {
            double sy = g.sy((uint32_t)j);
            // x-interval where every edge function >= -eps * len
            double lo = (double)i0, hi = (double)i1;
            bool empty = false;
            for (int k = 0; k < 3; ++k) {
                double c = B[k] * sy + Cc[k] + eps * Ln[k];  // A x + c >= 0
                if (A[k] == 0) {
                    if (c < 0) empty = true;
                } else {
                    double xr = -c / A[k];                  // boundary in x
                    double ir = (xr - g.x0) / g.dx - 0.5;   // in sample units
                    if (A[k] > 0) lo = std::max(lo, std::floor(ir));
                    else hi = std::min(hi, std::ceil(ir));
                }
            }
            if (empty || lo > hi) continue;
            for (long i = (long)lo; i <= (long)hi; ++i) {
                double sx = g.sx((uint32_t)i);
                size_t s = (size_t)j * g.n + (size_t)i;
                bool nr = false, in = true;
                for (int k = 0; k < 3; ++k) {
                    double e = A[k] * sx + B[k] * sy + Cc[k];
                    if (e <= 0) in = false;
                    if (std::fabs(e) <= eps * Ln[k] + 1e-300) {
                        const double* a = P[k];
                        const double* b = P[(k + 1) % 3];
                        if (seg_dist(sx, sy, a[0], a[1], b[0], b[1]) < eps) nr = true;
                    }
                }
                if (nr) near[s] = 1;
                else if (in && counts[s] < 255) counts[s]++;
            }
        }
    }
    if (area) *area += a_sum;
}

Grid make_grid(double x0, double y0, double x1, double y1, uint32_t n) {
    Grid g;
    g.x0 = x0;
    g.y0 = y0;
    g.n = n;
    g.dx = (x1 - x0) / n;
    g.dy = (y1 - y0) / n;
    if (!(g.dx > 0)) g.dx = 1e-30;
    if (!(g.dy > 0)) g.dy = 1e-30;
    return g;
}

template <class F>
void parallel_rows(uint32_t n, F&& f) {
    unsigned nt = std::max(1u, std::min(8u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t) {
        uint32_t r0 = (uint32_t)((uint64_t)n * t / nt), r1 = (uint32_t)((uint64_t)n * (t + 1) / nt);
        th.emplace_back([=, &f] { f(t, r0, r1); });
    }
    for (auto& x : th) x.join();
}

}  // namespace

extern "C" {

// out[0] mismatching samples, out[1] excluded samples, out[2] covered by a, out[3]
// covered by b, out[4] samples covered more than once by a, out[5] same for b.
// area[0], area[1]: f64 sums of triangle areas.
int orc_coverage_compare(const float* va, const uint32_t* ia, uint64_t nta, const float* vb,
                         const uint32_t* ib, uint64_t ntb, double x0, double y0, double x1,
                         double y1, uint32_t n, double eps, uint64_t* out, double* area) {
    Grid g = make_grid(x0, y0, x1, y1, n);
    size_t ns = (size_t)n * n;
    std::vector<uint8_t> ca(ns, 0), cb(ns, 0), near(ns, 0);
    double ar[2] = {0, 0};
    parallel_rows(n, [&](unsigned, uint32_t r0, uint32_t r1) {
        raster(va, ia, 0, nta, g, eps, r0, r1, ca.data(), near.data(), r0 == 0 ? &ar[0] : nullptr);
        raster(vb, ib, 0, ntb, g, eps, r0, r1, cb.data(), near.data(), r0 == 0 ? &ar[1] : nullptr);
    });
    uint64_t mis = 0, exc = 0, cova = 0, covb = 0, ova = 0, ovb = 0;
    for (size_t s = 0; s < ns; ++s) {
        if (near[s]) {
            ++exc;
            continue;
        }
        bool a = ca[s] > 0, b = cb[s] > 0;
        cova += a;
        covb += b;
        ova += ca[s] > 1;
        ovb += cb[s] > 1;
        mis += a != b;
    }
    out[0] = mis;
    out[1] = exc;
    out[2] = cova;
    out[3] = covb;
    out[4] = ova;
    out[5] = ovb;
    area[0] = ar[0];
    area[1] = ar[1];
    return 0;
}

// Compares triangles (v, idx) with the winding-number coverage of the path's flattened
// edges (the curves flattened as the fill's event builder flattens them).  out[] as above
// with "a" = triangles and "b" = winding; out[6] = the number of flattened edges.
int orc_winding_coverage_compare(const uint8_t* verbs, uint32_t n_verbs, const float* points,
                                 uint32_t n_points, float tolerance, int fill_rule,
                                 const float* v, const uint32_t* idx, uint64_t nt, double x0,
                                 double y0, double x1, double y1, uint32_t n, double eps,
                                 uint64_t* out, double* area) {
    refpath::PathView pv{verbs, n_verbs, points, n_points};
    std::vector<float> edges;  // x0 y0 x1 y1 per edge, oriented as in the path
    reffill::flattened_edges(pv, tolerance, edges);
    const size_t ne = edges.size() / 4;
    Grid g = make_grid(x0, y0, x1, y1, n);
    size_t ns = (size_t)n * n;
    std::vector<uint8_t> ca(ns, 0), near(ns, 0), wb(ns, 0);
    double ar = 0;
    // the edges each row of samples can meet (within eps), as lists per row
    std::vector<uint32_t> rstart((size_t)n + 1, 0);
    auto row_range = [&](size_t e, long* j0, long* j1) {
        const double a = std::min(edges[4 * e + 1], edges[4 * e + 3]) - eps;
        const double b = std::max(edges[4 * e + 1], edges[4 * e + 3]) + eps;
        *j0 = std::max(0L, (long)std::floor((a - g.y0) / g.dy - 0.5));
        *j1 = std::min((long)n - 1, (long)std::ceil((b - g.y0) / g.dy - 0.5));
    };
    for (size_t e = 0; e < ne; ++e) {
        long j0, j1;
        row_range(e, &j0, &j1);
        for (long j = j0; j <= j1; ++j) ++rstart[j + 1];
    }
    for (uint32_t j = 0; j < n; ++j) rstart[j + 1] += rstart[j];
    std::vector<uint32_t> rlist(rstart[n]), rcur(rstart.begin(), rstart.end() - 1);
    for (size_t e = 0; e < ne; ++e) {
        long j0, j1;
        row_range(e, &j0, &j1);
        for (long j = j0; j <= j1; ++j) rlist[rcur[j]++] = (uint32_t)e;
    }
    parallel_rows(n, [&](unsigned, uint32_t r0, uint32_t r1) {
        raster(v, idx, 0, nt, g, eps, r0, r1, ca.data(), near.data(), r0 == 0 ? &ar : nullptr);
        // winding per row: crossings with half-open y rule, then prefix over x
        std::vector<std::pair<double, int>> xs;
        for (uint32_t j = r0; j < r1; ++j) {
            double sy = g.sy(j);
            xs.clear();
            for (uint32_t q = rstart[j]; q < rstart[j + 1]; ++q) {
                const size_t e = rlist[q];
                double ax = edges[4 * e], ay = edges[4 * e + 1], bx = edges[4 * e + 2], by = edges[4 * e + 3];
                if ((ay <= sy) == (by <= sy)) continue;
                double t = (sy - ay) / (by - ay);
                double x = ax + t * (bx - ax);
                xs.push_back({x, by > ay ? 1 : -1});
            }
            std::sort(xs.begin(), xs.end());
            int w = 0;
            size_t k = 0;
            for (uint32_t i = 0; i < n; ++i) {
                double sx = g.sx(i);
                while (k < xs.size() && xs[k].first < sx) w += xs[k++].second;
                bool in = fill_rule == 0 ? (w % 2 != 0) : (w != 0);
                wb[(size_t)j * n + i] = in;
            }
            // exclusion around the path's edges
            for (uint32_t q = rstart[j]; q < rstart[j + 1]; ++q) {
                const size_t e = rlist[q];
                double ax = edges[4 * e], ay = edges[4 * e + 1], bx = edges[4 * e + 2], by = edges[4 * e + 3];
                if (sy < std::min(ay, by) - eps || sy > std::max(ay, by) + eps) continue;
                double mnx = std::min(ax, bx) - eps, mxx = std::max(ax, bx) + eps;
                long i0 = std::max(0L, (long)std::floor((mnx - g.x0) / g.dx - 0.5));
                long i1 = std::min((long)n - 1, (long)std::ceil((mxx - g.x0) / g.dx - 0.5));
                // only the samples within eps of the segment's line at this row
                double dxe = bx - ax, dye = by - ay;
                double len = std::sqrt(dxe * dxe + dye * dye);
                if (std::fabs(dye) > 1e-300 && len > 0) {
                    double xc = ax + (sy - ay) / dye * dxe;
                    double half = eps * len / std::fabs(dye);
                    i0 = std::max(i0, (long)std::floor((xc - half - g.x0) / g.dx - 0.5));
                    i1 = std::min(i1, (long)std::ceil((xc + half - g.x0) / g.dx - 0.5));
                }
                for (long i = i0; i <= i1; ++i)
                    if (seg_dist(g.sx((uint32_t)i), sy, ax, ay, bx, by) < eps) near[(size_t)j * n + i] = 1;
            }
        }
    });
    uint64_t mis = 0, exc = 0, cova = 0, covb = 0, ova = 0;
    for (size_t s = 0; s < ns; ++s) {
        if (near[s]) {
            ++exc;
            continue;
        }
        bool a = ca[s] > 0, b = wb[s] != 0;
        cova += a;
        covb += b;
        ova += ca[s] > 1;
        mis += a != b;
    }
    out[0] = mis;
    out[1] = exc;
    out[2] = cova;
    out[3] = covb;
    out[4] = ova;
    out[5] = 0;
    out[6] = ne;
    area[0] = ar;
    area[1] = 0;
    return 0;
}

}  // extern "C"
