// Structured starting configurations (host side).
//
// The reference seeds the chains of one `packing` run from a portfolio of
// layouts (cuda/packing.cu:304-452) and then jitters each chain by its place
// in its 8-temperature ladder (cuda/packing.cu:454-475):
//   - Goebel-strip topologies: triangular stacks of axis-aligned squares in
//     the corners and a band of 45-degree squares along the main diagonal,
//     enumerated over the stack sizes and the band width, kept when their
//     capacity is n..n+4, sorted by box side;
//   - wall-hugging grids with a column and a row of tilted squares;
//   - four corner squares and a tilted grid.
// This file builds the same families for a request and assigns them to
// replicas by ladder and temperature position.  The families follow
// cuda/packing.cu:304-452 step for step, because they define the starting
// chains the `packing` contract produces: the Goebel-strip enumeration (stack
// sizes, band widths, box-side formula, capacity window, square order, the 64
// smallest), the wall grid and the corner-tilted grid (square positions and
// order), and the assignment of families to the eight temperature classes.
// What differs: the assignment is written for ladders of any length (class =
// position x 8 / ladder), ties in the sort keep enumeration order, and the
// jitter draws from the request's Philox stream instead of the reference's
// generator.  Everything is FP64 and deterministic.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "../../include/sis_abi.h"

namespace sis {

namespace {

struct Pt {
    double x, y, a;
};

struct Layout {
    double L;      // box side the layout was built for
    int order;     // enumeration order (tie-breaker of the sort)
    std::vector<Pt> pts;
};

constexpr double kEps = 1e-7;

// Triangular stack of `m` rows in a corner: squares (i, j), i + j < m, at unit
// pitch (plus eps) from the corner (cx, cy) towards (sx, sy).
void corner_stack(int m, double cx, double cy, double sx, double sy, std::vector<Pt>& out) {
    for (int i = 0; i < m; ++i)
        for (int j = 0; i + j < m; ++j)
            out.push_back({cx + sx * (0.5 + i + kEps), cy + sy * (0.5 + j + kEps), 0.0});
}

// Band of 45-degree squares.  In the diagonal frame u = (x + y)/sqrt2,
// v = (y - x)/sqrt2 a tilted unit square is axis-aligned with half-side 1/2
// and the box is the diamond |v| <= u, |v| <= sqrt2 L - u; the band has `w`
// columns starting just past the lower-left stack of `a1` rows.  Stacks of
// b1 / b2 rows in the upper-left / lower-right corners cap the band's v range.
// Returns the count; fills `out` when given.
int tilted_band(double L, int w, int a1, int b1, int b2, std::vector<Pt>* out) {
    const double r2 = std::sqrt(2.0), h = 1.0 / r2;
    const double u0 = (a1 + 1.0) * h + 0.5 + 2 * kEps;
    int count = 0;
    for (int k = 0; k < w; ++k) {
        const double u = u0 + k * (1.0 + kEps);
        const double reach = std::min(u - 1.0 - kEps, r2 * L - u - 1.0 - kEps);
        double hi = reach, lo = -reach;
        if (b1 > 0) hi = std::min(hi, (L - b1 - 1.0) * h - 0.5 - 2 * kEps);
        if (b2 > 0) lo = std::max(lo, -(L - b2 - 1.0) * h + 0.5 + 2 * kEps);
        if (hi < lo) continue;
        const int m = (int)std::floor((hi - lo) / (1.0 + kEps)) + 1;
        count += m;
        if (!out) continue;
        const double v0 = 0.5 * (lo + hi) - 0.5 * (m - 1) * (1.0 + kEps);
        for (int t = 0; t < m; ++t) {
            const double v = v0 + t * (1.0 + kEps);
            out->push_back({(u - v) * h, (u + v) * h, 0.7853981633974483});
        }
    }
    return count;
}

// Goebel-strip topologies for n squares with box side in
// [sqrt(n) - 0.05, seed_L + 0.15], best (smallest side) first, at most 64.
std::vector<Layout> strip_topologies(int n, double seed_L) {
    const double h = 1.0 / std::sqrt(2.0);
    std::vector<Layout> out;
    int order = 0;
    auto keep = [&](double L, std::vector<Pt>&& pts) {
        pts.resize(n);
        out.push_back(Layout{L, order++, std::move(pts)});
    };
    for (int w = 1; w <= 24; ++w)
        for (int a1 = 0; a1 <= 32; ++a1)
            for (int a2 = a1; a2 <= 32; ++a2) {
                const double L = 0.5 * (a1 + a2) + 1.0 + w * h + (w + 4) * kEps;
                if (L < std::sqrt((double)n) - 0.05 || L > seed_L + 0.15) continue;
                const int bmax = std::max(0, (int)std::floor(L - std::max(a1, a2)));
                for (int b1 = 0; b1 <= bmax; ++b1)
                    for (int b2 = 0; b2 <= b1; ++b2) {
                        const int corners = a1 * (a1 + 1) / 2 + a2 * (a2 + 1) / 2 + b1 * (b1 + 1) / 2 +
                                            b2 * (b2 + 1) / 2;
                        if (corners > n + 3) continue;
                        const int cap = corners + tilted_band(L, w, a1, b1, b2, nullptr);
                        if (cap < n || cap > n + 4) continue;
                        std::vector<Pt> c, t;
                        corner_stack(a1, 0.0, 0.0, 1, 1, c);
                        corner_stack(a2, L, L, -1, -1, c);
                        corner_stack(b1, 0.0, L, 1, -1, c);
                        corner_stack(b2, L, 0.0, -1, 1, c);
                        tilted_band(L, w, a1, b1, b2, &t);
                        // axis-aligned squares kept first, then the band
                        std::vector<Pt> p1 = c;
                        p1.insert(p1.end(), t.begin(), t.end());
                        keep(L, std::move(p1));
                        if (cap > n && !c.empty()) {  // band kept whole, corners dropped
                            std::vector<Pt> p2 = t;
                            p2.insert(p2.end(), c.begin(), c.end());
                            keep(L, std::move(p2));
                        }
                    }
                // two stacks and the full band, one or two squares short: the
                // free corners take them at a slightly larger side
                std::vector<Pt> p;
                corner_stack(a1, 0.0, 0.0, 1, 1, p);
                corner_stack(a2, L, L, -1, -1, p);
                tilted_band(L, w, a1, 0, 0, &p);
                if (L > std::max(a1, a2) + 1.0) {
                    if ((int)p.size() == n - 1) {
                        p.push_back({0.5 + kEps, L - 0.5 - kEps, 0.0});
                        keep(L + 0.005, std::move(p));
                    } else if ((int)p.size() == n - 2) {
                        p.push_back({0.5 + kEps, L - 0.5 - kEps, 0.0});
                        p.push_back({L - 0.5 - kEps, 0.5 + kEps, 0.0});
                        keep(L + 0.01, std::move(p));
                    }
                }
            }
    std::stable_sort(out.begin(), out.end(), [](const Layout& a, const Layout& b) {
        return a.L < b.L || (a.L == b.L && a.order < b.order);
    });
    if (out.size() > 64) out.resize(64);
    return out;
}

// Wall-hugging grid of cx x cy squares (the far corner cut when asked), extra
// corner squares by ladder, the rest in a staggered column right of the grid
// and a staggered row above it, tilted by th.
std::vector<Pt> wall_grid(int n, int ladder, bool cut, double seed_L) {
    const int k = (int)std::floor(std::sqrt((double)n));
    const int cx = std::max(1, k - 1 - (ladder % 2)), cy = std::max(1, k - 1 - ((ladder / 2) % 2));
    const bool cut_corner = cut && cx >= 2 && cy >= 2;
    const double Lb = std::max(std::sqrt((double)n) + 0.05, seed_L - 0.03);
    std::vector<Pt> p;
    for (int i = 0; i < cx && (int)p.size() < n; ++i)
        for (int j = 0; j < cy && (int)p.size() < n; ++j)
            if (!(cut_corner && i == cx - 1 && j == cy - 1)) p.push_back({0.5 + i, 0.5 + j, 0.0});
    if (ladder % 3 != 0 && (int)p.size() < n) p.push_back({Lb - 0.5, Lb - 0.5, 0.0});
    if (ladder % 3 == 1 && (int)p.size() < n) {
        p.push_back({0.5, Lb - 0.5, 0.0});
        if ((int)p.size() < n) p.push_back({Lb - 0.5, 0.5, 0.0});
    }
    const double th = 0.25 + 0.08 * (ladder % 7);
    const int rem = n - (int)p.size(), right = (rem + 1) / 2, top = rem - right;
    auto clampb = [&](double v) { return std::min(std::max(v, 0.6), Lb - 0.6); };
    for (int m = 0; m < right; ++m) {
        const double y = 0.7 + (Lb - 1.4) * (m + 0.5) / std::max(1, right);
        const double x = 0.5 * (cx + Lb) + (m % 2 == 0 ? -0.18 : 0.18);
        p.push_back({clampb(x), clampb(y), th});
    }
    for (int m = 0; m < top; ++m) {
        const double x = 0.7 + (Lb - 1.4) * (m + 0.5) / std::max(1, top);
        const double y = 0.5 * (cy + Lb) + (m % 2 == 0 ? -0.18 : 0.18);
        p.push_back({clampb(x), clampb(y), th});
    }
    return p;
}

// Four corner squares and the rest on a tilted grid inside the box.
std::vector<Pt> corner_tilted_grid(int n, int ladder, double seed_L) {
    const double Lb = std::max(std::sqrt((double)n) + 0.05, seed_L - 0.02);
    std::vector<Pt> p;
    const Pt corners[4] = {{0.5, 0.5, 0.0}, {Lb - 0.5, 0.5, 0.0}, {0.5, Lb - 0.5, 0.0}, {Lb - 0.5, Lb - 0.5, 0.0}};
    for (int c = 0; c < 4 && c < n; ++c) p.push_back(corners[c]);
    const int rem = n - (int)p.size();
    const int cols = std::max(1, (int)std::ceil(std::sqrt((double)rem)));
    const double th = 0.35 + 0.07 * (ladder % 6);
    for (int m = 0; m < rem; ++m) {
        const double x = 0.8 + (Lb - 1.6) * ((m % cols) + 0.5) / cols;
        const double y = 0.8 + (Lb - 1.6) * ((m / cols) + 0.5) / cols;
        p.push_back({x, y, th});
    }
    return p;
}

}  // namespace

// Starting configuration of every replica (R x n x 3, x y angle) before
// jitter.  Replica r sits at temperature position r % ladder of ladder
// r / ladder; its family follows the position scaled to eight classes.
void portfolio_layouts(int n, int R, int ladder, double seed_L, const double* seed, std::vector<double>& out) {
    out.assign((size_t)R * n * 3, 0.0);
    auto put = [&](int r, const std::vector<Pt>& p) {
        for (int i = 0; i < n; ++i) {
            out[((size_t)r * n + i) * 3 + 0] = p[i].x;
            out[((size_t)r * n + i) * 3 + 1] = p[i].y;
            out[((size_t)r * n + i) * 3 + 2] = p[i].a;
        }
    };
    std::vector<Pt> s(n);
    for (int i = 0; i < n; ++i) s[i] = {seed[3 * i], seed[3 * i + 1], seed[3 * i + 2]};
    const std::vector<Layout> tops = n > 1 ? strip_topologies(n, seed_L) : std::vector<Layout>{};
    const int lad = std::max(1, ladder);
    for (int r = 0; r < R; ++r) {
        const int lid = r / lad, cls = (r % lad) * 8 / lad;
        if (n <= 1 || cls <= 1) {
            // the two coldest classes keep the seed on the first two ladders
            if (n > 1 && lid >= 2 && !tops.empty())
                put(r, tops[(lid - 2) % tops.size()].pts);
            else
                put(r, s);
        } else if (cls <= 4) {
            put(r, tops.empty() ? s : tops[lid % tops.size()].pts);
        } else if (cls <= 6) {
            put(r, wall_grid(n, lid, cls == 6, seed_L));
        } else {
            put(r, corner_tilted_grid(n, lid, seed_L));
        }
    }
}

// Jitter amplitude of the temperature class (cuda/packing.cu:459-461 keeps the
// coldest class exact, jitters the next by 1e-4 x 100^u, the rest by 0.03).
double portfolio_jitter(int cls, double u) {
    if (cls == 0) return 0.0;
    if (cls == 1) return 1e-4 * std::pow(100.0, u);
    return 0.03;
}

}  // namespace sis
