/*
 * CPU oracle for the PT-SA square-packing path (test infrastructure only; see
 * sis_oracle.h).  Built with g++ -O2 -ffp-contract=off so that every rounding
 * is the one written here.
 *
 * Parity status per group (what pins each one; the tests live under tests/):
 *  - Philox4x32-10: Random123 known-answer vectors + rocRAND's open header.
 *  - reference FP32 geometry (o_ref_*): the reference's own selftest corpus
 *    (mt19937(923)) against the FP64 axis oracle, tolerance 2e-6, and the
 *    cell-list vs all-pairs identity (1e-4), as cuda/tests/selftest.cu:53.
 *  - product FP32 formulation (o_*_cs): closed form R = 1/2(1+|cos|+|sin|)
 *    equals the reference's per-axis radius; axis-aligned reduction to
 *    1-D interval overlap; brute-force tiny cases.
 *  - FP64 pair/boundary/gradient: seeded_test corpus (mt19937(832)) vs the
 *    independent axis oracle (1e-12) and central finite differences (1e-7),
 *    as cuda/tests/seeded_test.cu:30.
 *  - certified side: every emitted configuration passes the reference's
 *    interval certifier (oracle/_ref); lambda >= 1 and L >= sqrt(n) - 1e-7.
 *  - CG polish: energy never increases (accepts only strictly lower energy).
 *  - PT-SA engine (CPU baseline): forced-exchange invariant and shrink
 *    bookkeeping; trajectories are parity unpinned against the reference
 *    (cuRAND streams cannot be reproduced).
 *  - sweep model (o_sweep32_model): a CPU restatement of the product's own
 *    colour-checkerboard sweep; pinned by FP64 recomputation of every
 *    proposal's energy change (1e-5 relative above an FP32 floor of 2e-6 x
 *    the move's overlap depths), by the stencil energy against all pairs,
 *    and by detailed balance: one square in a walled box samples
 *    exp(-E/T) (mean energy and position histogram against quadrature,
 *    cell lists kept for 1 and 4 sweeps).
 *  - exchange step (o_ladder_model): the fixed FP64 2^x against pow within
 *    2 ulp; the exchange criterion by hand cases that a sign error fails and
 *    by the stationary distribution of the slot permutation on a four-level
 *    toy (detailed balance of parallel tempering); pairing inside ladders;
 *    adaptation by its closed form; the improvement path and zero-overlap
 *    shrink by their bookkeeping.
 *  - certified side of an FP32 state (o_certify_xycs): the reference's
 *    all-pairs rule on orientations scaled to unit length.
 */
#include "sis_oracle.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

/* ------------------------------------------------------------------------ */
/* Philox4x32-10, from Salmon, Moraes, Dror, Shaw, "Parallel random numbers: */
/* as easy as 1, 2, 3", SC'11: ten rounds of two 32x32->64 multiplies with   */
/* a Weyl-sequence key schedule.                                             */
/* ------------------------------------------------------------------------ */
void o_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t x0 = ctr[0], x1 = ctr[1], x2 = ctr[2], x3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int round = 0; round < 10; ++round) {
        uint64_t p0 = (uint64_t)0xD2511F53u * x0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * x2;
        uint32_t y0 = (uint32_t)(p1 >> 32) ^ x1 ^ k0;
        uint32_t y1 = (uint32_t)p1;
        uint32_t y2 = (uint32_t)(p0 >> 32) ^ x3 ^ k1;
        uint32_t y3 = (uint32_t)p0;
        x0 = y0; x1 = y1; x2 = y2; x3 = y3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = x0; out[1] = x1; out[2] = x2; out[3] = x3;
}

/* (k + 1/2) * 2^-23 for the top 23 bits k: exact, symmetric, in (0, 1). */
float o_u01f(uint32_t r) { return ((float)(r >> 9) + 0.5f) * (1.0f / 8388608.0f); }

/* 2^x of the Metropolis test (x <= 0), the acceptance of cuda/packing.cu:80,123
 * written as FP32 operations that each round once or not at all: k = rint(x),
 * f = x - k (exact) in [-1/2, 1/2], p(f) = the degree-6 polynomial through
 * 2^f at the 7 Chebyshev nodes of that interval (Horner, fused multiply-adds),
 * then the exact scaling by 2^k.  Any IEEE FP32 machine with fma gives the
 * same bits, the device included; relative error below 1e-7 (pinned against
 * 2^x in tests/test_oracle_geometry.py).  0 below x = -100, where no accept
 * uniform (>= 2^-24) can pass. */
float o_accept_exp2(float x) {
    if (!(x >= -100.0f)) return 0.0f;
    const float k = rintf(x);
    const float f = x - k;
    float p = 0x1.444p-13f;
    p = fmaf(p, f, 0x1.5f48cp-10f);
    p = fmaf(p, f, 0x1.3b2a1cp-7f);
    p = fmaf(p, f, 0x1.c6aeccp-5f);
    p = fmaf(p, f, 0x1.ebfbep-3f);
    p = fmaf(p, f, 0x1.62e43p-1f);
    p = fmaf(p, f, 1.0f);
    return ldexpf(p, (int)k);
}

double o_u01d(uint32_t hi, uint32_t lo) {
    uint64_t k = ((uint64_t)hi << 21) ^ (uint64_t)(lo >> 11);
    k &= ((1ull << 52) - 1);
    return ((double)k + 0.5) * (1.0 / 4503599627370496.0);
}

/* ------------------------------------------------------------------------ */
/* Reference FP32 geometry.                                                 */
/* ------------------------------------------------------------------------ */

/* cuda/packing.cu:22-25: 16x16 grid of width 2, clamped to the edge cells. */
int o_ref_cell(float x, float y) {
    int cx = (int)floorf(x / 2);
    int cy = (int)floorf(y / 2);
    cx = cx < 0 ? 0 : (cx > 15 ? 15 : cx);
    cy = cy < 0 ? 0 : (cy > 15 ? 15 : cy);
    return cy * 16 + cx;
}

/* cuda/packing.cu:26-34: penetration depth along the four edge normals, each
 * with its own projected radius built from four |projections|. */
float o_ref_depth32(const float p[3], const float q[3]) {
    const float cp = cosf(p[2]), sp = sinf(p[2]);
    const float cq = cosf(q[2]), sq = sinf(q[2]);
    const float nx[4] = {cp, -sp, cq, -sq};
    const float ny[4] = {sp, cp, sq, cq};
    float best = 1e10f;
    for (int k = 0; k < 4; ++k) {
        float half = .5f * (fabsf(nx[k] * cp + ny[k] * sp) + fabsf(-nx[k] * sp + ny[k] * cp) +
                            fabsf(nx[k] * cq + ny[k] * sq) + fabsf(-nx[k] * sq + ny[k] * cq));
        float sep = fabsf((p[0] - q[0]) * nx[k] + (p[1] - q[1]) * ny[k]);
        best = fminf(best, half - sep);
    }
    return fmaxf(0.f, best);
}

/* cuda/packing.cu:35-39: squared excursion of the rotated square's extent
 * beyond the box [0, L]^2, per axis, measured from the box centre. */
float o_ref_wall32(const float p[3], float L) {
    float ext = .5f * (fabsf(cosf(p[2])) + fabsf(sinf(p[2])));
    float ox = fmaxf(0.f, fabsf(p[0] - L / 2) + ext - L / 2);
    float oy = fmaxf(0.f, fabsf(p[1] - L / 2) + ext - L / 2);
    return ox * ox + oy * oy;
}

/* cuda/packing.cu:40-47 through the list build at :62-64, and the all-pairs
 * sum of cuda/tests/selftest.cu:25-29.  Lists are built in ascending index
 * order with head insertion, which is one of the orders the reference's
 * atomicExch insertion can produce. */
void o_ref_local32(int n, const float* q, float L, float* local_out, float* full_out) {
    std::vector<int> head(256, -1), next(n, -1);
    for (int i = 0; i < n; ++i) {
        int c = o_ref_cell(q[3 * i], q[3 * i + 1]);
        next[i] = head[c];
        head[c] = i;
    }
    for (int i = 0; i < n; ++i) {
        const float* p = q + 3 * i;
        float e = o_ref_wall32(p, L);
        int c = o_ref_cell(p[0], p[1]);
        int cx = c % 16, cy = c / 16;
        for (int oy = -1; oy <= 1; ++oy)
            for (int ox = -1; ox <= 1; ++ox) {
                int xx = cx + ox, yy = cy + oy;
                if (xx < 0 || xx >= 16 || yy < 0 || yy >= 16) continue;
                for (int j = head[yy * 16 + xx]; j != -1; j = next[j])
                    if (j != i) {
                        float d = o_ref_depth32(p, q + 3 * j);
                        e += d * d;
                    }
            }
        local_out[i] = e;
        float full = o_ref_wall32(p, L);
        for (int j = 0; j < n; ++j)
            if (j != i) {
                float d = o_ref_depth32(p, q + 3 * j);
                full += d * d;
            }
        full_out[i] = full;
    }
}

/* ------------------------------------------------------------------------ */
/* The product's FP32 formulation.  The device kernels evaluate exactly these */
/* operations in this order, with explicit fma, so results agree bit for bit. */
/* ------------------------------------------------------------------------ */

/* clamp(floor((x + shift) * (1/w))) with the correctly rounded reciprocal;
 * w = 2 gives exactly 0.5 and the reference's floor(x / 2). */
int o_cellcoord(float x, float shift, float w, int lo, int hi) {
    const float inv_w = 1.0f / w;
    int k = (int)floorf((x + shift) * inv_w);
    return k < lo ? lo : (k > hi ? hi : k);
}

/* Linear SAT depth: the four edge normals of both squares share the projected
 * radius R = 1/2 (1 + |cos(ti - tj)| + |sin(ti - tj)|); the depth is
 * max(0, R - max_k |d . n_k|).  Same value as cuda/packing.cu:26-34 in exact
 * arithmetic; dx, dy = centre of j minus centre of i. */
static inline float sat_gap32(float dx, float dy, float ci, float si, float cj, float sj) {
    float cd = fmaf(ci, cj, si * sj);
    float sd = fmaf(si, cj, -(ci * sj));
    float R = fmaf(0.5f, fabsf(cd) + fabsf(sd), 0.5f);
    float a1 = fabsf(fmaf(dx, ci, dy * si));
    float a2 = fabsf(fmaf(dy, ci, -(dx * si)));
    float a3 = fabsf(fmaf(dx, cj, dy * sj));
    float a4 = fabsf(fmaf(dy, cj, -(dx * sj)));
    float m = fmaxf(fmaxf(a1, a2), fmaxf(a3, a4));
    return R - m;
}

float o_depth32_cs(float dx, float dy, float ci, float si, float cj, float sj) {
    float d2 = fmaf(dx, dx, dy * dy);
    if (!(d2 < 2.0f)) return 0.0f;
    float g = sat_gap32(dx, dy, ci, si, cj, sj);
    return g > 0.0f ? g : 0.0f;
}

float o_depth2_32_cs(float dx, float dy, float ci, float si, float cj, float sj) {
    float d2 = fmaf(dx, dx, dy * dy);
    float g = sat_gap32(dx, dy, ci, si, cj, sj);
    return (d2 < 2.0f && g > 0.0f) ? g * g : 0.0f;
}

float o_wall32_cs(float x, float y, float c, float s, float L) {
    float ext = 0.5f * (fabsf(c) + fabsf(s));
    float h = 0.5f * L;
    float ox = fmaxf(0.0f, (fabsf(x - h) + ext) - h);
    float oy = fmaxf(0.0f, (fabsf(y - h) + ext) - h);
    return fmaf(ox, ox, oy * oy);
}

/* Rotate (c, s) by t with Taylor polynomials (|t| <= 0.8: truncation below
 * 2e-10) and one Newton step back onto the unit circle. */
void o_rot32(float c, float s, float t, float* c_out, float* s_out) {
    float t2 = t * t;
    float sp = fmaf(t2, fmaf(t2, fmaf(t2, fmaf(t2, 2.7557319e-6f, -1.9841270e-4f), 8.3333333e-3f),
                             -1.6666667e-1f), 1.0f);
    float sn = t * sp;
    float cs = fmaf(t2, fmaf(t2, fmaf(t2, fmaf(t2, 2.4801587e-5f, -1.3888889e-3f), 4.1666667e-2f),
                             -0.5f), 1.0f);
    float c2 = fmaf(c, cs, -(s * sn));
    float s2 = fmaf(s, cs, c * sn);
    float nn = fmaf(c2, c2, s2 * s2);
    float k = fmaf(-0.5f, nn, 1.5f);
    *c_out = c2 * k;
    *s_out = s2 * k;
}

/* ------------------------------------------------------------------------ */
/* Reference FP64 geometry.                                                 */
/* ------------------------------------------------------------------------ */

static inline double sgn(double v) { return (double)((v > 0) - (v < 0)); }
static inline double sgn_dead(double v) { return fabs(v) > 1e-6 ? (v >= 0 ? 1.0 : -1.0) : 0.0; }

/* cuda/packing.cu:168-172: angle modulo pi/2 into [0, pi/2). */
double o_wrap_angle(double a) {
    const double quarter = 1.57079632679489661923;
    a = fmod(a, quarter);
    if (a < 0.0) a += quarter;
    return a;
}

/* Squared SAT penetration of p by q plus d/dp of it, with cached cos/sin.
 * Energy: cuda/packing.cu:177-189; gradient of the active axis:
 * cuda/seeded.cu:40-66 (hard sign) and cuda/packing.cu:191-214 (sign with a
 * 1e-6 dead zone, `smooth`). */
static double pair64_impl(const double p[3], const double q[3], double c, double s, double d,
                          double t, double grad[3], bool smooth) {
    const double dx = q[0] - p[0], dy = q[1] - p[1];
    if (dx * dx + dy * dy >= 2.000001) return 0.0;
    const double cd = c * d + s * t, sd = s * d - c * t;
    const double R = .5 + .5 * (fabs(cd) + fabs(sd));
    const double dR = smooth ? .5 * (-sgn_dead(cd) * sd + sgn_dead(sd) * cd)
                             : .5 * (-sgn(cd) * sd + sgn(sd) * cd);
    const double nx[4] = {c, -s, d, -t}, ny[4] = {s, c, t, d};
    double gap = -1e100, gx = 0, gy = 0, ga = 0;
    for (int k = 0; k < 4; ++k) {
        const double v = dx * nx[k] + dy * ny[k];
        const double z = fabs(v) - R;
        if (z > gap) {
            gap = z;
            const double sv = smooth ? (v >= 0.0 ? 1.0 : -1.0) : sgn(v);
            gx = -sv * nx[k];
            gy = -sv * ny[k];
            ga = -dR + (k < 2 ? sv * (-dx * ny[k] + dy * nx[k]) : 0.0);
        }
    }
    if (gap >= 0.0) return 0.0;
    if (grad) {
        grad[0] += 2.0 * gap * gx;
        grad[1] += 2.0 * gap * gy;
        grad[2] += 2.0 * gap * ga;
    }
    return gap * gap;
}

/* cuda/seeded.cu:88-90: pair() computes the trig itself. */
double o_pair64(const double p[3], const double q[3], double grad[3]) {
    return pair64_impl(p, q, cos(p[2]), sin(p[2]), cos(q[2]), sin(q[2]), grad, false);
}

double o_pair64_smooth_cs(const double p[3], const double q[3], double c, double s, double d,
                          double t, double grad[3]) {
    return pair64_impl(p, q, c, s, d, t, grad, true);
}

/* cuda/seeded.cu:92-113: four one-sided wall terms and their gradient. */
double o_boundary64(const double p[3], double L, double grad[3]) {
    const double c = cos(p[2]), s = sin(p[2]);
    const double r = .5 * (fabs(c) + fabs(s));
    const double v[4] = {r - p[0], p[0] + r - L, r - p[1], p[1] + r - L};
    const double dR = .5 * (-sgn_dead(c) * s + sgn_dead(s) * c);
    double e = 0;
    for (int k = 0; k < 4; ++k)
        if (v[k] > 0) {
            e += v[k] * v[k];
            if (grad) {
                const double f = 2 * v[k];
                grad[0] += f * (k == 0 ? -1 : k == 1 ? 1 : 0);
                grad[1] += f * (k == 2 ? -1 : k == 3 ? 1 : 0);
                grad[2] += f * dR;
            }
        }
    return e;
}

/* cuda/tests/seeded_test.cu:5-11: independent axis oracle, squared depth. */
double o_sat_oracle64(const double p[3], const double q[3]) {
    const double c = cos(p[2]), s = sin(p[2]), d = cos(q[2]), t = sin(q[2]);
    const double ax[4][2] = {{c, s}, {-s, c}, {d, t}, {-t, d}};
    double gap = -1e100;
    for (auto& a : ax) {
        double r = .5 * (fabs(a[0] * c + a[1] * s) + fabs(-a[0] * s + a[1] * c) +
                         fabs(a[0] * d + a[1] * t) + fabs(-a[0] * t + a[1] * d));
        gap = std::max(gap, fabs((p[0] - q[0]) * a[0] + (p[1] - q[1]) * a[1]) - r);
    }
    double depth = std::max(0.0, -gap);
    return depth * depth;
}

/* cuda/tests/selftest.cu:7-16: FP64 axis oracle on FP32 inputs, linear depth. */
double o_sat_oracle_depth_f(const float p[3], const float q[3]) {
    const double c = cos((double)p[2]), s = sin((double)p[2]);
    const double d = cos((double)q[2]), t = sin((double)q[2]);
    const double ax[4][2] = {{c, s}, {-s, c}, {d, t}, {-t, d}};
    double v = 1e9;
    for (auto& a : ax) {
        double r = .5 * (fabs(a[0] * c + a[1] * s) + fabs(-a[0] * s + a[1] * c) +
                         fabs(a[0] * d + a[1] * t) + fabs(-a[0] * t + a[1] * d));
        v = std::min(v, r - fabs(((double)p[0] - q[0]) * a[0] + ((double)p[1] - q[1]) * a[1]));
    }
    return std::max(0.0, v);
}

/* cuda/packing.cu:240-252: walls (one side per axis) plus all pairs i<j. */
double o_total64_cs(int n, const double* x, const double* y, const double* c, const double* s,
                    double L) {
    double e = 0.0;
    for (int i = 0; i < n; ++i) {
        const double r = 0.5 * (fabs(c[i]) + fabs(s[i]));
        if (x[i] < r) { double u = r - x[i]; e += u * u; }
        else if (x[i] > L - r) { double u = x[i] + r - L; e += u * u; }
        if (y[i] < r) { double u = r - y[i]; e += u * u; }
        else if (y[i] > L - r) { double u = y[i] + r - L; e += u * u; }
        const double pi[3] = {x[i], y[i], 0.0};
        for (int j = i + 1; j < n; ++j) {
            const double pj[3] = {x[j], y[j], 0.0};
            e += pair64_impl(pi, pj, c[i], s[i], c[j], s[j], nullptr, false);
        }
    }
    return e;
}

double o_total64(int n, const double* q, double L) {
    std::vector<double> x(n), y(n), c(n), s(n);
    for (int i = 0; i < n; ++i) {
        x[i] = q[3 * i]; y[i] = q[3 * i + 1];
        c[i] = cos(q[3 * i + 2]); s[i] = sin(q[3 * i + 2]);
    }
    return o_total64_cs(n, x.data(), y.data(), c.data(), s.data(), L);
}

/* cuda/packing.cu:254-297: smallest uniform centre dilation lambda that puts
 * every near pair (d^2 < 2.05) at least tau = 2.6e-8 beyond contact on its
 * best axis; reject lambda > 1.008; then the bounding side of the dilated
 * configuration, and (when it improves current_best) the translated output. */
double o_certified_side(int n, const double* q, double current_best, double* out,
                        double* lam_out) {
    const double tau = 2.6e-8;
    std::vector<double> c(n), s(n);
    for (int i = 0; i < n; ++i) { c[i] = cos(q[3 * i + 2]); s[i] = sin(q[3 * i + 2]); }
    double lam = 1.0;
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j) {
            const double dx = q[3 * i] - q[3 * j], dy = q[3 * i + 1] - q[3 * j + 1];
            if (dx * dx + dy * dy >= 2.05) continue;
            const double R = 0.5 + 0.5 * (fabs(c[i] * c[j] + s[i] * s[j]) +
                                          fabs(-s[i] * c[j] + c[i] * s[j]));
            const double v = std::max(std::max(fabs(dx * c[i] + dy * s[i]), fabs(-dx * s[i] + dy * c[i])),
                                      std::max(fabs(dx * c[j] + dy * s[j]), fabs(-dx * s[j] + dy * c[j])));
            const double need = v > 1e-12 ? (R + tau) / v : 1e9;
            if (need > lam) lam = need;
        }
    if (lam_out) *lam_out = lam;
    if (lam > 1.008) return 1e9;
    double lox = 1e9, hix = -1e9, loy = 1e9, hiy = -1e9;
    for (int i = 0; i < n; ++i) {
        const double r = 0.5 * (fabs(c[i]) + fabs(s[i]));
        const double x = q[3 * i] * lam, y = q[3 * i + 1] * lam;
        lox = std::min(lox, x - r); hix = std::max(hix, x + r);
        loy = std::min(loy, y - r); hiy = std::max(hiy, y + r);
    }
    const double side = std::max(hix - lox, hiy - loy);
    if (side < current_best && out)
        for (int i = 0; i < n; ++i) {
            out[3 * i] = q[3 * i] * lam - lox;
            out[3 * i + 1] = q[3 * i + 1] * lam - loy;
            out[3 * i + 2] = o_wrap_angle(q[3 * i + 2]);
        }
    return side;
}

/* cuda/packing.cu:624-687: Polak-Ribiere+ conjugate gradient on the overlap +
 * wall energy; restart every 25 iterations; five backtracking trials of
 * factor 0.45 from step (0.04 initially, x1.25 up to 0.15 on success, x0.5
 * down to 1e-7 on failure); only strictly lower energies are accepted. */
int o_polish64(int n, double* q, double L, int iters, double* energy_inout) {
    std::vector<double> g(3 * n), gold(3 * n, 0.0), dir(3 * n, 0.0), saved(3 * n);
    double energy = o_total64(n, q, L);
    double step = 0.04;
    int it = 0;
    for (; it < iters; ++it) {
        if (energy < 1e-24) break;
        std::vector<double> c(n), s(n);
        for (int i = 0; i < n; ++i) { c[i] = cos(q[3 * i + 2]); s[i] = sin(q[3 * i + 2]); }
        double gg = 0, dgg = 0;
        for (int i = 0; i < n; ++i) {
            double gi[3] = {0, 0, 0};
            const double ext = 0.5 * (fabs(c[i]) + fabs(s[i]));
            const double dR = .5 * (-sgn_dead(c[i]) * s[i] + sgn_dead(s[i]) * c[i]);
            const double px = q[3 * i], py = q[3 * i + 1];
            if (px < ext) { double u = ext - px; gi[0] -= 2 * u; gi[2] += 2 * u * dR; }
            else if (px > L - ext) { double u = px + ext - L; gi[0] += 2 * u; gi[2] += 2 * u * dR; }
            if (py < ext) { double u = ext - py; gi[1] -= 2 * u; gi[2] += 2 * u * dR; }
            else if (py > L - ext) { double u = py + ext - L; gi[1] += 2 * u; gi[2] += 2 * u * dR; }
            for (int j = 0; j < n; ++j)
                if (j != i) pair64_impl(q + 3 * i, q + 3 * j, c[i], s[i], c[j], s[j], gi, true);
            for (int k = 0; k < 3; ++k) {
                g[3 * i + k] = gi[k];
                gg += gold[3 * i + k] * gold[3 * i + k];
                dgg += (gi[k] - gold[3 * i + k]) * gi[k];
            }
        }
        const double beta = (it == 0 || gg < 1e-30 || it % 25 == 0) ? 0.0 : std::max(0.0, dgg / gg);
        double norm2 = 0;
        for (int k = 0; k < 3 * n; ++k) {
            dir[k] = -g[k] + beta * dir[k];
            gold[k] = g[k];
            saved[k] = q[k];
            norm2 += dir[k] * dir[k];
        }
        if (norm2 < 1e-28) break;
        double alpha = step;
        bool ok = false;
        for (int ls = 0; ls < 5; ++ls) {
            for (int i = 0; i < n; ++i) {
                q[3 * i] = saved[3 * i] + alpha * dir[3 * i];
                q[3 * i + 1] = saved[3 * i + 1] + alpha * dir[3 * i + 1];
                q[3 * i + 2] = o_wrap_angle(saved[3 * i + 2] + alpha * dir[3 * i + 2]);
            }
            const double ne = o_total64(n, q, L);
            if (ne < energy) {
                energy = ne;
                step = std::min(0.15, alpha * 1.25);
                ok = true;
                break;
            }
            alpha *= 0.45;
        }
        if (!ok) {
            for (int k = 0; k < 3 * n; ++k) { q[k] = saved[k]; dir[k] = 0.0; }
            step = std::max(1e-7, step * 0.5);
        }
    }
    if (energy_inout) *energy_inout = energy;
    return it;
}

/* ------------------------------------------------------------------------ */
/* CPU PT-SA engine (FP64).  Follows the reference's FP64 search semantics:  */
/* single-square moves with an O(n) energy change (cuda/packing.cu:542-563), */
/* geometric temperature ladder (:516-517), Metropolis on log u (:555),      */
/* even/odd exchange within a ladder (:750-765), shrink-and-rescale of the   */
/* box on a zero-overlap configuration (:833-848).  Sweeps visit the squares */
/* in index order; steps are uniform of half-width sqrt(3)*sigma, adapted    */
/* per temperature toward the target acceptance at every exchange point.    */
/* ------------------------------------------------------------------------ */
namespace {

struct Replica {
    std::vector<double> x, y, a, c, s;
    double E = 0;
    uint64_t acc = 0, prop = 0, acc_total = 0;
};

struct SpinBarrier {
    std::atomic<int> count{0};
    std::atomic<int> phase{0};
    int total;
    explicit SpinBarrier(int t) : total(t) {}
    void wait() {
        int ph = phase.load(std::memory_order_acquire);
        if (count.fetch_add(1, std::memory_order_acq_rel) + 1 == total) {
            count.store(0, std::memory_order_relaxed);
            phase.store(ph + 1, std::memory_order_release);
        } else {
            while (phase.load(std::memory_order_acquire) == ph) std::this_thread::yield();
        }
    }
};

double replica_energy(const Replica& r, int n, double L) {
    return o_total64_cs(n, r.x.data(), r.y.data(), r.c.data(), r.s.data(), L);
}

double wall64_cs(double x, double y, double c, double s, double L) {
    const double r = .5 * (fabs(c) + fabs(s));
    const double v[4] = {r - x, x + r - L, r - y, y + r - L};
    double e = 0;
    for (double t : v)
        if (t > 0) e += t * t;
    return e;
}

}  // namespace

int o_ptsa_run(const o_ptsa_cfg* cfg, o_ptsa_out* out) {
    const int n = cfg->n, R = cfg->replicas;
    const int ladder = cfg->ladder > 0 ? cfg->ladder : R;
    if (n < 1 || R < 1 || R % ladder != 0 || cfg->swap_every < 1 || cfg->sweeps < 0) return 2;
    const uint32_t key[2] = {(uint32_t)cfg->seed, (uint32_t)(cfg->seed >> 32)};
    double L = cfg->L0;
    std::vector<double> T(ladder);
    for (int k = 0; k < ladder; ++k)
        T[k] = ladder == 1 ? cfg->tmin
                           : cfg->tmin * pow(cfg->tmax / cfg->tmin, (double)k / (ladder - 1));
    std::vector<double> sig_xy(ladder, cfg->move_sigma), sig_th(ladder, cfg->rot_sigma);
    std::vector<int> slot_of(R), rep_of(R);
    for (int r = 0; r < R; ++r) { slot_of[r] = r % ladder; rep_of[r] = r; }

    std::vector<Replica> reps(R);
    for (int r = 0; r < R; ++r) {
        Replica& rp = reps[r];
        rp.x.resize(n); rp.y.resize(n); rp.a.resize(n); rp.c.resize(n); rp.s.resize(n);
        for (int i = 0; i < n; ++i) {
            if (cfg->init) {
                rp.x[i] = cfg->init[3 * ((size_t)r * n + i)];
                rp.y[i] = cfg->init[3 * ((size_t)r * n + i) + 1];
                rp.a[i] = cfg->init[3 * ((size_t)r * n + i) + 2];
            } else {
                uint32_t ctr[4] = {(uint32_t)i, (uint32_t)r, 0u, 3u}, o[4];
                o_philox(ctr, key, o);
                rp.x[i] = 0.5 + (L - 1.0) * o_u01d(o[0], o[1]);
                rp.y[i] = 0.5 + (L - 1.0) * o_u01d(o[2], o[3]);
                uint32_t ctr2[4] = {(uint32_t)i, (uint32_t)r, 1u, 3u};
                o_philox(ctr2, key, o);
                rp.a[i] = 1.5707963267948966 * o_u01d(o[0], o[1]);
            }
            rp.c[i] = cos(rp.a[i]);
            rp.s[i] = sin(rp.a[i]);
        }
        rp.E = replica_energy(rp, n, L);
    }

    const int threads = std::max(1, std::min(cfg->threads, R));
    const int64_t blocks = (cfg->sweeps + cfg->swap_every - 1) / cfg->swap_every;
    SpinBarrier bar(threads);
    std::atomic<bool> stop{false};
    double L_feasible = cfg->L0;
    bool have_feasible = false;
    int shrinks = 0;
    uint64_t ex_try = 0, ex_acc = 0;
    if (out && out->best_q)
        for (int i = 0; i < n; ++i) {
            out->best_q[3 * i] = reps[0].x[i];
            out->best_q[3 * i + 1] = reps[0].y[i];
            out->best_q[3 * i + 2] = reps[0].a[i];
        }

    auto worker = [&](int tid) {
        for (int64_t b = 0; b < blocks; ++b) {
            const int64_t s0 = b * cfg->swap_every;
            const int64_t s1 = std::min<int64_t>(cfg->sweeps, s0 + cfg->swap_every);
            for (int r = tid; r < R; r += threads) {
                Replica& rp = reps[r];
                const int slot = slot_of[r];
                const double a_xy = std::sqrt(3.0) * sig_xy[slot], a_th = std::sqrt(3.0) * sig_th[slot];
                const double invT = 1.0 / T[slot];
                for (int64_t sw = s0; sw < s1; ++sw) {
                    for (int i = 0; i < n; ++i) {
                        uint32_t ctr[4] = {(uint32_t)i, (uint32_t)r, (uint32_t)sw,
                                           (uint32_t)(sw >> 32) << 8 | 4u};
                        uint32_t o[4];
                        o_philox(ctr, key, o);
                        rp.prop++;
                        const double nx = rp.x[i] + a_xy * (2.0 * o_u01f(o[0]) - 1.0);
                        const double ny = rp.y[i] + a_xy * (2.0 * o_u01f(o[1]) - 1.0);
                        if (nx < 0 || nx > L || ny < 0 || ny > L) continue;
                        const double na = rp.a[i] + a_th * (2.0 * o_u01f(o[2]) - 1.0);
                        const double nc = cos(na), ns = sin(na);
                        double dE = wall64_cs(nx, ny, nc, ns, L) -
                                    wall64_cs(rp.x[i], rp.y[i], rp.c[i], rp.s[i], L);
                        const double pn[3] = {nx, ny, 0}, po[3] = {rp.x[i], rp.y[i], 0};
                        for (int j = 0; j < n; ++j) {
                            if (j == i) continue;
                            const double pj[3] = {rp.x[j], rp.y[j], 0};
                            dE += pair64_impl(pn, pj, nc, ns, rp.c[j], rp.s[j], nullptr, false) -
                                  pair64_impl(po, pj, rp.c[i], rp.s[i], rp.c[j], rp.s[j], nullptr, false);
                        }
                        if (dE <= 0.0 || log((double)o_u01f(o[3])) < -dE * invT) {
                            rp.x[i] = nx; rp.y[i] = ny; rp.a[i] = o_wrap_angle(na);
                            rp.c[i] = nc; rp.s[i] = ns;
                            rp.acc++;
                        }
                    }
                }
                rp.E = replica_energy(rp, n, L);
            }
            bar.wait();
            if (tid == 0) {
                /* adaptation per temperature slot from the replica that held it */
                for (int r = 0; r < R; ++r) {
                    Replica& rp = reps[r];
                    if (rp.prop) {
                        double rate = (double)rp.acc / (double)rp.prop;
                        int k = slot_of[r];
                        double f = exp((rate - cfg->target_accept) * 0.5);
                        sig_xy[k] = std::min(0.5, std::max(1e-7, sig_xy[k] * f));
                        sig_th[k] = std::min(0.45, std::max(1e-7, sig_th[k] * f));
                    }
                    rp.acc_total += rp.acc;
                    rp.acc = 0;
                }
                /* even/odd exchange on adjacent temperatures of each ladder */
                const int parity = (int)(b & 1);
                for (int base = 0; base < R; base += ladder)
                    for (int k = parity; k + 1 < ladder; k += 2) {
                        int ra = rep_of[base + k], rb = rep_of[base + k + 1];
                        uint32_t ctr[4] = {(uint32_t)(base + k), 0u, (uint32_t)b, 2u}, o[4];
                        o_philox(ctr, key, o);
                        ex_try++;
                        double lhs = log((double)o_u01f(o[0]));
                        if (lhs < (1.0 / T[k] - 1.0 / T[k + 1]) * (reps[ra].E - reps[rb].E)) {
                            std::swap(rep_of[base + k], rep_of[base + k + 1]);
                            slot_of[ra] = k + 1;
                            slot_of[rb] = k;
                            ex_acc++;
                        }
                    }
                /* shrink on a zero-overlap replica: keep it, then squeeze */
                int zero = -1;
                for (int k = 0; k < R && zero < 0; ++k)
                    if (reps[rep_of[k]].E == 0.0) zero = rep_of[k];
                if (zero >= 0) {
                    if (L < L_feasible || !have_feasible) {
                        L_feasible = L;
                        have_feasible = true;
                        if (out && out->best_q)
                            for (int i = 0; i < n; ++i) {
                                out->best_q[3 * i] = reps[zero].x[i];
                                out->best_q[3 * i + 1] = reps[zero].y[i];
                                out->best_q[3 * i + 2] = reps[zero].a[i];
                            }
                    }
                    double Ln = std::max(L - cfg->shrink, std::sqrt((double)n));
                    if (Ln < L) {
                        double f = Ln / L;
                        for (auto& rp : reps)
                            for (int i = 0; i < n; ++i) { rp.x[i] *= f; rp.y[i] *= f; }
                        L = Ln;
                        shrinks++;
                        for (auto& rp : reps) rp.E = replica_energy(rp, n, L);
                    }
                }
            }
            bar.wait();
            if (stop.load()) break;
        }
    };

    auto t0 = std::chrono::steady_clock::now();
    std::vector<std::thread> pool;
    for (int t = 1; t < threads; ++t) pool.emplace_back(worker, t);
    worker(0);
    for (auto& th : pool) th.join();
    double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();

    if (out) {
        uint64_t prop = 0, acc = 0;
        double emin = 1e300;
        for (auto& rp : reps) {
            prop += rp.prop;
            acc += rp.acc_total + rp.acc;
            emin = std::min(emin, rp.E);
        }
        out->L_final = L;
        out->L_feasible = L_feasible;
        out->proposals = prop;
        out->accepted = acc;
        out->exchanges_tried = ex_try;
        out->exchanges_accepted = ex_acc;
        out->shrinks = shrinks;
        out->seconds = secs;
        out->e_min = emin;
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* CPU model of the product's FP32 colour-checkerboard sweep.  The product  */
/* follows the reference's FP32 walk (cuda/packing.cu:52-95: uniform cell    */
/* list, simultaneous moves only for squares that cannot interact, Metropolis */
/* on the local energy change) with a 3x3-colour checkerboard of cells of    */
/* width w >= sqrt(2)/2 (Anderson et al., "Scalable Metropolis Monte Carlo   */
/* for simulation of hard shapes", Comput. Phys. Commun. 204 (2016) 21-30):  */
/* moves that leave their cell are rejected, so same-colour cells never      */
/* interact, and the grid origin is shifted at random every k sweeps         */
/* (sweeps_per_shift).                                                       */
/* ------------------------------------------------------------------------ */
namespace {

constexpr int kRankCap = 4;          /* ranks >= this go to the serial group */
constexpr int kGroups = 9 * kRankCap;

struct Grid {
    int m, M;
    float shx, shy;
    std::vector<int> kx, ky, cell, start, sorted, rank;
};

void build_grid(Grid& g, int n, float L, float w, float shx, float shy, const float* x,
                const float* y) {
    g.m = (int)floorf((L + w) / w) + 1;
    g.M = g.m + 4;
    g.shx = shx; g.shy = shy;
    g.kx.assign(n, 0); g.ky.assign(n, 0); g.cell.assign(n, 0); g.rank.assign(n, 0);
    g.start.assign(g.M * g.M + 1, 0);
    for (int i = 0; i < n; ++i) {
        g.kx[i] = o_cellcoord(x[i], shx, w, 0, g.m - 1) + 2;
        g.ky[i] = o_cellcoord(y[i], shy, w, 0, g.m - 1) + 2;
        g.cell[i] = g.ky[i] * g.M + g.kx[i];
        g.start[g.cell[i] + 1]++;
    }
    for (int c = 0; c < g.M * g.M; ++c) g.start[c + 1] += g.start[c];
    std::vector<int> fill(g.M * g.M, 0);
    g.sorted.assign(n, 0);
    /* canonical order: ascending square index inside each cell */
    for (int i = 0; i < n; ++i) {
        g.rank[i] = fill[g.cell[i]]++;
        g.sorted[g.start[g.cell[i]] + g.rank[i]] = i;
    }
}

/* FP64 energy of square i at pose (px, py, pc, ps) against every other
 * square and the walls of a box of side L, from the FP32 state converted
 * exactly; *depth accumulates the overlap depths of the nonzero terms. */
double move_energy64(int n, const float* x, const float* y, const float* c, const float* s, int i, float px,
                     float py, float pc, float ps, float L, double* depth) {
    double e = wall64_cs(px, py, pc, ps, L);
    const double r = .5 * (fabs((double)pc) + fabs((double)ps));
    const double v[4] = {r - px, px + r - L, r - py, py + r - L};
    for (double t : v)
        if (t > 0) *depth += t;
    const double pi[3] = {px, py, 0.0};
    for (int j = 0; j < n; ++j) {
        if (j == i) continue;
        const double pj[3] = {x[j], y[j], 0.0};
        const double t = pair64_impl(pi, pj, pc, ps, c[j], s[j], nullptr, false);
        e += t;
        *depth += sqrt(t);
    }
    return e;
}

}  // namespace

int o_sweep32_model(const o_sweep_cfg* cfg, float* x, float* y, float* c, float* s,
                    uint64_t counts[3]) {
    const int n = cfg->n;
    const float L = cfg->L, w = cfg->w;
    const uint32_t key[2] = {cfg->key0, cfg->key1};
    uint32_t o[4];
    {
        /* blocks of k sweeps share one shift (a move never leaves its cell, so
           the cell lists built here equal the ones the device keeps) */
        const uint32_t k = cfg->sweeps_per_shift > 1 ? cfg->sweeps_per_shift : 1u;
        uint32_t ctr[4] = {0u, cfg->replica, cfg->sweep / k * k, 1u};
        o_philox(ctr, key, o);
    }
    const float shx = w * o_u01f(o[0]), shy = w * o_u01f(o[1]);
    Grid g;
    build_grid(g, n, L, w, shx, shy, x, y);
    std::vector<std::vector<int>> groups(kGroups + 1);
    for (int i = 0; i < n; ++i) {
        int colour = (g.kx[i] % 3) + 3 * (g.ky[i] % 3);
        int key_g = g.rank[i] < kRankCap ? g.rank[i] * 9 + colour : kGroups;
        groups[key_g].push_back(i);
    }
    int accepted = 0;
    uint64_t C = 0, P = 0;
    for (int gi = 0; gi <= kGroups; ++gi) {
        /* members of one group cannot interact: any order gives the same result */
        for (int i : groups[gi]) {
            uint32_t ctr[4] = {(uint32_t)i, cfg->replica, cfg->sweep, 0u};
            o_philox(ctr, key, o);
            const float fx = fmaf(2.0f, o_u01f(o[0]), -1.0f);
            const float fy = fmaf(2.0f, o_u01f(o[1]), -1.0f);
            const float ft = fmaf(2.0f, o_u01f(o[2]), -1.0f);
            const float nx = x[i] + cfg->step_xy * fx;
            const float ny = y[i] + cfg->step_xy * fy;
            if (cfg->dE32) cfg->dE32[i] = NAN;
            if (cfg->dE64) cfg->dE64[i] = NAN;
            if (!(nx >= 0.0f && nx <= L && ny >= 0.0f && ny <= L)) continue;
            if (o_cellcoord(nx, g.shx, w, 0, g.m - 1) + 2 != g.kx[i] ||
                o_cellcoord(ny, g.shy, w, 0, g.m - 1) + 2 != g.ky[i])
                continue;
            float nc, ns;
            o_rot32(c[i], s[i], cfg->step_th * ft, &nc, &ns);
            const float wdiff = o_wall32_cs(nx, ny, nc, ns, L) - o_wall32_cs(x[i], y[i], c[i], s[i], L);
            /* stencil rows in order, CSR order inside a row; the flat index f
             * (self included) selects one of four partial sums */
            float part[4] = {0.0f, 0.0f, 0.0f, 0.0f};
            int f = 0;
            for (int oy = -2; oy <= 2; ++oy) {
                const int row = (g.ky[i] + oy) * g.M;
                const int a = g.start[row + g.kx[i] - 2], b = g.start[row + g.kx[i] + 3];
                for (int k = a; k < b; ++k, ++f) {
                    const int j = g.sorted[k];
                    if (j == i) continue;
                    const float dxn = x[j] - nx, dyn = y[j] - ny;
                    const float dxo = x[j] - x[i], dyo = y[j] - y[i];
                    const float en = o_depth2_32_cs(dxn, dyn, nc, ns, c[j], s[j]);
                    const float eo = o_depth2_32_cs(dxo, dyo, c[i], s[i], c[j], s[j]);
                    part[f & 3] += en - eo;
                    C++;
                    P += (fmaf(dxn, dxn, dyn * dyn) < 2.0f) + (fmaf(dxo, dxo, dyo * dyo) < 2.0f);
                }
            }
            const float dE = wdiff + ((part[0] + part[1]) + (part[2] + part[3]));
            if (cfg->dE32) cfg->dE32[i] = dE;
            if (cfg->dE64) {
                double dn = 0.0, d_o = 0.0;
                const double en = move_energy64(n, x, y, c, s, i, nx, ny, nc, ns, L, &dn);
                const double eo = move_energy64(n, x, y, c, s, i, x[i], y[i], c[i], s[i], L, &d_o);
                cfg->dE64[i] = en - eo;
                if (cfg->depth64) cfg->depth64[i] = dn + d_o;
            }
            const float u = o_u01f(o[3]);
            if (dE <= 0.0f || u < o_accept_exp2(-dE * cfg->inv_t_log2e)) {
                x[i] = nx; y[i] = ny; c[i] = nc; s[i] = ns;
                accepted++;
            }
        }
    }
    if (counts) {
        counts[0] += (uint64_t)n;
        counts[1] += C;
        counts[2] += P;
    }
    return accepted;
}

float o_energy32_model(int n, float L, float w, const float* x, const float* y, const float* c,
                       const float* s) {
    Grid g;
    build_grid(g, n, L, w, 0.0f, 0.0f, x, y);
    std::vector<float> e(n);
    for (int i = 0; i < n; ++i) {
        float ei = o_wall32_cs(x[i], y[i], c[i], s[i], L);
        for (int oy = -2; oy <= 2; ++oy) {
            const int row = (g.ky[i] + oy) * g.M;
            const int a = g.start[row + g.kx[i] - 2], b = g.start[row + g.kx[i] + 3];
            for (int k = a; k < b; ++k) {
                const int j = g.sorted[k];
                if (j <= i) continue;
                ei += o_depth2_32_cs(x[j] - x[i], y[j] - y[i], c[i], s[i], c[j], s[j]);
            }
        }
        e[i] = ei;
    }
    for (int stride = 1; stride < n; stride <<= 1)
        for (int i = 0; i + stride < n; i += 2 * stride) e[i] += e[i + stride];
    return n > 0 ? e[0] : 0.0f;
}

/* ------------------------------------------------------------------------ */
/* Collective moves and the kick of the product (CPU model): the FP32       */
/* restatement of what the device applies after a launch's sweeps, for      */
/* cuda/packing.cu:564-616 (modes 1-6) and :696-745 (worst-square void      */
/* relocation + jitter).  Same draws, same operation order.                 */
/* ------------------------------------------------------------------------ */
namespace {

struct Coll {
    float cx, cy, cc, cs, mv, rt, tc, ts, th;
    int mode;
};

bool coll_take(const Coll& w, float x, float y, float c, float s) {
    const float dx = x - w.cx, dy = y - w.cy;
    switch (w.mode) {
        case 1: return fabsf(dy) < 0.55f;
        case 2: return fabsf(dx) < 0.55f;
        case 3: return fmaf(dx, dx, dy * dy) < 2.5f;
        case 4: {
            const float cd = fmaf(c, w.cc, s * w.cs), sd = fmaf(s, w.cc, -(c * w.cs));
            return fmaxf(fabsf(cd), fabsf(sd)) > 0.99820054f;
        }
        case 5: return fabsf(dx + dy) < 0.85f || fabsf(dx - dy) < 0.85f;
        default: return fmaxf(fabsf(c), fabsf(s)) < 0.99920011f;
    }
}

}  // namespace

int o_collective32_model(const o_sweep_cfg* cfg, int moves, float* x, float* y, float* c, float* s,
                         float* energy) {
    const int n = cfg->n;
    const float L = cfg->L;
    const uint32_t key[2] = {cfg->key0, cfg->key1};
    float e_old = o_energy32_model(n, L, cfg->w, x, y, c, s);
    std::vector<float> bx(n), by(n), bc(n), bs(n);
    int accepted = 0;
    for (int k = 0; k < moves; ++k) {
        uint32_t o1[4], o2[4];
        {
            uint32_t ctr[4] = {(uint32_t)k, cfg->replica, cfg->sweep, 5u};
            o_philox(ctr, key, o1);
            uint32_t ctr2[4] = {(uint32_t)k, cfg->replica, cfg->sweep, 6u};
            o_philox(ctr2, key, o2);
        }
        Coll w;
        const int anchor = std::min(n - 1, (int)(o_u01f(o1[0]) * (float)n));
        w.mode = 1 + std::min(5, (int)(o_u01f(o1[1]) * 6.0f));
        w.cx = x[anchor]; w.cy = y[anchor]; w.cc = c[anchor]; w.cs = s[anchor];
        w.mv = cfg->step_xy * fmaf(2.0f, o_u01f(o1[2]), -1.0f);
        w.rt = cfg->step_xy * fmaf(2.0f, o_u01f(o1[3]), -1.0f);
        w.th = cfg->step_th * fmaf(2.0f, o_u01f(o2[0]), -1.0f);
        o_rot32(1.0f, 0.0f, w.th, &w.tc, &w.ts);
        const float uacc = o_u01f(o2[1]);
        for (int i = 0; i < n; ++i) { bx[i] = x[i]; by[i] = y[i]; bc[i] = c[i]; bs[i] = s[i]; }
        for (int i = 0; i < n; ++i) {
            if (!coll_take(w, bx[i], by[i], bc[i], bs[i])) continue;
            const float dx = bx[i] - w.cx, dy = by[i] - w.cy;
            float nx, ny;
            if (w.mode == 3 || w.mode == 4) {
                nx = (w.cx + fmaf(w.tc, dx, -(w.ts * dy))) + w.mv;
                ny = (w.cy + fmaf(w.ts, dx, w.tc * dy)) + w.rt;
            } else if (w.mode == 6) {
                const float sg = (i & 1) ? 1.0f : -1.0f;
                nx = bx[i] + sg * w.mv;
                ny = by[i] + sg * w.rt;
            } else {
                nx = bx[i] + w.mv;
                ny = by[i] + w.rt;
            }
            o_rot32(bc[i], bs[i], w.th, &c[i], &s[i]);
            x[i] = fminf(L - 0.45f, fmaxf(0.45f, nx));
            y[i] = fminf(L - 0.45f, fmaxf(0.45f, ny));
        }
        const float e_new = o_energy32_model(n, L, cfg->w, x, y, c, s);
        const float dE = e_new - e_old;
        if (dE <= 0.0f || uacc < o_accept_exp2(-dE * cfg->inv_t_log2e)) {
            e_old = e_new;
            accepted++;
        } else {
            for (int i = 0; i < n; ++i) { x[i] = bx[i]; y[i] = by[i]; c[i] = bc[i]; s[i] = bs[i]; }
        }
    }
    if (energy) *energy = e_old;
    return accepted;
}

int o_kick32_model(const o_sweep_cfg* cfg, uint32_t interval, int relocate, float* x, float* y, float* c,
                   float* s, float* energy) {
    const int n = cfg->n;
    const float L = cfg->L;
    const uint32_t key[2] = {cfg->key0, cfg->key1};
    const float e0 = o_energy32_model(n, L, cfg->w, x, y, c, s);
    if (energy) *energy = e0;
    if (!(e0 > 1e-12f)) return 0;
    Grid g;
    build_grid(g, n, L, cfg->w, 0.0f, 0.0f, x, y);
    std::vector<float> sq(n);
    int worst = 0;
    for (int i = 0; i < n; ++i) {
        float ei = o_wall32_cs(x[i], y[i], c[i], s[i], L);
        for (int oy = -2; oy <= 2; ++oy) {
            const int row = (g.ky[i] + oy) * g.M;
            const int a = g.start[row + g.kx[i] - 2], b = g.start[row + g.kx[i] + 3];
            for (int k = a; k < b; ++k) {
                const int j = g.sorted[k];
                if (j == i) continue;
                ei += o_depth2_32_cs(x[j] - x[i], y[j] - y[i], c[i], s[i], c[j], s[j]);
            }
        }
        sq[i] = ei;
        if (ei > sq[worst]) worst = i;  /* first maximum */
    }
    if (relocate && sq[worst] > 1e-5f) {
        int best = 0;
        float best_d2 = -1.0f;
        for (int p = 0; p < 64; ++p) {
            const float vx = fmaf(L - 1.1f, ((p & 7) + 0.5f) * 0.125f, 0.55f);
            const float vy = fmaf(L - 1.1f, ((p >> 3) + 0.5f) * 0.125f, 0.55f);
            float d2 = 1e9f;
            for (int j = 0; j < n; ++j) {
                if (j == worst) continue;
                const float dx = vx - x[j], dy = vy - y[j];
                d2 = fminf(d2, fmaf(dx, dx, dy * dy));
            }
            if (d2 > best_d2) { best_d2 = d2; best = p; }  /* first maximum */
        }
        const float vx = fmaf(L - 1.1f, ((best & 7) + 0.5f) * 0.125f, 0.55f);
        const float vy = fmaf(L - 1.1f, ((best >> 3) + 0.5f) * 0.125f, 0.55f);
        uint32_t o[4];
        uint32_t ctr[4] = {0u, cfg->replica, interval, 7u};
        o_philox(ctr, key, o);
        x[worst] = fminf(L - 0.5f, fmaxf(0.5f, vx + 0.05f * fmaf(2.0f, o_u01f(o[0]), -1.0f)));
        y[worst] = fminf(L - 0.5f, fmaxf(0.5f, vy + 0.05f * fmaf(2.0f, o_u01f(o[1]), -1.0f)));
        o_rot32(0.70710678f, 0.70710678f, o_u01f(o[2]) * 1.57079633f - 0.78539816f, &c[worst], &s[worst]);
    }
    for (int i = 0; i < n; ++i) {
        const bool hot = sq[i] > 1e-6f;
        const float axy = hot ? 0.08f : 0.008f, aa = hot ? 0.18f : 0.015f;
        uint32_t o[4];
        uint32_t ctr[4] = {(uint32_t)i, cfg->replica, interval, 8u};
        o_philox(ctr, key, o);
        x[i] = fminf(L - 0.48f, fmaxf(0.48f, x[i] + axy * fmaf(2.0f, o_u01f(o[0]), -1.0f)));
        y[i] = fminf(L - 0.48f, fmaxf(0.48f, y[i] + axy * fmaf(2.0f, o_u01f(o[1]), -1.0f)));
        const float c0 = c[i], s0 = s[i];
        o_rot32(c0, s0, aa * fmaf(2.0f, o_u01f(o[2]), -1.0f), &c[i], &s[i]);
    }
    if (energy) *energy = o_energy32_model(n, L, cfg->w, x, y, c, s);
    return 1;
}

/* src/squares/core.py:27-35: half-integer grid with integer pitch. */
int o_grid_packing(int n, double* q) {
    if (n < 1) return 2;
    int width = 1;
    while (width * width < n) ++width;
    for (int i = 0; i < n; ++i) {
        q[3 * i] = (i % width) + 0.5;
        q[3 * i + 1] = (i / width) + 0.5;
        q[3 * i + 2] = 0.0;
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Certified side of an FP32 state (cuda/packing.cu:254-297).  The engine   */
/* keeps orientations as FP32 (cos, sin); they are scaled to unit length in */
/* FP64 first.  All pairs i < j nearer than sqrt(2.05), as the reference.   */
/* ------------------------------------------------------------------------ */
namespace {
inline void unit64(float c, float s, double& cd, double& sd) {
    const double x = c, y = s;
    const double h = 1.0 / sqrt(x * x + y * y);
    cd = x * h;
    sd = y * h;
}
}  // namespace

void o_certify_xycs(int n, const float* q, double out[4]) {
    const double tau = 2.6e-8;
    std::vector<double> c(n), s(n);
    for (int i = 0; i < n; ++i) unit64(q[4 * i + 2], q[4 * i + 3], c[i], s[i]);
    double lam = 1.0;
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j) {
            const double dx = (double)q[4 * i] - (double)q[4 * j], dy = (double)q[4 * i + 1] - (double)q[4 * j + 1];
            if (dx * dx + dy * dy >= 2.05) continue;
            const double r = 0.5 + 0.5 * (fabs(c[i] * c[j] + s[i] * s[j]) + fabs(-s[i] * c[j] + c[i] * s[j]));
            const double v1 = fabs(dx * c[i] + dy * s[i]), v2 = fabs(-dx * s[i] + dy * c[i]);
            const double v3 = fabs(dx * c[j] + dy * s[j]), v4 = fabs(-dx * s[j] + dy * c[j]);
            const double vmax = std::max(std::max(v1, v2), std::max(v3, v4));
            const double req = vmax > 1e-12 ? (r + tau) / vmax : 1e9;
            if (req > lam) lam = req;
        }
    double lox = 1e9, hix = -1e9, loy = 1e9, hiy = -1e9;
    for (int i = 0; i < n; ++i) {
        const double r = 0.5 * (fabs(c[i]) + fabs(s[i]));
        const double x = (double)q[4 * i] * lam, y = (double)q[4 * i + 1] * lam;
        lox = std::min(lox, x - r); hix = std::max(hix, x + r);
        loy = std::min(loy, y - r); hiy = std::max(hiy, y + r);
    }
    out[0] = lam > 1.008 ? 1e9 : std::max(hix - lox, hiy - loy);
    out[1] = lam;
    out[2] = lox;
    out[3] = loy;
}

/* ------------------------------------------------------------------------ */
/* The exchange step: the engine's kernel run after every sweep launch,     */
/* restated on the host.  Reference semantics: acceptance adaptation        */
/* (cuda/packing.cu:618), replica exchange on adjacent temperatures of one  */
/* ladder with log u < (1/Ta - 1/Tb)(Ea - Eb) (:141-152, :750-765), the     */
/* reheating cycle (:516-517, :756-758) and the improvement path of the     */
/* host loop (:503-504, :831-851).  The engine swaps temperature slots      */
/* instead of coordinates, tests u < 2^x with x in log2 units and 2^x the   */
/* FP32 sequence of o_accept_exp2, and adapts by 2^((rate - target)         */
/* log2(e) / 2) with the same sequence.                                     */
/* ------------------------------------------------------------------------ */
/* 2^x in FP64: k = rint(x), f = x - k, the Taylor polynomial of 2^f of     */
/* degree 13 (coefficients (ln 2)^j / j!) by fused multiply-adds, ldexp.    */
double o_exp2_fixed64(double x) {
    const double k = rint(x);
    const double f = x - k;
    double p = 1.3691488853904128e-12;
    p = fma(p, f, 2.5678435993488206e-11);
    p = fma(p, f, 4.4455382718708116e-10);
    p = fma(p, f, 7.054911620801123e-09);
    p = fma(p, f, 1.01780860092397e-07);
    p = fma(p, f, 1.321548679014431e-06);
    p = fma(p, f, 1.5252733804059841e-05);
    p = fma(p, f, 0.0001540353039338161);
    p = fma(p, f, 0.0013333558146428443);
    p = fma(p, f, 0.009618129107628477);
    p = fma(p, f, 0.05550410866482158);
    p = fma(p, f, 0.24022650695910072);
    p = fma(p, f, 0.6931471805599453);
    p = fma(p, f, 1.0);
    return ldexp(p, (int)k);
}

void o_cool_factor(uint32_t k, int P, double* cool, double* root) {
    if (P < 2) {
        *cool = 1.0;
        *root = 1.0;
        return;
    }
    /* log2(0.002) = -8.965784284662087 */
    const double x = (double)(k % (uint32_t)P) / (double)(P - 1) * -8.965784284662087;
    *cool = o_exp2_fixed64(x);
    *root = o_exp2_fixed64(0.5 * x);
}

int o_ladder_model(o_ladder_desc* d, o_rep_state* reps, o_slot_state* slots, int* ros, float* xycs,
                   float* arch) {
    static const double mults[8] = {0.02, 0.1, 0.4, 1.0, 2.5, 5.0, 10.0, 20.0};
    const int n = d->n, R = d->R, lad = d->ladder;
    const int64_t left = d->sweeps_total - d->sweeps_done;
    if (left <= 0) return 0;
    const int ran = (int)std::min<int64_t>(d->spe, left);
    std::vector<float> eS(R), kS(R);
    std::vector<int> rS(R);
    uint64_t best_lc = ~0ull;
    /* 1. adaptation of every slot from its replica's acceptance */
    for (int r = 0; r < R; ++r) {
        const o_rep_state& rs = reps[r];
        const float rate = (float)rs.acc / (float)((int64_t)n * ran);
        const float f = o_accept_exp2((rate - d->target_acc) * 0.72134752f);
        o_slot_state& sl = slots[rs.slot];
        eS[rs.slot] = rs.E;
        kS[rs.slot] = sl.inv_t_log2e;
        rS[rs.slot] = r;
        sl.base_xy = fminf(fmaxf(sl.base_xy * f, 1.7320508e-7f), 0.8660254f);
        sl.base_th = fminf(fmaxf(sl.base_th * f, 1.7320508e-7f), 0.7794229f);
        uint64_t bits;
        memcpy(&bits, &rs.Lc, 8);
        if (d->improve && bits < best_lc) best_lc = bits;
        /* replicas with a certified side count as checkpoints; the sweep
           workgroups' counters move into the instance's totals */
        if (d->improve && rs.Lc < 1e9) d->certs += 1;
        reps[r].cand = 0;
        reps[r].near_pairs = 0;
        reps[r].coll_tried = reps[r].coll_acc = reps[r].kicks = 0;
    }
    const double L0 = d->L, lmin = sqrt((double)n) + 1e-6;
    bool improved = false;
    /* 2. improvement path */
    if (d->improve) {
        double lc;
        memcpy(&lc, &best_lc, 8);
        improved = lc < L0 - 1e-11;
        if (improved) {
            int rb = 0;
            for (int r = 0; r < R; ++r) {
                uint64_t bits;
                memcpy(&bits, &reps[r].Lc, 8);
                if (bits == best_lc) { rb = r; break; }
            }
            const double Lb = reps[rb].Lc, lam = reps[rb].lam, mx = reps[rb].minx, my = reps[rb].miny;
            const float box_b = slots[reps[rb].slot].L;
            memcpy(arch, xycs + (size_t)4 * rb * n, sizeof(float) * 4 * n);
            const double shrink = (double)d->shrink;
            for (int s = 0; s < R; ++s)
                slots[s].L = (float)std::max(lmin, Lb - shrink * mults[((s % lad) * 8) / lad]);
            const int heads = R / lad, extra = (L0 - Lb > 5e-4) ? lad - 1 : 0;
            for (int h = 0; h < heads + extra; ++h) {
                const int s = h < heads ? h * lad : h - heads + 1;
                const double tL = std::max(lmin, Lb - shrink * mults[((s % lad) * 8) / lad]);
                const double sc = tL / Lb;
                float* dst = xycs + (size_t)4 * rS[s] * n;
                for (int i = 0; i < n; ++i) {
                    dst[4 * i] = (float)(((double)arch[4 * i] * lam - mx) * sc);
                    dst[4 * i + 1] = (float)(((double)arch[4 * i + 1] * lam - my) * sc);
                    dst[4 * i + 2] = arch[4 * i + 2];
                    dst[4 * i + 3] = arch[4 * i + 3];
                }
            }
            d->L = Lb;
            d->Lf = (float)Lb;
            d->L_arch = Lb;
            d->L_box_arch = (double)box_b;
            d->arch_valid = 1;
            d->feasible += 1;
            d->shrinks += 1;
        }
    }
    /* 3. even/odd exchange inside each ladder */
    if (!improved) {
        const int parity = (int)(d->interval & 1u);
        const uint32_t key[2] = {d->key0, d->key1};
        for (int s = parity; s + 1 < R; s += 2) {
            if ((s % lad) + 1 >= lad) continue;
            const int ra = rS[s], rb = rS[s + 1];
            const float Ea = eS[s], Eb = eS[s + 1];
            const float x = (kS[s] - kS[s + 1]) * (Ea - Eb);
            uint32_t ctr[4] = {(uint32_t)s, 0u, d->interval, 2u}, o[4];
            o_philox(ctr, key, o);
            d->ex_tried++;
            if (x >= 0.0f || o_u01f(o[0]) < o_accept_exp2(x)) {
                rS[s] = rb; rS[s + 1] = ra;
                eS[s] = Eb; eS[s + 1] = Ea;
                ros[s] = rb; ros[s + 1] = ra;
                reps[ra].slot = s + 1;
                reps[rb].slot = s;
                d->ex_acc++;
            }
        }
    }
    /* 4. improve = 0: FP64 zero-overlap test of the coldest zero-energy replica */
    double scale = 1.0, Ln = L0;
    bool archive = false;
    if (!d->improve) {
        int zero = -1;
        for (int s = 0; s < R && zero < 0; ++s)
            if (eS[s] == 0.0f) zero = s;
        if (zero >= 0) {
            const int r = rS[zero];
            const float* src = xycs + (size_t)4 * r * n;
            std::vector<double> x(n), y(n), c(n), s(n);
            for (int i = 0; i < n; ++i) {
                x[i] = src[4 * i];
                y[i] = src[4 * i + 1];
                unit64(src[4 * i + 2], src[4 * i + 3], c[i], s[i]);
            }
            if (o_total64_cs(n, x.data(), y.data(), c.data(), s.data(), L0) != 0.0) {
                d->fp64_rejects += 1;
            } else {
                archive = !d->arch_valid || L0 < d->L_arch;
                if (archive) memcpy(arch, src, sizeof(float) * 4 * n);
                Ln = std::max(L0 - (double)d->shrink, sqrt((double)n));
                if (Ln < L0) scale = Ln / L0;
            }
        }
        if (Ln < L0)
            for (int s = 0; s < R; ++s) slots[s].L = (float)Ln;
    }
    /* 5. next interval's temperatures and step sizes */
    double cool, root;
    o_cool_factor(d->interval + 1u, d->cool_period, &cool, &root);
    for (int s = 0; s < R; ++s) {
        o_slot_state& sl = slots[s];
        sl.inv_t_log2e = (float)(1.4426950408889634 / ((double)sl.T * cool));
        sl.step_xy = (float)((double)sl.base_xy * root);
        sl.step_th = (float)((double)sl.base_th * root);
    }
    if (archive) {
        d->arch_valid = 1;
        d->L_arch = L0;
        d->L_box_arch = L0;
        d->feasible += 1;
    }
    if (Ln < L0) {
        d->L = Ln;
        d->Lf = (float)Ln;
        d->shrinks += 1;
    }
    d->scale = (float)scale;
    d->sweeps_done += ran;
    d->interval += 1u;
    return improved ? 1 : 0;
}
