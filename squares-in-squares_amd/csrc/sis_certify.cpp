// Exact, conservative certificate of a packing given as decimal strings.
//
// Semantics follow the reference's certifier (cuda/certifier.cpp:738-912, the
// C++ twin of the official metric): the literal grammar and limits
// (<= 100 characters, |x|,|y| <= 1e4, |angle| <= 100), contact accepted and
// any interior overlap rejected, L = max(x-span, y-span) of the exact corners
// printed with 50 decimals (half-to-even), and rejection when L falls below
// sqrt(n) - 1e-7.  The arithmetic is this file's own: fixed point in base
// 1e18 limbs with directed rounding, interval trigonometry by Taylor series
// with alternating-series remainders, pi from Machin's formula
// (pi/4 = 4 atan(1/5) - atan(1/239)), and a precision ladder of 36, 72, 144,
// 288 and 576 fractional digits for pairs a double-precision test cannot
// decide.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <unordered_map>
#include <vector>

#include "sis_certify.h"

namespace sis {
namespace cert {

using u64 = uint64_t;
using u128 = unsigned __int128;
constexpr u64 kBase = 1000000000000000000ull;  // 1e18

// Magnitude with K fractional limbs: d[0] integer part (< 1e18), d[1..K] fraction.
template <int K>
struct Fx {
    int neg = 0;
    u64 d[K + 1] = {};
    bool zero() const {
        for (int i = 0; i <= K; ++i)
            if (d[i]) return false;
        return true;
    }
    double to_double() const {
        double v = 0, s = 1;
        for (int i = 0; i <= K && i < 3; ++i) {
            v += (double)d[i] * s;
            s *= 1e-18;
        }
        return neg ? -v : v;
    }
    static Fx from_int(long long v) {
        Fx r;
        r.neg = v < 0;
        r.d[0] = (u64)(v < 0 ? -v : v);
        return r;
    }
};

template <int K>
int cmp_mag(const Fx<K>& a, const Fx<K>& b) {
    for (int i = 0; i <= K; ++i)
        if (a.d[i] != b.d[i]) return a.d[i] < b.d[i] ? -1 : 1;
    return 0;
}
template <int K>
int cmp(const Fx<K>& a, const Fx<K>& b) {
    const bool az = a.zero(), bz = b.zero();
    const int sa = az ? 0 : (a.neg ? -1 : 1), sb = bz ? 0 : (b.neg ? -1 : 1);
    if (sa != sb) return sa < sb ? -1 : 1;
    if (sa == 0) return 0;
    const int m = cmp_mag(a, b);
    return sa > 0 ? m : -m;
}
template <int K>
void add_mag(const Fx<K>& a, const Fx<K>& b, Fx<K>& r) {
    u64 carry = 0;
    for (int i = K; i >= 0; --i) {
        u64 s = a.d[i] + b.d[i] + carry;
        if (i > 0 && s >= kBase) {
            s -= kBase;
            carry = 1;
        } else {
            carry = 0;
        }
        r.d[i] = s;
    }
}
template <int K>
void sub_mag(const Fx<K>& a, const Fx<K>& b, Fx<K>& r) {  // |a| >= |b|
    u64 borrow = 0;
    for (int i = K; i >= 0; --i) {
        u64 bi = b.d[i] + borrow;
        if (a.d[i] >= bi) {
            r.d[i] = a.d[i] - bi;
            borrow = 0;
        } else {
            r.d[i] = a.d[i] + kBase - bi;
            borrow = 1;
        }
    }
}
// Exact signed sum / difference (fixed point addition never rounds).
template <int K>
Fx<K> add(const Fx<K>& a, const Fx<K>& b) {
    Fx<K> r;
    if (a.neg == b.neg) {
        add_mag(a, b, r);
        r.neg = a.neg;
    } else if (cmp_mag(a, b) >= 0) {
        sub_mag(a, b, r);
        r.neg = a.neg;
    } else {
        sub_mag(b, a, r);
        r.neg = b.neg;
    }
    if (r.zero()) r.neg = 0;
    return r;
}
template <int K>
Fx<K> neg(Fx<K> a) {
    if (!a.zero()) a.neg ^= 1;
    return a;
}
template <int K>
Fx<K> sub(const Fx<K>& a, const Fx<K>& b) { return add(a, neg(b)); }

// Add one unit of the last place to the magnitude.
template <int K>
void bump(Fx<K>& a) {
    for (int i = K; i >= 0; --i) {
        if (++a.d[i] < kBase || i == 0) return;
        a.d[i] = 0;
    }
}

// Product rounded toward +inf (up = 1) or -inf (up = 0).
template <int K>
Fx<K> mul(const Fx<K>& a, const Fx<K>& b, int up) {
    u128 acc[2 * K + 2] = {};
    for (int i = 0; i <= K; ++i) {
        if (!a.d[i]) continue;
        for (int j = 0; j <= K; ++j) acc[i + j] += (u128)a.d[i] * b.d[j];
    }
    for (int k = 2 * K; k > 0; --k) {
        acc[k - 1] += acc[k] / kBase;
        acc[k] %= kBase;
    }
    Fx<K> r;
    bool tail = false;
    for (int k = K + 1; k <= 2 * K; ++k)
        if (acc[k]) tail = true;
    for (int k = 0; k <= K; ++k) r.d[k] = (u64)acc[k];
    r.neg = a.neg ^ b.neg;
    // magnitude truncated: rounds toward zero; move away when the direction asks
    if (tail && (r.neg ? !up : up)) bump(r);
    if (r.zero()) r.neg = 0;
    return r;
}

// Quotient by a small positive integer, rounded toward +inf / -inf.
template <int K>
Fx<K> div_small(const Fx<K>& a, u64 m, int up) {
    Fx<K> r;
    u128 rem = 0;
    for (int i = 0; i <= K; ++i) {
        u128 cur = rem * kBase + a.d[i];
        r.d[i] = (u64)(cur / m);
        rem = cur % m;
    }
    r.neg = a.neg;
    if (rem && (r.neg ? !up : up)) bump(r);
    if (r.zero()) r.neg = 0;
    return r;
}

template <int K>
struct Iv {
    Fx<K> lo, hi;
    static Iv point(const Fx<K>& v) { return Iv{v, v}; }
};
template <int K>
Iv<K> iadd(const Iv<K>& a, const Iv<K>& b) { return {add(a.lo, b.lo), add(a.hi, b.hi)}; }
template <int K>
Iv<K> isub(const Iv<K>& a, const Iv<K>& b) { return {sub(a.lo, b.hi), sub(a.hi, b.lo)}; }
template <int K>
Iv<K> ineg(const Iv<K>& a) { return {neg(a.hi), neg(a.lo)}; }
template <int K>
Iv<K> imul(const Iv<K>& a, const Iv<K>& b) {
    const Fx<K>* xs[2] = {&a.lo, &a.hi};
    const Fx<K>* ys[2] = {&b.lo, &b.hi};
    Iv<K> r;
    bool first = true;
    for (auto x : xs)
        for (auto y : ys) {
            Fx<K> l = mul(*x, *y, 0), h = mul(*x, *y, 1);
            if (first || cmp(l, r.lo) < 0) r.lo = l;
            if (first || cmp(h, r.hi) > 0) r.hi = h;
            first = false;
        }
    return r;
}
template <int K>
Iv<K> idiv_small(const Iv<K>& a, u64 m) { return {div_small(a.lo, m, 0), div_small(a.hi, m, 1)}; }
template <int K>
Iv<K> ihull(const Iv<K>& a, const Iv<K>& b) {
    return {cmp(a.lo, b.lo) <= 0 ? a.lo : b.lo, cmp(a.hi, b.hi) >= 0 ? a.hi : b.hi};
}

// ---------------------------------------------------------------------------
// Decimal literals.
// ---------------------------------------------------------------------------
// Grammar of cuda/certifier.cpp:505-528: [+-] digits [. digits] [e|E [+-] digits],
// at least one mantissa digit, surrounding whitespace ignored, <= 100 chars.
bool valid_literal(const std::string& raw, std::string& s) {
    size_t a = 0, b = raw.size();
    while (a < b && isspace((unsigned char)raw[a])) ++a;
    while (b > a && isspace((unsigned char)raw[b - 1])) --b;
    s = raw.substr(a, b - a);
    if (s.empty() || s.size() > 100) return false;
    size_t i = 0;
    const size_t n = s.size();
    if (s[i] == '+' || s[i] == '-') ++i;
    bool digits = false;
    while (i < n && isdigit((unsigned char)s[i])) { digits = true; ++i; }
    if (i < n && s[i] == '.') {
        ++i;
        while (i < n && isdigit((unsigned char)s[i])) { digits = true; ++i; }
    }
    if (!digits) return false;
    if (i < n && (s[i] == 'e' || s[i] == 'E')) {
        ++i;
        if (i < n && (s[i] == '+' || s[i] == '-')) ++i;
        bool ed = false;
        while (i < n && isdigit((unsigned char)s[i])) { ed = true; ++i; }
        if (!ed) return false;
    }
    return i == n;
}

// Exact digit string and decimal exponent: value = +-0.DIGITS * 10^point.
struct Decimal {
    int neg = 0;
    std::string digits;  // no leading zeros
    long long point = 0; // digits before the decimal point
};

Decimal to_decimal(const std::string& s) {
    Decimal d;
    size_t i = 0;
    if (s[i] == '+' || s[i] == '-') d.neg = s[i++] == '-';
    std::string mant;
    long long frac = 0;
    bool dot = false;
    for (; i < s.size() && s[i] != 'e' && s[i] != 'E'; ++i) {
        if (s[i] == '.') { dot = true; continue; }
        mant.push_back(s[i]);
        if (dot) ++frac;
    }
    long long ex = 0;
    if (i < s.size()) {
        ++i;
        int es = 1;
        if (s[i] == '+' || s[i] == '-') es = s[i++] == '-' ? -1 : 1;
        for (; i < s.size(); ++i) ex = std::min<long long>(ex * 10 + (s[i] - '0'), 1000000);
        ex *= es;
    }
    size_t lead = 0;
    while (lead < mant.size() && mant[lead] == '0') ++lead;
    d.digits = mant.substr(lead);
    d.point = (long long)mant.size() - lead - frac + ex;
    while (!d.digits.empty() && d.digits.back() == '0') d.digits.pop_back();
    if (d.digits.empty()) d.neg = 0;
    return d;
}

// Interval enclosing the decimal (a point when it fits K limbs exactly).
template <int K>
Iv<K> to_iv(const Decimal& dec) {
    Fx<K> v;
    bool rest = false;
    const long long nd = (long long)dec.digits.size();
    for (long long k = 0; k < nd; ++k) {
        const long long pos = dec.point - 1 - k;  // power of ten of this digit
        const int dig = dec.digits[(size_t)k] - '0';
        if (pos >= 0) {
            if (pos >= 18) return {Fx<K>::from_int(-1), Fx<K>::from_int(-1)};  // out of range, caller rejects
            static const u64 p10[19] = {1ull, 10ull, 100ull, 1000ull, 10000ull, 100000ull, 1000000ull,
                                        10000000ull, 100000000ull, 1000000000ull, 10000000000ull,
                                        100000000000ull, 1000000000000ull, 10000000000000ull,
                                        100000000000000ull, 1000000000000000ull, 10000000000000000ull,
                                        100000000000000000ull, 1000000000000000000ull};
            v.d[0] += (u64)dig * p10[pos];
        } else {
            const long long f = -pos - 1;  // 0-based fractional digit
            const long long limb = f / 18 + 1;
            if (limb > K) {
                if (dig) rest = true;
                continue;
            }
            u64 scale = 1;
            for (long long t = 0; t < 17 - f % 18; ++t) scale *= 10;
            v.d[limb] += (u64)dig * scale;
        }
    }
    v.neg = dec.neg;
    if (v.zero()) v.neg = 0;
    if (!rest) return Iv<K>::point(v);
    Fx<K> lo = v, hi = v;  // truncated magnitude: widen by one unit away from zero
    if (dec.neg) bump(lo); else bump(hi);
    if (lo.zero()) lo.neg = 0;
    return {lo, hi};
}

// Half-to-even at the 50th decimal (18 + 18 + 14 digits); needs K >= 4.
template <int K>
std::string format50(const Fx<K>& v) {
    static_assert(K >= 4, "need 72 digits");
    u64 d0 = v.d[0], d1 = v.d[1], d2 = v.d[2], d3 = v.d[3] / 10000ull;
    const u64 tail4 = v.d[3] % 10000ull;
    bool beyond = false;
    for (int i = 4; i <= K; ++i)
        if (v.d[i]) beyond = true;
    const bool up = tail4 > 5000 || (tail4 == 5000 && (beyond || (d3 & 1)));
    if (up && ++d3 == 100000000000000ull) {
        d3 = 0;
        if (++d2 == kBase) {
            d2 = 0;
            if (++d1 == kBase) {
                d1 = 0;
                ++d0;
            }
        }
    }
    char buf[96];
    const bool z = !d0 && !d1 && !d2 && !d3;
    std::snprintf(buf, sizeof buf, "%s%llu.%018llu%018llu%014llu", (v.neg && !z) ? "-" : "",
                  (unsigned long long)d0, (unsigned long long)d1, (unsigned long long)d2,
                  (unsigned long long)d3);
    return buf;
}

// ---------------------------------------------------------------------------
// Trigonometry.
// ---------------------------------------------------------------------------
// atan(1/k) = sum (-1)^m / ((2m+1) k^(2m+1)), alternating with decreasing terms.
template <int K>
Iv<K> atan_inv(u64 k) {
    Iv<K> power = idiv_small(Iv<K>::point(Fx<K>::from_int(1)), k);  // 1/k^(2m+1)
    Iv<K> sum = power;
    const u64 k2 = k * k;
    for (u64 m = 1;; ++m) {
        power = idiv_small(power, k2);
        Iv<K> term = idiv_small(power, 2 * m + 1);
        if (power.hi.zero() || term.hi.zero()) {
            Fx<K> r = term.hi;
            bump(r);
            sum = {sub(sum.lo, r), add(sum.hi, r)};
            break;
        }
        sum = (m & 1) ? isub(sum, term) : iadd(sum, term);
        if (m > 4000) break;
    }
    return sum;
}

template <int K>
const Iv<K>& half_pi() {
    static const Iv<K> v = [] {
        Iv<K> a5 = atan_inv<K>(5), a239 = atan_inv<K>(239);
        Iv<K> quarter = isub(imul(Iv<K>::point(Fx<K>::from_int(4)), a5), a239);  // pi/4
        return imul(Iv<K>::point(Fx<K>::from_int(2)), quarter);                    // pi/2
    }();
    return v;
}

// sin/cos of an interval angle (width w): values at the lower end widened by w
// (both are 1-Lipschitz).  Reduction by the nearest multiple of pi/2, then
// Taylor series with the first omitted term as the remainder bound.
template <int K>
void sincos(const Iv<K>& ang, Iv<K>& s, Iv<K>& c) {
    if (ang.lo.zero() && ang.hi.zero()) {
        s = Iv<K>::point(Fx<K>());
        c = Iv<K>::point(Fx<K>::from_int(1));
        return;
    }
    const Iv<K>& hp = half_pi<K>();
    const double mid = 0.5 * (ang.lo.to_double() + ang.hi.to_double());
    const long long q = (long long)std::llround(mid / 1.5707963267948966);
    Iv<K> x = Iv<K>::point(ang.lo);
    Iv<K> r = isub(x, imul(Iv<K>::point(Fx<K>::from_int(q)), hp));
    Iv<K> r2 = imul(r, r);
    Iv<K> ssum = r, csum = Iv<K>::point(Fx<K>::from_int(1));
    Iv<K> sterm = r, cterm = csum;
    // a term is negligible once both ends are within 4 units of the last place
    auto tiny = [](const Iv<K>& t) {
        for (const Fx<K>* e : {&t.lo, &t.hi}) {
            for (int i = 0; i < K; ++i)
                if (e->d[i]) return false;
            if (e->d[K] > 4) return false;
        }
        return true;
    };
    for (u64 m = 1; m < 400; ++m) {
        cterm = idiv_small(imul(cterm, r2), (2 * m - 1) * (2 * m));
        sterm = idiv_small(imul(sterm, r2), (2 * m) * (2 * m + 1));
        csum = (m & 1) ? isub(csum, cterm) : iadd(csum, cterm);
        ssum = (m & 1) ? isub(ssum, sterm) : iadd(ssum, sterm);
        if (tiny(cterm) && tiny(sterm)) break;
    }
    // remainder of an alternating series with decreasing terms (|r| <= pi/4 < 1):
    // at most the first omitted term, itself below the last included one
    auto widen = [](Iv<K>& v) {
        Fx<K> u;
        u.d[K] = 8;
        v = {sub(v.lo, u), add(v.hi, u)};
    };
    widen(csum);
    widen(ssum);
    Iv<K> sr, cr;
    switch (((q % 4) + 4) % 4) {
        case 0: sr = ssum; cr = csum; break;
        case 1: sr = csum; cr = ineg(ssum); break;
        case 2: sr = ineg(ssum); cr = ineg(csum); break;
        default: sr = ineg(csum); cr = ssum; break;
    }
    Fx<K> w = sub(ang.hi, ang.lo);
    s = {sub(sr.lo, w), add(sr.hi, w)};
    c = {sub(cr.lo, w), add(cr.hi, w)};
}

// ---------------------------------------------------------------------------
// Squares.
// ---------------------------------------------------------------------------
struct Sq {
    Decimal x, y, a;
    std::string a_str;
    double xd, yd, ad, c, s;
    double box[4];  // xmin, ymin, xmax, ymax
};

template <int K>
struct SqK {
    Iv<K> x, y, c, s;
    Iv<K> cx[4], cy[4];
};

template <int K>
void corners(SqK<K>& q) {
    Iv<K> hc = idiv_small(q.c, 2), hs = idiv_small(q.s, 2);
    Iv<K> ax = isub(hc, hs), ay = iadd(hs, hc);      // (c - s, s + c) / 2
    Iv<K> bx = ineg(iadd(hc, hs)), by = isub(hc, hs); // (-(c + s), c - s) / 2
    q.cx[0] = iadd(q.x, ax); q.cy[0] = iadd(q.y, ay);
    q.cx[1] = iadd(q.x, bx); q.cy[1] = iadd(q.y, by);
    q.cx[2] = isub(q.x, ax); q.cy[2] = isub(q.y, ay);
    q.cx[3] = isub(q.x, bx); q.cy[3] = isub(q.y, by);
}

template <int K>
void build(const std::vector<Sq>& sq, const std::vector<int>& ids, std::unordered_map<int, SqK<K>>& out) {
    std::unordered_map<std::string, std::pair<Iv<K>, Iv<K>>> trig;
    for (int i : ids) {
        if (out.count(i)) continue;
        SqK<K> q;
        q.x = to_iv<K>(sq[i].x);
        q.y = to_iv<K>(sq[i].y);
        auto it = trig.find(sq[i].a_str);
        if (it == trig.end()) {
            Iv<K> s, c;
            sincos<K>(to_iv<K>(sq[i].a), s, c);
            it = trig.emplace(sq[i].a_str, std::make_pair(s, c)).first;
        }
        q.s = it->second.first;
        q.c = it->second.second;
        corners(q);
        out.emplace(i, q);
    }
}

// 1 separated (touching allowed), -1 overlapping, 0 undecided at this precision.
template <int K>
int sat_iv(const SqK<K>& A, const SqK<K>& B) {
    const Iv<K>* axes[4][2] = {{&A.c, &A.s}, {nullptr, nullptr}, {&B.c, &B.s}, {nullptr, nullptr}};
    Iv<K> nA[2] = {ineg(A.s), A.c}, nB[2] = {ineg(B.s), B.c};
    axes[1][0] = &nA[0]; axes[1][1] = &nA[1];
    axes[3][0] = &nB[0]; axes[3][1] = &nB[1];
    bool all_overlap = true;
    for (auto& ax : axes) {
        Iv<K> pa[4], pb[4];
        for (int k = 0; k < 4; ++k) {
            pa[k] = iadd(imul(A.cx[k], *ax[0]), imul(A.cy[k], *ax[1]));
            pb[k] = iadd(imul(B.cx[k], *ax[0]), imul(B.cy[k], *ax[1]));
        }
        Iv<K> amin = pa[0], amax = pa[0], bmin = pb[0], bmax = pb[0];
        for (int k = 1; k < 4; ++k) {
            amin = {cmp(pa[k].lo, amin.lo) < 0 ? pa[k].lo : amin.lo, cmp(pa[k].hi, amin.hi) < 0 ? pa[k].hi : amin.hi};
            amax = {cmp(pa[k].lo, amax.lo) > 0 ? pa[k].lo : amax.lo, cmp(pa[k].hi, amax.hi) > 0 ? pa[k].hi : amax.hi};
            bmin = {cmp(pb[k].lo, bmin.lo) < 0 ? pb[k].lo : bmin.lo, cmp(pb[k].hi, bmin.hi) < 0 ? pb[k].hi : bmin.hi};
            bmax = {cmp(pb[k].lo, bmax.lo) > 0 ? pb[k].lo : bmax.lo, cmp(pb[k].hi, bmax.hi) > 0 ? pb[k].hi : bmax.hi};
        }
        if (cmp(amax.hi, bmin.lo) <= 0 || cmp(bmax.hi, amin.lo) <= 0) return 1;
        const bool overlap = cmp(amax.lo, bmin.hi) > 0 && cmp(bmax.lo, amin.hi) > 0;
        if (!overlap) all_overlap = false;
    }
    return all_overlap ? -1 : 0;
}

template <int K>
int decide_pairs(const std::vector<Sq>& sq, std::vector<std::pair<int, int>>& pairs) {
    std::vector<int> ids;
    for (auto& p : pairs) { ids.push_back(p.first); ids.push_back(p.second); }
    std::unordered_map<int, SqK<K>> cache;
    build<K>(sq, ids, cache);
    std::vector<std::pair<int, int>> left;
    for (auto& p : pairs) {
        int r = sat_iv<K>(cache[p.first], cache[p.second]);
        if (r < 0) return -1;
        if (r == 0) left.push_back(p);
    }
    pairs.swap(left);
    return 0;
}

// Double-precision SAT on corner projections: separation on the best axis.
double sep_double(const Sq& a, const Sq& b) {
    const double ax[4][2] = {{a.c, a.s}, {-a.s, a.c}, {b.c, b.s}, {-b.s, b.c}};
    auto corners_d = [](const Sq& q, double cx[4], double cy[4]) {
        const double hc = 0.5 * q.c, hs = 0.5 * q.s;
        const double px = hc - hs, py = hs + hc, qx = -(hc + hs), qy = hc - hs;
        cx[0] = q.xd + px; cy[0] = q.yd + py;
        cx[1] = q.xd + qx; cy[1] = q.yd + qy;
        cx[2] = q.xd - px; cy[2] = q.yd - py;
        cx[3] = q.xd - qx; cy[3] = q.yd - qy;
    };
    double acx[4], acy[4], bcx[4], bcy[4];
    corners_d(a, acx, acy);
    corners_d(b, bcx, bcy);
    double best = -1e300;
    for (auto& u : ax) {
        double amin = 1e300, amax = -1e300, bmin = 1e300, bmax = -1e300;
        for (int k = 0; k < 4; ++k) {
            const double pa = acx[k] * u[0] + acy[k] * u[1], pb = bcx[k] * u[0] + bcy[k] * u[1];
            amin = std::min(amin, pa); amax = std::max(amax, pa);
            bmin = std::min(bmin, pb); bmax = std::max(bmax, pb);
        }
        best = std::max(best, std::max(bmin - amax, amin - bmax));
    }
    return best;
}

template <int K>
bool side_at(const std::vector<Sq>& sq, std::string& L50, Fx<K>& Llo) {
    std::vector<int> ids(sq.size());
    for (size_t i = 0; i < sq.size(); ++i) ids[i] = (int)i;
    std::unordered_map<int, SqK<K>> all;
    build<K>(sq, ids, all);
    Iv<K> xlo, xhi, ylo, yhi;
    bool first = true;
    for (auto& kv : all)
        for (int k = 0; k < 4; ++k) {
            const Iv<K>& cx = kv.second.cx[k];
            const Iv<K>& cy = kv.second.cy[k];
            if (first) {
                xlo = xhi = cx;
                ylo = yhi = cy;
                first = false;
                continue;
            }
            xlo = {cmp(cx.lo, xlo.lo) < 0 ? cx.lo : xlo.lo, cmp(cx.hi, xlo.hi) < 0 ? cx.hi : xlo.hi};
            xhi = {cmp(cx.lo, xhi.lo) > 0 ? cx.lo : xhi.lo, cmp(cx.hi, xhi.hi) > 0 ? cx.hi : xhi.hi};
            ylo = {cmp(cy.lo, ylo.lo) < 0 ? cy.lo : ylo.lo, cmp(cy.hi, ylo.hi) < 0 ? cy.hi : ylo.hi};
            yhi = {cmp(cy.lo, yhi.lo) > 0 ? cy.lo : yhi.lo, cmp(cy.hi, yhi.hi) > 0 ? cy.hi : yhi.hi};
        }
    Iv<K> sx = isub(xhi, xlo), sy = isub(yhi, ylo);
    Iv<K> L = {cmp(sx.lo, sy.lo) >= 0 ? sx.lo : sy.lo, cmp(sx.hi, sy.hi) >= 0 ? sx.hi : sy.hi};
    const std::string a = format50(L.lo), b = format50(L.hi);
    Llo = L.lo;
    if (a != b) return false;
    L50 = a;
    return true;
}

}  // namespace cert

int certify_packing(const std::vector<std::string>& xs, const std::vector<std::string>& ys,
                    const std::vector<std::string>& as, std::string& L50, std::string& reason) {
    using namespace cert;
    const int n = (int)xs.size();
    if (n < 1 || n > 1000 || (int)ys.size() != n || (int)as.size() != n) {
        reason = "bad square count";
        return 2;
    }
    std::vector<Sq> sq(n);
    for (int i = 0; i < n; ++i) {
        std::string x, y, a;
        if (!valid_literal(xs[i], x) || !valid_literal(ys[i], y) || !valid_literal(as[i], a)) {
            reason = "square " + std::to_string(i) + ": cannot parse coordinates";
            return 2;
        }
        Sq& q = sq[i];
        q.xd = std::strtod(x.c_str(), nullptr);
        q.yd = std::strtod(y.c_str(), nullptr);
        q.ad = std::strtod(a.c_str(), nullptr);
        if (!std::isfinite(q.xd) || !std::isfinite(q.yd) || !std::isfinite(q.ad)) {
            reason = "square " + std::to_string(i) + ": value out of range for a double";
            return 2;
        }
        if (std::fabs(q.xd) > 1e4 || std::fabs(q.yd) > 1e4) {
            reason = "square " + std::to_string(i) + ": coordinate out of bounds";
            return 2;
        }
        if (std::fabs(q.ad) > 100.0) {
            reason = "square " + std::to_string(i) + ": angle out of bounds";
            return 2;
        }
        q.x = to_decimal(x);
        q.y = to_decimal(y);
        q.a = to_decimal(a);
        q.a_str = a;
        q.c = std::cos(q.ad);
        q.s = std::sin(q.ad);
        const double e = 0.5 * (std::fabs(q.c) + std::fabs(q.s)) + 1e-7;
        q.box[0] = q.xd - e; q.box[1] = q.yd - e; q.box[2] = q.xd + e; q.box[3] = q.yd + e;
    }
    // broad phase: sort and sweep on padded boxes
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i) order[i] = i;
    std::sort(order.begin(), order.end(), [&](int a, int b) { return sq[a].box[0] < sq[b].box[0]; });
    std::vector<std::pair<int, int>> undecided;
    for (int ii = 0; ii < n; ++ii) {
        const Sq& A = sq[order[ii]];
        for (int jj = ii + 1; jj < n; ++jj) {
            const Sq& B = sq[order[jj]];
            if (B.box[0] > A.box[2]) break;
            if (B.box[1] > A.box[3] || A.box[1] > B.box[3]) continue;
            const double s = sep_double(A, B);
            if (s > 1e-8) continue;
            if (s < -1e-8) {
                reason = "squares " + std::to_string(std::min(order[ii], order[jj])) + " and " +
                         std::to_string(std::max(order[ii], order[jj])) + " overlap";
                return 2;
            }
            undecided.emplace_back(order[ii], order[jj]);
        }
    }
    if (!undecided.empty()) {
        int r = decide_pairs<2>(sq, undecided);
        if (!r && !undecided.empty()) r = decide_pairs<4>(sq, undecided);
        if (!r && !undecided.empty()) r = decide_pairs<8>(sq, undecided);
        if (!r && !undecided.empty()) r = decide_pairs<16>(sq, undecided);
        if (!r && !undecided.empty()) r = decide_pairs<32>(sq, undecided);
        if (r < 0) {
            reason = "squares overlap (certified)";
            return 2;
        }
        if (!undecided.empty()) {
            reason = "overlap undecidable at maximum precision";
            return 2;
        }
    }
    Fx<5> lo5;
    bool ok = side_at<5>(sq, L50, lo5);
    if (!ok) {
        Fx<10> lo10;
        ok = side_at<10>(sq, L50, lo10);
        if (!ok) {
            Fx<32> lo32;
            ok = side_at<32>(sq, L50, lo32);
        }
        if (!ok) {
            reason = "side not resolved to 50 decimals";
            return 2;
        }
    }
    const double Ld = std::strtod(L50.c_str(), nullptr);
    if (Ld < std::sqrt((double)n) - 1e-7) {
        reason = "L is below the area lower bound";
        return 2;
    }
    reason = "ok";
    return 0;
}

std::string shortest_repr(double v) {
    char buf[64];
    for (int p = 1; p <= 17; ++p) {
        std::snprintf(buf, sizeof buf, "%.*g", p, v);
        if (std::strtod(buf, nullptr) == v) break;
    }
    std::string s(buf);
    if (s.find_first_of(".eEn") == std::string::npos) s += ".0";
    return s;
}

int export_certify(int n, const double* q, double* dilation, std::string& L50, std::string& reason) {
    static const double ladder[3] = {1.000002, 1.00001, 1.0001};
    for (int i = 0; i < 3 * n; ++i)
        if (!std::isfinite(q[i])) {
            reason = "non-finite coordinate";
            return 2;
        }
    std::vector<std::string> xs(n), ys(n), as(n), rx(n), ry(n), ra(n);
    char buf[64];
    for (double d : ladder) {
        for (int i = 0; i < n; ++i) {
            std::snprintf(buf, sizeof buf, "%.17g", q[3 * i] * d);
            xs[i] = buf;
            std::snprintf(buf, sizeof buf, "%.17g", q[3 * i + 1] * d);
            ys[i] = buf;
            std::snprintf(buf, sizeof buf, "%.17g", q[3 * i + 2]);
            as[i] = buf;
            rx[i] = shortest_repr(std::strtod(xs[i].c_str(), nullptr));
            ry[i] = shortest_repr(std::strtod(ys[i].c_str(), nullptr));
            ra[i] = shortest_repr(std::strtod(as[i].c_str(), nullptr));
        }
        std::string L, Lr, why;
        if (certify_packing(xs, ys, as, L, why) != 0) {
            reason = why;
            continue;
        }
        if (certify_packing(rx, ry, ra, Lr, why) != 0) {
            reason = "round trip: " + why;
            continue;
        }
        if (dilation) *dilation = d;
        L50 = L;
        reason = "ok";
        return 0;
    }
    return 2;
}

}  // namespace sis

// ---------------------------------------------------------------------------
// C ABI.
// ---------------------------------------------------------------------------
#include "../../include/sis_abi.h"

namespace sis {
int fail(int code, const std::string& msg);
}

extern "C" int sis_certify_strings(int n, const char* const* xs, const char* const* ys,
                                   const char* const* angles, char* L50, int L50_len) {
    if (n < 1 || n > 1000 || !xs || !ys || !angles) return sis::fail(SIS_EINVAL, "bad arguments");
    std::vector<std::string> x(n), y(n), a(n);
    for (int i = 0; i < n; ++i) {
        if (!xs[i] || !ys[i] || !angles[i]) return sis::fail(SIS_EINVAL, "null string");
        x[i] = xs[i];
        y[i] = ys[i];
        a[i] = angles[i];
    }
    std::string L, why;
    if (sis::certify_packing(x, y, a, L, why) != 0) return sis::fail(SIS_EINVAL, why);
    if (L50 && L50_len > 0) {
        std::strncpy(L50, L.c_str(), (size_t)L50_len - 1);
        L50[L50_len - 1] = 0;
    }
    return SIS_OK;
}

extern "C" int sis_export_certify(int n, const double* q, double* dilation, char* L50, int L50_len) {
    if (n < 1 || n > 1000 || !q) return sis::fail(SIS_EINVAL, "bad arguments");
    std::string L, why;
    if (sis::export_certify(n, q, dilation, L, why) != 0) return sis::fail(SIS_EINVAL, "rejected: " + why);
    if (L50 && L50_len > 0) {
        std::strncpy(L50, L.c_str(), (size_t)L50_len - 1);
        L50[L50_len - 1] = 0;
    }
    return SIS_OK;
}
