// gfx950 kernels of the PT-SA square-packing path.
//
//  k_sweep     FP32 Metropolis sweeps of replica tiles staged in LDS, colour
//              checkerboard of cells so that non-interacting squares move at
//              once (reference semantics: cuda/packing.cu:52-95, `walk`).
//              Optionally ends with every replica's FP64 certified side.
//  k_ladder    per instance: step adaptation, replica exchange by swapping
//              temperature slots, then either the FP64 zero-overlap test and
//              box shrink or the certified-side improvement path, and the
//              reheating cycle (cuda/packing.cu:141-152, 503-517, 750-765,
//              823-853).
//  k_finalize  FP64 conjugate-gradient polish, certified-side dilation and
//              export of one configuration per instance (cuda/packing.cu:
//              254-297, 624-697).
//  k_depth32 / k_cellcoord   parity probes of the geometry on caller data.
//  k_accept_exp2             parity probe of the Metropolis test's 2^x.
//
// All arithmetic whose bits are checked against the CPU oracle is written
// with explicit fmaf and compiled with contraction off.
#pragma clang fp contract(off)

#include "sis_internal.h"
#include "sis_kernels.h"

namespace sis {

// ---------------------------------------------------------------------------
// Counter-based RNG: Philox4x32-10 (Salmon et al., SC'11).
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint4 philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                        uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        // full 64-bit products: one v_mad_u64_u32 each instead of a low and
        // a high multiply
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0, n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c0 = n0; c1 = (uint32_t)p1; c2 = n2; c3 = (uint32_t)p0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return make_uint4(c0, c1, c2, c3);
}
__device__ __forceinline__ float u01(uint32_t r) {
    return ((float)(r >> 9) + 0.5f) * (1.0f / 8388608.0f);
}
// 2^x for the Metropolis test (x <= 0), in FP32 operations that each round
// once or not at all (round to nearest integer, an exact subtraction, fused
// multiply-adds, an exact power-of-two scaling), so that any IEEE FP32 machine
// with a fused multiply-add gives the same bits.  k = rint(x), f = x - k in
// [-1/2, 1/2], and p(f) interpolates 2^f at the 7 Chebyshev nodes of that
// interval (relative error below 1e-7).  Below x = -100 the value is 0: no
// accept uniform is smaller than 2^-24.
__device__ __forceinline__ float accept_exp2(float x) {
    if (!(x >= -100.0f)) return 0.0f;
    const float k = rintf(x);
    const float f = x - k;
    float p = 0x1.444p-13f;
    p = __builtin_fmaf(p, f, 0x1.5f48cp-10f);
    p = __builtin_fmaf(p, f, 0x1.3b2a1cp-7f);
    p = __builtin_fmaf(p, f, 0x1.c6aeccp-5f);
    p = __builtin_fmaf(p, f, 0x1.ebfbep-3f);
    p = __builtin_fmaf(p, f, 0x1.62e43p-1f);
    p = __builtin_fmaf(p, f, 1.0f);
    return ldexpf(p, (int)k);
}
// Move variates of square i of replica `rep` in sweep `sweep`: fx, fy, ftheta
// uniform in [-1, 1) and the accept uniform (Philox counter: square, replica,
// sweep, 0).
__device__ __forceinline__ float4 move_variates(uint32_t i, uint32_t rep, uint32_t sweep, uint32_t k0,
                                                uint32_t k1) {
    const uint4 o = philox(i, rep, sweep, 0u, k0, k1);
    return make_float4(__builtin_fmaf(2.0f, u01(o.x), -1.0f), __builtin_fmaf(2.0f, u01(o.y), -1.0f),
                       __builtin_fmaf(2.0f, u01(o.z), -1.0f), u01(o.w));
}

// ---------------------------------------------------------------------------
// FP32 geometry.  (dx, dy) = centre j - centre i; (c, s) = cos/sin of angles.
// ---------------------------------------------------------------------------
// Cell coordinate clamp(floor((x + shift) * (1/w)), lo, hi).  The reciprocal
// is the correctly rounded 1/w; for w = 2 it is exactly 0.5, which makes this
// the reference's floor(x / 2) bit for bit.
__device__ __forceinline__ int cellcoord(float x, float shift, float inv_w, int lo, int hi) {
    int k = (int)floorf((x + shift) * inv_w);
    return k < lo ? lo : (k > hi ? hi : k);
}

__device__ __forceinline__ float sat_gap(float dx, float dy, float ci, float si, float cj,
                                         float sj) {
    const float cd = __builtin_fmaf(ci, cj, si * sj);
    const float sd = __builtin_fmaf(si, cj, -(ci * sj));
    const float R = __builtin_fmaf(0.5f, fabsf(cd) + fabsf(sd), 0.5f);
    const float a1 = fabsf(__builtin_fmaf(dx, ci, dy * si));
    const float a2 = fabsf(__builtin_fmaf(dy, ci, -(dx * si)));
    const float a3 = fabsf(__builtin_fmaf(dx, cj, dy * sj));
    const float a4 = fabsf(__builtin_fmaf(dy, cj, -(dx * sj)));
    const float m = fmaxf(fmaxf(a1, a2), fmaxf(a3, a4));
    return R - m;
}

__device__ __forceinline__ float depth2(float dx, float dy, float ci, float si, float cj,
                                        float sj) {
    const float d2 = __builtin_fmaf(dx, dx, dy * dy);
    const float g = sat_gap(dx, dy, ci, si, cj, sj);
    return (d2 < 2.0f && g > 0.0f) ? g * g : 0.0f;
}

__device__ __forceinline__ float wall32(float x, float y, float c, float s, float L) {
    const float ext = 0.5f * (fabsf(c) + fabsf(s));
    const float h = 0.5f * L;
    const float ox = fmaxf(0.0f, (fabsf(x - h) + ext) - h);
    const float oy = fmaxf(0.0f, (fabsf(y - h) + ext) - h);
    return __builtin_fmaf(ox, ox, oy * oy);
}

// Rotate (c, s) by t: Taylor polynomials (|t| <= 0.8) and one Newton step of
// the norm, so the state carries cos/sin and never calls a transcendental.
__device__ __forceinline__ void rot32(float c, float s, float t, float& co, float& so) {
    const float t2 = t * t;
    const float sp = __builtin_fmaf(
        t2, __builtin_fmaf(t2, __builtin_fmaf(t2, __builtin_fmaf(t2, 2.7557319e-6f, -1.9841270e-4f),
                                              8.3333333e-3f), -1.6666667e-1f), 1.0f);
    const float sn = t * sp;
    const float cs = __builtin_fmaf(
        t2, __builtin_fmaf(t2, __builtin_fmaf(t2, __builtin_fmaf(t2, 2.4801587e-5f, -1.3888889e-3f),
                                              4.1666667e-2f), -0.5f), 1.0f);
    const float c2 = __builtin_fmaf(c, cs, -(s * sn));
    const float s2 = __builtin_fmaf(s, cs, c * sn);
    const float nn = __builtin_fmaf(c2, c2, s2 * s2);
    const float k = __builtin_fmaf(-0.5f, nn, 1.5f);
    co = c2 * k;
    so = s2 * k;
}

// ---------------------------------------------------------------------------
// LDS helpers.
// ---------------------------------------------------------------------------
// A sweep workgroup is one wavefront (kSweepThreads) or four (kSweepThreadsBig,
// instances of kBigN squares or more).  A wavefront's LDS instructions execute
// in issue order, so in a one-wavefront workgroup an LDS read sees every
// earlier LDS write or atomic without a wait (which is also why the compiler
// inserts none between them): its phase boundaries only stop the compiler
// moving LDS accesses across them.  Four wavefronts need the workgroup barrier.
static_assert(kSweepThreads == 64, "the small sweep workgroup must be one wavefront");
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_wave_barrier();
    asm volatile("" ::: "memory");
}
template <int NT>
__device__ __forceinline__ void tile_sync() {
    if constexpr (NT == 64)
        wave_sync();
    else
        __syncthreads();
}

// Lane exchange inside aligned groups of four lanes (DPP quad permutes, no LDS
// round trip): xor 1 swaps neighbours, xor 2 swaps pairs.
__device__ __forceinline__ float quad_xor1(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xF, 0xF, true));
}
__device__ __forceinline__ float quad_xor2(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x4E, 0xF, 0xF, true));
}
// Inclusive prefix sum over the 64 lanes of a wavefront in DPP moves (row
// shifts by 1, 2, 4, 8 inside each row of 16 lanes, then the row totals
// broadcast to the following rows), no LDS round trip.
__device__ __forceinline__ int wave_incl_scan(int v) {
    v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xF, 0xF, false);  // row_shr:1
    v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xF, 0xF, false);  // row_shr:2
    v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xF, 0xF, false);  // row_shr:4
    v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xF, 0xF, false);  // row_shr:8
    v += __builtin_amdgcn_update_dpp(0, v, 0x142, 0xA, 0xF, false);  // row_bcast:15 into rows 1, 3
    v += __builtin_amdgcn_update_dpp(0, v, 0x143, 0xC, 0xF, false);  // row_bcast:31 into rows 2, 3
    return v;
}

// Sum and minimum of a value over the 64 lanes of a wavefront by the same
// DPP scan (lanes a shift or a row mask leaves out take the identity), lane
// 63's result read back to every lane.  Every lane of the wavefront must be
// active.  64-bit values move as two 32-bit halves.
__device__ __forceinline__ unsigned wave_sum_u32(unsigned v) {
    return (unsigned)__builtin_amdgcn_readlane(wave_incl_scan((int)v), 63);
}
template <int kCtrl, int kRows, bool kMin>
__device__ __forceinline__ unsigned long long dpp_step_u64(unsigned long long v) {
    constexpr unsigned long long id = kMin ? ~0ull : 0ull;
    const unsigned lo =
        (unsigned)__builtin_amdgcn_update_dpp((int)(unsigned)id, (int)(unsigned)v, kCtrl, kRows, 0xF, false);
    const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp((int)(unsigned)(id >> 32), (int)(unsigned)(v >> 32),
                                                              kCtrl, kRows, 0xF, false);
    const unsigned long long o = (unsigned long long)hi << 32 | lo;
    return kMin ? (o < v ? o : v) : v + o;
}
template <bool kMin>
__device__ __forceinline__ unsigned long long wave_reduce_u64(unsigned long long v) {
    v = dpp_step_u64<0x111, 0xF, kMin>(v);  // row_shr:1
    v = dpp_step_u64<0x112, 0xF, kMin>(v);  // row_shr:2
    v = dpp_step_u64<0x114, 0xF, kMin>(v);  // row_shr:4
    v = dpp_step_u64<0x118, 0xF, kMin>(v);  // row_shr:8
    v = dpp_step_u64<0x142, 0xA, kMin>(v);  // row_bcast:15 into rows 1, 3
    v = dpp_step_u64<0x143, 0xC, kMin>(v);  // row_bcast:31 into rows 2, 3
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)v, 63);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(v >> 32), 63);
    return (unsigned long long)hi << 32 | lo;
}

// Value of the lane N places up inside a row of 16 lanes (DPP row shift).
template <int N>
__device__ __forceinline__ float row_up(float v) {
    static_assert(N >= 1 && N <= 15, "row shift of 1..15 lanes");
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x100 + N, 0xF, 0xF, true));
}

// 16-bit counters packed two per 32-bit LDS word (counts never exceed 1000).
__device__ __forceinline__ unsigned add_u16(uint16_t* base, int idx, unsigned v) {
    unsigned* w = reinterpret_cast<unsigned*>(base) + (idx >> 1);
    const unsigned sh = (idx & 1) << 4;
    return (atomicAdd(w, v << sh) >> sh) & 0xFFFFu;
}

// Inclusive prefix sum of the u16 array a[0..K) (K a multiple of 8, 16-byte
// aligned) by one wavefront: each lane owns a contiguous run of 16-byte
// chunks, sums them, the lane totals are scanned across the wave, and a
// second pass writes the prefixes.  Totals stay below 65536.  Called by one
// wavefront of the workgroup.
__device__ void wave_scan_u16(uint16_t* a, int K) {
    const int lane = threadIdx.x & 63;
    uint4* v = reinterpret_cast<uint4*>(a);
    const int nv = K >> 3;
    const int per = (nv + kSweepThreads - 1) / kSweepThreads;
    const int b = min(nv, lane * per), e = min(nv, b + per);
    unsigned sum = 0;
    for (int k = b; k < e; ++k) {
        const uint4 w = v[k];
        sum += (w.x & 0xFFFFu) + (w.x >> 16) + (w.y & 0xFFFFu) + (w.y >> 16) + (w.z & 0xFFFFu) +
               (w.z >> 16) + (w.w & 0xFFFFu) + (w.w >> 16);
    }
    const unsigned incl = (unsigned)wave_incl_scan((int)sum);
    unsigned run = incl - sum;
    for (int k = b; k < e; ++k) {
        uint4 w = v[k];
        unsigned p0 = run + (w.x & 0xFFFFu), p1 = p0 + (w.x >> 16);
        unsigned p2 = p1 + (w.y & 0xFFFFu), p3 = p2 + (w.y >> 16);
        unsigned p4 = p3 + (w.z & 0xFFFFu), p5 = p4 + (w.z >> 16);
        unsigned p6 = p5 + (w.w & 0xFFFFu), p7 = p6 + (w.w >> 16);
        run = p7;
        w.x = p0 | (p1 << 16);
        w.y = p2 | (p3 << 16);
        w.z = p4 | (p5 << 16);
        w.w = p6 | (p7 << 16);
        v[k] = w;
    }
}

// A mover record: five 16-byte words, read in one LDS round trip.
//   w0  position at the start of the sweep (x, y, cos, sin)
//   w1  proposed x, y, cos, sin (cos = kNoMove: rejected before the stencil)
//   w2  wall-energy change, accept uniform, log2(e)/T, square | list slot << 16
//   w3  stencil a0 | d1 << 16, d2 | d3 << 16, d4 | c1 << 16, c2 | c3 << 16
//   w4  stencil c4 | total << 16, tile replica, 0, 0
constexpr int kRecWords = 5;
static_assert(kRecWords * 16 == kMoverRecBytes, "mover record size");

struct TileLds {
    float4* pos;
    float4* prop;     // per level item: proposed x, y, cos, sin of the sweep (cos = kNoMove:
                      // rejected before the stencil); null when the tile proposes inline
    float4* pw;       // per level item: wall-energy change of the proposal, accept uniform,
                      // square | list slot << 16, cell kx | ky << 8 (the last two as bits)
    float4* spos;     // positions in cell-list order: spos[k] is the square sorted[k];
                      // null when the tile reads positions through `sorted`
    uint4* rec;       // mover records in level order, kRecWords per item (replace prop
                      // and pw); null unless sweep_records()
    uint16_t* A;      // per replica grid_words(M) cell starts
    uint16_t* sorted;
    uint16_t* cellk;  // kx | ky << 8
    uint16_t* rank;   // rank, then group key
    uint16_t* items;
    uint16_t* slotq;  // cell-list position of each square
    float* shx;       // grid shift of the current sweep
    float* shy;
    float* shx_next;  // grid shift of the next sweep
    float* shy_next;
    float* sxy;
    float* sth;
    float* ik;
    float* Lr;        // box side of each replica (its temperature slot's)
    float* eout;      // canonical energy per replica at the end of the launch
    unsigned* maxd;   // per replica: the largest pair depth^2 of the last energy pass (float bits)
    CollWork* cw;     // per replica: collective-move / kick scratch
    int* acc;
    int* gcnt;        // kGroups + 2: level counts, then scatter cursors
    int* gstart;      // kGroups + 2: level starts (read by every wavefront of a large tile)
    unsigned long long* gmask;  // per replica: which colour x rank groups are occupied
    unsigned divn;    // ceil(2^26 / n): replica of a tile square by a multiply
};

// Replica of tile square q, q / n, as a multiply by m = ceil(2^26 / n): exact
// while q (m n - 2^26) < 2^26, which q < 2^16 (G n <= 65535) and n < 2^10
// guarantee.
__device__ __forceinline__ int rep_of(const TileLds& t, int q) {
    return (int)__umulhi((unsigned)q << 6, t.divn);
}

__device__ inline TileLds carve(unsigned char* smem, int n, int M, int G, bool wide) {
    TileLds t;
    const int gn = G * n;
    t.divn = (67108864u + (unsigned)n - 1u) / (unsigned)n;
    t.pos = reinterpret_cast<float4*>(smem);
    float4* next = t.pos + gn;
    t.prop = nullptr;
    t.pw = nullptr;
    t.rec = nullptr;
    const bool recs = sweep_records(n, gn, wide);
    if (!recs && sweep_prop_in_lds(gn, wide)) {
        t.prop = next;
        next += gn;
    }
    t.spos = nullptr;
    if (sweep_spos_in_lds(gn)) {
        t.spos = next;
        next += gn;
    }
    if (recs) {
        t.rec = reinterpret_cast<uint4*>(next);
        next += kRecWords * gn;
    }
    if (t.prop) {
        t.pw = next;
        next += gn;
    }
    t.A = reinterpret_cast<uint16_t*>(next);
    t.sorted = t.A + (size_t)G * grid_words(M);
    t.cellk = t.sorted + gn;
    t.rank = t.cellk + gn;
    t.items = t.rank + gn;
    t.slotq = t.items + gn;
    size_t off = (size_t)(reinterpret_cast<unsigned char*>(t.slotq + gn) - smem);
    off = (off + 15) & ~(size_t)15;
    t.gmask = reinterpret_cast<unsigned long long*>(smem + off);
    off += 8 * (size_t)G;
    t.cw = reinterpret_cast<CollWork*>(smem + off);
    off += sizeof(CollWork) * (size_t)G;
    float* f = reinterpret_cast<float*>(smem + off);
    t.shx = f;
    t.shy = f + G;
    t.shx_next = f + 2 * G;
    t.shy_next = f + 3 * G;
    t.sxy = f + 4 * G;
    t.sth = f + 5 * G;
    t.ik = f + 6 * G;
    t.Lr = f + 7 * G;
    t.eout = f + 8 * G;
    t.acc = reinterpret_cast<int*>(f + 9 * G);
    t.maxd = reinterpret_cast<unsigned*>(f + 10 * G);
    t.gcnt = reinterpret_cast<int*>(t.maxd + G);
    t.gstart = t.gcnt + (kGroups + 2);
    return t;
}

// Cell lists of every replica of the tile in canonical order (ascending
// square index inside a cell), with the occupancy rank of each square.
// One prefix sum runs over the cell counts of all G replicas back to back, so
// A holds tile-global positions into `sorted` (replica g's squares occupy
// sorted[g*n, (g+1)*n), each replica contributing exactly n counts).
template <int NT>
__device__ void bin_tile(const TileLds& t, int n, int G, int M, int m, float inv_w) {
    const int tid = threadIdx.x;
    const int gw = grid_words(M);
    uint4* A128 = reinterpret_cast<uint4*>(t.A);
    for (int k = tid; k < (G * gw) / 8; k += NT) A128[k] = make_uint4(0u, 0u, 0u, 0u);
    tile_sync<NT>();
    for (int q = tid; q < G * n; q += NT) {
        const int g = rep_of(t, q);
        const float4 p = t.pos[q];
        const int kx = cellcoord(p.x, t.shx[g], inv_w, 0, m - 1) + 2;
        const int ky = cellcoord(p.y, t.shy[g], inv_w, 0, m - 1) + 2;
        t.cellk[q] = (uint16_t)(kx | (ky << 8));
        add_u16(t.A + g * gw, ky * M + kx + 2, 1u);
    }
    tile_sync<NT>();
    if (tid < 64) wave_scan_u16(t.A, G * gw);
    tile_sync<NT>();
    for (int q = tid; q < G * n; q += NT) {
        const int g = rep_of(t, q), i = q - g * n;
        const int ck = t.cellk[q];
        const int c = (ck >> 8) * M + (ck & 255);
        const unsigned slot = add_u16(t.A + g * gw, c + 1, 1u);
        t.sorted[slot] = (uint16_t)i;
    }
    tile_sync<NT>();
    for (int q = tid; q < G * n; q += NT) {
        const int g = rep_of(t, q), i = q - g * n;
        const int ck = t.cellk[q];
        const int c = (ck >> 8) * M + (ck & 255);
        const uint16_t* Ag = t.A + g * gw;
        const int a = Ag[c], b = Ag[c + 1];
        int r = 0;
        for (int k = a; k < b; ++k) r += t.sorted[k] < i;
        t.rank[q] = (uint16_t)r;
    }
    tile_sync<NT>();
    for (int q = tid; q < G * n; q += NT) {
        const int g = rep_of(t, q), i = q - g * n;
        const int ck = t.cellk[q];
        const int c = (ck >> 8) * M + (ck & 255);
        const int k = t.A[g * gw + c] + t.rank[q];
        t.sorted[k] = (uint16_t)i;
        if (t.spos) t.spos[k] = t.pos[q];
        t.slotq[q] = (uint16_t)k;
    }
    tile_sync<NT>();
}

// The 5x5-cell stencil around (kx, ky) as one flat candidate sequence: five
// row ranges of the CSR cell list, visited row by row and, inside a row, in
// ascending CSR position.  One flat loop per lane costs the wave the largest
// per-lane total instead of the sum of per-row maxima.
struct StencilRows {
    int a[5], e[5];  // CSR start of each row range and its end
    __device__ __forceinline__ void load(const uint16_t* A, int M, int kx, int ky) {
#pragma unroll
        for (int r = 0; r < 5; ++r) {
            const int row = (ky + r - 2) * M + kx;
            a[r] = A[row - 2];
            e[r] = A[row + 3];
        }
    }
};
struct Stencil {
    int a0, d1, d2, d3, d4, c1, c2, c3, c4, total;
    __device__ __forceinline__ Stencil(const uint16_t* A, int M, int kx, int ky) {
        StencilRows rw;
        rw.load(A, M, kx, ky);
        init(rw);
    }
    __device__ __forceinline__ explicit Stencil(const StencilRows& rw) { init(rw); }
    __device__ __forceinline__ void init(const StencilRows& rw) {
        int a[5], l[5];
#pragma unroll
        for (int r = 0; r < 5; ++r) {
            a[r] = rw.a[r];
            l[r] = rw.e[r] - rw.a[r];
        }
        c1 = l[0];
        c2 = c1 + l[1];
        c3 = c2 + l[2];
        c4 = c3 + l[3];
        total = c4 + l[4];
        // row r's offset is o_r = a[r] - c_r (o_0 = a[0]); keep the steps
        // o_r - o_{r-1}
        a0 = a[0];
        d1 = (a[1] - c1) - a0;
        d2 = (a[2] - c2) - (a[1] - c1);
        d3 = (a[3] - c3) - (a[2] - c2);
        d4 = (a[4] - c4) - (a[3] - c3);
    }
    // CSR position of flat candidate f: f + o_K with K the last row whose
    // cumulative start c_K <= f (the c are nondecreasing, so the steps of
    // rows 1..K telescope to o_K and empty rows add nothing).  Written as
    // conditional increments rather than a select among the row offsets,
    // which the compiler would turn into an indexed table in scratch memory.
    Stencil() = default;
    // the packed form of a mover record (all fields below 2^16)
    __device__ __forceinline__ void pack(uint4& s0, unsigned& s1x) const {
        s0 = make_uint4((unsigned)a0 | (unsigned)d1 << 16, (unsigned)d2 | (unsigned)d3 << 16,
                        (unsigned)d4 | (unsigned)c1 << 16, (unsigned)c2 | (unsigned)c3 << 16);
        s1x = (unsigned)c4 | (unsigned)total << 16;
    }
    __device__ __forceinline__ static Stencil unpack(const uint4& s0, unsigned s1x) {
        Stencil sc;
        sc.a0 = (int)(s0.x & 0xFFFFu);
        sc.d1 = (int)(s0.x >> 16);
        sc.d2 = (int)(s0.y & 0xFFFFu);
        sc.d3 = (int)(s0.y >> 16);
        sc.d4 = (int)(s0.z & 0xFFFFu);
        sc.c1 = (int)(s0.z >> 16);
        sc.c2 = (int)(s0.w & 0xFFFFu);
        sc.c3 = (int)(s0.w >> 16);
        sc.c4 = (int)(s1x & 0xFFFFu);
        sc.total = (int)(s1x >> 16);
        return sc;
    }
    __device__ __forceinline__ int at(int f) const {
        int off = a0;
        off += f >= c1 ? d1 : 0;
        off += f >= c2 ? d2 : 0;
        off += f >= c3 ? d3 : 0;
        off += f >= c4 ? d4 : 0;
        return f + off;
    }
};

// Proposal of tile square q (replica g) at state p with variates r: the new
// centre and orientation and the wall-energy change, or false when the move
// leaves the box or its cell (rejected before any neighbour is examined).
struct Move {
    float nx, ny, nc, ns, wdiff;
};
constexpr float kNoMove = 8.0f;  // cos field of a staged proposal rejected early
// (replica step sizes sxy, sth, grid shift shx, shy and box side L given)
__device__ __forceinline__ bool make_move(float sxy, float sth, float shx, float shy, const float4& p,
                                          const float4& r, int kx, int ky, int m, float inv_w, float L,
                                          Move& mv) {
    mv.nx = p.x + sxy * r.x;
    mv.ny = p.y + sxy * r.y;
    if (!(mv.nx >= 0.0f && mv.nx <= L && mv.ny >= 0.0f && mv.ny <= L)) return false;
    if (cellcoord(mv.nx, shx, inv_w, 0, m - 1) + 2 != kx || cellcoord(mv.ny, shy, inv_w, 0, m - 1) + 2 != ky)
        return false;
    rot32(p.z, p.w, sth * r.z, mv.nc, mv.ns);
    mv.wdiff = wall32(mv.nx, mv.ny, mv.nc, mv.ns, L) - wall32(p.x, p.y, p.z, p.w, L);
    return true;
}
__device__ __forceinline__ bool make_move(const TileLds& t, int g, const float4& p, const float4& r, int kx,
                                          int ky, int m, float inv_w, Move& mv) {
    return make_move(t.sxy[g], t.sth[g], t.shx[g], t.shy[g], p, r, kx, ky, m, inv_w, t.Lr[g], mv);
}

// Stencil energy change and Metropolis test of one proposal of tile square q
// (replica g, cell-list slot `self`, position p, move mv, accept uniform u,
// log2(e)/T ik), computed by `lpi` consecutive lanes (1, 2, 4, 8 or 16; this
// lane is `sub`).  The stencil energy change is four partial sums over the
// flat candidate index f (f mod 4), each in increasing f, combined as
// (s0 + s1) + (s2 + s3): the same bits whatever the number of lanes sharing
// the work.  kNear: also count the near pairs (centres closer than sqrt 2)
// before and after the move.
template <bool kQuads, bool kNear>
__device__ __forceinline__ void evaluate(const TileLds& t, int q, int g, int self, const float4& p,
                                         const Stencil& sc, const Move& mv, float u, float ik, int n,
                                         int lpi, int sub, unsigned long long& cC,
                                         unsigned long long& cP) {
    const float nx = mv.nx, ny = mv.ny, nc = mv.nc, ns = mv.ns, wdiff = mv.wdiff;
    const int last = sc.total - 1;  // >= 0: the mover is in its own cell
    unsigned near = 0;
    // Energy change against the square at cell-list position k.  The mover
    // itself and positions past the end of the stencil contribute an exact
    // +0 (no partial sum can hold -0, so this equals skipping them).
    auto delta = [&](const float4& v, int k, bool valid) -> float {
        const float dxn = v.x - nx, dyn = v.y - ny;
        const float dxo = v.x - p.x, dyo = v.y - p.y;
        const float en = depth2(dxn, dyn, nc, ns, v.z, v.w);
        const float eo = depth2(dxo, dyo, p.z, p.w, v.z, v.w);
        const bool other = valid && k != self;
        // near pairs: two compare masks combined with `other` on the scalar
        // unit, two conditional increments
        if constexpr (kNear) {
            near += (unsigned)(other && __builtin_fmaf(dxn, dxn, dyn * dyn) < 2.0f);
            near += (unsigned)(other && __builtin_fmaf(dxo, dxo, dyo * dyo) < 2.0f);
        }
        return other ? en - eo : 0.0f;
    };
    // Four candidates at once: their positions are read from the cell-list
    // ordered copy in one go (or through their square indices, four and
    // four), so four candidates cost one or two LDS round trips.
    const float4* pg = t.pos + (size_t)g * n;
    auto quad = [&](int f0, int f1, int f2, int f3, float& d0, float& d1, float& d2, float& d3) {
        const int k0 = sc.at(min(f0, last));
        const int k1 = sc.at(min(f1, last));
        const int k2 = sc.at(min(f2, last));
        const int k3 = sc.at(min(f3, last));
        float4 v0, v1, v2, v3;
        if (t.spos) {
            v0 = t.spos[k0]; v1 = t.spos[k1]; v2 = t.spos[k2]; v3 = t.spos[k3];
        } else {
            const int j0 = t.sorted[k0], j1 = t.sorted[k1], j2 = t.sorted[k2], j3 = t.sorted[k3];
            v0 = pg[j0]; v1 = pg[j1]; v2 = pg[j2]; v3 = pg[j3];
        }
        d0 = delta(v0, k0, f0 <= last);
        d1 = delta(v1, k1, f1 <= last);
        d2 = delta(v2, k2, f2 <= last);
        d3 = delta(v3, k3, f3 <= last);
    };
    // partial k accumulates f = k (mod 4); one loop per lane count
    float sum, d0, d1, d2, d3;
    if (!kQuads && lpi < 8) {
        // One-wavefront tiles: one candidate per lane and step, f = sub +
        // lpi j, without the four-candidate batches (their registers and code
        // buy latency hiding that a lone wavefront per SIMD does not use).
        // Partial f mod 4 of each step is the same on every lane of a mover
        // for lpi 1 (all four by j) and 4 (always sub); for lpi 2 it is sub
        // or sub + 2 by the parity of j.
        float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f, s3 = 0.0f;
        int j = 0;
        for (int f = sub; f <= last; f += lpi, ++j) {
            const int k = sc.at(f);
            const float4 v = t.spos ? t.spos[k] : pg[t.sorted[k]];
            const float d = delta(v, k, true);
            const int slot = lpi == 1 ? (j & 3) : lpi == 2 ? 2 * (j & 1) : 0;
            if (slot == 0)
                s0 += d;
            else if (slot == 1)
                s1 += d;
            else if (slot == 2)
                s2 += d;
            else
                s3 += d;
        }
        if (lpi == 1) {
            sum = (s0 + s1) + (s2 + s3);
        } else if (lpi == 2) {
            // s0: partial sub, s2: partial sub + 2
            sum = (s0 + quad_xor1(s0)) + (s2 + quad_xor1(s2));
        } else {
            const float h = s0 + quad_xor1(s0);
            sum = h + quad_xor2(h);
        }
    } else if (lpi == 1) {
        // (every four-candidate branch runs its first round unguarded: quad()
        // clamps the positions past `last` and masks them to an exact +0,
        // which leaves a partial as a skipped round would)
        float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f, s3 = 0.0f;
        for (int f = 0;; f += 4) {
            quad(f, f + 1, f + 2, f + 3, d0, d1, d2, d3);
            s0 += d0;
            s1 += d1;
            s2 += d2;
            s3 += d3;
            if (f + 4 > last) break;
        }
        sum = (s0 + s1) + (s2 + s3);
    } else if (lpi == 2) {
        // sub 0 owns partials 0, 1; sub 1 owns 2, 3
        float sa = 0.0f, sb = 0.0f;
        for (int f = 2 * sub;; f += 8) {
            quad(f, f + 1, f + 4, f + 5, d0, d1, d2, d3);
            sa += d0;
            sb += d1;
            sa += d2;
            sb += d3;
            if (f + 8 > last) break;
        }
        const float h = sa + sb;
        sum = h + quad_xor1(h);
    } else if (lpi == 4) {
        // sub k owns partial k
        float sa = 0.0f;
        for (int f = sub;; f += 16) {
            quad(f, f + 4, f + 8, f + 12, d0, d1, d2, d3);
            sa += d0;
            sa += d1;
            sa += d2;
            sa += d3;
            if (f + 16 > last) break;
        }
        const float h = sa + quad_xor1(sa);
        sum = h + quad_xor2(h);
    } else {
        // 8 or 16 lanes, one candidate each per round of lpi candidates: lane
        // sub holds f = round + sub, and partial k (f = k mod 4) folds the
        // lanes k, k + 4, (k + 8, k + 12) of each round in that (increasing f)
        // order into lane k; the lanes of a mover are aligned inside one row
        // of 16 lanes.
        float sa = 0.0f;
        for (int base = 0;; base += lpi) {
            const int f = base + sub;
            const int k = sc.at(min(f, last));
            const float4 v = t.spos ? t.spos[k] : pg[t.sorted[k]];
            const float d = delta(v, k, f <= last);
            const float d4 = row_up<4>(d);
            if (lpi == 8) {
                sa = (sa + d) + d4;
            } else {
                const float d8 = row_up<8>(d), d12 = row_up<12>(d);
                sa = (((sa + d) + d4) + d8) + d12;
            }
            if (base + lpi > last) break;  // the first round runs unguarded (last >= 0)
        }
        const float h = sa + quad_xor1(sa);
        sum = h + quad_xor2(h);
    }
    // every stencil position but the mover's own is a candidate: `last` of
    // them per proposal, counted once by the mover's first lane
    cC += sub == 0 ? (unsigned)last : 0u;
    cP += near;
    const float dE = wdiff + sum;
    if (sub == 0 && (dE <= 0.0f || u < accept_exp2(-dE * ik))) {
        const float4 np = make_float4(nx, ny, nc, ns);
        t.pos[q] = np;
        if (t.spos) t.spos[self] = np;
        atomicAdd(&t.acc[g], 1);
    }
}

// One Metropolis proposal for square q of its tile replica: the move comes
// from t.prop (staged for the whole sweep before the first level) or is drawn
// here.
template <bool kQuads, bool kNear>
__device__ __forceinline__ void propose(const TileLds& t, int q, int n, int M, int m, float inv_w,
                                        int r0, uint32_t sweep, uint32_t k0, uint32_t k1,
                                        int lpi, int sub, unsigned long long& cC,
                                        unsigned long long& cP) {
    // squares outside the levels (the serial group) draw their move here
    const int g = rep_of(t, q);
    const int gw = grid_words(M);
    const float4 p = t.pos[q];
    const int self = t.slotq[q];
    const int ck = t.cellk[q];
    const int kx = ck & 255, ky = ck >> 8;
    // the stencil depends only on the current cell: its loads go out first
    const Stencil sc(t.A + g * gw, M, kx, ky);
    const float ik = t.ik[g];
    Move mv;
    const float4 r = move_variates((uint32_t)(q - g * n), (uint32_t)(r0 + g), sweep, k0, k1);
    if (!make_move(t, g, p, r, kx, ky, m, inv_w, mv)) return;
    evaluate<kQuads, kNear>(t, q, g, self, p, sc, mv, r.w, ik, n, lpi, sub, cC, cP);
}

// The staged proposal of level item j (tiles with t.prop): the move, wall
// change, uniform, square, list slot and cell in one LDS round trip; then
// the square's position, 1/T and stencil.
template <bool kQuads, bool kNear>
__device__ __forceinline__ void propose_item(const TileLds& t, int j, int n, int M, int lpi, int sub,
                                             unsigned long long& cC, unsigned long long& cP) {
    const float4 a = t.prop[j];
    const float4 b = t.pw[j];
    if (a.z == kNoMove) return;
    const unsigned qs = __float_as_uint(b.z), ck = __float_as_uint(b.w);
    const int q = (int)(qs & 0xFFFFu), g = rep_of(t, q);
    const Stencil sc(t.A + g * grid_words(M), M, (int)(ck & 255u), (int)(ck >> 8));
    evaluate<kQuads, kNear>(t, q, g, (int)(qs >> 16), t.pos[q], sc, Move{a.x, a.y, a.z, a.w, b.x}, b.y, t.ik[g], n,
                            lpi, sub, cC, cP);
}

// Level passes of a tile with staged proposals: each mover's data is read one
// level ahead, its staged proposal and indices when the level before starts
// (in flight while that level computes) and its stencil rows, position and
// 1/T once that level's moves are written (in flight across the barrier), so
// that a level waits on one LDS round trip, its candidates.  A square moves
// only in its own level and no cell list changes during a sweep, so nothing
// read ahead is written in between.  A level with more movers than one round
// of lanes takes rounds of propose_item instead.  gs: lane l holds the start
// of level l (lane kGroups: the end of the last).
template <int NT, bool kNear>
__device__ __forceinline__ void staged_levels(const TileLds& t, int n, int M, int gs, unsigned long long& cC,
                                              unsigned long long& cP) {
    const int tid = threadIdx.x;
    const int gw = grid_words(M);
    auto lsh_of = [](int cnt) {
        return cnt > NT / 2 ? 0 : cnt > NT / 4 ? 1 : cnt > NT / 8 ? 2 : cnt > NT / 16 ? 3 : 4;
    };
    struct Ahead {
        float4 mv, side, p;
        float ik;
        StencilRows rw;
    };
    auto stage1 = [&](int j, Ahead& x) {
        x.mv = t.prop[j];
        x.side = t.pw[j];
    };
    auto stage2 = [&](Ahead& x) {
        const unsigned qs = __float_as_uint(x.side.z), ck = __float_as_uint(x.side.w);
        const int q = (int)(qs & 0xFFFFu), g = rep_of(t, q);
        x.p = t.pos[q];
        x.ik = t.ik[g];
        x.rw.load(t.A + g * gw, M, (int)(ck & 255u), (int)(ck >> 8));
    };
    int a = __builtin_amdgcn_readlane(gs, 0), b = __builtin_amdgcn_readlane(gs, 1);
    int lsh = lsh_of(b - a);
    Ahead cur;
    bool cur_ok = false;
    {
        const int j = a + (tid >> lsh);
        if (j < b && b - a <= (NT >> lsh)) {
            stage1(j, cur);
            stage2(cur);
            cur_ok = true;
        }
    }
    for (int lv = 0; lv < kGroups && a < b; ++lv) {
        const int cnt = b - a;
        const int lpi = 1 << lsh, sub = tid & (lpi - 1);
        const int a2 = b, b2 = lv + 1 < kGroups ? __builtin_amdgcn_readlane(gs, lv + 2) : b;
        const int lsh2 = lsh_of(b2 - a2);
        const int j2 = a2 + (tid >> lsh2);
        const bool ok2 = j2 < b2 && b2 - a2 <= (NT >> lsh2);
        Ahead nxt;
        if (ok2) stage1(j2, nxt);
        if (cnt <= (NT >> lsh)) {
            if (cur_ok && cur.mv.z != kNoMove) {
                const unsigned qs = __float_as_uint(cur.side.z);
                const int q = (int)(qs & 0xFFFFu);
                evaluate<(NT > 64), kNear>(t, q, rep_of(t, q), (int)(qs >> 16), cur.p, Stencil(cur.rw),
                                           Move{cur.mv.x, cur.mv.y, cur.mv.z, cur.mv.w, cur.side.x}, cur.side.y,
                                           cur.ik, n, lpi, sub, cC, cP);
            }
        } else {
            const int per = NT >> lsh;
            for (int base = a; base < b; base += per)
                if (base + (tid >> lsh) < b)
                    propose_item<(NT > 64), kNear>(t, base + (tid >> lsh), n, M, lpi, sub, cC, cP);
        }
        if (ok2) stage2(nxt);
        tile_sync<NT>();
        cur = nxt;
        cur_ok = ok2;
        a = a2;
        b = b2;
        lsh = lsh2;
    }
}

// Mover record j of a level (one-wavefront tiles): square q (replica g,
// index i in its replica, CSR slot `self`) at position p in cell (kx, ky),
// with its stencil, staged move (variates r), wall change, uniform and 1/T.
__device__ __forceinline__ void write_record(const TileLds& t, int j, int q, int g, int self, const float4& p,
                                             int kx, int ky, const Stencil& sc, const float4& r, float sxy,
                                             float sth, float shx, float shy, float ik, int m, float inv_w,
                                             float L) {
    Move mv;
    const bool ok = make_move(sxy, sth, shx, shy, p, r, kx, ky, m, inv_w, L, mv);
    uint4 s0;
    unsigned s1x;
    sc.pack(s0, s1x);
    uint4* w = t.rec + (size_t)j * kRecWords;
    w[0] = make_uint4(__float_as_uint(p.x), __float_as_uint(p.y), __float_as_uint(p.z), __float_as_uint(p.w));
    w[1] = ok ? make_uint4(__float_as_uint(mv.nx), __float_as_uint(mv.ny), __float_as_uint(mv.nc),
                           __float_as_uint(mv.ns))
              : make_uint4(0u, 0u, __float_as_uint(kNoMove), 0u);
    w[2] = make_uint4(__float_as_uint(ok ? mv.wdiff : 0.0f), __float_as_uint(r.w), __float_as_uint(ik),
                      (unsigned)q | (unsigned)self << 16);
    w[3] = s0;
    w[4] = make_uint4(s1x, (unsigned)g, 0u, 0u);
}
// The same from the tile's LDS arrays (square q of level item j), written
// once per sweep by one lane per square when the level items are scattered.
__device__ __forceinline__ void write_record(const TileLds& t, int j, int q, int n, int M, int m,
                                             float inv_w, int r0, uint32_t sweep, uint32_t k0,
                                             uint32_t k1) {
    const int g = rep_of(t, q), i = q - g * n;
    const int ck = t.cellk[q];
    const int kx = ck & 255, ky = ck >> 8;
    const Stencil sc(t.A + g * grid_words(M), M, kx, ky);
    const float4 r = move_variates((uint32_t)i, (uint32_t)(r0 + g), sweep, k0, k1);
    write_record(t, j, q, g, t.slotq[q], t.pos[q], kx, ky, sc, r, t.sxy[g], t.sth[g], t.shx[g], t.shy[g],
                 t.ik[g], m, inv_w, t.Lr[g]);
}

// The proposal of mover record j: one LDS round trip for everything but the
// candidates.
template <bool kNear>
__device__ __forceinline__ void propose_rec(const TileLds& t, int j, int n, int lpi, int sub,
                                            unsigned long long& cC, unsigned long long& cP) {
    const uint4* w = t.rec + (size_t)j * kRecWords;
    uint4 w0 = w[0], w1 = w[1], w2 = w[2], w3 = w[3];
    uint2 w4 = *reinterpret_cast<const uint2*>(w + 4);
    // all five words in flight before the flag test
    asm volatile("" : "+v"(w0.x), "+v"(w0.y), "+v"(w0.z), "+v"(w0.w), "+v"(w1.x), "+v"(w1.y), "+v"(w1.z),
                 "+v"(w1.w), "+v"(w2.x), "+v"(w2.y), "+v"(w2.z), "+v"(w2.w));
    asm volatile("" : "+v"(w3.x), "+v"(w3.y), "+v"(w3.z), "+v"(w3.w), "+v"(w4.x), "+v"(w4.y));
    if (__uint_as_float(w1.z) == kNoMove) return;
    const float4 p = make_float4(__uint_as_float(w0.x), __uint_as_float(w0.y), __uint_as_float(w0.z),
                                 __uint_as_float(w0.w));
    const Move mv{__uint_as_float(w1.x), __uint_as_float(w1.y), __uint_as_float(w1.z), __uint_as_float(w1.w),
                  __uint_as_float(w2.x)};
    const Stencil sc = Stencil::unpack(w3, w4.x);
    evaluate<false, kNear>(t, (int)(w2.w & 0xFFFFu), (int)w4.y, (int)(w2.w >> 16), p, sc, mv, __uint_as_float(w2.y),
                    __uint_as_float(w2.z), n, lpi, sub, cC, cP);
}

// Lanes per mover of a level of cnt movers in one wavefront: 2^lsh, the most
// that fit (at most 16, the lanes of one DPP row).
__device__ __forceinline__ int level_lsh(int cnt) {
    return cnt > 32 ? 0 : cnt > 16 ? 1 : cnt > 8 ? 2 : cnt > 4 ? 3 : 4;
}

struct MoverRec {
    uint4 w0, w1, w2, w3;
    uint2 w4;
};
__device__ __forceinline__ MoverRec load_rec(const TileLds& t, int j) {
    const uint4* w = t.rec + (size_t)j * kRecWords;
    MoverRec r;
    r.w0 = w[0];
    r.w1 = w[1];
    r.w2 = w[2];
    r.w3 = w[3];
    r.w4 = *reinterpret_cast<const uint2*>(w + 4);
    return r;
}

// 2^x of the Metropolis test without a branch: accept_exp2's value for
// x >= -100 (the same operations on the same argument) and 0 below; x is
// clamped to [-127, 0] first, which changes nothing where the value is used
// (the test is made only when the energy rises, so x < 0) and keeps the
// integer conversion in range elsewhere.
__device__ __forceinline__ float accept_exp2_sel(float x) {
    const float xc = fminf(fmaxf(x, -127.0f), 0.0f);
    const float k = rintf(xc);
    const float f = xc - k;
    float p = 0x1.444p-13f;
    p = __builtin_fmaf(p, f, 0x1.5f48cp-10f);
    p = __builtin_fmaf(p, f, 0x1.3b2a1cp-7f);
    p = __builtin_fmaf(p, f, 0x1.c6aeccp-5f);
    p = __builtin_fmaf(p, f, 0x1.ebfbep-3f);
    p = __builtin_fmaf(p, f, 0x1.62e43p-1f);
    p = __builtin_fmaf(p, f, 1.0f);
    const float v = ldexpf(p, (int)k);
    return x >= -100.0f ? v : 0.0f;
}

// One level of a one-wavefront tile with mover records when all its movers
// fit the wavefront at once, for 2^LSH lanes per mover: lane l serves record
// a + (l >> LSH) as its (l & (lpi - 1))-th lane.  The arithmetic, the order
// of the four partial sums and the Metropolis test are those of evaluate();
// a mover rejected before the stencil (or a lane past the level's end) runs
// the same instructions with its contributions masked, so the level has no
// divergent exit.  The lane count is a template argument so that the loop
// and the partial-sum order carry no run-time branches, and the Metropolis
// test is evaluated without one.  Counts are 32-bit per lane.
template <int LSH, bool kNear>
__device__ __forceinline__ void rec_level(const TileLds& t, const MoverRec& r, bool act, int sub, unsigned& cC,
                                          unsigned& cP) {
    constexpr int lpi = 1 << LSH;
    const bool ok = act && __uint_as_float(r.w1.z) != kNoMove;
    const float px = __uint_as_float(r.w0.x), py = __uint_as_float(r.w0.y);
    const float pc = __uint_as_float(r.w0.z), ps = __uint_as_float(r.w0.w);
    const float nx = __uint_as_float(r.w1.x), ny = __uint_as_float(r.w1.y);
    const float nc = __uint_as_float(r.w1.z), ns = __uint_as_float(r.w1.w);
    const int self = (int)(r.w2.w >> 16);
    const Stencil sc = Stencil::unpack(r.w3, r.w4.x);
    const int last = ok ? sc.total - 1 : -1;  // the whole group of a skipped mover runs no candidate
    unsigned near = 0;
    auto delta = [&](const float4& v, int k, bool valid) -> float {
        const float dxn = v.x - nx, dyn = v.y - ny;
        const float dxo = v.x - px, dyo = v.y - py;
        const float en = depth2(dxn, dyn, nc, ns, v.z, v.w);
        const float eo = depth2(dxo, dyo, pc, ps, v.z, v.w);
        const bool other = valid && k != self;
        if constexpr (kNear) {
            near += (unsigned)(other && __builtin_fmaf(dxn, dxn, dyn * dyn) < 2.0f);
            near += (unsigned)(other && __builtin_fmaf(dxo, dxo, dyo * dyo) < 2.0f);
        }
        return other ? en - eo : 0.0f;
    };
    // Every branch runs its first round on every lane, unguarded: a position
    // past `last` reads an in-bounds candidate (the clamp) and adds an exact
    // +0 (never -0), which leaves a partial sum as a skipped round would; the
    // rounds after the first loop as before.  Most levels need one round.
    const int lastc = max(last, 0);
    float sum;
    if constexpr (lpi >= 8) {
        // rounds of lpi candidates, f = round + sub; partial k (f = k mod 4)
        // folds lanes k, k + 4 (k + 8, k + 12) of each round into lane k
        float sa = 0.0f;
        auto round = [&](int base) {
            const int f = base + sub;
            const int k = sc.at(min(f, lastc));
            const float d = delta(t.spos[k], k, f <= last);
            const float d4 = row_up<4>(d);
            if constexpr (lpi == 8) {
                sa = (sa + d) + d4;
            } else {
                const float d8 = row_up<8>(d), d12 = row_up<12>(d);
                sa = (((sa + d) + d4) + d8) + d12;
            }
        };
        round(0);
        for (int base = lpi; base <= last; base += lpi) round(base);
        const float h = sa + quad_xor1(sa);
        sum = h + quad_xor2(h);
    } else if constexpr (lpi == 4) {
        // f = sub + 4 j: always partial sub
        float s0 = 0.0f;
        {
            const int k = sc.at(min(sub, lastc));
            s0 += delta(t.spos[k], k, sub <= last);
        }
        for (int f = sub + 4; f <= last; f += 4) {
            const int k = sc.at(f);
            s0 += delta(t.spos[k], k, true);
        }
        const float h = s0 + quad_xor1(s0);
        sum = h + quad_xor2(h);
    } else if constexpr (lpi == 2) {
        // f = sub + 2 j: partial sub for even j, sub + 2 for odd j
        float s0 = 0.0f, s2 = 0.0f;
        {
            const int k = sc.at(min(sub, lastc));
            const int k2 = sc.at(min(sub + 2, lastc));
            s0 += delta(t.spos[k], k, sub <= last);
            s2 += delta(t.spos[k2], k2, sub + 2 <= last);
        }
        for (int f = sub + 4; f <= last; f += 4) {
            const int k = sc.at(f);
            s0 += delta(t.spos[k], k, true);
            if (f + 2 <= last) {
                const int k2 = sc.at(f + 2);
                s2 += delta(t.spos[k2], k2, true);
            }
        }
        sum = (s0 + quad_xor1(s0)) + (s2 + quad_xor1(s2));
    } else {
        // one lane: partial f mod 4 of candidate f
        float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f, s3 = 0.0f;
        {
            const int k0 = sc.at(0), k1 = sc.at(min(1, lastc)), k2 = sc.at(min(2, lastc)),
                      k3 = sc.at(min(3, lastc));
            s0 += delta(t.spos[k0], k0, 0 <= last);
            s1 += delta(t.spos[k1], k1, 1 <= last);
            s2 += delta(t.spos[k2], k2, 2 <= last);
            s3 += delta(t.spos[k3], k3, 3 <= last);
        }
        for (int f = 4; f <= last; f += 4) {
            const int k0 = sc.at(f);
            s0 += delta(t.spos[k0], k0, true);
            if (f + 1 <= last) {
                const int k1 = sc.at(f + 1);
                s1 += delta(t.spos[k1], k1, true);
            }
            if (f + 2 <= last) {
                const int k2 = sc.at(f + 2);
                s2 += delta(t.spos[k2], k2, true);
            }
            if (f + 3 <= last) {
                const int k3 = sc.at(f + 3);
                s3 += delta(t.spos[k3], k3, true);
            }
        }
        sum = (s0 + s1) + (s2 + s3);
    }
    const bool lead = ok && sub == 0;
    cC += lead ? (unsigned)last : 0u;
    cP += near;
    const float dE = __uint_as_float(r.w2.x) + sum;
    const float pacc = accept_exp2_sel(-dE * __uint_as_float(r.w2.z));
    if (lead && (dE <= 0.0f || __uint_as_float(r.w2.y) < pacc)) {
        const float4 np = make_float4(nx, ny, nc, ns);
        t.pos[r.w2.w & 0xFFFFu] = np;
        t.spos[self] = np;
        atomicAdd(&t.acc[r.w4.y], 1);
    }
}

// One level of mover records with the lane count of its mover count.
template <bool kNear>
__device__ __forceinline__ void rec_level_any(const TileLds& t, const MoverRec& r, int a, int b, int lsh,
                                              unsigned& cC, unsigned& cP) {
    const int lane = threadIdx.x;
    const bool act = a + (lane >> lsh) < b;
    const int sub = lane & ((1 << lsh) - 1);
    switch (lsh) {
        case 4: rec_level<4, kNear>(t, r, act, sub, cC, cP); break;
        case 3: rec_level<3, kNear>(t, r, act, sub, cC, cP); break;
        case 2: rec_level<2, kNear>(t, r, act, sub, cC, cP); break;
        case 1: rec_level<1, kNear>(t, r, act, sub, cC, cP); break;
        default: rec_level<0, kNear>(t, r, act, sub, cC, cP); break;
    }
}

// The levels of one sweep of a one-wavefront tile with mover records (gs:
// lane l holds the start of level l, lane kGroups the end of the last).
// Records do not change during a sweep (a square moves once, and a move never
// leaves its cell), so the next level's records are read while this level
// computes; only the candidates' positions wait for the level before.  Two
// levels per iteration, alternating between two record registers, so the
// read-ahead needs no copy.  Levels of more than 64 movers (tiles of more
// than 64 squares) take rounds of the general pass.
template <bool kNear>
__device__ __forceinline__ void rec_levels(const TileLds& t, int n, int gs, unsigned long long& cC,
                                           unsigned long long& cP) {
    const int lane = threadIdx.x;
    unsigned c32 = 0, p32 = 0;
    int a = __builtin_amdgcn_readlane(gs, 0), b = __builtin_amdgcn_readlane(gs, 1);
    int lsh = level_lsh(b - a);
    int lv = 0;
    MoverRec A, B;
    if (a < b && b - a <= kSweepThreads) A = load_rec(t, min(a + (lane >> lsh), b - 1));
    // level lv from `cur`, with the next level's records read into `nxt`
    auto level = [&](const MoverRec& cur, MoverRec& nxt) {
        const int a2 = b, b2 = lv + 1 < kGroups ? __builtin_amdgcn_readlane(gs, lv + 2) : b;
        const int lsh2 = level_lsh(b2 - a2);
        const bool ahead = a2 < b2 && b2 - a2 <= kSweepThreads;
        if (b - a <= kSweepThreads) {
            if (ahead) nxt = load_rec(t, min(a2 + (lane >> lsh2), b2 - 1));
            rec_level_any<kNear>(t, cur, a, b, lsh, c32, p32);
        } else {
            const int lpi = 1 << lsh, sub = lane & (lpi - 1), per = kSweepThreads >> lsh;
            for (int base = a; base < b; base += per)
                if (base + (lane >> lsh) < b) propose_rec<kNear>(t, base + (lane >> lsh), n, lpi, sub, cC, cP);
            if (ahead) nxt = load_rec(t, min(a2 + (lane >> lsh2), b2 - 1));
        }
        wave_sync();
        a = a2;
        b = b2;
        lsh = lsh2;
        ++lv;
    };
    while (lv < kGroups && a < b) {
        level(A, B);
        if (!(lv < kGroups && a < b)) break;
        level(B, A);
    }
    cC += c32;
    cP += p32;
}

// Canonical FP32 energy of every tile replica: per-square wall + pairs with
// larger index inside the 5x5 stencil, then a fixed pairwise tree over the
// square index, so the bits do not depend on the tile shape.
// With kMaxDepth (uniform), also the largest pair depth^2 of every replica
// (t.maxd).
template <int NT>
__device__ void tile_energy(const TileLds& t, int n, int G, int M, int m, float inv_w,
                            float* out /* G floats, LDS or global */, bool kMaxDepth = false) {
    const int tid = threadIdx.x;
    for (int g = tid; g < G; g += NT) {
        t.shx[g] = 0.0f;
        t.shy[g] = 0.0f;
        if (kMaxDepth) t.maxd[g] = 0u;
    }
    tile_sync<NT>();
    bin_tile<NT>(t, n, G, M, m, inv_w);
    float* e = reinterpret_cast<float*>(t.rank);   // rank + items: 4 bytes per square
    const int gw = grid_words(M);
    for (int q = tid; q < G * n; q += NT) {
        const int g = rep_of(t, q), i = q - g * n;
        const float4 p = t.pos[q];
        const int ck = t.cellk[q];
        const int kx = ck & 255, ky = ck >> 8;
        const Stencil sc(t.A + g * gw, M, kx, ky);
        const float4* pg = t.pos + g * n;
        float ei = wall32(p.x, p.y, p.z, p.w, t.Lr[g]);
        float md = 0.0f;
        // four candidates' index and position reads in flight at once, then
        // their terms added in candidate order
        for (int f0 = 0; f0 < sc.total; f0 += 4) {
            int kk[4], jj[4];
            float4 vv[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) kk[u] = sc.at(min(f0 + u, sc.total - 1));
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                jj[u] = t.sorted[kk[u]];
                // with the cell-ordered copy the position does not wait for the index
                if (t.spos) vv[u] = t.spos[kk[u]];
            }
            if (!t.spos) {
#pragma unroll
                for (int u = 0; u < 4; ++u) vv[u] = pg[jj[u]];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (f0 + u >= sc.total || jj[u] <= i) continue;
                const float4 v = vv[u];
                const float d2 = depth2(v.x - p.x, v.y - p.y, p.z, p.w, v.z, v.w);
                ei += d2;
                if (kMaxDepth) md = fmaxf(md, d2);
            }
        }
        e[q] = ei;
        if (kMaxDepth && md > 0.0f) atomicMax(&t.maxd[g], __float_as_uint(md));
    }
    tile_sync<NT>();
    if (NT == 64 && G * n <= 64) {
        // one square per lane (q = lane): the same tree by lane shifts
        const int q = tid, g = q < G * n ? rep_of(t, q) : 0, i = q - g * n;
        float v = q < G * n ? e[q] : 0.0f;
        for (int s = 1; s < n; s <<= 1) {
            const float o = __shfl_down(v, (unsigned)s, 64);
            if (q < G * n && (i & (2 * s - 1)) == 0 && i + s < n) v += o;
        }
        if (q < G * n && i == 0) out[g] = v;
        tile_sync<NT>();
        return;
    }
    for (int s = 1; s < n; s <<= 1) {
        for (int q = tid; q < G * n; q += NT) {
            const int g = rep_of(t, q), i = q - g * n;
            if ((i & (2 * s - 1)) == 0 && i + s < n) e[q] += e[q + s];
        }
        tile_sync<NT>();
    }
    for (int g = tid; g < G; g += NT) out[g] = e[g * n];
    tile_sync<NT>();
}

// ---------------------------------------------------------------------------
// Collective moves (cuda/packing.cu:564-616) and the kick (:696-745), applied
// to a tile after its sweeps.  Every draw is Philox on (index, replica,
// sweep or interval, purpose 5..8); the energies are the canonical FP32 ones.
// ---------------------------------------------------------------------------
// (cos a, sin a) for a in [0, pi/2) without a transcendental: the polynomial
// rotation of (cos pi/4, sin pi/4) by a - pi/4.
__device__ __forceinline__ void angle_cs(float a, float& c, float& s) {
    rot32(0.70710678f, 0.70710678f, a - 0.78539816f, c, s);
}

// Does square z join the group of a collective move anchored at w?  Modes:
// 1 row, 2 column, 3 cluster, 4 same orientation (within 0.06 rad of the
// anchor's modulo pi/2), 5 the two diagonals, 6 tilted squares (more than
// 0.04 rad from axis-aligned).
__device__ __forceinline__ bool coll_take(const CollWork& w, float4 z) {
    const float dx = z.x - w.cx, dy = z.y - w.cy;
    switch (w.mode) {
        case 1: return fabsf(dy) < 0.55f;
        case 2: return fabsf(dx) < 0.55f;
        case 3: return __builtin_fmaf(dx, dx, dy * dy) < 2.5f;
        case 4: {
            const float cd = __builtin_fmaf(z.z, w.cc, z.w * w.cs);
            const float sd = __builtin_fmaf(z.w, w.cc, -(z.z * w.cs));
            return fmaxf(fabsf(cd), fabsf(sd)) > 0.99820054f;  // cos 0.06
        }
        case 5: return fabsf(dx + dy) < 0.85f || fabsf(dx - dy) < 0.85f;
        default: return fmaxf(fabsf(z.z), fabsf(z.w)) < 0.99920011f;  // cos 0.04
    }
}

// One collective proposal for every replica of the tile.  t.eout holds the
// canonical energy on entry and on exit.  The configuration before the move
// is kept in the replica's own HBM slot, written and read back by the same
// thread.
template <int NT>
__device__ void collective_step(const TileLds& t, int n, int G, int M, int m, float inv_w, int r0, uint32_t sweep_end, uint32_t c, uint32_t k0, uint32_t k1,
                                float4* gstate, unsigned& tried, unsigned& accd) {
    const int tid = threadIdx.x;
    for (int g = tid; g < G; g += NT) {
        CollWork& w = t.cw[g];
        const uint4 o1 = philox(c, (uint32_t)(r0 + g), sweep_end, 5u, k0, k1);
        const uint4 o2 = philox(c, (uint32_t)(r0 + g), sweep_end, 6u, k0, k1);
        w.anchor = min(n - 1, (int)(u01(o1.x) * (float)n));
        w.mode = 1 + min(5, (int)(u01(o1.y) * 6.0f));
        const float4 a = t.pos[g * n + w.anchor];
        w.cx = a.x; w.cy = a.y; w.cc = a.z; w.cs = a.w;
        w.mv = t.sxy[g] * __builtin_fmaf(2.0f, u01(o1.z), -1.0f);
        w.rt = t.sxy[g] * __builtin_fmaf(2.0f, u01(o1.w), -1.0f);
        w.th = t.sth[g] * __builtin_fmaf(2.0f, u01(o2.x), -1.0f);
        rot32(1.0f, 0.0f, w.th, w.tc, w.ts);
        w.uacc = u01(o2.y);
        w.eold = t.eout[g];
    }
    tile_sync<NT>();
    for (int q = tid; q < G * n; q += NT) {
        const int g = rep_of(t, q), i = q - g * n;
        const CollWork& w = t.cw[g];
        const float4 z = t.pos[q];
        gstate[q] = z;
        if (!coll_take(w, z)) continue;
        const float dx = z.x - w.cx, dy = z.y - w.cy;
        float nx, ny;
        if (w.mode == 3 || w.mode == 4) {  // rotate the group about the anchor, then shift
            nx = (w.cx + __builtin_fmaf(w.tc, dx, -(w.ts * dy))) + w.mv;
            ny = (w.cy + __builtin_fmaf(w.ts, dx, w.tc * dy)) + w.rt;
        } else if (w.mode == 6) {          // split: alternate squares move apart
            const float sg = (i & 1) ? 1.0f : -1.0f;
            nx = z.x + sg * w.mv;
            ny = z.y + sg * w.rt;
        } else {
            nx = z.x + w.mv;
            ny = z.y + w.rt;
        }
        float nc, ns;
        rot32(z.z, z.w, w.th, nc, ns);
        const float L = t.Lr[g];
        t.pos[q] = make_float4(fminf(L - 0.45f, fmaxf(0.45f, nx)), fminf(L - 0.45f, fmaxf(0.45f, ny)), nc, ns);
    }
    tile_sync<NT>();
    tile_energy<NT>(t, n, G, M, m, inv_w, t.eout);
    for (int g = tid; g < G; g += NT) {
        CollWork& w = t.cw[g];
        const float en = t.eout[g], dE = en - w.eold;
        w.accept = (dE <= 0.0f || w.uacc < accept_exp2(-dE * t.ik[g])) ? 1 : 0;
        t.eout[g] = w.accept ? en : w.eold;
        tried += 1;
        accd += (unsigned)w.accept;
    }
    tile_sync<NT>();
    for (int q = tid; q < G * n; q += NT)
        if (!t.cw[rep_of(t, q)].accept) t.pos[q] = gstate[q];
    tile_sync<NT>();
}

// Kick of every replica with energy (cuda/packing.cu:696-745): per-square
// energies (wall + every stencil pair), the worst square (first on ties); on
// a relocating kick, when its energy exceeds 1e-5, it moves to the probe of
// an 8 x 8 grid farthest from every other centre (first on ties) with a
// +-0.05 jitter and a uniform angle; then every square is jittered, squares
// with energy above 1e-6 by +-0.08 / +-0.18 rad, the rest by +-0.008 /
// +-0.015 rad, centres clamped to [0.48, L - 0.48].  t.eout holds the
// canonical energy on entry and on exit.  With skip_certified, replicas that
// the certified-side pass just certified (t.maxd marked) keep their state:
// the exchange step may archive it as the best configuration.
constexpr unsigned kCertifiedMark = 0xFFFFFFFFu;
template <int NT>
__device__ void kick(const TileLds& t, int n, int G, int M, int m, float inv_w, int r0,
                     uint32_t interval, bool relocate, uint32_t k0, uint32_t k1, unsigned& kicked,
                     bool skip_certified) {
    const int tid = threadIdx.x, gw = grid_words(M);
    auto live = [&](int g) { return t.eout[g] > 1e-12f && !(skip_certified && t.maxd[g] == kCertifiedMark); };
    float* sq = reinterpret_cast<float*>(t.rank);  // rank + items: 4 bytes per square
    for (int g = tid; g < G; g += NT) {
        t.gmask[g] = 0ull;
        t.shx[g] = 0.0f;
        t.shy[g] = 0.0f;
    }
    tile_sync<NT>();
    bin_tile<NT>(t, n, G, M, m, inv_w);  // cell lists of the current positions
    for (int q = tid; q < G * n; q += NT) {
        const int g = rep_of(t, q), i = q - g * n;
        if (!live(g)) continue;
        const float4 p = t.pos[q];
        const int ck = t.cellk[q];
        const Stencil sc(t.A + g * gw, M, ck & 255, ck >> 8);
        const float4* pg = t.pos + (size_t)g * n;
        float ei = wall32(p.x, p.y, p.z, p.w, t.Lr[g]);
        for (int f = 0; f < sc.total; ++f) {
            const int j = t.sorted[sc.at(f)];
            if (j == i) continue;
            const float4 v = pg[j];
            ei += depth2(v.x - p.x, v.y - p.y, p.z, p.w, v.z, v.w);
        }
        sq[q] = ei;
        atomicMax(&t.gmask[g], ((unsigned long long)__float_as_uint(ei) << 32) | (0xFFFFFFFFu - (unsigned)i));
    }
    tile_sync<NT>();
    for (int g = tid; g < G; g += NT) {
        CollWork& w = t.cw[g];
        const unsigned long long k = t.gmask[g];
        w.worst = (int)(0xFFFFFFFFu - (unsigned)(k & 0xFFFFFFFFull));
        w.eold = __uint_as_float((unsigned)(k >> 32));
        w.accept = relocate && live(g) && w.eold > 1e-5f;  // relocate this replica
        t.gmask[g] = 0ull;
        if (live(g)) kicked += 1;
    }
    tile_sync<NT>();
    if (relocate) {
        for (int k = tid; k < G * 64; k += NT) {
            const int g = k >> 6, pidx = k & 63;
            const CollWork& w = t.cw[g];
            if (!w.accept) continue;
            const float L = t.Lr[g];
            const float vx = __builtin_fmaf(L - 1.1f, ((pidx & 7) + 0.5f) * 0.125f, 0.55f);
            const float vy = __builtin_fmaf(L - 1.1f, ((pidx >> 3) + 0.5f) * 0.125f, 0.55f);
            const float4* pg = t.pos + (size_t)g * n;
            float d2 = 1e9f;
            for (int j = 0; j < n; ++j) {
                if (j == w.worst) continue;
                const float4 v = pg[j];
                const float dx = vx - v.x, dy = vy - v.y;
                d2 = fminf(d2, __builtin_fmaf(dx, dx, dy * dy));
            }
            atomicMax(&t.gmask[g], ((unsigned long long)__float_as_uint(d2) << 32) | (0xFFFFFFFFu - (unsigned)pidx));
        }
        tile_sync<NT>();
        for (int g = tid; g < G; g += NT) {
            const CollWork& w = t.cw[g];
            if (!w.accept) continue;
            const float L = t.Lr[g];
            const int pidx = (int)(0xFFFFFFFFu - (unsigned)(t.gmask[g] & 0xFFFFFFFFull));
            const float vx = __builtin_fmaf(L - 1.1f, ((pidx & 7) + 0.5f) * 0.125f, 0.55f);
            const float vy = __builtin_fmaf(L - 1.1f, ((pidx >> 3) + 0.5f) * 0.125f, 0.55f);
            const uint4 o = philox(0u, (uint32_t)(r0 + g), interval, 7u, k0, k1);
            const float x = fminf(L - 0.5f, fmaxf(0.5f, vx + 0.05f * __builtin_fmaf(2.0f, u01(o.x), -1.0f)));
            const float y = fminf(L - 0.5f, fmaxf(0.5f, vy + 0.05f * __builtin_fmaf(2.0f, u01(o.y), -1.0f)));
            float c, s;
            angle_cs(u01(o.z) * 1.57079633f, c, s);
            t.pos[g * n + w.worst] = make_float4(x, y, c, s);
        }
        tile_sync<NT>();
    }
    for (int q = tid; q < G * n; q += NT) {
        const int g = rep_of(t, q), i = q - g * n;
        if (!live(g)) continue;
        const bool hot = sq[q] > 1e-6f;
        const float axy = hot ? 0.08f : 0.008f, aa = hot ? 0.18f : 0.015f;
        const uint4 o = philox((uint32_t)i, (uint32_t)(r0 + g), interval, 8u, k0, k1);
        const float4 p = t.pos[q];
        const float L = t.Lr[g];
        const float x = fminf(L - 0.48f, fmaxf(0.48f, p.x + axy * __builtin_fmaf(2.0f, u01(o.x), -1.0f)));
        const float y = fminf(L - 0.48f, fmaxf(0.48f, p.y + axy * __builtin_fmaf(2.0f, u01(o.y), -1.0f)));
        float c, s;
        rot32(p.z, p.w, aa * __builtin_fmaf(2.0f, u01(o.z), -1.0f), c, s);
        t.pos[q] = make_float4(x, y, c, s);
    }
    tile_sync<NT>();
    tile_energy<NT>(t, n, G, M, m, inv_w, t.eout);
}

// ---------------------------------------------------------------------------
// FP64 certified side of every tile replica (cuda/packing.cu:254-297), from
// the cell lists of the final positions.  The dilation is the largest
// (R + tau) / vmax over pairs nearer than sqrt(2.05); a pair in cells more
// than two apart (2w > sqrt(2) + 6e-4) cannot overlap and needs a dilation
// below 1, so the 5 x 5 stencil finds every pair that raises it.
// ---------------------------------------------------------------------------
// Order-preserving map of a double onto 64-bit unsigned integers, so LDS
// integer min / max atomics order doubles.
__device__ __forceinline__ unsigned long long dkey(double v) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
__device__ __forceinline__ double dval(unsigned long long k) {
    return __longlong_as_double((long long)((k >> 63) ? (k & 0x7FFFFFFFFFFFFFFFull) : ~k));
}
// FP64 cos / sin of a stored FP32 orientation, scaled to unit length.
__device__ __forceinline__ void unit64(float c, float s, double& cd, double& sd) {
    const double x = c, y = s;
    const double h = 1.0 / sqrt(x * x + y * y);
    cd = x * h;
    sd = y * h;
}
// Dilation a pair needs to be separated by tau (cuda/packing.cu:266-271).
__device__ __forceinline__ double cert_need(double dx, double dy, double ci, double si, double cj, double sj) {
    const double r = 0.5 + 0.5 * (fabs(ci * cj + si * sj) + fabs(-si * cj + ci * sj));
    const double v1 = fabs(dx * ci + dy * si), v2 = fabs(-dx * si + dy * ci);
    const double v3 = fabs(dx * cj + dy * sj), v4 = fabs(-dx * sj + dy * cj);
    const double vmax = fmax(fmax(v1, v2), fmax(v3, v4));
    return vmax > 1e-12 ? (r + kCertTau) / vmax : 1e9;
}
// Bounds that keep the FP64 work to the pairs that can matter, both exact:
//  - a replica with an FP32 pair depth above kCertSkipDepth2 (depth 0.00975)
//    needs a dilation above 1 + 0.00975 / R >= 1 + 0.00975 / 1.2072 > 1.008
//    (an FP32 depth is within 1e-6 of the FP64 one), so it has no certified
//    side;
//  - a pair whose FP32 gap is below -1e-5 is separated by more than
//    1e-5 - 1e-6 on its best axis, so its dilation (R + tau) / vmax is below 1.
constexpr float kCertSkipDepth2 = 9.5e-5f;
constexpr float kCertPairGap = -1e-5f;

// Order-preserving keys of a running min / max, flushed to LDS atomics when
// the thread's next square belongs to another replica.
struct SpanAcc {
    int g = -1;
    double lox = 1e300, hix = -1e300, loy = 1e300, hiy = -1e300;
    __device__ __forceinline__ void flush(unsigned long long* acc, int stride) {
        if (g < 0) return;
        unsigned long long* a = acc + stride * g;
        atomicMin(a + 1, dkey(lox));
        atomicMax(a + 2, dkey(hix));
        atomicMin(a + 3, dkey(loy));
        atomicMax(a + 4, dkey(hiy));
    }
};

// rs: the tile's first replica record.  Needs the cell lists (shift 0) and
// t.maxd of the final positions.  Writes Lc (1e9 when the dilation exceeds
// 1.008), and for certified replicas lam, minx, miny; certs counts them.
template <int NT>
__device__ void tile_certify(const TileLds& t, int n, int G, int M, RepState* rs, unsigned& certs) {
    static_assert(sizeof(CollWork) >= 5 * 8, "the per-replica scratch holds five 64-bit keys");
    const int tid = threadIdx.x, gw = grid_words(M);
    // per replica (the collective-move scratch, free now): dilation, min x,
    // max x, min y, max y as ordered keys
    unsigned long long* acc = reinterpret_cast<unsigned long long*>(t.cw);
    constexpr int kStride = (int)(sizeof(CollWork) / 8);
    for (int g = tid; g < G; g += NT) {
        unsigned long long* a = acc + kStride * g;
        a[0] = dkey(1.0);
        a[1] = ~0ull;
        a[2] = 0ull;
        a[3] = ~0ull;
        a[4] = 0ull;
    }
    tile_sync<NT>();
    for (int q = tid; q < G * n; q += NT) {
        const int g = rep_of(t, q), i = q - g * n;
        if (__uint_as_float(t.maxd[g]) > kCertSkipDepth2) continue;
        const float4 p = t.pos[q];
        const int ck = t.cellk[q];
        const Stencil sc(t.A + g * gw, M, ck & 255, ck >> 8);
        const float4* pg = t.pos + (size_t)g * n;
        double ci = 0.0, si = 0.0, lam = 1.0;
        bool have = false;
        // four candidates' reads in flight at once (the maximum does not
        // depend on the order)
        for (int f0 = 0; f0 < sc.total; f0 += 4) {
            int kk[4], jj[4];
            float4 vv[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) kk[u] = sc.at(min(f0 + u, sc.total - 1));
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                jj[u] = t.sorted[kk[u]];
                if (t.spos) vv[u] = t.spos[kk[u]];
            }
            if (!t.spos) {
#pragma unroll
                for (int u = 0; u < 4; ++u) vv[u] = pg[jj[u]];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (f0 + u >= sc.total || jj[u] <= i) continue;
                const float4 v = vv[u];
                const float fx = v.x - p.x, fy = v.y - p.y;
                if (!(__builtin_fmaf(fx, fx, fy * fy) < 2.06f)) continue;
                if (!(sat_gap(fx, fy, p.z, p.w, v.z, v.w) > kCertPairGap)) continue;
                const double dx = (double)p.x - (double)v.x, dy = (double)p.y - (double)v.y;
                if (dx * dx + dy * dy >= kCertNear2) continue;
                if (!have) {
                    unit64(p.z, p.w, ci, si);
                    have = true;
                }
                double cj, sj;
                unit64(v.z, v.w, cj, sj);
                lam = fmax(lam, cert_need(dx, dy, ci, si, cj, sj));
            }
        }
        if (lam > 1.0) atomicMax(acc + kStride * g, dkey(lam));
    }
    tile_sync<NT>();
    SpanAcc sp;
    for (int q = tid; q < G * n; q += NT) {
        const int g = rep_of(t, q);
        if (__uint_as_float(t.maxd[g]) > kCertSkipDepth2) continue;
        const double lam = dval(acc[kStride * g]);
        if (lam > kCertMaxLam) continue;
        if (g != sp.g) {
            sp.flush(acc, kStride);
            sp = SpanAcc{};
            sp.g = g;
        }
        const float4 p = t.pos[q];
        double ci, si;
        unit64(p.z, p.w, ci, si);
        const double r = 0.5 * (fabs(ci) + fabs(si));
        const double x = (double)p.x * lam, y = (double)p.y * lam;
        sp.lox = fmin(sp.lox, x - r);
        sp.hix = fmax(sp.hix, x + r);
        sp.loy = fmin(sp.loy, y - r);
        sp.hiy = fmax(sp.hiy, y + r);
    }
    sp.flush(acc, kStride);
    tile_sync<NT>();
    for (int g = tid; g < G; g += NT) {
        const unsigned long long* a = acc + kStride * g;
        const bool skip = __uint_as_float(t.maxd[g]) > kCertSkipDepth2;
        const double lam = skip ? kCertNone : dval(a[0]);
        const bool ok = lam <= kCertMaxLam;
        RepState* r = rs + g;
        r->Lc = ok ? fmax(dval(a[2]) - dval(a[1]), dval(a[4]) - dval(a[3])) : kCertNone;
        r->lam = lam;
        r->minx = ok ? dval(a[1]) : 0.0;
        r->miny = ok ? dval(a[3]) : 0.0;
        certs += ok ? 1u : 0u;
        if (ok) t.maxd[g] = kCertifiedMark;  // the kick after this pass leaves it alone
    }
    tile_sync<NT>();
}

// ---------------------------------------------------------------------------
// k_sweep: `spe` sweeps (or what is left of the budget) of every replica of
// one tile.  One sweep = every square proposes exactly once.  NT threads: one
// wavefront, or four for large instances (the same arithmetic either way).
// ---------------------------------------------------------------------------
// Moves of a sweep that keeps the cell lists of the previous one (same grid
// shift): every square's staged move is drawn again from its current position,
// either into prop / pw or into its mover record j (one-wavefront tiles with
// records; square, CSR slot, replica, 1/T and stencil stay).
// The staged proposal of square q as level item j.
__device__ __forceinline__ void stage_item(const TileLds& t, int j, int q, int n, int m, float inv_w, int r0,
                                           uint32_t sweep, uint32_t k0, uint32_t k1) {
    const int g = rep_of(t, q), i = q - g * n;
    const int ck = t.cellk[q];
    const float4 r = move_variates((uint32_t)i, (uint32_t)(r0 + g), sweep, k0, k1);
    Move mv;
    const bool ok = make_move(t, g, t.pos[q], r, ck & 255, ck >> 8, m, inv_w, mv);
    t.prop[j] = ok ? make_float4(mv.nx, mv.ny, mv.nc, mv.ns) : make_float4(0.0f, 0.0f, kNoMove, 0.0f);
    t.pw[j] = make_float4(ok ? mv.wdiff : 0.0f, r.w, __uint_as_float((unsigned)q | (unsigned)t.slotq[q] << 16),
                          __uint_as_float((unsigned)ck));
}
__device__ __forceinline__ void restage_record(const TileLds& t, int j, int n, int m, float inv_w, int r0, uint32_t sweep, uint32_t k0, uint32_t k1) {
    uint4* w = t.rec + (size_t)j * kRecWords;
    const uint4 w2 = w[2];
    const int g = (int)w[4].y;
    const int q = (int)(w2.w & 0xFFFFu);
    const float4 p = t.pos[q];
    const int ck = t.cellk[q];
    const float4 r = move_variates((uint32_t)(q - g * n), (uint32_t)(r0 + g), sweep, k0, k1);
    Move mv;
    const bool ok = make_move(t, g, p, r, ck & 255, ck >> 8, m, inv_w, mv);
    w[0] = make_uint4(__float_as_uint(p.x), __float_as_uint(p.y), __float_as_uint(p.z), __float_as_uint(p.w));
    w[1] = ok ? make_uint4(__float_as_uint(mv.nx), __float_as_uint(mv.ny), __float_as_uint(mv.nc),
                           __float_as_uint(mv.ns))
              : make_uint4(0u, 0u, __float_as_uint(kNoMove), 0u);
    w[2] = make_uint4(__float_as_uint(ok ? mv.wdiff : 0.0f), __float_as_uint(r.w), w2.z, w2.w);
}

// Sweep set-up of a one-wavefront tile of at most 64 squares, one square per
// lane (q = lane): the binning of bin_tile, the group keys, the levels, the
// level starts and the mover records of the general set-up in k_sweep, with
// each square's position, cell, CSR slot and rank kept in registers between
// the phases instead of read back from LDS, and its move's variates drawn
// while the first reads are in flight.  Leaves the same LDS state and returns
// the level starts (lane l: start of level l; lane kGroups + 1: the total).
__device__ int setup_lane_squares(const TileLds& t, int n, int G, int M, int m, float inv_w, float w,
                                  int r0, uint32_t sweep, bool next_shift, uint32_t next_sweep, uint32_t k0,
                                  uint32_t k1) {
    const int lane = threadIdx.x;
    const bool has = lane < G * n;
    const int gw = grid_words(M);
    const int q = lane;
    const int g = has ? rep_of(t, q) : 0;
    const int i = q - g * n;
    uint16_t* Ag = t.A + g * gw;
    float4 p = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    float shx = 0.0f, shy = 0.0f, sxy = 0.0f, sth = 0.0f, ik = 0.0f, L = 0.0f;
    if (has) {
        p = t.pos[q];
        shx = t.shx[g];
        shy = t.shy[g];
        sxy = t.sxy[g];
        sth = t.sth[g];
        ik = t.ik[g];
        L = t.Lr[g];
    }
    const float4 r = move_variates((uint32_t)i, (uint32_t)(r0 + g), sweep, k0, k1);
    if (lane < kGroups + 2) t.gcnt[lane] = 0;
    if (lane < G) t.gmask[lane] = 0ull;
    uint4* A128 = reinterpret_cast<uint4*>(t.A);
    for (int k = lane; k < (G * gw) / 8; k += kSweepThreads) A128[k] = make_uint4(0u, 0u, 0u, 0u);
    int kx = 0, ky = 0, c = 0;
    if (has) {
        kx = cellcoord(p.x, shx, inv_w, 0, m - 1) + 2;
        ky = cellcoord(p.y, shy, inv_w, 0, m - 1) + 2;
        c = ky * M + kx;
        t.cellk[q] = (uint16_t)(kx | (ky << 8));
    }
    wave_sync();
    if (has) add_u16(Ag, c + 2, 1u);
    wave_sync();
    wave_scan_u16(t.A, G * gw);
    wave_sync();
    if (has) t.sorted[add_u16(Ag, c + 1, 1u)] = (uint16_t)i;
    wave_sync();
    // the cell starts are final: the stencil's row ranges are read with the
    // cell's members
    Stencil sc;
    int a = 0, rk = 0;
    if (has) {
        sc = Stencil(Ag, M, kx, ky);
        a = Ag[c];
        const int b = Ag[c + 1];
        for (int k = a; k < b; ++k) rk += t.sorted[k] < i;
    }
    wave_sync();
    const int self = a + rk;
    int key = kGroups;
    if (has) {
        t.sorted[self] = (uint16_t)i;
        t.spos[self] = p;
        t.slotq[q] = (uint16_t)self;
        const int colour = (kx % 3) + 3 * (ky % 3);
        key = rk < kRankCap ? rk * 9 + colour : kGroups;
        atomicOr(&t.gmask[g], 1ull << key);
    }
    if (next_shift && lane < G) {
        const uint4 o = philox(0u, (uint32_t)(r0 + lane), next_sweep, 1u, k0, k1);
        t.shx_next[lane] = w * u01(o.x);
        t.shy_next[lane] = w * u01(o.y);
    }
    wave_sync();
    int lvl = kGroups;
    if (has) {
        if (key < kGroups) lvl = __popcll(t.gmask[g] & ((1ull << key) - 1ull));
        t.rank[q] = (uint16_t)lvl;
        atomicAdd(&t.gcnt[lvl], 1);
    }
    wave_sync();
    const int cnt = lane <= kGroups ? t.gcnt[lane] : 0;
    const int gs = wave_incl_scan(cnt) - cnt;
    if (lane <= kGroups) t.gcnt[lane] = gs;
    wave_sync();
    if (has && lvl < kGroups)
        write_record(t, atomicAdd(&t.gcnt[lvl], 1), q, g, self, p, kx, ky, sc, r, sxy, sth, shx, shy, ik, m,
                     inv_w, L);
    wave_sync();
    return gs;
}

// W = 4 waves per SIMD: at most 128 VGPRs (the tiles under load need the
// residency more than the kernel needs registers).  W = 1 for launches whose
// tiles are planned for an idle device (one workgroup per compute unit): the
// registers of a whole SIMD, so nothing spills (the same arithmetic).
template <int NT, bool kNear, int W>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(W))) void k_sweep(const Tile* __restrict__ tiles,
                                              const InstDesc* __restrict__ insts,
                                              RepState* __restrict__ reps,
                                              const SlotState* __restrict__ slots,
                                              float4* __restrict__ state) {
    extern __shared__ __align__(16) unsigned char smem[];
    const Tile tl = tiles[blockIdx.x];
    const InstDesc* D = insts + tl.inst;
    const long long left = D->sweeps_total - D->sweeps_done;
    const int nsw = (int)min((long long)D->spe, left);
    if (nsw <= 0) return;
    const int n = D->n, G = tl.g, M = D->M, tid = threadIdx.x, lane = tid & 63;
    const float w = D->w, scale = D->scale;
    const float inv_w = 1.0f / w;
    const int m = grid_cells(D->Lf, w);  // every replica's box is at most Lf
    const uint32_t k0 = D->key0, k1 = D->key1;
    const uint32_t sweep0 = (uint32_t)D->sweeps_done;
    const int ncoll = D->collective, kick_every = D->kick_every;
    const uint32_t interval = D->interval;
    const uint32_t spsh = (uint32_t)max(1, D->spsh);
    // every descriptor field the launch uses is read here: the sweep never
    // writes the descriptor
    const bool improve = D->improve != 0;
    const uint32_t certify_every = (uint32_t)max(1, D->certify_every);
    const int rep_base = D->rep_off + tl.r0;
    TileLds t = carve(smem, n, M, G, tl.wide != 0);
    float4* gstate = state + D->state_off + (size_t)tl.r0 * n;

    // the state and the replica records load together; the clamp into each
    // replica's box follows in LDS
    for (int q = tid; q < G * n; q += NT) t.pos[q] = gstate[q];
    for (int g = tid; g < G; g += NT) {
        const int slot = reps[rep_base + g].slot;
        const float4 ss = *reinterpret_cast<const float4*>(slots + D->rep_off + slot);
        t.ik[g] = ss.x;
        t.sxy[g] = ss.y;
        t.sth[g] = ss.z;
        t.Lr[g] = ss.w;  // the box of the slot the replica holds
        t.acc[g] = 0;
        // grid shift of the first sweep's block (blocks of spsh sweeps share
        // one shift, drawn at the block's first sweep counter; the next
        // block's is drawn ahead)
        const uint4 o = philox(0u, (uint32_t)(tl.r0 + g), sweep0 / spsh * spsh, 1u, k0, k1);
        t.shx[g] = w * u01(o.x);
        t.shy[g] = w * u01(o.y);
    }
    tile_sync<NT>();
    // staged state: centres rescaled (after a box shrink) and clamped into the
    // replica's box
    for (int q = tid; q < G * n; q += NT) {
        const float L = t.Lr[rep_of(t, q)];
        float4 v = t.pos[q];
        v.x = fminf(fmaxf(v.x * scale, 0.0f), L);
        v.y = fminf(fmaxf(v.y * scale, 0.0f), L);
        t.pos[q] = v;
    }
    tile_sync<NT>();

    unsigned long long cC = 0, cP = 0;
    int gs = 0;  // level starts (lane l: level l), kept with the cell lists
    for (int sw = 0; sw < nsw; ++sw) {
        const uint32_t sweep = sweep0 + (uint32_t)sw;
        const uint32_t next_block = (sweep / spsh + 1u) * spsh;
        const bool next_in_launch = (long long)next_block < (long long)sweep0 + nsw;
        if (sw > 0 && sweep % spsh != 0u) {
            // same grid shift as the previous sweep: no move left its cell, so
            // the cell lists, ranks and levels still hold; only the moves are
            // drawn again
            if (NT == 64 && t.rec) {
                const int items = __builtin_amdgcn_readlane(gs, kGroups);
                for (int j = tid; j < items; j += NT) restage_record(t, j, n, m, inv_w, tl.r0, sweep, k0, k1);
            } else if (t.prop) {
                const int items = __builtin_amdgcn_readlane(gs, kGroups);
                for (int j = tid; j < items; j += NT)
                    stage_item(t, j, t.items[j], n, m, inv_w, tl.r0, sweep, k0, k1);
            }
            tile_sync<NT>();
        } else if (NT == 64 && t.rec && G * n <= kSweepThreads) {
            if constexpr (NT == 64)
                gs = setup_lane_squares(t, n, G, M, m, inv_w, w, tl.r0, sweep, next_in_launch, next_block, k0,
                                        k1);
        } else {
            for (int k = tid; k < kGroups + 2; k += NT) t.gcnt[k] = 0;
            for (int g = tid; g < G; g += NT) t.gmask[g] = 0ull;
            tile_sync<NT>();
            bin_tile<NT>(t, n, G, M, m, inv_w);
            // group key: occupancy rank x colour; ranks >= kRankCap go serial.
            // Tiles with staged proposals or mover records draw each square's
            // move (one Philox block per square; a square's own state does not
            // change before its level runs) when its level item is written.
            for (int q = tid; q < G * n; q += NT) {
                const int g = rep_of(t, q);
                const int ck = t.cellk[q];
                const int colour = ((ck & 255) % 3) + 3 * ((ck >> 8) % 3);
                const int r = t.rank[q];
                const int key = r < kRankCap ? r * 9 + colour : kGroups;
                t.rank[q] = (uint16_t)key;
                atomicOr(&t.gmask[g], 1ull << key);
            }
            if (next_in_launch)
                for (int g = tid; g < G; g += NT) {
                    const uint4 o = philox(0u, (uint32_t)(tl.r0 + g), next_block, 1u, k0, k1);
                    t.shx_next[g] = w * u01(o.x);
                    t.shy_next[g] = w * u01(o.y);
                }
            tile_sync<NT>();
            // level = position of the square's group among its replica's occupied
            // groups: each replica walks only its own non-empty groups, in key
            // order, and one pass serves level l of every replica of the tile
            for (int q = tid; q < G * n; q += NT) {
                const int key = t.rank[q];
                const int lvl =
                    key < kGroups ? __popcll(t.gmask[rep_of(t, q)] & ((1ull << key) - 1ull)) : kGroups;
                t.rank[q] = (uint16_t)lvl;
                atomicAdd(&t.gcnt[lvl], 1);
            }
            tile_sync<NT>();
            // level starts: one wave-wide exclusive scan of the level counts; lane
            // l keeps the start of level l (lane kGroups + 1: the total), and the
            // counts become the scatter cursors.  The other wavefronts of a large
            // tile read the starts back from LDS.
            static_assert(kGroups + 2 <= 64, "level starts must fit one wavefront");
            if (tid < 64) {
                const int c = lane <= kGroups ? t.gcnt[lane] : 0;
                gs = wave_incl_scan(c) - c;
                if (lane <= kGroups) t.gcnt[lane] = gs;
                if (NT > 64 && lane <= kGroups + 1) t.gstart[lane] = gs;
            }
            tile_sync<NT>();
            if (NT > 64) gs = lane <= kGroups + 1 ? t.gstart[lane] : 0;
            for (int q = tid; q < G * n; q += NT) {
                const int lvl = t.rank[q];
                if (lvl < kGroups) {
                    const int j = atomicAdd(&t.gcnt[lvl], 1);
                    if (NT == 64 && t.rec) {
                        write_record(t, j, q, n, M, m, inv_w, tl.r0, sweep, k0, k1);
                    } else {
                        t.items[j] = (uint16_t)q;
                        if (t.prop) stage_item(t, j, q, n, m, inv_w, tl.r0, sweep, k0, k1);
                    }
                }
            }
            tile_sync<NT>();
        }
        if constexpr (NT == 64) {
            if (t.rec) rec_levels<kNear>(t, n, gs, cC, cP);
        }
        // four-wavefront tiles with staged proposals read each level ahead (a
        // lone wavefront has no registers to spare for it)
        if constexpr (NT > 64) {
            if (t.prop) staged_levels<NT, kNear>(t, n, M, gs, cC, cP);
        }
        for (int lv = 0; lv < kGroups && !(NT == 64 && t.rec) && !(NT > 64 && t.prop); ++lv) {
            const int a = __builtin_amdgcn_readlane(gs, lv);
            const int b = __builtin_amdgcn_readlane(gs, lv + 1);
            if (a == b) break;  // levels are dense: the first empty one is the last
            // sparse levels give each mover 2, 4, 8 or 16 lanes (aligned
            // inside one row of 16 lanes of one wavefront)
            const int cnt = b - a;
            const int lsh = cnt > NT / 2 ? 0 : cnt > NT / 4 ? 1 : cnt > NT / 8 ? 2 : cnt > NT / 16 ? 3 : 4;
            const int lpi = 1 << lsh;  // a power of two: shifts, no division
            const int per = NT >> lsh;
            const int slot = tid >> lsh, sub = tid & (lpi - 1);
            for (int base = a; base < b; base += per) {
                if (base + slot >= b) continue;
                if (t.prop)
                    propose_item<(NT > 64), kNear>(t, base + slot, n, M, lpi, sub, cC, cP);
                else
                    propose<(NT > 64), kNear>(t, t.items[base + slot], n, M, m, inv_w, tl.r0, sweep, k0, k1, lpi,
                                              sub, cC, cP);
            }
            tile_sync<NT>();
        }
        if (__builtin_amdgcn_readlane(gs, kGroups + 1) > __builtin_amdgcn_readlane(gs, kGroups)) {
            if (tid == 0)
                for (int q = 0; q < G * n; ++q)
                    if (t.rank[q] == kGroups)
                        propose<(NT > 64), kNear>(t, q, n, M, m, inv_w, tl.r0, sweep, k0, k1, 1, 0, cC, cP);
            tile_sync<NT>();
        }
        if ((sweep + 1u) % spsh == 0u) {  // the next sweep starts a block: its shift becomes current
            float* tmp = t.shx;
            t.shx = t.shx_next;
            t.shx_next = tmp;
            tmp = t.shy;
            t.shy = t.shy_next;
            t.shy_next = tmp;
        }
    }

    // improve = 1: certified sides every certify_every launches and at the
    // last one; the other launches mark them absent
    const bool certify_now = improve && ((interval + 1u) % certify_every == 0u || (long long)nsw >= left);
    tile_energy<NT>(t, n, G, M, m, inv_w, t.eout, certify_now);
    unsigned ctried = 0, caccd = 0, kicked = 0;
    for (int c = 0; c < ncoll; ++c)
        collective_step<NT>(t, n, G, M, m, inv_w, tl.r0, sweep0 + (uint32_t)nsw, (uint32_t)c, k0, k1, gstate, ctried,
                            caccd);
    // The certified sides come before the kick, as the reference checkpoints
    // each chain before kicking it (cuda/packing.cu:689-745); a replica they
    // certify is not kicked, so that its state is the configuration they
    // certified.
    if (certify_now) {
        // the cell lists and largest overlaps must be those of the positions
        // after the collective moves (the canonical energy written again is
        // the same)
        if (ncoll > 0) tile_energy<NT>(t, n, G, M, m, inv_w, t.eout, true);
        unsigned certs = 0;
        tile_certify<NT>(t, n, G, M, reps + rep_base, certs);
    }
    const bool kick_now = kick_every > 0 && (interval + 1u) % (uint32_t)kick_every == 0u;
    if (kick_now)
        kick<NT>(t, n, G, M, m, inv_w, tl.r0, interval, (((interval + 1u) / (uint32_t)kick_every) & 1u) != 0u, k0, k1,
                 kicked, certify_now);
    for (int q = tid; q < G * n; q += NT) gstate[q] = t.pos[q];
    for (int g = tid; g < G; g += NT) {
        RepState* rs = reps + rep_base + g;
        rs->E = t.eout[g];
        rs->acc = (unsigned)t.acc[g];
        if (improve && !certify_now) rs->Lc = kCertNone;
    }
    // The tile's counters go to its first replica's record, stored (not
    // added: the exchange step after every launch moves them into the
    // instance's totals) by one thread, so that the launch makes no atomics
    // on the shared descriptor and waits on no global load at its end.
    cC = wave_reduce_u64<false>(cC);
    cP = wave_reduce_u64<false>(cP);
    ctried = wave_sum_u32(ctried);
    caccd = wave_sum_u32(caccd);
    kicked = wave_sum_u32(kicked);
    if constexpr (NT > 64) {
        // wavefront totals through the collective-move scratch (free now)
        unsigned long long* red = reinterpret_cast<unsigned long long*>(t.cw);
        static_assert(sizeof(CollWork) >= 5 * 8, "five 64-bit sums fit one replica's scratch");
        tile_sync<NT>();
        if (tid < 5) red[tid] = 0ull;
        tile_sync<NT>();
        if (lane == 0) {
            atomicAdd(red + 0, cC);
            atomicAdd(red + 1, cP);
            atomicAdd(red + 2, (unsigned long long)ctried);
            atomicAdd(red + 3, (unsigned long long)caccd);
            atomicAdd(red + 4, (unsigned long long)kicked);
        }
        tile_sync<NT>();
        cC = red[0];
        cP = red[1];
        ctried = (unsigned)red[2];
        caccd = (unsigned)red[3];
        kicked = (unsigned)red[4];
    }
    if (tid == 0) {
        RepState* r0s = reps + rep_base;
        r0s->cand = cC;
        r0s->near_pairs = cP;
        r0s->coll_tried = ctried;
        r0s->coll_acc = caccd;
        r0s->kicks = kicked;
    }
}

// ---------------------------------------------------------------------------
// Block reductions (fixed tree, deterministic).
// ---------------------------------------------------------------------------
template <int NT>
__device__ double block_sum_d(double v, double* tmp) {
    tmp[threadIdx.x] = v;
    __syncthreads();
    for (int s = NT / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) tmp[threadIdx.x] += tmp[threadIdx.x + s];
        __syncthreads();
    }
    double r = tmp[0];
    __syncthreads();
    return r;
}
template <int NT>
__device__ double block_max_d(double v, double* tmp) {
    tmp[threadIdx.x] = v;
    __syncthreads();
    for (int s = NT / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) tmp[threadIdx.x] = fmax(tmp[threadIdx.x], tmp[threadIdx.x + s]);
        __syncthreads();
    }
    double r = tmp[0];
    __syncthreads();
    return r;
}
template <int NT>
__device__ double block_min_d(double v, double* tmp) {
    tmp[threadIdx.x] = v;
    __syncthreads();
    for (int s = NT / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) tmp[threadIdx.x] = fmin(tmp[threadIdx.x], tmp[threadIdx.x + s]);
        __syncthreads();
    }
    double r = tmp[0];
    __syncthreads();
    return r;
}

// ---------------------------------------------------------------------------
// FP64 geometry (cuda/packing.cu:177-222).
// ---------------------------------------------------------------------------
__device__ __forceinline__ double pair64(double dx, double dy, double c, double s, double d,
                                         double t) {
    if (dx * dx + dy * dy >= 2.000001) return 0.0;
    const double cd = c * d + s * t, sd = s * d - c * t;
    const double R = .5 + .5 * (fabs(cd) + fabs(sd));
    const double v1 = fabs(dx * c + dy * s), v2 = fabs(-dx * s + dy * c);
    const double v3 = fabs(dx * d + dy * t), v4 = fabs(-dx * t + dy * d);
    const double gap = fmax(fmax(v1, v2), fmax(v3, v4)) - R;
    return gap >= 0.0 ? 0.0 : gap * gap;
}

__device__ __forceinline__ double sgn_dead(double v) {
    return fabs(v) > 1e-6 ? (v >= 0 ? 1.0 : -1.0) : 0.0;
}

// Energy and gradient w.r.t. square i of its overlap with j (active axis).
__device__ __forceinline__ void pair64_grad(double dx, double dy, double c, double s, double d,
                                            double t, double& gx, double& gy, double& ga) {
    if (dx * dx + dy * dy >= 2.000001) return;
    const double cd = c * d + s * t, sd = s * d - c * t;
    const double R = .5 + .5 * (fabs(cd) + fabs(sd));
    const double dR = .5 * (-sgn_dead(cd) * sd + sgn_dead(sd) * cd);
    const double nx[4] = {c, -s, d, -t}, ny[4] = {s, c, t, d};
    double gap = -1e100, ax = 0, ay = 0, aa = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const double v = dx * nx[k] + dy * ny[k];
        const double z = fabs(v) - R;
        if (z > gap) {
            gap = z;
            const double sv = v >= 0.0 ? 1.0 : -1.0;
            ax = -sv * nx[k];
            ay = -sv * ny[k];
            aa = -dR + (k < 2 ? sv * (-dx * ny[k] + dy * nx[k]) : 0.0);
        }
    }
    if (gap >= 0.0) return;
    gx += 2.0 * gap * ax;
    gy += 2.0 * gap * ay;
    ga += 2.0 * gap * aa;
}

__device__ __forceinline__ double wall64(double x, double y, double c, double s, double L) {
    const double r = 0.5 * (fabs(c) + fabs(s));
    double e = 0.0;
    if (x < r) { const double u = r - x; e += u * u; }
    else if (x > L - r) { const double u = x + r - L; e += u * u; }
    if (y < r) { const double u = r - y; e += u * u; }
    else if (y > L - r) { const double u = y + r - L; e += u * u; }
    return e;
}

__device__ __forceinline__ double wrap_quarter(double a) {
    const double quarter = 1.57079632679489661923;
    a = fmod(a, quarter);
    if (a < 0.0) a += quarter;
    return a;
}

// Total FP64 energy of the configuration held in LDS (all pairs i<j).
template <int NT>
__device__ double total64(const double* x, const double* y, const double* c, const double* s,
                          int n, double L, double* tmp) {
    double e = 0.0;
    for (int i = threadIdx.x; i < n; i += NT) {
        e += wall64(x[i], y[i], c[i], s[i], L);
        for (int j = i + 1; j < n; ++j) e += pair64(x[j] - x[i], y[j] - y[i], c[i], s[i], c[j], s[j]);
    }
    return block_sum_d<NT>(e, tmp);
}

// Exclusive prefix sum of one int per thread over the workgroup (NT a
// multiple of 64, at most 4096): wavefront 0 scans the NT values, NT / 64 per
// lane.
template <int NT>
__device__ int block_excl_scan(int v, int* part) {
    static_assert(NT % 64 == 0 && NT <= 4096, "NT / 64 values per lane of one wavefront");
    part[threadIdx.x] = v;
    __syncthreads();
    if (threadIdx.x < 64) {
        constexpr int K = NT / 64;
        const int l = threadIdx.x;
        int loc[K], sum = 0;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            loc[k] = part[l * K + k];
            sum += loc[k];
        }
        int run = wave_incl_scan(sum) - sum;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            part[l * K + k] = run;
            run += loc[k];
        }
    }
    __syncthreads();
    const int r = part[threadIdx.x];
    __syncthreads();
    return r;
}

// Zero-overlap test in FP64 of the configuration held in LDS: true when every
// wall term and every pair term that total64 adds is exactly zero, i.e. when
// total64 returns 0.  Terms are never negative, so the sum is zero exactly
// when each term is.  A pair term is zero unless d^2 < 2.000001, so only pairs
// in the same or adjacent cells of a uniform grid of width 1.5 can be
// nonzero; each such pair i < j is evaluated with the arguments total64
// passes.  O(n) instead of the O(n^2) of total64.
constexpr int kZeroAllPairsMaxN = 96;  // below this the test visits all pairs
struct ZeroGrid {
    int* start;       // zero_grid_side(L)^2 + 1
    int* part;        // kLadderThreads
    int* flag;
    uint16_t* idx;    // n
};
// The exchange kernel's LDS (ladder_lds_bytes) for n squares at side L.
struct LadderLds {
    double *x, *y, *c, *s;
    ZeroGrid zg;
};
__device__ inline LadderLds carve_ladder(unsigned char* smem, int n, double L) {
    LadderLds t;
    const int mc = zero_grid_side(L);
    t.x = reinterpret_cast<double*>(smem);
    t.y = t.x + n;
    t.c = t.y + n;
    t.s = t.c + n;
    t.zg.start = reinterpret_cast<int*>(t.s + n);
    t.zg.part = t.zg.start + (mc * mc + 1);
    t.zg.flag = t.zg.part + kLadderThreads;
    t.zg.idx = reinterpret_cast<uint16_t*>(t.zg.flag + 1);
    return t;
}
template <int NT>
__device__ bool overlap_free64(const double* x, const double* y, const double* c, const double* s, int n,
                               double L, const ZeroGrid& zg) {
    const int tid = threadIdx.x;
    if (n <= kZeroAllPairsMaxN) {
        // few squares: every pair i < j directly, with the arguments total64
        // uses, spread evenly over the threads (flat index over the n x n
        // grid) and without the grid's passes and barriers
        // (the result through the LDS flag, as below: __syncthreads_or
        // would reserve static LDS in every kernel that can reach here)
        if (tid == 0) *zg.flag = 0;
        __syncthreads();
        bool bad = false;
        for (int i = tid; i < n; i += NT) bad |= wall64(x[i], y[i], c[i], s[i], L) > 0.0;
        for (int p = tid; p < n * n; p += NT) {
            const int i = p / n, j = p - i * n;
            if (j > i) bad |= pair64(x[j] - x[i], y[j] - y[i], c[i], s[i], c[j], s[j]) > 0.0;
        }
        if (bad) *zg.flag = 1;
        __syncthreads();
        const bool ok = *zg.flag == 0;
        __syncthreads();
        return ok;
    }
    const int mc = zero_grid_side(L);
    const int ncell = mc * mc;
    auto cell = [&](double v) {
        const int k = (int)floor(v / kZeroCell);
        return k < 0 ? 0 : (k > mc - 1 ? mc - 1 : k);
    };
    for (int k = tid; k <= ncell; k += NT) zg.start[k] = 0;
    if (tid == 0) *zg.flag = 0;
    __syncthreads();
    for (int i = tid; i < n; i += NT) atomicAdd(&zg.start[cell(y[i]) * mc + cell(x[i]) + 1], 1);
    __syncthreads();
    // start[c] = squares in cells before c (counts were placed one cell up)
    const int per = (ncell + 1 + NT - 1) / NT;
    const int b = min(ncell + 1, tid * per), e = min(ncell + 1, b + per);
    int sum = 0;
    for (int k = b; k < e; ++k) sum += zg.start[k];
    int run = block_excl_scan<NT>(sum, zg.part);
    for (int k = b; k < e; ++k) {
        run += zg.start[k];
        zg.start[k] = run;
    }
    __syncthreads();
    // scatter; start[c] then holds the end of cell c (the start of c + 1)
    for (int i = tid; i < n; i += NT) zg.idx[atomicAdd(&zg.start[cell(y[i]) * mc + cell(x[i])], 1)] = (uint16_t)i;
    __syncthreads();
    bool bad = false;
    for (int i = tid; i < n; i += NT) {
        bad |= wall64(x[i], y[i], c[i], s[i], L) > 0.0;
        const int cx = cell(x[i]), cy = cell(y[i]);
        for (int yy = max(cy - 1, 0); yy <= min(cy + 1, mc - 1); ++yy)
            for (int xx = max(cx - 1, 0); xx <= min(cx + 1, mc - 1); ++xx) {
                const int cc = yy * mc + xx;
                for (int k = cc ? zg.start[cc - 1] : 0; k < zg.start[cc]; ++k) {
                    const int j = zg.idx[k];
                    if (j > i) bad |= pair64(x[j] - x[i], y[j] - y[i], c[i], s[i], c[j], s[j]) > 0.0;
                }
            }
    }
    if (bad) *zg.flag = 1;
    __syncthreads();
    const bool ok = *zg.flag == 0;
    __syncthreads();
    return ok;
}

// ---------------------------------------------------------------------------
// k_ladder: one workgroup per instance (insts[list[b]], or insts[b] without a
// list), after every sweep launch.  In order:
//  1. step adaptation of every temperature slot from the acceptance of the
//     replica that held it: base x e^((rate - target) / 2), clamped;
//  2. improve = 1: the smallest certified side over the replicas (the lowest
//     replica index on ties); when it beats the best side by 1e-11
//     (cuda/packing.cu:831-833) it becomes the best side, the replica's state
//     the archive, every slot the box max(sqrt n + 1e-6, best - shrink *
//     mults[8 position / ladder]) (:503-504, :841), and the coldest slot of
//     every ladder -- every slot of the first ladder when the gain exceeds
//     5e-4 -- restarts from the archive dilated, translated and scaled to its
//     box (:836-846); such an interval does no exchange (:849-851);
//  3. otherwise (and always with improve = 0) even/odd exchange of adjacent
//     slots inside each ladder: take when log u < (1/Ta - 1/Tb)(Ea - Eb)
//     (cuda/packing.cu:141-152, 750-765), written u < 2^x with
//     x = (log2e/Ta - log2e/Tb)(Ea - Eb) and the sweep's FP32 2^x;
//  4. improve = 0: the coldest replica (lowest slot) with FP32 energy exactly
//     zero goes through the FP64 zero-overlap test; if it passes it becomes
//     the archive when its box is the smallest yet, and the box shrinks by
//     `shrink` (every centre rescaled on the next staging);
//  5. the next interval's temperatures and step sizes (cooling cycle).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kLadderThreads) void k_ladder(InstDesc* __restrict__ insts,
                                                           const int* __restrict__ list,
                                                           RepState* __restrict__ reps,
                                                           SlotState* __restrict__ slots,
                                                           int* __restrict__ rep_of_slot,
                                                           float4* __restrict__ state,
                                                           float4* __restrict__ arch,
                                                           unsigned* __restrict__ seq) {
    extern __shared__ __align__(16) unsigned char lsm[];
    __shared__ int zero_slot;
    __shared__ int best_r;
    __shared__ unsigned long long best_lc;
    // accepted, certified, candidates, near pairs, collective tried / accepted,
    // kicks, exchanges tried / accepted
    __shared__ unsigned long long tot[9];
    const int di = list ? list[blockIdx.x] : (int)blockIdx.x;
    InstDesc* D = insts + di;
    // Every descriptor field in one round of loads; the kernel's updates are
    // made to this copy and stored by one thread at the end (no other kernel
    // writes the descriptor meanwhile, and each instance has one workgroup).
    InstDesc d = *D;
    const long long left = d.sweeps_total - d.sweeps_done;
    if (left <= 0) return;
    const int ran = (int)min((long long)d.spe, left);
    const int tid = threadIdx.x;
    const int R = d.R, n = d.n, lad = d.ladder, base = d.rep_off;
    const int improve = d.improve;

    // The ladder of the instance in LDS, by temperature slot: energy and
    // 1/T (log2 units) of the slot, replica holding it and that replica's
    // accepted moves.  The replica and the slot records load together.
    float* eS = reinterpret_cast<float*>(lsm);
    float* kS = eS + R;
    int* rS = reinterpret_cast<int*>(kS + R);
    unsigned* aS = reinterpret_cast<unsigned*>(rS + R);
    // The thread's first replica and slot records (a clamped index keeps the
    // loads in bounds) go out before anything waits, and the uniform of the
    // first exchange pair the thread owns is drawn while they are in flight.
    const RepState rs_first = reps[base + min(tid, R - 1)];
    const SlotState ss_first = slots[base + min(tid, R - 1)];
    const uint32_t interval = d.interval, key0 = d.key0, key1 = d.key1;
    const int parity = (int)(interval & 1u);
    const uint32_t ex_bits = philox((uint32_t)(2 * tid + parity), 0u, interval, 2u, key0, key1).x;
    if (tid == 0) {
        zero_slot = R;
        best_r = R;
        best_lc = ~0ull;
    }
    if (tid < 9) tot[tid] = 0ull;
    __syncthreads();

    // 1. the replicas' energies, acceptances and certified sides by the slot
    //    they held (the smallest side: per thread, per wavefront, then one LDS
    //    atomic per wavefront); the interval's counters of the replicas and of
    //    the sweep workgroups move into the instance's totals
    unsigned long long my_lc = ~0ull;
    unsigned long long c_acc = 0, c_cert = 0, c_cand = 0, c_near = 0, c_ct = 0, c_ca = 0, c_kick = 0;
    // slot record `tid` stays in registers for step 4 (no kernel step before
    // step 4 writes slot records)
    for (int r = tid; r < R; r += kLadderThreads) {
        RepState rs = rs_first;
        SlotState sr = ss_first;
        if (r != tid) {
            rs = reps[base + r];
            sr = slots[base + r];
        }
        kS[r] = sr.inv_t_log2e;  // slot r
        c_acc += rs.acc;
        c_cert += (improve && rs.Lc < kCertNone) ? 1u : 0u;
        if (rs.cand | rs.near_pairs | rs.coll_tried | rs.coll_acc | rs.kicks) {
            c_cand += rs.cand;
            c_near += rs.near_pairs;
            c_ct += rs.coll_tried;
            c_ca += rs.coll_acc;
            c_kick += rs.kicks;
            RepState* w = reps + base + r;
            w->cand = 0ull;
            w->near_pairs = 0ull;
            w->coll_tried = 0u;
            w->coll_acc = 0u;
            w->kicks = 0u;
        }
        eS[rs.slot] = rs.E;
        rS[rs.slot] = r;
        aS[rs.slot] = rs.acc;
        if (improve) my_lc = min(my_lc, (unsigned long long)__double_as_longlong(rs.Lc));
    }
    if (improve) {
        my_lc = wave_reduce_u64<true>(my_lc);
        if ((tid & 63) == 0) atomicMin(&best_lc, my_lc);
    }
    c_acc = wave_reduce_u64<false>(c_acc);
    c_cert = wave_reduce_u64<false>(c_cert);
    c_cand = wave_reduce_u64<false>(c_cand);
    c_near = wave_reduce_u64<false>(c_near);
    c_ct = wave_reduce_u64<false>(c_ct);
    c_ca = wave_reduce_u64<false>(c_ca);
    c_kick = wave_reduce_u64<false>(c_kick);
    if ((tid & 63) == 0 && (c_acc | c_cert | c_cand | c_near | c_ct | c_ca | c_kick)) {
        atomicAdd(tot + 0, c_acc);
        atomicAdd(tot + 1, c_cert);
        atomicAdd(tot + 2, c_cand);
        atomicAdd(tot + 3, c_near);
        atomicAdd(tot + 4, c_ct);
        atomicAdd(tot + 5, c_ca);
        atomicAdd(tot + 6, c_kick);
    }
    __syncthreads();

    // 2. improvement: the replica with the smallest certified side
    const double L0 = d.L;
    const double shrink = (double)d.shrink, lmin = d.Lmin + 1e-6;
    bool improved = false;
    double Lb = 0.0;
    if (improve) {
        const double lc = __longlong_as_double((long long)best_lc);
        improved = lc < L0 - 1e-11;
        if (improved) {
            for (int r = tid; r < R; r += kLadderThreads)
                if ((unsigned long long)__double_as_longlong(reps[base + r].Lc) == best_lc) atomicMin(&best_r, r);
            __syncthreads();
            const int rb = best_r;
            const RepState rs = reps[base + rb];
            const double lam = rs.lam, mx = rs.minx, my = rs.miny;
            Lb = rs.Lc;
            const float box_b = slots[base + rs.slot].L;
            const float4* src = state + d.state_off + (size_t)rb * n;
            // restarts: slot 0 of every ladder, and the whole first ladder on
            // a large gain (every slot's box below the new best side is set
            // with the slot records in step 4).  The scale of each restarted
            // slot goes to LDS (the energies and 1/T by slot are not read
            // again on an improving interval); every square of the best
            // configuration is read once, archived, and written to the
            // restarted replicas, which the threads holding copies of that
            // square share out.
            const int heads = R / lad, extra = (L0 - Lb > 5e-4) ? lad - 1 : 0;
            const int H = heads + extra;
            double* scS = reinterpret_cast<double*>(eS);  // eS and kS: 8 R bytes, H <= R
            auto slot_of = [&](int h) { return h < heads ? h * lad : h - heads + 1; };
            for (int h = tid; h < H; h += kLadderThreads) {
                const int s = slot_of(h);
                const double tL = fmax(lmin, Lb - shrink * kTargetMults[((s % lad) * 8) / lad]);
                scS[h] = tL / Lb;
            }
            auto restart = [&](int i, const float4& v, int h0, int step) {
                for (int h = h0; h < H; h += step) {
                    const double sc = scS[h];
                    float4* dst = state + d.state_off + (size_t)rS[slot_of(h)] * n;
                    dst[i] = make_float4((float)(((double)v.x * lam - mx) * sc),
                                         (float)(((double)v.y * lam - my) * sc), v.z, v.w);
                }
            };
            if (n <= kLadderThreads) {
                // kLadderThreads / n copies of every square; all reads of the
                // best replica precede the barrier, since it may be restarted
                const int copies = kLadderThreads / n, c = tid / n, i = tid - c * n;
                float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                if (c < copies) v = src[i];
                __syncthreads();
                if (c < copies) {
                    if (c == 0) arch[d.arch_off + i] = v;
                    restart(i, v, c, copies);
                }
            } else {
                // each square read and written by one thread only
                __syncthreads();
                for (int i = tid; i < n; i += kLadderThreads) {
                    const float4 v = src[i];
                    arch[d.arch_off + i] = v;
                    restart(i, v, 0, 1);
                }
            }
            d.L = Lb;
            d.Lf = (float)Lb;
            d.L_arch = Lb;
            d.L_box_arch = (double)box_b;
            d.E_arch = rs.E;
            d.arch_valid = 1;
            d.feasible += 1;
            d.shrinks += 1;
        }
    }

    // 3. even/odd exchange of adjacent temperatures inside each ladder (the
    //    pairs are disjoint, so each thread owns the slots it swaps)
    if (!improved) {
        unsigned tried = 0, accd = 0;
        for (int p = tid; 2 * p + parity + 1 < R; p += kLadderThreads) {
            const int s = 2 * p + parity;
            if ((s % lad) + 1 >= lad) continue;
            const int ra = rS[s], rb = rS[s + 1];
            const float Ea = eS[s], Eb = eS[s + 1];
            const float x = (kS[s] - kS[s + 1]) * (Ea - Eb);
            uint32_t bits = ex_bits;
            if (p != tid) bits = philox((uint32_t)s, 0u, interval, 2u, key0, key1).x;
            ++tried;
            if (x >= 0.0f || u01(bits) < accept_exp2(x)) {
                rS[s] = rb;
                rS[s + 1] = ra;
                eS[s] = Eb;
                eS[s + 1] = Ea;
                rep_of_slot[base + s] = rb;
                rep_of_slot[base + s + 1] = ra;
                reps[base + ra].slot = s + 1;
                reps[base + rb].slot = s;
                ++accd;
            }
        }
        tried = wave_sum_u32(tried);
        accd = wave_sum_u32(accd);
        if ((tid & 63) == 0 && tried) {
            atomicAdd(tot + 7, (unsigned long long)tried);
            atomicAdd(tot + 8, (unsigned long long)accd);
        }
    }

    // 4. every slot record, written once: step adaptation from the acceptance
    //    of the replica that held the slot, base x e^((rate - target) / 2)
    //    clamped; the box (below the new best side after an improvement); the
    //    next interval's temperature and step sizes (cooling cycle).  The
    //    thread's first record is the one it read in step 1.
    {
        const float target = d.target_acc;
        double cool, root;
        cool_factor(interval + 1u, d.cool_period, cool, root);
        for (int s = tid; s < R; s += kLadderThreads) {
            SlotState ss = ss_first;
            if (s != tid) ss = slots[base + s];
            const float rate = (float)aS[s] / (float)((long long)n * ran);
            // e^((rate - target) / 2) as 2^((rate - target) log2(e) / 2)
            const float f = accept_exp2((rate - target) * 0.72134752f);
            ss.base_xy = fminf(fmaxf(ss.base_xy * f, 1.7320508e-7f), 0.8660254f);
            ss.base_th = fminf(fmaxf(ss.base_th * f, 1.7320508e-7f), 0.7794229f);
            if (improved) ss.L = (float)fmax(lmin, Lb - shrink * kTargetMults[((s % lad) * 8) / lad]);
            slot_effective(ss, cool, root);
            slots[base + s] = ss;
        }
    }
    __syncthreads();

    // 5. zero-overlap test: coldest replica with FP32 energy exactly zero
    double scale = 1.0, Ln = L0;
    bool archive = false, rejected = false;
    if (!improve) {
        for (int s = tid; s < R; s += kLadderThreads)
            if (eS[s] == 0.0f) atomicMin(&zero_slot, s);
        __syncthreads();
        if (zero_slot < R) {
            const int r = rS[zero_slot];
            __syncthreads();  // the ladder's LDS is reused below
            const float4* src = state + d.state_off + (size_t)r * n;
            const LadderLds t = carve_ladder(lsm, n, L0);
            for (int i = tid; i < n; i += kLadderThreads) {
                const float4 v = src[i];
                double c, s;
                unit64(v.z, v.w, c, s);
                t.x[i] = v.x; t.y[i] = v.y; t.c[i] = c; t.s[i] = s;
            }
            __syncthreads();
            const bool zero64 = overlap_free64<kLadderThreads>(t.x, t.y, t.c, t.s, n, L0, t.zg);
            rejected = !zero64;
            if (zero64) {
                archive = !d.arch_valid || L0 < d.L_arch;
                if (archive)
                    for (int i = tid; i < n; i += kLadderThreads) arch[d.arch_off + i] = src[i];
                Ln = fmax(L0 - shrink, d.Lmin);
                if (Ln < L0) scale = Ln / L0;
            }
        }
        if (Ln < L0)
            for (int s = tid; s < R; s += kLadderThreads) slots[base + s].L = (float)Ln;
    }
    __syncthreads();

    if (tid == 0) {
        if (rejected) d.fp64_rejects += 1;
        if (archive) {
            d.arch_valid = 1;
            d.L_arch = L0;
            d.L_box_arch = L0;
            d.E_arch = 0.0f;
            d.feasible += 1;
        }
        if (Ln < L0) {
            d.L = Ln;
            d.Lf = (float)Ln;
            d.shrinks += 1;
        }
        d.scale = (float)scale;
        d.sweeps_done += ran;
        d.interval += 1u;
        d.proposals += (unsigned long long)n * (unsigned long long)R * (unsigned long long)ran;
        d.accepted += tot[0];
        d.certs += tot[1];
        d.cand += tot[2];
        d.near += tot[3];
        d.coll_tried += tot[4];
        d.coll_acc += tot[5];
        d.kicks += (int)tot[6];
        d.ex_tried += tot[7];
        d.ex_acc += tot[8];
        *D = d;
        // the interval count the host polls (mapped host memory); the host
        // reads nothing else of the instance while a run is in flight, so
        // the store needs no ordering with the descriptor's
        if (seq) __hip_atomic_store(seq + di, d.interval, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

// Certified side of the configuration held in LDS by one workgroup
// (cuda/packing.cu:254-297): all pairs i < j nearer than sqrt(2.05).  Returns
// the side (1e9 when the dilation exceeds 1.008) and the dilation and lower
// corner of the dilated configuration.
template <int NT>
__device__ double certified_side_block(const double* x, const double* y, const double* c, const double* s, int n,
                                       double* tmp, double& lam_out, double& minx_out, double& miny_out) {
    double lam = 1.0;
    for (int i = threadIdx.x; i < n; i += NT)
        for (int j = i + 1; j < n; ++j) {
            const double dx = x[i] - x[j], dy = y[i] - y[j];
            if (dx * dx + dy * dy >= kCertNear2) continue;
            lam = fmax(lam, cert_need(dx, dy, c[i], s[i], c[j], s[j]));
        }
    lam = block_max_d<NT>(lam, tmp);
    double lox = 1e9, hix = -1e9, loy = 1e9, hiy = -1e9;
    for (int i = threadIdx.x; i < n; i += NT) {
        const double r = 0.5 * (fabs(c[i]) + fabs(s[i]));
        const double xi = x[i] * lam, yi = y[i] * lam;
        lox = fmin(lox, xi - r); hix = fmax(hix, xi + r);
        loy = fmin(loy, yi - r); hiy = fmax(hiy, yi + r);
    }
    lox = block_min_d<NT>(lox, tmp);
    hix = block_max_d<NT>(hix, tmp);
    loy = block_min_d<NT>(loy, tmp);
    hiy = block_max_d<NT>(hiy, tmp);
    lam_out = lam;
    minx_out = lox;
    miny_out = loy;
    return lam > kCertMaxLam ? kCertNone : fmax(hix - lox, hiy - loy);
}

// ---------------------------------------------------------------------------
// k_finalize: FP64 polish + certified side + export, one instance per block.
// src: n float4 (x, y, cos, sin) per instance at FP64 box side Lsrc.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kPolishThreads) void k_finalize(const FinalizeJob* __restrict__ jobs,
                                                             FinalizeOut* __restrict__ outs,
                                                             double* __restrict__ q_out) {
    __shared__ double x[kMaxN], y[kMaxN], a[kMaxN], cc[kMaxN], ss[kMaxN];
    __shared__ double tmp[kPolishThreads];
    __shared__ double step_sh;
    __shared__ float best_e[kPolishThreads];
    __shared__ int best_r[kPolishThreads];
    constexpr int NT = kPolishThreads, KPT = kPolishKPT;
    const FinalizeJob job = jobs[blockIdx.x];
    FinalizeOut* out = outs + blockIdx.x;
    const int n = job.n, tid = threadIdx.x;
    const double L = job.L;
    const float4* src32 = job.src32;
    if (job.reps) {
        // lowest FP32 energy over the replicas, the lowest index on ties
        float e = INFINITY;
        int r = 0x7fffffff;
        for (int k = tid; k < job.R; k += NT) {
            const float ek = job.reps[k].E;
            if (ek < e) { e = ek; r = k; }
        }
        best_e[tid] = e;
        best_r[tid] = r;
        __syncthreads();
        for (int s = NT / 2; s > 0; s >>= 1) {
            if (tid < s) {
                const float eo = best_e[tid + s];
                const int ro = best_r[tid + s];
                if (eo < best_e[tid] || (eo == best_e[tid] && ro < best_r[tid])) {
                    best_e[tid] = eo;
                    best_r[tid] = ro;
                }
            }
            __syncthreads();
        }
        src32 += (size_t)best_r[0] * n;
        if (tid == 0) {
            out->best = best_r[0];
            out->e32 = best_e[0];
        }
    }
    for (int i = tid; i < n; i += NT) {
        double xi, yi, ai;
        if (src32) {
            const float4 v = src32[i];
            xi = v.x; yi = v.y;
            ai = wrap_quarter(atan2((double)v.w, (double)v.z));
        } else {
            xi = job.src64[3 * i]; yi = job.src64[3 * i + 1]; ai = job.src64[3 * i + 2];
        }
        x[i] = xi; y[i] = yi; a[i] = ai;
        if (src32) {
            // the orientation as the sweeps' certified side reads it
            unit64(src32[i].z, src32[i].w, cc[i], ss[i]);
        } else {
            cc[i] = cos(ai); ss[i] = sin(ai);
        }
    }
    __syncthreads();
    double* qo = q_out + job.out_off;
    // certified side before the polish, exported; the polish's wins only when
    // smaller (the reference keeps the smaller of its checkpoints,
    // cuda/packing.cu:506-507, 689-693)
    double lam0, mx0, my0;
    const double side0 = certified_side_block<NT>(x, y, cc, ss, n, tmp, lam0, mx0, my0);
    {
        const double lx = lam0 > kCertMaxLam ? 1.0 : lam0;
        for (int i = tid; i < n; i += NT) {
            qo[3 * i] = x[i] * lx - (lam0 > kCertMaxLam ? 0.0 : mx0);
            qo[3 * i + 1] = y[i] * lx - (lam0 > kCertMaxLam ? 0.0 : my0);
            qo[3 * i + 2] = a[i];
        }
    }
    if (src32) {
        for (int i = tid; i < n; i += NT) { cc[i] = cos(a[i]); ss[i] = sin(a[i]); }
        __syncthreads();
    }
    double energy = total64<NT>(x, y, cc, ss, n, L, tmp);
    const double e_before = energy;

    // Polak-Ribiere+ conjugate gradient, cuda/packing.cu:624-687.
    double g[KPT][3], gold[KPT][3], dir[KPT][3], saved[KPT][3];
#pragma unroll
    for (int k = 0; k < KPT; ++k)
        for (int d = 0; d < 3; ++d) { gold[k][d] = 0.0; dir[k][d] = 0.0; g[k][d] = 0.0; saved[k][d] = 0.0; }
    if (tid == 0) step_sh = 0.04;
    __syncthreads();
    int iters = 0;
    for (int it = 0; it < job.iters; ++it) {
        if (energy < 1e-24) break;
        double gg = 0.0, dgg = 0.0;
#pragma unroll
        for (int k = 0; k < KPT; ++k) {
            const int i = tid + k * NT;
            if (i >= n) break;
            double gx = 0, gy = 0, ga = 0;
            const double ci = cc[i], si = ss[i], xi = x[i], yi = y[i];
            const double ext = 0.5 * (fabs(ci) + fabs(si));
            const double dR = .5 * (-sgn_dead(ci) * si + sgn_dead(si) * ci);
            if (xi < ext) { const double u = ext - xi; gx -= 2 * u; ga += 2 * u * dR; }
            else if (xi > L - ext) { const double u = xi + ext - L; gx += 2 * u; ga += 2 * u * dR; }
            if (yi < ext) { const double u = ext - yi; gy -= 2 * u; ga += 2 * u * dR; }
            else if (yi > L - ext) { const double u = yi + ext - L; gy += 2 * u; ga += 2 * u * dR; }
            for (int j = 0; j < n; ++j)
                if (j != i) pair64_grad(x[j] - xi, y[j] - yi, ci, si, cc[j], ss[j], gx, gy, ga);
            g[k][0] = gx; g[k][1] = gy; g[k][2] = ga;
            for (int d = 0; d < 3; ++d) {
                gg += gold[k][d] * gold[k][d];
                dgg += (g[k][d] - gold[k][d]) * g[k][d];
            }
        }
        gg = block_sum_d<NT>(gg, tmp);
        dgg = block_sum_d<NT>(dgg, tmp);
        const double beta = (it == 0 || gg < 1e-30 || it % 25 == 0) ? 0.0 : fmax(0.0, dgg / gg);
        double nrm = 0.0;
#pragma unroll
        for (int k = 0; k < KPT; ++k) {
            const int i = tid + k * NT;
            if (i >= n) break;
            for (int d = 0; d < 3; ++d) {
                dir[k][d] = -g[k][d] + beta * dir[k][d];
                gold[k][d] = g[k][d];
                nrm += dir[k][d] * dir[k][d];
            }
            saved[k][0] = x[i]; saved[k][1] = y[i]; saved[k][2] = a[i];
        }
        if (block_sum_d<NT>(nrm, tmp) < 1e-28) break;
        iters = it + 1;
        double alpha = step_sh;
        bool ok = false;
        for (int ls = 0; ls < 5; ++ls) {
#pragma unroll
            for (int k = 0; k < KPT; ++k) {
                const int i = tid + k * NT;
                if (i >= n) break;
                x[i] = saved[k][0] + alpha * dir[k][0];
                y[i] = saved[k][1] + alpha * dir[k][1];
                a[i] = wrap_quarter(saved[k][2] + alpha * dir[k][2]);
                cc[i] = cos(a[i]);
                ss[i] = sin(a[i]);
            }
            __syncthreads();
            const double ne = total64<NT>(x, y, cc, ss, n, L, tmp);
            if (ne < energy) {
                energy = ne;
                if (tid == 0) step_sh = fmin(0.15, alpha * 1.25);
                ok = true;
                break;
            }
            alpha *= 0.45;
        }
        if (!ok) {
#pragma unroll
            for (int k = 0; k < KPT; ++k) {
                const int i = tid + k * NT;
                if (i >= n) break;
                x[i] = saved[k][0]; y[i] = saved[k][1]; a[i] = saved[k][2];
                cc[i] = cos(a[i]); ss[i] = sin(a[i]);
                dir[k][0] = dir[k][1] = dir[k][2] = 0.0;
            }
            if (tid == 0) step_sh = fmax(1e-7, step_sh * 0.5);
        }
        __syncthreads();
    }

    // Certified side after the polish.
    double lam1, mx1, my1;
    const double side1 = certified_side_block<NT>(x, y, cc, ss, n, tmp, lam1, mx1, my1);
    const bool post = side1 < side0 || (side0 >= kCertNone && side1 >= kCertNone);
    if (post) {
        // translated, dilated export of the polished configuration (undilated
        // when neither checkpoint certifies)
        const double lx = lam1 > kCertMaxLam ? 1.0 : lam1;
        double lox = mx1, loy = my1;
        if (lam1 > kCertMaxLam) {
            lox = 1e9; loy = 1e9;
            for (int i = tid; i < n; i += NT) {
                const double r = 0.5 * (fabs(cc[i]) + fabs(ss[i]));
                lox = fmin(lox, x[i] - r); loy = fmin(loy, y[i] - r);
            }
            lox = block_min_d<NT>(lox, tmp);
            loy = block_min_d<NT>(loy, tmp);
        }
        for (int i = tid; i < n; i += NT) {
            qo[3 * i] = x[i] * lx - lox;
            qo[3 * i + 1] = y[i] * lx - loy;
            qo[3 * i + 2] = a[i];
        }
    }
    if (tid == 0) {
        out->e64_before = e_before;
        out->e64_after = energy;
        out->lambda = post ? lam1 : lam0;
        out->L_cert = post ? side1 : side0;
        out->span = post ? side1 : side0;
        out->iters = iters;
        out->pre = post ? 0 : 1;
    }
}

// ---------------------------------------------------------------------------
// Parity probes.
// ---------------------------------------------------------------------------
// Probe: the ladder's grid zero-overlap test beside the all-pairs FP64 energy
// of one configuration (x, y, cos, sin as float4, normalised as k_ladder does).
__global__ __launch_bounds__(kLadderThreads) void k_zero64(int n, double L, const float4* __restrict__ q,
                                                           int* __restrict__ zero, double* __restrict__ e64) {
    extern __shared__ __align__(16) unsigned char lsm[];
    __shared__ double tmp[kLadderThreads];
    const LadderLds t = carve_ladder(lsm, n, L);
    for (int i = threadIdx.x; i < n; i += kLadderThreads) {
        const float4 v = q[i];
        const double c = v.z, s = v.w, h = 1.0 / sqrt(c * c + s * s);
        t.x[i] = v.x; t.y[i] = v.y; t.c[i] = c * h; t.s[i] = s * h;
    }
    __syncthreads();
    const bool z = overlap_free64<kLadderThreads>(t.x, t.y, t.c, t.s, n, L, t.zg);
    const double e = total64<kLadderThreads>(t.x, t.y, t.c, t.s, n, L, tmp);
    if (threadIdx.x == 0) {
        *zero = z ? 1 : 0;
        *e64 = e;
    }
}

hipError_t launch_zero64(int n, double L, const float4* q, int* zero, double* e64) {
    hipLaunchKernelGGL(k_zero64, dim3(1), dim3(kLadderThreads), ladder_lds_bytes(n, L, 0), nullptr, n, L, q, zero,
                       e64);
    return hipGetLastError();
}

__global__ void k_depth32(int npairs, const float* __restrict__ in, float* __restrict__ depth,
                          float* __restrict__ d2out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= npairs) return;
    const float* p = in + 8 * (size_t)k;
    const float dx = p[0], dy = p[1];
    const float d2 = __builtin_fmaf(dx, dx, dy * dy);
    const float g = sat_gap(dx, dy, p[2], p[3], p[4], p[5]);
    depth[k] = (d2 < 2.0f && g > 0.0f) ? g : 0.0f;
    d2out[k] = depth2(dx, dy, p[2], p[3], p[4], p[5]);
}

__global__ void k_accept_exp2(int count, const float* __restrict__ x, float* __restrict__ out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < count) out[k] = accept_exp2(x[k]);
}

__global__ void k_cellcoord(int count, const float* __restrict__ x, float shift, float w, int lo,
                            int hi, int* __restrict__ out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < count) out[k] = cellcoord(x[k], shift, 1.0f / w, lo, hi);
}

// ---------------------------------------------------------------------------
// Launch wrappers.
// ---------------------------------------------------------------------------
template <int NT, bool kNear, int W>
static void launch_sweep_t(hipStream_t st, int ntiles, size_t lds, const Tile* tiles, InstDesc* insts,
                           RepState* reps, const SlotState* slots, float4* state) {
    hipLaunchKernelGGL((k_sweep<NT, kNear, W>), dim3(ntiles), dim3(NT), lds, st, tiles, insts, reps, slots, state);
}

hipError_t launch_sweep(hipStream_t st, int ntiles, size_t lds, const Tile* tiles,
                        InstDesc* insts, RepState* reps, const SlotState* slots, float4* state,
                        int threads, bool near_pairs, bool alone) {
    if (threads != kSweepThreadsBig && threads != kSweepThreads) return hipErrorInvalidValue;
    const bool big = threads == kSweepThreadsBig;
    if (big && near_pairs && alone) launch_sweep_t<kSweepThreadsBig, true, 1>(st, ntiles, lds, tiles, insts, reps, slots, state);
    else if (big && near_pairs) launch_sweep_t<kSweepThreadsBig, true, 4>(st, ntiles, lds, tiles, insts, reps, slots, state);
    else if (big && alone) launch_sweep_t<kSweepThreadsBig, false, 1>(st, ntiles, lds, tiles, insts, reps, slots, state);
    else if (big) launch_sweep_t<kSweepThreadsBig, false, 4>(st, ntiles, lds, tiles, insts, reps, slots, state);
    else if (near_pairs && alone) launch_sweep_t<kSweepThreads, true, 1>(st, ntiles, lds, tiles, insts, reps, slots, state);
    else if (near_pairs) launch_sweep_t<kSweepThreads, true, 4>(st, ntiles, lds, tiles, insts, reps, slots, state);
    else if (alone) launch_sweep_t<kSweepThreads, false, 1>(st, ntiles, lds, tiles, insts, reps, slots, state);
    else launch_sweep_t<kSweepThreads, false, 4>(st, ntiles, lds, tiles, insts, reps, slots, state);
    return hipGetLastError();
}

hipError_t launch_ladder(hipStream_t st, int ninst, size_t lds, InstDesc* insts, RepState* reps,
                         SlotState* slots, int* rep_of_slot, float4* state, float4* arch,
                         const int* list, unsigned* seq) {
    hipLaunchKernelGGL(k_ladder, dim3(ninst), dim3(kLadderThreads), lds, st, insts, list, reps, slots,
                       rep_of_slot, state, arch, seq);
    return hipGetLastError();
}

// Copy of the descriptors named by `list` into out[0..count), 4-byte words.
__global__ void k_gather_desc(const InstDesc* __restrict__ insts, const int* __restrict__ list,
                              InstDesc* __restrict__ out) {
    constexpr int kWords = (int)(sizeof(InstDesc) / 4);
    static_assert(sizeof(InstDesc) % 4 == 0, "descriptor copied in 4-byte words");
    const uint32_t* src = reinterpret_cast<const uint32_t*>(insts + list[blockIdx.x]);
    uint32_t* dst = reinterpret_cast<uint32_t*>(out + blockIdx.x);
    for (int k = threadIdx.x; k < kWords; k += blockDim.x) dst[k] = src[k];
}

hipError_t launch_gather_desc(hipStream_t st, int count, const InstDesc* insts, const int* list,
                              InstDesc* out) {
    hipLaunchKernelGGL(k_gather_desc, dim3(count), dim3(64), 0, st, insts, list, out);
    return hipGetLastError();
}

hipError_t launch_finalize(hipStream_t st, int njobs, const FinalizeJob* jobs, FinalizeOut* outs,
                           double* q_out) {
    hipLaunchKernelGGL(k_finalize, dim3(njobs), dim3(kPolishThreads), 0, st, jobs, outs, q_out);
    return hipGetLastError();
}

hipError_t launch_depth32(hipStream_t st, int npairs, const float* in, float* depth, float* d2) {
    hipLaunchKernelGGL(k_depth32, dim3((npairs + 255) / 256), dim3(256), 0, st, npairs, in, depth,
                       d2);
    return hipGetLastError();
}

hipError_t launch_cellcoord(hipStream_t st, int count, const float* x, float shift, float w, int lo,
                            int hi, int* out) {
    hipLaunchKernelGGL(k_cellcoord, dim3((count + 255) / 256), dim3(256), 0, st, count, x, shift, w,
                       lo, hi, out);
    return hipGetLastError();
}

hipError_t launch_accept_exp2(hipStream_t st, int count, const float* x, float* out) {
    hipLaunchKernelGGL(k_accept_exp2, dim3((count + 255) / 256), dim3(256), 0, st, count, x, out);
    return hipGetLastError();
}

const void* sweep_kernel_ptr(int threads, bool near_pairs, bool alone) {
    const bool big = threads == kSweepThreadsBig;
    if (alone) {
        if (near_pairs)
            return big ? reinterpret_cast<const void*>(k_sweep<kSweepThreadsBig, true, 1>)
                       : reinterpret_cast<const void*>(k_sweep<kSweepThreads, true, 1>);
        return big ? reinterpret_cast<const void*>(k_sweep<kSweepThreadsBig, false, 1>)
                   : reinterpret_cast<const void*>(k_sweep<kSweepThreads, false, 1>);
    }
    if (near_pairs)
        return big ? reinterpret_cast<const void*>(k_sweep<kSweepThreadsBig, true, 4>)
                   : reinterpret_cast<const void*>(k_sweep<kSweepThreads, true, 4>);
    return big ? reinterpret_cast<const void*>(k_sweep<kSweepThreadsBig, false, 4>)
               : reinterpret_cast<const void*>(k_sweep<kSweepThreads, false, 4>);
}

}  // namespace sis
