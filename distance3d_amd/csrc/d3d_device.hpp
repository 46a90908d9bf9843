// d3d_device.hpp -- device-side geometry for the gfx950 collision kernels:
// FP64 vector math, collider records, support functions (analytic in FP64,
// convex-hull argmax as an FP32 wavefront sweep with an exactness
// certificate), AABBs, and the Jolt/van den Bergen simplex sub-routines.
//
// Precision contract: every quantity that reaches the simplex solve, the
// termination tests or the closest-point reconstruction is FP64, evaluated
// left to right with one rounding per operation (the library is compiled with
// -ffp-contract=off).  Only the hull argmax scan runs in FP32, and its winner is
// certified against the FP64 argmax (see hull_argmax).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "d3d.h"

namespace d3d {

constexpr double kEps = 2.220446049250313e-16;  // utils.py:8 (np.finfo(float).eps)
constexpr double kEpsSqr = kEps * kEps;         // gjk/_gjk_jolt.py:18
constexpr double kMaxFloat = 1.7976931348623157e308;  // utils.py:7
constexpr int kWave = 64;

struct v3 {
    double x, y, z;
};

__device__ __forceinline__ v3 mk(double x, double y, double z) { return v3{x, y, z}; }
__device__ __forceinline__ v3 operator+(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ v3 operator-(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ v3 operator-(v3 a) { return mk(-a.x, -a.y, -a.z); }
__device__ __forceinline__ v3 operator*(v3 a, double s) { return mk(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ double dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ v3 cross(v3 a, v3 b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
// Component-wise select.  A ternary on the aggregate would become a select
// between pointers plus a copy, which keeps every v3 it touches in scratch
// memory; scalar selects keep them in VGPRs.
__device__ __forceinline__ v3 vsel(bool c, v3 a, v3 b) {
    return mk(c ? a.x : b.x, c ? a.y : b.y, c ? a.z : b.z);
}
__device__ __forceinline__ v3 sel4(int k, v3 a, v3 b, v3 c, v3 d) {
    v3 r = a;
    r = vsel(k == 1, b, r);
    r = vsel(k == 2, c, r);
    r = vsel(k == 3, d, r);
    return r;
}

// scalar_triple_product (utils.py:33-73)
__device__ __forceinline__ double stp(v3 a, v3 b, v3 c) { return dot(a, cross(b, c)); }

// norm_vector (utils.py:12-30)
__device__ __forceinline__ v3 norm_vector(v3 v) {
    double n = sqrt(dot(v, v));
    if (n == 0.0) return v;
    return mk(v.x / n, v.y / n, v.z / n);
}

// ------------------------------------------------------------ wave helpers

__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ double readlane_f64(double v, int lane) {
    unsigned long long b = (unsigned long long)__double_as_longlong(v);
    unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, lane);
    unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), lane);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Hides a value from the optimiser so loads addressed by it are redone where
// they appear instead of being hoisted out of a loop (keeps loop-carried
// register state small; the reloads hit L1/L2).
__device__ __forceinline__ int opaque(int v) {
    asm volatile("" : "+v"(v));
    return v;
}

__device__ __forceinline__ int readlane_i32(int v, int lane) {
    return __builtin_amdgcn_readlane(v, lane);
}

// For values the caller knows to be equal in every active lane: moves them to
// scalar registers (readfirstlane), so loads they address become scalar loads.
__device__ __forceinline__ int wave_uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ double wave_uniform(double v) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)b);
    const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(b >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Butterfly exchange on the VALU (DPP and permlane swaps, no LDS round trip).
// xpair<OFF>(x, a, b) leaves in (a, b) the values x of the two halves of the
// lane's aligned 2*OFF-lane block that a reduction step at distance OFF
// merges: for OFF <= 8, a = the lane's own x and b = x of lane ^ OFF (OFF = 1,
// 2: quad permutes) or of the mirrored lane (OFF = 4: row_half_mirror, lane
// 7 - i of the 8-lane block; OFF = 8: row_mirror, lane 15 - i of the row),
// which holds the same value as lane ^ OFF once every aligned OFF-lane block
// is uniform; for OFF = 16, 32, a = x of the lower half and b = x of the upper
// half (permlane16/32 swap of x with itself).  Reductions therefore step OFF =
// 1, 2, 4, ... upwards and merge (a, b) with a symmetric, exact selection
// (max, min, argmax with index ties, top-2), after which every lane of the
// group holds the group's result.  Every lane of the group must be active.
template <int N> struct IntC { static constexpr int value = N; };

template <int OFF>
__device__ __forceinline__ void xpair(unsigned x, unsigned &a, unsigned &b) {
    static_assert(OFF == 1 || OFF == 2 || OFF == 4 || OFF == 8 || OFF == 16 || OFF == 32, "");
    if constexpr (OFF <= 8) {
        constexpr int ctrl = OFF == 1 ? 0xB1 : OFF == 2 ? 0x4E : OFF == 4 ? 0x141 : 0x140;
        a = x;
        b = (unsigned)__builtin_amdgcn_update_dpp((int)x, (int)x, ctrl, 0xF, 0xF, false);
    } else if constexpr (OFF == 16) {
        const auto p = __builtin_amdgcn_permlane16_swap(x, x, false, false);
        a = p[0];
        b = p[1];
    } else {
        const auto p = __builtin_amdgcn_permlane32_swap(x, x, false, false);
        a = p[0];
        b = p[1];
    }
}
template <int OFF>
__device__ __forceinline__ void xpair(int x, int &a, int &b) {
    unsigned ua, ub;
    xpair<OFF>((unsigned)x, ua, ub);
    a = (int)ua;
    b = (int)ub;
}
template <int OFF>
__device__ __forceinline__ void xpair(float x, float &a, float &b) {
    unsigned ua, ub;
    xpair<OFF>(__float_as_uint(x), ua, ub);
    a = __uint_as_float(ua);
    b = __uint_as_float(ub);
}
template <int OFF>
__device__ __forceinline__ void xpair(double x, double &a, double &b) {
    const unsigned long long u = (unsigned long long)__double_as_longlong(x);
    unsigned la, lb, ha, hb;
    xpair<OFF>((unsigned)u, la, lb);
    xpair<OFF>((unsigned)(u >> 32), ha, hb);
    a = __longlong_as_double((long long)(((unsigned long long)ha << 32) | la));
    b = __longlong_as_double((long long)(((unsigned long long)hb << 32) | lb));
}

// Calls f(IntC<OFF>) for OFF = 1, 2, 4, ... < G: the steps of a G-lane
// butterfly reduction in the order xpair requires.
template <int G, int OFF = 1, class F>
__device__ __forceinline__ void for_butterfly(F &&f) {
    if constexpr (OFF < G) {
        f(IntC<OFF>{});
        for_butterfly<G, OFF * 2>(f);
    }
}

// Merge of two FP32 (largest m1 at i1, largest other m2) summaries: exact,
// symmetric in its operands; ties in m1 go to the smaller index.
__device__ __forceinline__ void top2_merge(float am1, float am2, int ai, float bm1, float bm2, int bi,
                                           float &m1, float &m2, int &i1) {
    m2 = fmaxf(fmaxf(am2, bm2), fminf(am1, bm1));
    const bool tb = (bm1 > am1) || (bm1 == am1 && bi < ai);
    m1 = tb ? bm1 : am1;
    i1 = tb ? bi : ai;
}

template <int G>
__device__ __forceinline__ void group_top2(float &m1, float &m2, int &i1) {
    for_butterfly<G>([&](auto o) {
        constexpr int OFF = decltype(o)::value;
        float am1, bm1, am2, bm2;
        int ai, bi;
        xpair<OFF>(m1, am1, bm1);
        xpair<OFF>(m2, am2, bm2);
        xpair<OFF>(i1, ai, bi);
        top2_merge(am1, am2, ai, bm1, bm2, bi, m1, m2, i1);
    });
}

// (largest value, smallest index among its ties) over the G lanes of a group.
template <int G>
__device__ __forceinline__ void group_argmaxf(float &v, int &k) {
    for_butterfly<G>([&](auto o) {
        constexpr int OFF = decltype(o)::value;
        float av, bv;
        int ak, bk;
        xpair<OFF>(v, av, bv);
        xpair<OFF>(k, ak, bk);
        const bool tb = (bv > av) || (bv == av && bk < ak);
        v = tb ? bv : av;
        k = tb ? bk : ak;
    });
}

// (smallest value, smallest index among its ties) over the G lanes of a group.
template <int G>
__device__ __forceinline__ void group_argmin(double &v, int &k) {
    for_butterfly<G>([&](auto o) {
        constexpr int OFF = decltype(o)::value;
        double av, bv;
        int ak, bk;
        xpair<OFF>(v, av, bv);
        xpair<OFF>(k, ak, bk);
        const bool tb = (bv < av) || (bv == av && bk < ak);
        v = tb ? bv : av;
        k = tb ? bk : ak;
    });
}

// ------------------------------------------------------------- colliders

// Collider record as laid out in the scene (see d3d.h).
struct Shape {
    double R[9];
    double t[3];
    double p[3];
    double margin;
    int type, nv, voff, voff64;
};

__device__ __forceinline__ Shape load_shape(const d3d_scene &s, int k) {
    Shape h;
    const double *f = s.shape_f64 + 16 * (int64_t)k;
#pragma unroll
    for (int j = 0; j < 9; ++j) h.R[j] = f[j];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        h.t[j] = f[9 + j];
        h.p[j] = f[12 + j];
    }
    h.margin = f[15];
    const int4 I = *reinterpret_cast<const int4 *>(s.shape_i32 + 4 * (int64_t)k);
    h.type = I.x;
    h.nv = I.y;
    h.voff = I.z;
    h.voff64 = I.w;
    return h;
}

// The FP64 block of record k (pose, parameters, margin) with a type already
// known: for analytic records, whose supports read nothing else.
__device__ __forceinline__ Shape load_shape_f64(const d3d_scene &s, int k, int type) {
    Shape h;
    const double *f = s.shape_f64 + 16 * (int64_t)k;
#pragma unroll
    for (int j = 0; j < 9; ++j) h.R[j] = f[j];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        h.t[j] = f[9 + j];
        h.p[j] = f[12 + j];
    }
    h.margin = f[15];
    h.type = type;
    h.nv = 0;
    h.voff = 0;
    h.voff64 = -1;
    return h;
}

__device__ __forceinline__ Shape empty_shape() {
    Shape h;
#pragma unroll
    for (int j = 0; j < 9; ++j) h.R[j] = 0.0;
    h.R[0] = h.R[4] = h.R[8] = 1.0;
    h.t[0] = h.t[1] = h.t[2] = 0.0;
    h.p[0] = h.p[1] = h.p[2] = 0.0;
    h.margin = 0.0;
    h.type = D3D_SPHERE;
    h.nv = 0;
    h.voff = 0;
    h.voff64 = -1;
    return h;
}

// Index and vertex-range checks so a malformed table cannot make a lane read
// outside the pools.
__device__ __forceinline__ bool load_checked(const d3d_scene &s, int k, Shape &h) {
    if (k < 0 || (int64_t)k >= s.n_shapes) {
        h = empty_shape();
        return false;
    }
    h = load_shape(s, k);
    if (h.type < D3D_HULL || h.type > D3D_ELLIPSE) {
        h = empty_shape();
        return false;
    }
    if (h.type == D3D_HULL) {
        const bool ok32 = h.nv >= 1 && h.voff >= 0 && (int64_t)h.voff + h.nv <= s.n_vertices;
        const bool ok64 =
            h.voff64 < 0 || ((int64_t)h.voff64 + h.nv <= s.n_vertices64 && s.v64 != nullptr);
        // cluster fields: p1 = first record, p2 = count (0 = unclustered)
        const double c0 = h.p[1], cn = h.p[2];
        const bool okc =
            (cn == 0.0) ||
            (s.clusters != nullptr && c0 >= 0.0 && c0 == floor(c0) &&
             cn == (double)((h.nv + D3D_CLUSTER_SIZE - 1) / D3D_CLUSTER_SIZE) &&
             c0 + cn <= (double)s.n_clusters && h.voff % D3D_CLUSTER_SIZE == 0);
        if (!ok32 || !ok64 || !okc) {
            h = empty_shape();
            return false;
        }
    }
    return true;
}

// R^T d, i.e. np.dot(A2B[:3, :3].T, d) (geometry.py:202)
__device__ __forceinline__ v3 to_local(const Shape &h, v3 d) {
    return mk(h.R[0] * d.x + h.R[3] * d.y + h.R[6] * d.z,
              h.R[1] * d.x + h.R[4] * d.y + h.R[7] * d.z,
              h.R[2] * d.x + h.R[5] * d.y + h.R[8] * d.z);
}

// transform_point: t + R p (utils.py:125-143)
__device__ __forceinline__ v3 to_world(const Shape &h, v3 p) {
    return mk(h.t[0] + (h.R[0] * p.x + h.R[1] * p.y + h.R[2] * p.z),
              h.t[1] + (h.R[3] * p.x + h.R[4] * p.y + h.R[5] * p.z),
              h.t[2] + (h.R[6] * p.x + h.R[7] * p.y + h.R[8] * p.z));
}

__device__ __forceinline__ v3 hull_vertex(const d3d_scene &s, int voff, int voff64, int idx) {
    if (voff64 >= 0) {
        const double *q = s.v64 + 3 * ((int64_t)voff64 + idx);
        return mk(q[0], q[1], q[2]);
    }
    const int64_t j = (int64_t)voff + idx;
    return mk((double)s.vx[j], (double)s.vy[j], (double)s.vz[j]);
}

// Copies the nv FP32 vertices of the hull at FP32-pool offset voff into the
// LDS arrays cx, cy, cz at position at + (original index), by the N lanes of
// a group (this lane: sub).  The pools are 16-byte aligned (checked by every
// entry point), so the body of the range from the first multiple of four is
// read with 16-byte loads, four vertices and their original indices per lane
// per round; the up to three vertices before it and after it one at a time.
// An original index outside [0, nv) (a malformed table) writes nothing.
template <int N>
__device__ __forceinline__ void fill_orig(const d3d_scene &s, int64_t voff, int nv, float *cx,
                                          float *cy, float *cz, int at, int sub) {
    auto put = [&](int o, float x, float y, float z) {
        if ((unsigned)o < (unsigned)nv) {
            cx[at + o] = x;
            cy[at + o] = y;
            cz[at + o] = z;
        }
    };
    const int head = min(nv, (int)((4 - (voff & 3)) & 3));
    const int nq = (nv - head) >> 2;
    const int tail = head + 4 * nq;
    // the head and tail: at most six single vertices
    for (int k = sub; k < head + (nv - tail); k += N) {
        const int j = k < head ? k : tail + (k - head);
        const int64_t p = voff + j;
        put(s.vorig ? s.vorig[p] : j, s.vx[p], s.vy[p], s.vz[p]);
    }
    const int64_t b = voff + head;
    const float4 *__restrict__ qx = reinterpret_cast<const float4 *>(s.vx + b);
    const float4 *__restrict__ qy = reinterpret_cast<const float4 *>(s.vy + b);
    const float4 *__restrict__ qz = reinterpret_cast<const float4 *>(s.vz + b);
    const int4 *__restrict__ qo = s.vorig ? reinterpret_cast<const int4 *>(s.vorig + b) : nullptr;
#pragma unroll 2
    for (int q = sub; q < nq; q += N) {
        const float4 x = qx[q], y = qy[q], z = qz[q];
        const int j = head + 4 * q;
        const int4 o = qo ? qo[q] : make_int4(j, j + 1, j + 2, j + 3);
        put(o.x, x.x, y.x, z.x);
        put(o.y, x.y, y.y, z.y);
        put(o.z, x.z, y.z, z.z);
        put(o.w, x.w, y.w, z.w);
    }
}

// A hull's vertex range and sweep metadata (see d3d.h for the clustered
// layout: p1 = first cluster record, p2 = number of clusters, 0 = none).
struct HullRef {
    int voff, voff64, nv, coff, nc;
    double L;
};

__device__ __forceinline__ HullRef hull_ref(const Shape &h) {
    return HullRef{h.voff, h.voff64, h.nv, (int)h.p[1], (int)h.p[2], h.p[0]};
}

__device__ __forceinline__ HullRef readlane_hull(const HullRef &H, int lane) {
    return HullRef{readlane_i32(H.voff, lane), readlane_i32(H.voff64, lane),
                   readlane_i32(H.nv, lane),   readlane_i32(H.coff, lane),
                   readlane_i32(H.nc, lane),   readlane_f64(H.L, lane)};
}

// Bits of a wavefront ballot that belong to the G-lane group of `lane`.
template <int G>
__device__ __forceinline__ unsigned long long group_bits(unsigned long long m, int lane) {
    if (G == kWave) return m;
    return (m >> (lane & ~(G - 1))) & ((1ull << G) - 1ull);
}

// Exclusive prefix sum of x over the G lanes of a group (lane order); total
// receives the group's sum in every lane.
template <int G>
__device__ __forceinline__ int group_excl_scan(int x, int &total) {
    int pre = 0, tot = x;
    for_butterfly<G>([&](auto o) {
        constexpr int OFF = decltype(o)::value;
        int a, b;
        xpair<OFF>(tot, a, b);
        const bool upper = (lane_id() & OFF) != 0;
        int lo, hi;
        if constexpr (OFF <= 8) {
            lo = upper ? b : a;
            hi = upper ? a : b;
        } else {
            lo = a;
            hi = b;
        }
        if (upper) pre += lo;
        tot = lo + hi;
    });
    total = tot;
    return pre;
}

template <int G>
__device__ __forceinline__ float group_maxf(float x) {
    for_butterfly<G>([&](auto o) {
        float a, b;
        xpair<decltype(o)::value>(x, a, b);
        x = fmaxf(a, b);
    });
    return x;
}

// One FP32 score into the running (largest m1 at i1, largest other m2).
__device__ __forceinline__ void top2(float sc, int j, float &m1, float &m2, int &i1) {
    m2 = fmaxf(m2, fminf(sc, m1));
    if (sc > m1) {
        m1 = sc;
        i1 = j;
    }
}

// A lane's sweep state: its two best FP32 scores with their vertices and its
// third best score (ties are kept: a score equal to m1 becomes m2, one equal
// to m2 becomes m3).  When the certificate fails, the lanes whose m3 lies
// below the rescan threshold hold all their candidates in i1, i2, so only
// those are evaluated in FP64 instead of sweeping the hull again.
struct Top3 {
    float m1, m2, m3;
    int i1, i2;
    __device__ __forceinline__ void init() {
        m1 = m2 = m3 = -INFINITY;
        i1 = i2 = 0x7fffffff;
    }
    __device__ __forceinline__ void take(float sc, int j) {
        const bool c1 = sc > m1, c2 = sc > m2;
        m3 = c2 ? m2 : fmaxf(m3, sc);
        i2 = c1 ? i1 : (c2 ? j : i2);
        m2 = fmaxf(m2, fminf(sc, m1));
        if (c1) {
            m1 = sc;
            i1 = j;
        }
    }
    // true when the group's lanes hold every vertex scoring >= thr in i1, i2
    // (a NaN or -inf threshold, or a NaN score, never qualifies)
    template <int G>
    __device__ __forceinline__ bool group_holds(float thr) const {
        const bool over = !(m3 < thr) || (!(m1 < thr) && i1 == 0x7fffffff) ||
                          (!(m2 < thr) && i2 == 0x7fffffff);
        return group_bits<G>(__ballot(over), lane_id()) == 0ull;
    }
};

// One score into a sweep state: with kKeep as Top3, else as top2 (m3 and i2
// are not kept).
template <bool kKeep>
__device__ __forceinline__ void sweep_take(Top3 &T, float sc, int j) {
    if constexpr (kKeep)
        T.take(sc, j);
    else
        top2(sc, j, T.m1, T.m2, T.i1);
}

// Argmax over one hull, vertices[argmax(vertices . l)] with the first
// (original) index on ties (colliders.py:131-132, np.argmax), evaluated by
// the G lanes of a group (G a power of two <= 64; all lanes of the group
// call it with the same arguments, `sub` = lane index in the group).
//
// FP32 sweep with an exactness certificate.  The FP32 score of vertex v
// differs from the exact v . l by at most
//     E = 8 * 2^-24 * L * |l|_inf,   L = max_i (|x_i| + |y_i| + |z_i|)
// (rounding of l to FP32, of the vertex when an FP64 pool backs it, and of
// the three FMA steps, with a factor-of-two margin).  The lanes track the
// largest score m1 (vertex i1) and the largest score m2 among all other
// scanned vertices; if m1 - m2 > 2E after the group merge, i1 is the unique
// exact argmax.  Otherwise the hull is rescanned in FP64 and ties go to the
// smallest original index (vorig).
//
// Clustered hulls (nc > 0) are stored as consecutive clusters of
// D3D_CLUSTER_SIZE vertices, each with a sphere (c, r) such that
// fp32(c . d) + r * |l| (one FMA chain, |l| rounded up) is at least the
// exact score of every vertex of the cluster.  The sweep scans the cluster
// with the largest bound first (cooperatively), then every other cluster whose
// bound is not below m1 - 2.5E, each by one lane (few dependent memory
// rounds per sweep): a skipped vertex has exact score below m1 - 2.5E, hence
// below the exact score of the winner (>= m1 - E), so the certificate keeps
// its meaning.  The result equals the full sweep's bit for bit.
// Running FP64 argmax with ties to the smallest original index.
struct ArgmaxF64 {
    double best;
    int bi, bo;
    __device__ __forceinline__ void init() {
        best = -INFINITY;
        bi = 0x7fffffff;
        bo = 0x7fffffff;
    }
    __device__ __forceinline__ void take(double val, int j, int oj) {
        if (val > best || (val == best && oj < bo)) {
            best = val;
            bi = j;
            bo = oj;
        }
    }
    // merge over the G lanes of a group; the index (0 when nothing was taken)
    template <int G>
    __device__ __forceinline__ int group_result() {
        for_butterfly<G>([&](auto o) {
            constexpr int OFF = decltype(o)::value;
            double ab, bb;
            int ai, bi2, ao, bo2;
            xpair<OFF>(best, ab, bb);
            xpair<OFF>(bi, ai, bi2);
            xpair<OFF>(bo, ao, bo2);
            const bool t = (bb > ab) || (bb == ab && bo2 < ao);
            best = t ? bb : ab;
            bi = t ? bi2 : ai;
            bo = t ? bo2 : ao;
        });
        return bi == 0x7fffffff ? 0 : bi;
    }
};

// FP64 score of vertex j with its original index, into the running argmax.
__device__ __forceinline__ void rescan_take(const d3d_scene &s, const HullRef &H, v3 l, int j,
                                            ArgmaxF64 &A) {
    const v3 v = hull_vertex(s, H.voff, H.voff64, j);
    A.take(v.x * l.x + v.y * l.y + v.z * l.z, j, s.vorig ? s.vorig[H.voff + j] : j);
}

// FP64 argmax over the hull by the G lanes of a group, ties to the smallest
// original index (vorig): the exact fallback of the FP32 certificate.  Only
// vertices whose FP32 score reaches thr = rd(m1 - 2.5E) are evaluated in FP64
// (clusters whose bound is below thr are skipped): the FP64 argmax j* and
// every vertex tied with it score at least m1 - E - 2d exactly (d: FP64
// rounding, far below E), hence at least m1 - 2E - 2d in FP32, so no
// candidate is missed.  A NaN or -inf threshold evaluates every vertex.
template <int G, bool kBounds = false>
__device__ __forceinline__ int hull_rescan(const d3d_scene &s, const HullRef &H, v3 l, int sub,
                                           float thr, const float4 *bounds = nullptr) {
    constexpr int K = D3D_CLUSTER_SIZE;
    const float dx = (float)l.x, dy = (float)l.y, dz = (float)l.z;
    const float *__restrict__ px = s.vx + H.voff;
    const float *__restrict__ py = s.vy + H.voff;
    const float *__restrict__ pz = s.vz + H.voff;
    const int nv = H.nv;
    if (thr != thr) thr = -INFINITY;
    ArgmaxF64 A;
    A.init();
    auto score = [&](int j) {
        return __builtin_fmaf(pz[j], dz, __builtin_fmaf(py[j], dy, px[j] * dx));
    };
    if (H.nc == 0) {
        for (int j = sub; j < nv; j += G)
            if (!(score(j) < thr)) rescan_take(s, H, l, j, A);
    } else {
        const float4 *__restrict__ cl =
            kBounds ? bounds : reinterpret_cast<const float4 *>(s.clusters) + H.coff;
        const float nl = (float)(sqrt(l.x * l.x + l.y * l.y + l.z * l.z) * (1.0 + 0x1p-19));
        for (int k = sub; k < H.nc; k += G) {
            const float4 c = cl[k];
            if (__builtin_fmaf(c.w, nl, __builtin_fmaf(c.z, dz, __builtin_fmaf(c.y, dy, c.x * dx))) < thr)
                continue;
            const int j0 = k * K;
            if (j0 + K <= nv) {
                // aligned cluster: four vertices per load, loads issued together
                const float4 *__restrict__ qx = reinterpret_cast<const float4 *>(px + j0);
                const float4 *__restrict__ qy = reinterpret_cast<const float4 *>(py + j0);
                const float4 *__restrict__ qz = reinterpret_cast<const float4 *>(pz + j0);
#pragma unroll 2
                for (int q = 0; q < K / 4; ++q) {
                    const float4 x = qx[q], y = qy[q], z = qz[q];
                    const int j = j0 + 4 * q;
                    if (!(__builtin_fmaf(z.x, dz, __builtin_fmaf(y.x, dy, x.x * dx)) < thr)) rescan_take(s, H, l, j, A);
                    if (!(__builtin_fmaf(z.y, dz, __builtin_fmaf(y.y, dy, x.y * dx)) < thr)) rescan_take(s, H, l, j + 1, A);
                    if (!(__builtin_fmaf(z.z, dz, __builtin_fmaf(y.z, dy, x.z * dx)) < thr)) rescan_take(s, H, l, j + 2, A);
                    if (!(__builtin_fmaf(z.w, dz, __builtin_fmaf(y.w, dy, x.w * dx)) < thr)) rescan_take(s, H, l, j + 3, A);
                }
            } else {
                for (int j = j0; j < nv; ++j)
                    if (!(score(j) < thr)) rescan_take(s, H, l, j, A);
            }
        }
    }
    return A.group_result<G>();
}

// With kBounds, `bounds` is a copy of the hull's cluster records (in LDS)
// read instead of the scene's; the result is the same.  kKeep: the lanes keep
// their candidates (Top3) so a failed certificate usually evaluates only
// those; without it (one-lane sweeps of whole hulls, where the extra work per
// vertex costs more than the rescans save) a failed certificate always
// sweeps again.
template <int G, bool kBounds = false, bool kKeep = true>
__device__ __forceinline__ int hull_argmax(const d3d_scene &s, const HullRef &H, v3 l, int sub,
                                           const float4 *bounds = nullptr, int *list = nullptr,
                                           int list_cap = 0) {
    constexpr int K = D3D_CLUSTER_SIZE;
    const float dx = (float)l.x, dy = (float)l.y, dz = (float)l.z;
    const float *__restrict__ px = s.vx + H.voff;
    const float *__restrict__ py = s.vy + H.voff;
    const float *__restrict__ pz = s.vz + H.voff;
    const int nv = H.nv;
    const double linf = fmax(fabs(l.x), fmax(fabs(l.y), fabs(l.z)));
    const double E = 8.0 * 5.9604644775390625e-08 * H.L * linf;
    Top3 T;
    T.init();
    auto score = [&](int j) {
        return __builtin_fmaf(pz[j], dz, __builtin_fmaf(py[j], dy, px[j] * dx));
    };
    if (H.nc == 0) {
        // strided sweep, four vertices per lane per step (loads issued together)
        int k = sub;
        for (; k + 3 * G < nv; k += 4 * G) {
            const float s0 = score(k), s1 = score(k + G), s2 = score(k + 2 * G), s3 = score(k + 3 * G);
            sweep_take<kKeep>(T, s0, k);
            sweep_take<kKeep>(T, s1, k + G);
            sweep_take<kKeep>(T, s2, k + 2 * G);
            sweep_take<kKeep>(T, s3, k + 3 * G);
        }
        for (; k < nv; k += G) sweep_take<kKeep>(T, score(k), k);
    } else {
        const float4 *__restrict__ cl =
            kBounds ? bounds : reinterpret_cast<const float4 *>(s.clusters) + H.coff;
        const int nc = H.nc;
        const float nl = (float)(sqrt(l.x * l.x + l.y * l.y + l.z * l.z) * (1.0 + 0x1p-19));
        auto bound = [&](int k) {
            const float4 c = cl[k];
            return __builtin_fmaf(c.w, nl, __builtin_fmaf(c.z, dz, __builtin_fmaf(c.y, dy, c.x * dx)));
        };
        // the cluster with the largest bound (first on ties)
        float bu = -INFINITY;
        int bk = 0x7fffffff;
#pragma unroll 4
        for (int k = sub; k < nc; k += G) {
            const float ub = bound(k);
            if (ub > bu) {
                bu = ub;
                bk = k;
            }
        }
        group_argmaxf<G>(bu, bk);
        const int ks = bk < nc ? bk : 0;
        // the best cluster, scanned by the group (lane sub takes slots sub, sub+G, ...)
        {
            const int j0 = ks * K;
            const int cnt = min(K, nv - j0);
#pragma unroll
            for (int o = sub % K; o < K; o += (G < K ? G : K))
                if ((G <= K || sub < K) && o < cnt) sweep_take<kKeep>(T, score(j0 + o), j0 + o);
        }
        float thr = __double2float_rd((double)group_maxf<G>(T.m1) - 2.5 * E);
        bool coop = false;
        if constexpr (kBounds && G >= 8) coop = list != nullptr && nc <= list_cap && nc <= 32 * G;
        if (coop) {
            // every other cluster whose bound reaches the threshold goes into the
            // group's list (lane order), then the group's lanes take its
            // clusters' vertex quads, so all their loads go out at once
            unsigned mask = 0u;
            int cnt = 0;
            for (int i = 0, k = sub; k < nc; ++i, k += G) {
                if (k != ks && !(bound(k) < thr)) {
                    mask |= 1u << i;
                    ++cnt;
                }
            }
            int total;
            int pos = group_excl_scan<G>(cnt, total);
            for (unsigned m = mask; m != 0u; m &= m - 1u) list[pos++] = sub + (__ffs((int)m) - 1) * G;
            // the group's lanes read what the others wrote
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll 2
            for (int t = sub; t < 4 * total; t += G) {
                const int j = list[t >> 2] * K + 4 * (t & 3);
                if (j + 4 <= nv) {
                    const float4 x = *reinterpret_cast<const float4 *>(px + j);
                    const float4 y = *reinterpret_cast<const float4 *>(py + j);
                    const float4 z = *reinterpret_cast<const float4 *>(pz + j);
                    sweep_take<kKeep>(T, __builtin_fmaf(z.x, dz, __builtin_fmaf(y.x, dy, x.x * dx)), j);
                    sweep_take<kKeep>(T, __builtin_fmaf(z.y, dz, __builtin_fmaf(y.y, dy, x.y * dx)), j + 1);
                    sweep_take<kKeep>(T, __builtin_fmaf(z.z, dz, __builtin_fmaf(y.z, dy, x.z * dx)), j + 2);
                    sweep_take<kKeep>(T, __builtin_fmaf(z.w, dz, __builtin_fmaf(y.w, dy, x.w * dx)), j + 3);
                } else {
                    for (int e = j; e < nv; ++e) sweep_take<kKeep>(T, score(e), e);
                }
            }
        }
        // otherwise every other cluster whose bound reaches the threshold is
        // scanned whole by the lane that owns it (16-byte aligned, four vertices
        // per load); the threshold only rises, with the lane's own running maximum
        for (int k = sub; !coop && k < nc; k += G) {
            if (k == ks || bound(k) < thr) continue;
            const int j0 = k * K;
            if (j0 + K <= nv) {
                const float4 *__restrict__ qx = reinterpret_cast<const float4 *>(px + j0);
                const float4 *__restrict__ qy = reinterpret_cast<const float4 *>(py + j0);
                const float4 *__restrict__ qz = reinterpret_cast<const float4 *>(pz + j0);
#pragma unroll 2
                for (int c = 0; c < K / 4; ++c) {
                    const float4 x = qx[c], y = qy[c], z = qz[c];
                    const int j = j0 + 4 * c;
                    sweep_take<kKeep>(T, __builtin_fmaf(z.x, dz, __builtin_fmaf(y.x, dy, x.x * dx)), j);
                    sweep_take<kKeep>(T, __builtin_fmaf(z.y, dz, __builtin_fmaf(y.y, dy, x.y * dx)), j + 1);
                    sweep_take<kKeep>(T, __builtin_fmaf(z.z, dz, __builtin_fmaf(y.z, dy, x.z * dx)), j + 2);
                    sweep_take<kKeep>(T, __builtin_fmaf(z.w, dz, __builtin_fmaf(y.w, dy, x.w * dx)), j + 3);
                }
            } else {
                for (int j = j0; j < nv; ++j) sweep_take<kKeep>(T, score(j), j);
            }
            thr = fmaxf(thr, __double2float_rd((double)T.m1 - 2.5 * E));
        }
    }
    float m1 = T.m1, m2 = T.m2;
    int i1 = T.i1;
    group_top2<G>(m1, m2, i1);
    if ((double)m1 - (double)m2 > 2.0 * E) return i1;
    float thr = __double2float_rd((double)m1 - 2.5 * E);
    if (thr != thr) thr = -INFINITY;
    if (kKeep && T.group_holds<G>(thr)) {
        // every candidate is some lane's i1 or i2: FP64 scores of those only
        ArgmaxF64 A;
        A.init();
        if (!(T.m1 < thr)) rescan_take(s, H, l, T.i1, A);
        if (!(T.m2 < thr)) rescan_take(s, H, l, T.i2, A);
        return A.group_result<G>();
    }
    return hull_rescan<G, kBounds>(s, H, l, sub, thr, bounds);
}

// Argmax over a hull whose FP32 vertices sit in LDS in their ORIGINAL order
// (cx/cy/cz[off + j] = vertex j; the hull's vertices are exact in FP32), by
// the G lanes of a group: the sweep and certificate of hull_argmax, and, when
// the certificate fails, the FP64 rescan of the candidates from the same LDS
// copy with ties to the smallest (original) index.  Returns the original
// index of vertices[argmax(vertices . l)] (colliders.py:131-132, first index
// on ties), the same vertex hull_argmax finds in the pools.
template <int G>
__device__ __forceinline__ int lds_orig_argmax(const float *cx, const float *cy, const float *cz,
                                               int off, int nv, double Lmax, v3 l, int sub) {
    const float dx = (float)l.x, dy = (float)l.y, dz = (float)l.z;
    const double linf = fmax(fabs(l.x), fmax(fabs(l.y), fabs(l.z)));
    const double E = 8.0 * 5.9604644775390625e-08 * Lmax * linf;
    Top3 T;
    T.init();
    for (int j = sub; j < nv; j += G)
        T.take(__builtin_fmaf(cz[off + j], dz, __builtin_fmaf(cy[off + j], dy, cx[off + j] * dx)), j);
    float m1 = T.m1, m2 = T.m2;
    int i1 = T.i1;
    group_top2<G>(m1, m2, i1);
    if ((double)m1 - (double)m2 > 2.0 * E) return i1;
    float thr = __double2float_rd((double)m1 - 2.5 * E);
    if (thr != thr) thr = -INFINITY;
    ArgmaxF64 A;
    A.init();
    if (T.group_holds<G>(thr)) {
        // every candidate is some lane's i1 or i2: FP64 scores of those only
        if (!(T.m1 < thr))
            A.take((double)cx[off + T.i1] * l.x + (double)cy[off + T.i1] * l.y +
                       (double)cz[off + T.i1] * l.z,
                   T.i1, T.i1);
        if (!(T.m2 < thr))
            A.take((double)cx[off + T.i2] * l.x + (double)cy[off + T.i2] * l.y +
                       (double)cz[off + T.i2] * l.z,
                   T.i2, T.i2);
        return A.group_result<G>();
    }
    for (int j = sub; j < nv; j += G) {
        const float x = cx[off + j], y = cy[off + j], z = cz[off + j];
        if (!(__builtin_fmaf(z, dz, __builtin_fmaf(y, dy, x * dx)) < thr))
            A.take((double)x * l.x + (double)y * l.y + (double)z * l.z, j, j);
    }
    return A.group_result<G>();
}

// Values of x in the lower and in the upper H-lane half of the lane's aligned
// 2H-lane block, in every lane of the block (x uniform within each half).
template <int H>
__device__ __forceinline__ void halves(double x, double &lo, double &hi) {
    double a, b;
    xpair<H>(x, a, b);
    if constexpr (H <= 8) {
        const bool lower = (lane_id() & H) == 0;
        lo = lower ? a : b;
        hi = lower ? b : a;
    } else {
        lo = a;
        hi = b;
    }
}

// Closed-form supports, FP64, in the same operation order as the reference.
__device__ __forceinline__ v3 support_analytic(const Shape &h, v3 d) {
    switch (h.type) {
    case D3D_SPHERE: {  // geometry.py:314-346
        const double r = h.p[0];
        const double sn = sqrt(dot(d, d));
        if (sn == 0.0) return mk(h.t[0] + 0.0, h.t[1] + 0.0, h.t[2] + r);
        return mk(h.t[0] + d.x / sn * r, h.t[1] + d.y / sn * r, h.t[2] + d.z / sn * r);
    }
    case D3D_CAPSULE: {  // geometry.py:209-256
        const v3 l = to_local(h, d);
        const double r = h.p[0], ht = h.p[1];
        const double s = sqrt(l.x * l.x + l.y * l.y + l.z * l.z);
        v3 p;
        if (s == 0.0) {
            p = mk(r, 0.0, 0.0);
        } else {
            const double k = r / s;
            p = mk(l.x * k, l.y * k, l.z * k);
        }
        if (l.z > 0.0) p.z += 0.5 * ht;
        else p.z -= 0.5 * ht;
        return to_world(h, p);
    }
    case D3D_ELLIPSOID: {  // geometry.py:259-284
        const v3 l = to_local(h, d);
        const v3 q = norm_vector(mk(l.x * h.p[0], l.y * h.p[1], l.z * h.p[2]));
        return to_world(h, mk(q.x * h.p[0], q.y * h.p[1], q.z * h.p[2]));
    }
    case D3D_CYLINDER: {  // geometry.py:160-206
        const v3 l = to_local(h, d);
        const double r = h.p[0], L = h.p[1];
        const double s = sqrt(l.x * l.x + l.y * l.y);
        const double z = (l.z < 0.0) ? -0.5 * L : 0.5 * L;
        v3 p;
        if (s == 0.0) {
            p = mk(r, 0.0, z);
        } else {
            const double k = r / s;
            p = mk(l.x * k, l.y * k, z);
        }
        return to_world(h, p);
    }
    case D3D_BOX: {  // convert_box_to_vertices + argmax, geometry.py:135-157
        const v3 l = to_local(h, d);
        return to_world(h, mk((l.x > 0.0 ? 0.5 : -0.5) * h.p[0], (l.y > 0.0 ? 0.5 : -0.5) * h.p[1],
                              (l.z > 0.0 ? 0.5 : -0.5) * h.p[2]));
    }
    case D3D_CONE: {  // geometry.py:417-454
        const v3 l = to_local(h, d);
        const double r = h.p[0], ht = h.p[1];
        v3 dp = mk(l.x, l.y, 0.0);
        const double n = sqrt(dot(dp, dp));
        if (n == 0.0) {
            dp = mk(0.0, 0.0, 0.0);
        } else {
            const double k = r / n;
            dp = mk(dp.x * k, dp.y * k, dp.z * k);
        }
        const v3 p = vsel(dot(l, dp) >= l.z * ht, dp, mk(0.0, 0.0, ht));
        return to_world(h, p);
    }
    case D3D_DISK: {  // geometry.py:349-383
        v3 l = to_local(h, d);
        l.z = 0.0;
        const double n = sqrt(dot(l, l));
        if (n == 0.0) return mk(h.t[0], h.t[1], h.t[2]);
        const double k = h.p[0] / n;
        l = mk(l.x * k, l.y * k, l.z * k);
        return to_world(h, l);
    }
    case D3D_ELLIPSE: {  // geometry.py:386-414
        const double l0 = h.R[0] * d.x + h.R[3] * d.y + h.R[6] * d.z;
        const double l1 = h.R[1] * d.x + h.R[4] * d.y + h.R[7] * d.z;
        double q0 = h.p[0] * l0, q1 = h.p[1] * l1;
        const double n = sqrt(q0 * q0 + q1 * q1);
        if (n != 0.0) {
            q0 = q0 / n;
            q1 = q1 / n;
        }
        const double a0 = q0 * h.p[0], a1 = q1 * h.p[1];
        return mk(h.t[0] + (a0 * h.R[0] + a1 * h.R[1]), h.t[1] + (a0 * h.R[3] + a1 * h.R[4]),
                  h.t[2] + (a0 * h.R[6] + a1 * h.R[7]));
    }
    default:
        return mk(NAN, NAN, NAN);
    }
}

// Margin adds margin * norm_vector(d) (colliders.py:629-631).
__device__ __forceinline__ v3 apply_margin(const Shape &h, v3 p, v3 d) {
    if (h.margin != 0.0) p = p + norm_vector(d) * h.margin;
    return p;
}

// Supports of record ia along d and of record ib along -d (the two halves of
// a Minkowski-difference support), with a single copy of the sweep compiled.  The G lanes of a group call it
// together with the same arguments; records are reloaded by index and must
// have been validated (load_checked) by the caller.  With G == 64 the
// arguments are wave-uniform and are held in scalar registers.
struct SupportPair {
    v3 a, b;
};

template <int G, bool kHull>
__device__ __forceinline__ SupportPair support_pair(const d3d_scene &s, int ia, int ib, v3 d,
                                                    int sub) {
    v3 a = mk(0.0, 0.0, 0.0), b = a;
    if (G == kWave) d = mk(wave_uniform(d.x), wave_uniform(d.y), wave_uniform(d.z));
#pragma unroll 1
    for (int k = 0; k < 2; ++k) {
        const int idx = k == 0 ? ia : ib;
        const Shape h = load_shape(s, G == kWave ? wave_uniform(idx) : opaque(idx));
        const v3 dir = vsel(k == 0, d, -d);
        v3 out;
        if (kHull && h.type == D3D_HULL) {
            const v3 l = to_local(h, dir);
            const int hi = hull_argmax<G>(s, hull_ref(h), l, sub);
            out = to_world(h, hull_vertex(s, h.voff, h.voff64, hi));
        } else {
            out = support_analytic(h, dir);
        }
        out = apply_margin(h, out, dir);
        a = vsel(k == 0, out, a);
        b = vsel(k == 0, b, out);
    }
    return SupportPair{a, b};
}

// support_pair for two analytic records of known types ta, tb: only their
// FP64 blocks are reloaded per call, both records' loads issued together.
__device__ __forceinline__ SupportPair support_pair_analytic(const d3d_scene &s, int ia, int ib,
                                                             int ta, int tb, v3 d) {
    v3 a = mk(0.0, 0.0, 0.0), b = a;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const Shape h = load_shape_f64(s, opaque(k == 0 ? ia : ib), k == 0 ? ta : tb);
        const v3 dir = vsel(k == 0, d, -d);
        const v3 out = apply_margin(h, support_analytic(h, dir), dir);
        a = vsel(k == 0, out, a);
        b = vsel(k == 0, b, out);
    }
    return SupportPair{a, b};
}

// Support of one collider per lane along d (per lane).  Hulls with more than
// serial_max vertices are swept by the whole wavefront, one lane's hull at a
// time.  Must be called by all 64 lanes (wave-uniform control flow); lanes
// with want == false take part in the sweeps of the others.
__device__ __forceinline__ v3 support_wave(const d3d_scene &s, const Shape &h, v3 d, bool want,
                                           int serial_max) {
    const bool coop = want && h.type == D3D_HULL && h.nv > serial_max;
    v3 out = mk(0.0, 0.0, 0.0);
    int idx = 0;
    v3 l = mk(0.0, 0.0, 0.0);
    if (want && h.type == D3D_HULL) l = to_local(h, d);
    if (want && !coop) {
        if (h.type == D3D_HULL) {
            idx = hull_argmax<1, false, false>(s, hull_ref(h), l, 0);
        } else {
            out = support_analytic(h, d);
        }
    }
    unsigned long long m = __ballot(coop);
    const HullRef H = hull_ref(h);
    while (m) {
        const int j = __ffsll((long long)m) - 1;
        m &= m - 1;
        const v3 lj = mk(readlane_f64(l.x, j), readlane_f64(l.y, j), readlane_f64(l.z, j));
        const int best = hull_argmax<kWave, false, false>(s, readlane_hull(H, j), lj, lane_id());
        if (lane_id() == j) idx = best;
    }
    if (want && h.type == D3D_HULL) out = to_world(h, hull_vertex(s, h.voff, h.voff64, idx));
    if (want) out = apply_margin(h, out, d);
    return out;
}

// ----------------------------------------------------- simplex routines
// Straight restatements of gjk/_gjk_jolt.py; arrays are indexed with
// compile-time constants only so the simplex stays in VGPRs.

// get_barycentric_coordinates_line (_gjk_jolt.py:291-312)
__device__ __forceinline__ void bary_line(v3 a, v3 b, double &u, double &v) {
    const v3 ab = b - a;
    const double den = dot(ab, ab);
    if (den < kEpsSqr) {
        if (dot(a, a) < dot(b, b)) {
            u = 1.0;
            v = 0.0;
        } else {
            u = 0.0;
            v = 1.0;
        }
    } else {
        v = -dot(a, ab) / den;
        u = 1.0 - v;
    }
}

// get_barycentric_coordinates_plane (_gjk_jolt.py:315-372)
__device__ __forceinline__ void bary_plane(v3 a, v3 b, v3 c, double &u, double &v, double &w) {
    const v3 v0 = b - a, v1 = c - a, v2 = c - b;
    const double d00 = dot(v0, v0), d11 = dot(v1, v1), d22 = dot(v2, v2);
    if (d00 <= d22) {
        const double d01 = dot(v0, v1);
        const double den = d00 * d11 - d01 * d01;
        if (fabs(den) < kEps) {
            if (d00 > d11) {
                bary_line(a, b, u, v);
                w = 0.0;
            } else {
                bary_line(a, c, u, w);
                v = 0.0;
            }
        } else {
            const double a0 = dot(a, v0), a1 = dot(a, v1);
            v = (d01 * a1 - d11 * a0) / den;
            w = (d01 * a0 - d00 * a1) / den;
            u = 1.0 - v - w;
        }
    } else {
        const double d12 = dot(v1, v2);
        const double den = d11 * d22 - d12 * d12;
        if (fabs(den) < kEps) {
            if (d11 > d22) {
                bary_line(a, c, u, w);
                v = 0.0;
            } else {
                bary_line(b, c, v, w);
                u = 0.0;
            }
        } else {
            const double c1 = dot(c, v1), c2 = dot(c, v2);
            u = (d22 * c1 - d12 * c2) / den;
            v = (d11 * c2 - d12 * c1) / den;
            w = 1.0 - u - v;
        }
    }
}

// get_barycentric_coordinates_tetrahedron (_gjk_jolt.py:375-390)
__device__ __forceinline__ void bary_tetra(v3 a, v3 b, v3 c, v3 d, double &u, double &v,
                                           double &w, double &x) {
    const v3 vab = b - a, vac = c - a, vad = d - a;
    const double va6 = -stp(b, d - b, c - b);
    const double vb6 = -stp(a, vac, vad);
    const double vc6 = -stp(a, vad, vab);
    const double vd6 = -stp(a, vab, vac);
    const double v6 = 1.0 / stp(vab, vac, vad);
    u = va6 * v6;
    v = vb6 * v6;
    w = vc6 * v6;
    x = vd6 * v6;
}

// closest_point_line (_gjk_jolt.py:393-412)
__device__ __forceinline__ v3 cp_line(v3 a, v3 b, int &set) {
    double u, v;
    bary_line(a, b, u, v);
    if (v <= 0.0) {
        set = 1;
        return a;
    }
    if (u <= 0.0) {
        set = 2;
        return b;
    }
    set = 3;
    return a * u + b * v;
}

// closest_point_triangle (_gjk_jolt.py:415-523)
__device__ __forceinline__ v3 cp_triangle(v3 a, v3 b, v3 c, int &set) {
    const v3 ab = b - a, ac = c - a, bc = c - b;
    // the normal from the shorter of bc, ac (one cross product of the chosen edge)
    const v3 n = cross(ab, vsel(dot(bc, bc) < dot(ac, ac), bc, ac));
    const double nn = dot(n, n);
    if (nn < kEpsSqr) {
        int s, ns;
        v3 cp = cp_line(a, b, s);
        double best = dot(cp, cp);
        v3 q = cp_line(a, c, ns);
        double dd = dot(q, q);
        if (dd < best) {
            cp = q;
            best = dd;
            s = (ns & 1) + ((ns & 2) << 1);
        }
        q = cp_line(b, c, ns);
        dd = dot(q, q);
        if (dd < best) {
            cp = q;
            s = ns << 1;
        }
        set = s;
        return cp;
    }
    const v3 ap = -a;
    const double d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0.0 && d2 <= 0.0) {
        set = 1;
        return a;
    }
    const v3 bp = -b;
    const double d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0.0 && d4 <= d3) {
        set = 2;
        return b;
    }
    const double vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0 && 0.0 <= d1 && d3 <= 0.0) {
        const double v = d1 / (d1 - d3);
        set = 3;
        return a + ab * v;
    }
    const v3 cpt = -c;
    const double d5 = dot(ab, cpt), d6 = dot(ac, cpt);
    if (d6 >= 0.0 && d5 <= d6) {
        set = 4;
        return c;
    }
    const double vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0 && 0.0 <= d2 && d6 <= 0.0) {
        const double w = d2 / (d2 - d6);
        set = 5;
        return a + ac * w;
    }
    const double va = d3 * d6 - d5 * d4;
    const double d4_d3 = d4 - d3, d5_d6 = d5 - d6;
    if (va <= 0.0 && 0.0 <= d4_d3 && d5_d6 >= 0.0) {
        const double w = d4_d3 / (d4_d3 + d5_d6);
        set = 6;
        return b + bc * w;
    }
    // n * (a + b + c).dot(n) / (3.0 * n_len_sq)
    const double s = dot((a + b) + c, n);
    const double den = 3.0 * nn;
    set = 7;
    return mk(n.x * s / den, n.y * s / den, n.z * s / den);
}

// origin_outside_of_tetrahedron_planes (_gjk_jolt.py:526-570) as a bit mask
// over the faces ABC, ACD, ADB, BDC.
__device__ __forceinline__ int outside_planes_mask(v3 a, v3 b, v3 c, v3 d) {
    const v3 ab = b - a, ac = c - a, ad = d - a, bd = d - b, bc = c - b;
    const v3 abc = cross(ab, ac), acd = cross(ac, ad), adb = cross(ad, ab), bdc = cross(bd, bc);
    const double sp0 = dot(a, abc), sp1 = dot(a, acd), sp2 = dot(a, adb), sp3 = dot(b, bdc);
    const double sd0 = dot(ad, abc), sd1 = dot(ab, acd), sd2 = dot(ac, adb), sd3 = -dot(ab, bdc);
    if (sd0 > 0.0 && sd1 > 0.0 && sd2 > 0.0 && sd3 > 0.0)
        return (sp0 >= -kEps ? 1 : 0) | (sp1 >= -kEps ? 2 : 0) | (sp2 >= -kEps ? 4 : 0) |
               (sp3 >= -kEps ? 8 : 0);
    if (sd0 < 0.0 && sd1 < 0.0 && sd2 < 0.0 && sd3 < 0.0)
        return (sp0 <= kEps ? 1 : 0) | (sp1 <= kEps ? 2 : 0) | (sp2 <= kEps ? 4 : 0) |
               (sp3 <= kEps ? 8 : 0);
    return 15;  // mixed signs: degenerate tetrahedron, test every face
}

// closest_point_tetrahedron (_gjk_jolt.py:573-631) and, for three points,
// closest_point_triangle: the faces ABC, ACD, ADB, BDC are visited in order,
// each by its own inlined copy of the triangle code with its vertices named
// directly (a loop over the faces picked them with a dozen selects per
// coordinate and was slower); face 0 always sets the result, later faces only
// when strictly closer; the triangle's feature bits are remapped to the
// tetrahedron's vertex order.
__device__ __forceinline__ v3 cp_faces(v3 y0, v3 y1, v3 y2, v3 y3, int mask, int &set) {
    int s = 15;
    v3 cp = mk(0.0, 0.0, 0.0);
    double best = kMaxFloat;
    auto face = [&](int f, v3 a, v3 b, v3 c) {
        int ns;
        const v3 q = cp_triangle(a, b, c, ns);
        const double dd = dot(q, q);
        if (f == 0 || dd < best) {
            best = dd;
            cp = q;
            s = (f == 0) ? ns
              : (f == 1) ? (ns & 1) + ((ns & 6) << 1)
              : (f == 2) ? (ns & 1) + ((ns & 2) << 2) + ((ns & 4) >> 1)
                         : ((ns & 1) << 1) + ((ns & 2) << 2) + (ns & 4);
        }
    };
    if (mask & 1) face(0, y0, y1, y2);
    if (mask & 2) face(1, y0, y2, y3);
    if (mask & 4) face(2, y0, y3, y1);
    if (mask & 8) face(3, y1, y3, y2);
    set = s;
    return cp;
}

// Minkowski-difference simplex Y of up to four points, held in named members
// so it stays in VGPRs.  Rows at index >= n keep stale contents exactly as
// the reference's in-place compaction leaves them.  The support points on A
// and B that go with each row (P, Q) are kept by the caller (see
// d3d_gjk_queue.hip), moved with the same compaction map.
struct Simplex {
    v3 Y0, Y1, Y2, Y3;
    int n;
};

__device__ __forceinline__ void simplex_zero(Simplex &S) {
    S.Y0 = mk(0.0, 0.0, 0.0);
    S.Y1 = mk(0.0, 0.0, 0.0);
    S.Y2 = mk(0.0, 0.0, 0.0);
    S.Y3 = mk(0.0, 0.0, 0.0);
    S.n = 0;
}

// Y[n] = y; n += 1
__device__ __forceinline__ void simplex_append(Simplex &S, v3 y) {
    S.Y0 = vsel(S.n == 0, y, S.Y0);
    S.Y1 = vsel(S.n == 1, y, S.Y1);
    S.Y2 = vsel(S.n == 2, y, S.Y2);
    S.Y3 = vsel(S.n >= 3, y, S.Y3);
    S.n += 1;
}

// Order-preserving compaction map of update_simplex_ypq (_gjk_jolt.py:654-664):
// new row j (< m) takes old row k_j.
struct CompactMap {
    int k0, k1, k2, m;
    bool moved;
};

// The compaction map of update_simplex_ypq (_gjk_jolt.py:654-664) for the
// rows of `set` among the first n.
__device__ __forceinline__ CompactMap compact_map(int n, int set) {
    set &= (1 << n) - 1;
    CompactMap c;
    c.m = __popc(set);
    c.k0 = 0;
    c.k1 = 1;
    c.k2 = 2;
    c.moved = set != (1 << c.m) - 1;
    if (c.moved) {
        c.k0 = __ffs(set) - 1;
        const int r1 = set & (set - 1);
        c.k1 = r1 ? __ffs(r1) - 1 : 3;
        const int r2 = r1 & (r1 - 1);
        c.k2 = r2 ? __ffs(r2) - 1 : 3;
    }
    return c;
}

// update_simplex_y / update_simplex_ypq on Y (_gjk_jolt.py:643-664); slots >=
// the new count are left as they were.  Returns the map to apply to P, Q.
__device__ __forceinline__ CompactMap simplex_compact(Simplex &S, int set) {
    const CompactMap c = compact_map(S.n, set);
    if (c.moved) {
        const v3 y0 = sel4(c.k0, S.Y0, S.Y1, S.Y2, S.Y3);
        const v3 y1 = sel4(c.k1, S.Y0, S.Y1, S.Y2, S.Y3);
        const v3 y2 = sel4(c.k2, S.Y0, S.Y1, S.Y2, S.Y3);
        S.Y0 = y0;
        S.Y1 = vsel(c.m > 1, y1, S.Y1);
        S.Y2 = vsel(c.m > 2, y2, S.Y2);
    }
    S.n = c.m;
    return c;
}

// max_y_length_squared (_gjk_jolt.py:634-640) of the simplex that compacting
// S to the rows of `set` leaves, read from S before the compaction: the kept
// rows in order (compaction keeps their order), row 0 alone when none is kept
// (it stays in place then).
__device__ __forceinline__ double max_y_len_sq_kept(const Simplex &S, int set) {
    set &= (1 << S.n) - 1;
    if (set == 0) return dot(S.Y0, S.Y0);
    double m = 0.0;
    bool any = false;
    auto acc = [&](bool keep, v3 y) {
        if (keep) {
            const double l = dot(y, y);
            m = !any ? l : ((l > m) ? l : m);
            any = true;
        }
    };
    acc(set & 1, S.Y0);
    acc(set & 2, S.Y1);
    acc(set & 4, S.Y2);
    acc(set & 8, S.Y3);
    return m;
}

// max_y_length_squared (_gjk_jolt.py:634-640)
__device__ __forceinline__ double max_y_len_sq(const Simplex &S) {
    double m = dot(S.Y0, S.Y0);
    if (S.n > 1) { const double l = dot(S.Y1, S.Y1); m = (l > m) ? l : m; }
    if (S.n > 2) { const double l = dot(S.Y2, S.Y2); m = (l > m) ? l : m; }
    if (S.n > 3) { const double l = dot(S.Y3, S.Y3); m = (l > m) ? l : m; }
    return m;
}

// get_closest_point_to_origin (_gjk_jolt.py:690-711)
__device__ __forceinline__ bool closest_to_origin(const Simplex &S, double prev, v3 &v, double &vv,
                                                  int &set) {
    v3 p;
    int s;
    if (S.n == 1) {
        s = 1;
        p = S.Y0;
    } else if (S.n == 2) {
        p = cp_line(S.Y0, S.Y1, s);
    } else {
        const int mask = (S.n == 3) ? 1 : outside_planes_mask(S.Y0, S.Y1, S.Y2, S.Y3);
        p = cp_faces(S.Y0, S.Y1, S.Y2, S.Y3, mask, s);
    }
    const double l = dot(p, p);
    if (l < prev) {
        v = p;
        vv = l;
        set = s;
        return true;
    }
    return false;
}

}  // namespace d3d
