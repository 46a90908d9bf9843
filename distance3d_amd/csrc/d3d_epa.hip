// d3d_epa.hip -- batched EPA penetration depth with contact points (gfx950).
//
// One wavefront (one 64-thread workgroup) per pair.  The polytope lives in
// LDS: vertex records (Minkowski point and the support points on A and B) and
// up to 128 faces (vertex indices, unit normal, distance); lane l owns faces l
// and l + 64 in every face-parallel step.  Each iteration:
//   closest face    wavefront argmin over face distances (first index on ties)
//   support         both colliders along the face normal (hull sweeps use all
//                   64 lanes, see hull_argmax)
//   visible faces   per-face test n . (w - v0) > eps
//   horizon         edges of removed faces whose reverse edge is not on
//                   another removed face (vertex-index matching)
//   rebuild         surviving faces compacted in order, new faces (horizon
//                   edge, w) appended in horizon order, via ballot prefixes
// The algorithm and its ordering rules are those of the FP64 checker's EPA
// (epa.py:9-202 with an oriented seed, index-matched horizon edges, contact
// points and a face cap that ends the loop instead of raising).
#include "d3d_device.hpp"
#include "d3d_runtime.hpp"

namespace d3d {

constexpr int kEpaMaxFaces = 128;
// polytope vertex rows: the 4 seed rows (plus scratch) and one per iteration;
// when they run out, rows no face refers to are recycled (epa_compact_rows).
// A closed triangulated polytope with F faces has F/2 + 2 vertices (66 for
// 128 faces), so recycling always frees rows unless the polytope is not a
// closed surface; then the loop ends like the face cap (status 2).
constexpr int kEpaMaxVerts = 72;
// FP32 vertex cache: a pair whose hulls have at most this many vertices in
// total is swept from LDS (most EPA pairs; larger ones use the global sweep)
constexpr int kEpaCache = 672;
constexpr int kEpaBlock = 64;

struct EpaLds {
    double wx[kEpaMaxVerts], wy[kEpaMaxVerts], wz[kEpaMaxVerts];
    double ax[kEpaMaxVerts], ay[kEpaMaxVerts], az[kEpaMaxVerts];
    double bx[kEpaMaxVerts], by[kEpaMaxVerts], bz[kEpaMaxVerts];
    double fnx[kEpaMaxFaces], fny[kEpaMaxFaces], fnz[kEpaMaxFaces], fd[kEpaMaxFaces];
    int fv0[kEpaMaxFaces], fv1[kEpaMaxFaces], fv2[kEpaMaxFaces];
    // directed edges (a -> b, bit a * kEpaMaxVerts + b) of the faces removed in
    // the current iteration; all zero between iterations
    unsigned eb[(kEpaMaxVerts * kEpaMaxVerts + 31) / 32];
    // the pair's hull vertices (pool order) and the offsets of A's and B's
    // copies (-1: not cached)
    float cx[kEpaCache], cy[kEpaCache], cz[kEpaCache];
    int coff[2];
    int split;  // A and B evaluated by the two halves at once
    // the two collider records (scene layout), staged once per pair
    double recf[2][16];
    int32_t reci[2][4];
};

// Collider record k (0: A, 1: B) of the wavefront's pair from LDS.
__device__ __forceinline__ Shape lds_shape(const EpaLds &L, int k) {
    Shape h;
#pragma unroll
    for (int j = 0; j < 9; ++j) h.R[j] = L.recf[k][j];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        h.t[j] = L.recf[k][9 + j];
        h.p[j] = L.recf[k][12 + j];
    }
    h.margin = L.recf[k][15];
    h.type = L.reci[k][0];
    h.nv = L.reci[k][1];
    h.voff = L.reci[k][2];
    h.voff64 = L.reci[k][3];
    return h;
}

__device__ __forceinline__ v3 readlane_v3(v3 v, int lane) {
    return mk(readlane_f64(v.x, lane), readlane_f64(v.y, lane), readlane_f64(v.z, lane));
}

// Supports of the pair's record A along d and of record B along -d for the
// EPA wavefront (see support_pair): the records come from LDS, staged once per
// pair, and LDS-cached hulls are swept from LDS.
__device__ __forceinline__ SupportPair epa_support_pair(const d3d_scene &s, const EpaLds &L, v3 d) {
    v3 a = mk(0.0, 0.0, 0.0), b = a;
    d = mk(wave_uniform(d.x), wave_uniform(d.y), wave_uniform(d.z));
    if (L.split) {
        // each collider analytic or a cached hull: lanes 0..31 evaluate A along
        // d, lanes 32..63 B along -d, at the same time (a hull is swept from
        // LDS by its half); the supports are read from lanes 0 and 32
        const int half = lane_id() >> 5;
        const Shape h = lds_shape(L, half);
        const v3 dir = vsel(half == 0, d, -d);
        v3 out;
        if (h.type == D3D_HULL) {
            const v3 l = to_local(h, dir);
            const int off = L.coff[half];
            const int idx = lds_orig_argmax<32>(L.cx, L.cy, L.cz, off, h.nv, h.p[0], l, lane_id() & 31);
            out = to_world(h, mk((double)L.cx[off + idx], (double)L.cy[off + idx], (double)L.cz[off + idx]));
        } else {
            out = support_analytic(h, dir);
        }
        out = apply_margin(h, out, dir);
        return SupportPair{readlane_v3(out, 0), readlane_v3(out, 32)};
    }
#pragma unroll 1
    for (int k = 0; k < 2; ++k) {
        const Shape h = lds_shape(L, k);
        const v3 dir = vsel(k == 0, d, -d);
        v3 out;
        if (h.type == D3D_HULL) {
            const v3 l = to_local(h, dir);
            const HullRef H = hull_ref(h);
            const int off = L.coff[k];
            v3 v;
            if (off >= 0) {
                const int idx = lds_orig_argmax<kWave>(L.cx, L.cy, L.cz, off, h.nv, h.p[0], l, lane_id());
                v = mk((double)L.cx[off + idx], (double)L.cy[off + idx], (double)L.cz[off + idx]);
            } else {
                v = hull_vertex(s, h.voff, h.voff64, hull_argmax<kWave>(s, H, l, lane_id()));
            }
            out = to_world(h, v);
        } else {
            out = support_analytic(h, dir);
        }
        out = apply_margin(h, out, dir);
        a = vsel(k == 0, out, a);
        b = vsel(k == 0, b, out);
    }
    return SupportPair{a, b};
}

__device__ __forceinline__ void edge_set(EpaLds &L, int a, int b) {
    const int bit = a * kEpaMaxVerts + b;
    atomicOr(&L.eb[bit >> 5], 1u << (bit & 31));
}
__device__ __forceinline__ void edge_clear(EpaLds &L, int a, int b) {
    const int bit = a * kEpaMaxVerts + b;
    atomicAnd(&L.eb[bit >> 5], ~(1u << (bit & 31)));
}
__device__ __forceinline__ bool edge_has(const EpaLds &L, int a, int b) {
    const int bit = a * kEpaMaxVerts + b;
    return (L.eb[bit >> 5] >> (bit & 31)) & 1u;
}

__device__ __forceinline__ v3 lds_w(const EpaLds &L, int v) { return mk(L.wx[v], L.wy[v], L.wz[v]); }

// Exclusive prefix sum over the wavefront (all lanes active) and its total:
// Hillis-Steele steps 1, 2, 4, 8 inside each 16-lane row by DPP row shifts
// (a lane shifted in from outside the row reads 0), then the row totals
// (lanes 15, 31, 47, 63) read into scalars and added to the later rows.
__device__ __forceinline__ int wave_excl_scan(int c, int lane, int &total) {
    int incl = c;
    incl += __builtin_amdgcn_update_dpp(0, incl, 0x111, 0xF, 0xF, false);  // row_shr:1
    incl += __builtin_amdgcn_update_dpp(0, incl, 0x112, 0xF, 0xF, false);  // row_shr:2
    incl += __builtin_amdgcn_update_dpp(0, incl, 0x114, 0xF, 0xF, false);  // row_shr:4
    incl += __builtin_amdgcn_update_dpp(0, incl, 0x118, 0xF, 0xF, false);  // row_shr:8
    const int r0 = readlane_i32(incl, 15), r1 = readlane_i32(incl, 31);
    const int r2 = readlane_i32(incl, 47), r3 = readlane_i32(incl, 63);
    const int row = lane >> 4;
    incl += (row > 0 ? r0 : 0) + (row > 1 ? r1 : 0) + (row > 2 ? r2 : 0);
    total = r0 + r1 + r2 + r3;
    return incl - c;
}

// make_face of the checker: unit normal of (w1 - w0) x (w2 - w0) and the
// distance w0 . n; false when the normal has zero length.
__device__ __forceinline__ bool face_normal(v3 w0, v3 w1, v3 w2, v3 &n, double &dist) {
    const v3 c = cross(w1 - w0, w2 - w0);
    const double l = sqrt(dot(c, c));
    if (!(l > 0.0)) return false;
    n = mk(c.x / l, c.y / l, c.z / l);
    dist = dot(w0, n);
    return true;
}

// face_normal's verdict alone: sqrt(x) > 0 exactly when x > 0 (x = c . c is
// never negative; NaN fails both), so no square root or division is needed to
// know whether the face is kept.
__device__ __forceinline__ bool face_valid(v3 w0, v3 w1, v3 w2) {
    const v3 c = cross(w1 - w0, w2 - w0);
    return dot(c, c) > 0.0;
}

// Polytope vertex row r of the LDS arrays: Minkowski point w and the support
// points on A and B.  Every lane of the wavefront computes the same values and
// writes them to the same addresses, so rows need no broadcast.
__device__ __forceinline__ v3 row_w(const EpaLds &L, int r) { return mk(L.wx[r], L.wy[r], L.wz[r]); }
__device__ __forceinline__ v3 row_a(const EpaLds &L, int r) { return mk(L.ax[r], L.ay[r], L.az[r]); }
__device__ __forceinline__ v3 row_b(const EpaLds &L, int r) { return mk(L.bx[r], L.by[r], L.bz[r]); }
__device__ __forceinline__ void put_row(EpaLds &L, int r, v3 w, v3 a, v3 b) {
    L.wx[r] = w.x; L.wy[r] = w.y; L.wz[r] = w.z;
    L.ax[r] = a.x; L.ay[r] = a.y; L.az[r] = a.z;
    L.bx[r] = b.x; L.by[r] = b.y; L.bz[r] = b.z;
}

// Seed tetrahedron containing the origin from a GJK simplex of n points held
// in rows 0..n-1 (uniform across the wavefront); completes rows n..3 with
// extra supports: n == 1, the six axis directions, keeping the point farthest
// from row 0; n == 2, four directions perpendicular to the segment, keeping
// the point farthest from its line; n == 3, both triangle normals, keeping
// the higher side.  One support call site (the loop below) keeps a single
// copy of the sweep; candidates go straight to their LDS rows (n == 3 uses
// row 4 as scratch).  Returns false for a degenerate seed.
__device__ bool epa_seed(const d3d_scene &s, EpaLds &L, int n) {
    if (n < 1 || n > 4) return false;
    while (n < 4) {
        const v3 w0 = row_w(L, 0);
        v3 ab = mk(0.0, 0.0, 0.0), p1 = ab, p2 = ab, nrm = ab;
        if (n == 2) {
            ab = row_w(L, 1) - w0;
            const double ax = fabs(ab.x), ay = fabs(ab.y), az = fabs(ab.z);
            const bool ex = ax <= ay && ax <= az, ey = !ex && ay <= az;
            const v3 e = mk(ex ? 1.0 : 0.0, ey ? 1.0 : 0.0, (!ex && !ey) ? 1.0 : 0.0);
            p1 = cross(ab, e);
            p2 = cross(ab, p1);
        } else if (n == 3) {
            nrm = cross(row_w(L, 1) - w0, row_w(L, 2) - w0);
            if (!(dot(nrm, nrm) > 0.0)) return false;
        }
        const int nk = n == 1 ? 6 : n == 2 ? 4 : 2;
        int best = -1;
        double bl = 0.0, h0 = 0.0, h1 = 0.0;
#pragma unroll 1
        for (int k = 0; k < nk; ++k) {
            v3 d;
            if (n == 1) {
                const double sg = (k & 1) ? -1.0 : 1.0;
                d = mk(k < 2 ? sg : 0.0, (k >= 2 && k < 4) ? sg : 0.0, k >= 4 ? sg : 0.0);
            } else if (n == 2) {
                d = sel4(k, p1, -p1, p2, -p2);
            } else {
                d = vsel(k == 0, nrm, -nrm);
            }
            const SupportPair sp = epa_support_pair(s, L, d);
            const v3 w = sp.a - sp.b;
            if (n == 3) {
                put_row(L, 3 + k, w, sp.a, sp.b);
                if (k == 0) h0 = dot(w - w0, nrm);
                else h1 = -dot(w - w0, nrm);
            } else {
                double l;
                if (n == 1) {
                    const v3 e = w - w0;
                    l = dot(e, e);
                } else {
                    const v3 cr = cross(ab, w - w0);
                    l = dot(cr, cr);
                }
                if (l > bl) {
                    bl = l;
                    best = k;
                    put_row(L, n, w, sp.a, sp.b);
                }
            }
        }
        if (n == 1) {
            if (best < 0 || !(bl > kEps)) return false;
        } else if (n == 2) {
            if (best < 0 || !(bl > kEps * dot(ab, ab))) return false;
        } else {
            if (!(h0 >= h1)) put_row(L, 3, row_w(L, 4), row_a(L, 4), row_b(L, 4));
            if (!((h0 > h1 ? h0 : h1) > 0.0)) return false;
        }
        ++n;
    }
    return true;
}

// Drops the polytope's vertex rows that no face refers to any more (vertices
// buried by earlier expansions) and renumbers the others in order, remapping
// the faces.  Only row numbers change, so the expansion goes on exactly as it
// would with more rows; a polytope of at most 128 faces refers to at most 66
// vertices.  The edge bitmap is all zero between iterations; its first three
// words hold the reference mask meanwhile.  Returns the new row count.
__device__ int epa_compact_rows(EpaLds &L, int nv, int nf, int lane) {
    static_assert(kEpaMaxVerts <= 96, "reference mask holds 96 rows");
#pragma unroll
    for (int sl = 0; sl < 2; ++sl) {
        const int k = sl * kWave + lane;
        if (k < nf) {
            const int v0 = L.fv0[k], v1 = L.fv1[k], v2 = L.fv2[k];
            atomicOr(&L.eb[v0 >> 5], 1u << (v0 & 31));
            atomicOr(&L.eb[v1 >> 5], 1u << (v1 & 31));
            atomicOr(&L.eb[v2 >> 5], 1u << (v2 & 31));
        }
    }
    __syncthreads();
    const unsigned m0 = L.eb[0], m1 = L.eb[1], m2 = L.eb[2];
    __syncthreads();
    if (lane < 3) L.eb[lane] = 0u;
    auto below = [&](int v) -> int {  // referenced rows before row v
        if (v < 32) return __popc(m0 & ((1u << v) - 1u));
        if (v < 64) return __popc(m0) + __popc(m1 & ((1u << (v - 32)) - 1u));
        return __popc(m0) + __popc(m1) + __popc(m2 & ((1u << (v - 64)) - 1u));
    };
    auto used = [&](int v) -> bool {
        const unsigned m = v < 32 ? m0 : v < 64 ? m1 : m2;
        return (m >> (v & 31)) & 1u;
    };
    double *const cols[9] = {L.wx, L.wy, L.wz, L.ax, L.ay, L.az, L.bx, L.by, L.bz};
    const int v1 = kWave + lane;
#pragma unroll
    for (int c = 0; c < 9; ++c) {
        // rows only move down: every lane reads before any lane writes
        const double x0 = lane < nv ? cols[c][lane] : 0.0;
        const double x1 = v1 < nv ? cols[c][v1] : 0.0;
        __syncthreads();
        if (lane < nv && used(lane)) cols[c][below(lane)] = x0;
        if (v1 < nv && used(v1)) cols[c][below(v1)] = x1;
        __syncthreads();
    }
#pragma unroll
    for (int sl = 0; sl < 2; ++sl) {
        const int k = sl * kWave + lane;
        if (k < nf) {
            L.fv0[k] = below(L.fv0[k]);
            L.fv1[k] = below(L.fv1[k]);
            L.fv2[k] = below(L.fv2[k]);
        }
    }
    __syncthreads();
    return __popc(m0) + __popc(m1) + __popc(m2);
}

__global__ __launch_bounds__(kEpaBlock) void epa_kernel(
    d3d_scene s, const int32_t *__restrict__ pa, const int32_t *__restrict__ pb, int64_t n,
    const int64_t *__restrict__ n_dev, const int32_t *__restrict__ work,
    const int32_t *__restrict__ work_off, const double *__restrict__ simplex,
    const int32_t *__restrict__ npts, int max_iter, int max_faces, double eps,
    double *__restrict__ out, int32_t *__restrict__ status_out, int32_t *__restrict__ iters_out,
    unsigned long long *__restrict__ next, double *__restrict__ faces_out,
    int32_t *__restrict__ nfaces_out) {
    __shared__ EpaLds L;
    const int lane = threadIdx.x;
    const int64_t count = n_dev ? *n_dev : n;
    if (work && work_off) work += *work_off;
    for (int k = lane; k < (int)(sizeof(L.eb) / sizeof(L.eb[0])); k += kEpaBlock) L.eb[k] = 0u;
    __syncthreads();
    // items: each wavefront's first is its block index; later ones are
    // claimed one at a time from `next`, past the first items (pairs differ
    // widely in cost, so a fixed stride leaves a long tail), or follow in a
    // fixed stride without it.  A wavefront whose first item is past the end
    // leaves without an atomic (a small batch starts far more wavefronts than
    // it has EPA pairs).
    auto claim = [&](int64_t prev) -> int64_t {
        if (prev < 0) return (int64_t)blockIdx.x;
        if (!next) return prev + gridDim.x;
        unsigned long long v = 0;
        if (lane == 0) v = gridDim.x + atomicAdd(next, 1ull);
        const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)v);
        const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(v >> 32));
        return (int64_t)(((unsigned long long)hi << 32) | lo);
    };
    for (int64_t item = claim(-1); item < count; item = claim(item)) {
        const int64_t pid = work ? (int64_t)work[item] : item;
        Shape SA, SB;  // validation only; supports reload the records by index
        const bool okA = load_checked(s, pa[pid], SA);
        const bool okB = load_checked(s, pb[pid], SB);
        double *o = out + 13 * pid;
        if (!(okA && okB)) {
            if (lane < 13) o[lane] = NAN;
            if (lane == 0) {
                status_out[pid] = 0x200;
                if (iters_out) iters_out[pid] = 0;
                if (nfaces_out) nfaces_out[pid] = 0;
            }
            continue;
        }

        // ---- the two collider records into LDS
        {
            const int64_t ra = pa[pid], rb = pb[pid];
            if (lane < 32) L.recf[lane >> 4][lane & 15] = s.shape_f64[16 * (lane < 16 ? ra : rb) + (lane & 15)];
            else if (lane < 40) L.reci[(lane - 32) >> 2][lane & 3] = s.shape_i32[4 * (lane < 36 ? ra : rb) + (lane & 3)];
        }
        // ---- hull vertices into the LDS cache when both fit, in their
        // original order (so sweeps and their FP64 rescans read LDS only and
        // ties go to the smallest index as in the pools); hulls whose vertices
        // are not exact in FP32 are swept from the pools
        {
            const int na = SA.type == D3D_HULL && SA.voff64 < 0 ? SA.nv : 0;
            const int nb = SB.type == D3D_HULL && SB.voff64 < 0 ? SB.nv : 0;
            const bool fit = na + nb <= kEpaCache;
            if (fit) {
                if (na > 0) fill_orig<kEpaBlock>(s, SA.voff, na, L.cx, L.cy, L.cz, 0, lane);
                if (nb > 0) fill_orig<kEpaBlock>(s, SB.voff, nb, L.cx, L.cy, L.cz, na, lane);
            }
            if (lane == 0) {
                L.coff[0] = fit && na > 0 ? 0 : -1;
                L.coff[1] = fit && nb > 0 ? na : -1;
                // both halves can work alone: analytic, or a hull in the cache
                L.split = (SA.type != D3D_HULL || (fit && na > 0)) &&
                          (SB.type != D3D_HULL || (fit && nb > 0));
            }
        }
        // ---- seed: the GJK simplex rows (w, support on A, support on B) go
        // straight into polytope rows 0..3
        const double *S = simplex + 36 * pid;
        if (lane < 36) {
            const int r = (lane % 12) / 3, c = lane % 3, part = lane / 12;
            double *dst = part == 0 ? (c == 0 ? L.wx : c == 1 ? L.wy : L.wz)
                        : part == 1 ? (c == 0 ? L.ax : c == 1 ? L.ay : L.az)
                                    : (c == 0 ? L.bx : c == 1 ? L.by : L.bz);
            dst[r] = S[lane];
        }
        __syncthreads();
        int status = 0;
        bool seed_ok = epa_seed(s, L, npts[pid]);
        int nf = 0;
        if (seed_ok) {
            const v3 w0 = row_w(L, 0);
            const v3 nrm = cross(row_w(L, 1) - w0, row_w(L, 2) - w0);
            if (dot(row_w(L, 3) - w0, nrm) > 0.0) {
                const v3 w1 = row_w(L, 1), a1 = row_a(L, 1), b1 = row_b(L, 1);
                put_row(L, 1, row_w(L, 2), row_a(L, 2), row_b(L, 2));
                put_row(L, 2, w1, a1, b1);
            }
            // seed faces {0,1,2}, {0,3,1}, {0,2,3}, {1,3,2}; a degenerate one
            // makes the seed degenerate
            bool fok = true;
            int i0 = 0, i1 = 0, i2 = 0;
            if (lane < 4) {
                i0 = (lane == 3) ? 1 : 0;
                i1 = (lane == 0) ? 1 : (lane == 1) ? 3 : (lane == 2) ? 2 : 3;
                i2 = (lane == 0) ? 2 : (lane == 1) ? 1 : (lane == 2) ? 3 : 2;
                const v3 w0 = row_w(L, i0);
                const v3 w1 = row_w(L, i1);
                const v3 w2 = row_w(L, i2);
                v3 nn;
                double dd;
                fok = face_normal(w0, w1, w2, nn, dd);
                L.fv0[lane] = i0; L.fv1[lane] = i1; L.fv2[lane] = i2;
                L.fnx[lane] = nn.x; L.fny[lane] = nn.y; L.fnz[lane] = nn.z; L.fd[lane] = dd;
            }
            if (__ballot(lane < 4 && !fok) != 0ull) seed_ok = false;
            nf = 4;
        }
        __syncthreads();
        if (!seed_ok) {
            if (lane < 13) o[lane] = 0.0;
            if (lane == 0) {
                status_out[pid] = 3;
                if (iters_out) iters_out[pid] = 0;
                if (nfaces_out) nfaces_out[pid] = 0;
            }
            __syncthreads();
            continue;
        }

        int nv = 4;
        int it = 0;
        int kb = 0;
        double depth = 0.0;
        v3 normal = mk(0.0, 0.0, 0.0);
        for (it = 0; it < max_iter; ++it) {
            // closest face: smallest distance, first index on ties
            double bd = INFINITY;
            int bk = 0x7fffffff;
#pragma unroll
            for (int sl = 0; sl < 2; ++sl) {
                const int k = sl * kWave + lane;
                if (k < nf) {
                    const double d = L.fd[k];
                    if (d < bd) { bd = d; bk = k; }
                }
            }
            group_argmin<kWave>(bd, bk);
            kb = bk;
            const v3 cn = mk(L.fnx[kb], L.fny[kb], L.fnz[kb]);
            const double cdist = L.fd[kb];
            const SupportPair sp = epa_support_pair(s, L, cn);
            const v3 spa = sp.a, spb = sp.b;
            const v3 w = spa - spb;
            const double wn = dot(w, cn);
            if (wn - cdist < eps) {
                status = 1;
                depth = wn;
                normal = cn;
                break;
            }
            if (nv >= kEpaMaxVerts) {
                nv = epa_compact_rows(L, nv, nf, lane);
                if (nv >= kEpaMaxVerts) {
                    status = 2;
                    break;
                }
            }
            const int wi = nv;
            if (lane == 0) {
                L.wx[wi] = w.x; L.wy[wi] = w.y; L.wz[wi] = w.z;
                L.ax[wi] = spa.x; L.ay[wi] = spa.y; L.az[wi] = spa.z;
                L.bx[wi] = spb.x; L.by[wi] = spb.y; L.bz[wi] = spb.z;
            }
            ++nv;
            // faces that see w
            bool rem[2];
            int fv[2][3];
            v3 fn[2];
            double fdist[2];
#pragma unroll
            for (int sl = 0; sl < 2; ++sl) {
                const int k = sl * kWave + lane;
                rem[sl] = false;
                fv[sl][0] = fv[sl][1] = fv[sl][2] = 0;
                fn[sl] = mk(0.0, 0.0, 0.0);
                fdist[sl] = 0.0;
                if (k < nf) {
                    fv[sl][0] = L.fv0[k]; fv[sl][1] = L.fv1[k]; fv[sl][2] = L.fv2[k];
                    fn[sl] = mk(L.fnx[k], L.fny[k], L.fnz[k]);
                    fdist[sl] = L.fd[k];
                    rem[sl] = dot(fn[sl], w - lds_w(L, fv[sl][0])) > eps;
                }
            }
            const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (kWave - lane));
#pragma unroll
            for (int sl = 0; sl < 2; ++sl) {
                if (rem[sl]) {
                    edge_set(L, fv[sl][0], fv[sl][1]);
                    edge_set(L, fv[sl][1], fv[sl][2]);
                    edge_set(L, fv[sl][2], fv[sl][0]);
                }
            }
            __syncthreads();
            // horizon edges of my removed faces (hz): no other removed face
            // holds the reverse edge (faces have three distinct vertices, so
            // a face never holds the reverse of its own edge); hv: the new
            // face on the edge is non-degenerate.  Order: (removed face, edge).
            // bit 3*sl+e of hzm: horizon edge; of hvm: its new face is valid
            int hzm = 0, hvm = 0;
            int cnt[2] = {0, 0};
#pragma unroll
            for (int sl = 0; sl < 2; ++sl) {
#pragma unroll
                for (int e = 0; e < 3; ++e) {
                    if (!rem[sl]) continue;
                    const int a = fv[sl][e], b = fv[sl][(e + 1) % 3];
                    if (edge_has(L, b, a)) continue;
                    hzm |= 1 << (3 * sl + e);
                    if (face_valid(lds_w(L, a), lds_w(L, b), w)) {
                        hvm |= 1 << (3 * sl + e);
                        ++cnt[sl];
                    }
                }
            }
            // both counts (at most 3 each) in one scan, slot 1 in the high half
            int tot;
            const int pre = wave_excl_scan(cnt[0] | (cnt[1] << 16), lane, tot);
            const int tot0 = tot & 0xffff, tot1 = tot >> 16;
            const int pre0 = pre & 0xffff, pre1 = tot0 + (pre >> 16);
            const bool kept0 = (lane < nf) && !rem[0];
            const bool kept1 = (kWave + lane < nf) && !rem[1];
            const unsigned long long k0 = __ballot(kept0), k1 = __ballot(kept1);
            const int nkept = __popcll(k0) + __popcll(k1);
            const int kpos0 = __popcll(k0 & lt), kpos1 = __popcll(k0) + __popcll(k1 & lt);
            // face cap: the checker appends new faces edge by edge and stops
            // at the first horizon edge (degenerate or not) met with
            // n_kept + (valid new faces so far) >= max_faces
            bool cap = false;
#pragma unroll
            for (int sl = 0; sl < 2; ++sl) {
                int before = sl == 0 ? pre0 : pre1;
#pragma unroll
                for (int e = 0; e < 3; ++e) {
                    if (!((hzm >> (3 * sl + e)) & 1)) continue;
                    if (nkept + before >= max_faces) cap = true;
                    before += (hvm >> (3 * sl + e)) & 1;
                }
            }
            cap = __ballot(cap) != 0ull;
            __syncthreads();  // all reads of the old face list and edge bits are done
#pragma unroll
            for (int sl = 0; sl < 2; ++sl) {
                if (rem[sl]) {
                    edge_clear(L, fv[sl][0], fv[sl][1]);
                    edge_clear(L, fv[sl][1], fv[sl][2]);
                    edge_clear(L, fv[sl][2], fv[sl][0]);
                }
            }
            if (kept0) {
                L.fv0[kpos0] = fv[0][0]; L.fv1[kpos0] = fv[0][1]; L.fv2[kpos0] = fv[0][2];
                L.fnx[kpos0] = fn[0].x; L.fny[kpos0] = fn[0].y; L.fnz[kpos0] = fn[0].z;
                L.fd[kpos0] = fdist[0];
            }
            if (kept1) {
                L.fv0[kpos1] = fv[1][0]; L.fv1[kpos1] = fv[1][1]; L.fv2[kpos1] = fv[1][2];
                L.fnx[kpos1] = fn[1].x; L.fny[kpos1] = fn[1].y; L.fnz[kpos1] = fn[1].z;
                L.fd[kpos1] = fdist[1];
            }
#pragma unroll
            for (int sl = 0; sl < 2; ++sl) {
                int pos = nkept + (sl == 0 ? pre0 : pre1);
#pragma unroll
                for (int e = 0; e < 3; ++e) {
                    if (!((hvm >> (3 * sl + e)) & 1)) continue;
                    if (pos < max_faces) {
                        L.fv0[pos] = fv[sl][e];
                        L.fv1[pos] = fv[sl][(e + 1) % 3];
                        L.fv2[pos] = wi;
                    }
                    ++pos;
                }
            }
            const int added = tot0 + tot1;
            nf = nkept + (added < max_faces - nkept ? added : (max_faces - nkept > 0 ? max_faces - nkept : 0));
            __syncthreads();
            // the new faces' unit normals and distances, one face per lane
            // (face_normal of the vertex rows the indices name: the same
            // operands as the verdict above, so the same bits as computing
            // them there)
#pragma unroll
            for (int sl = 0; sl < 2; ++sl) {
                const int k = nkept + sl * kWave + lane;
                if (k < nf) {
                    v3 nn = mk(0.0, 0.0, 0.0);
                    double dd = 0.0;
                    face_normal(lds_w(L, L.fv0[k]), lds_w(L, L.fv1[k]), lds_w(L, L.fv2[k]), nn, dd);
                    L.fnx[k] = nn.x; L.fny[k] = nn.y; L.fnz[k] = nn.z;
                    L.fd[k] = dd;
                }
            }
            __syncthreads();
            if (cap || nf == 0) {
                status = 2;
                break;
            }
        }
        if (status != 1) {
            double bd = INFINITY;
            int bk = 0x7fffffff;
#pragma unroll
            for (int sl = 0; sl < 2; ++sl) {
                const int k = sl * kWave + lane;
                if (k < nf) {
                    const double d = L.fd[k];
                    if (d < bd) { bd = d; bk = k; }
                }
            }
            group_argmin<kWave>(bd, bk);
            kb = bk == 0x7fffffff ? 0 : bk;
            if (nf > 0) {
                depth = L.fd[kb];
                normal = mk(L.fnx[kb], L.fny[kb], L.fnz[kb]);
            }
        }
        const int iterations = it + (status == 1 ? 1 : 0);
        v3 ca = mk(0.0, 0.0, 0.0), cb = mk(0.0, 0.0, 0.0);
        if (nf > 0) {
            const int i0 = L.fv0[kb], i1 = L.fv1[kb], i2 = L.fv2[kb];
            const v3 pt = normal * depth;
            const v3 a = lds_w(L, i0), b = lds_w(L, i1), c = lds_w(L, i2);
            const v3 v0 = b - a, v1 = c - a, v2 = pt - a;
            const double d00 = dot(v0, v0), d01 = dot(v0, v1), d11 = dot(v1, v1);
            const double d20 = dot(v2, v0), d21 = dot(v2, v1);
            const double den = d00 * d11 - d01 * d01;
            double bv = 0.0, bw = 0.0;
            if (den != 0.0) {
                bv = (d11 * d20 - d01 * d21) / den;
                bw = (d00 * d21 - d01 * d20) / den;
            }
            const double bu = 1.0 - bv - bw;
            const v3 pa0 = mk(L.ax[i0], L.ay[i0], L.az[i0]), pa1 = mk(L.ax[i1], L.ay[i1], L.az[i1]);
            const v3 pa2 = mk(L.ax[i2], L.ay[i2], L.az[i2]);
            const v3 pb0 = mk(L.bx[i0], L.by[i0], L.bz[i0]), pb1 = mk(L.bx[i1], L.by[i1], L.bz[i1]);
            const v3 pb2 = mk(L.bx[i2], L.by[i2], L.bz[i2]);
            ca = (pa0 * bu + pa1 * bv) + pa2 * bw;
            cb = (pb0 * bu + pb1 * bv) + pb2 * bw;
        }
        if (lane == 0) {
            const v3 mtv = normal * depth;
            o[0] = depth;
            o[1] = normal.x; o[2] = normal.y; o[3] = normal.z;
            o[4] = mtv.x; o[5] = mtv.y; o[6] = mtv.z;
            o[7] = ca.x; o[8] = ca.y; o[9] = ca.z;
            o[10] = cb.x; o[11] = cb.y; o[12] = cb.z;
            status_out[pid] = status;
            if (iters_out) iters_out[pid] = iterations;
            if (nfaces_out) nfaces_out[pid] = nf;
        }
        // the final polytope, Polytope.get_all_faces (epa.py:70,78,130-131):
        // per face its three vertices and its unit normal, (max_faces, 4, 3)
        // per pair; rows past nf are left as they are
        if (faces_out) {
            double *fo = faces_out + (int64_t)12 * max_faces * pid;
#pragma unroll
            for (int sl = 0; sl < 2; ++sl) {
                const int k = sl * kWave + lane;
                if (k < nf) {
                    const v3 w0 = lds_w(L, L.fv0[k]), w1 = lds_w(L, L.fv1[k]), w2 = lds_w(L, L.fv2[k]);
                    double *f = fo + 12 * k;
                    f[0] = w0.x; f[1] = w0.y; f[2] = w0.z;
                    f[3] = w1.x; f[4] = w1.y; f[5] = w1.z;
                    f[6] = w2.x; f[7] = w2.y; f[8] = w2.z;
                    f[9] = L.fnx[k]; f[10] = L.fny[k]; f[11] = L.fnz[k];
                }
            }
        }
        __syncthreads();
    }
}

int launch_epa(const d3d_scene &s, const int32_t *pa, const int32_t *pb, int64_t n_max,
               const int64_t *n_dev, const int32_t *work, const int32_t *work_off,
               const double *simplex, const int32_t *npts, int max_iter, int max_faces,
               double eps, double *out, int32_t *status, int32_t *iters,
               unsigned long long *next, hipStream_t stream, double *faces, int32_t *n_faces,
               int blocks_per_cu) {
    if (n_max <= 0) return D3D_OK;
    int64_t grid = n_max;
    if (n_dev) {
        int cus = 256, dev = 0;
        if (hipGetDevice(&dev) == hipSuccess)
            (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        const int64_t cap = (int64_t)cus * blocks_per_cu;
        if (grid > cap) grid = cap;
    } else if (grid > 65536) {
        grid = 65536;
    }
    hipLaunchKernelGGL(epa_kernel, dim3((unsigned)grid), dim3(kEpaBlock), 0, stream, s, pa, pb, n_max,
                       n_dev, work, work_off, simplex, npts, max_iter, max_faces, eps, out, status,
                       iters, next, faces, n_faces);
    D3D_CHECK_LAUNCH("epa_kernel");
    return D3D_OK;
}

}  // namespace d3d

using namespace d3d;

extern "C" int d3d_epa_batch(const d3d_scene *scene, const int32_t *pair_a, const int32_t *pair_b,
                             int64_t n, const double *simplex, const int32_t *n_points,
                             int32_t max_iterations, int32_t max_faces, double epsilon, double *out,
                             int32_t *status, int32_t *iterations, double *faces,
                             int32_t *n_faces, void *stream) {
    clear_error();
    if (int rc = check_scene(scene)) return rc;
    D3D_CHECK_ARG(n >= 0, "n must be >= 0");
    if (n == 0) return D3D_OK;
    D3D_CHECK_ARG(pair_a && pair_b && simplex && n_points && out && status, "NULL buffer");
    D3D_CHECK_ARG(max_iterations >= 1, "max_iterations must be >= 1");
    D3D_CHECK_ARG(max_faces >= 4 && max_faces <= kEpaMaxFaces, "max_faces must be in [4, 128]");
    return launch_epa(*scene, pair_a, pair_b, n, nullptr, nullptr, nullptr, simplex, n_points,
                      (int)max_iterations, (int)max_faces, epsilon, out, status, iterations,
                      nullptr, (hipStream_t)stream, faces, n_faces);
}
