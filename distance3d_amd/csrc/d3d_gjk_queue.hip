// d3d_gjk_queue.hip -- batched GJK distance as a sorted, queue-fed pipeline.
//
// 1. classify + counting sort (device): every pair gets a key
//        work class (total hull vertices of the pair, 8 classes)
//        x collider type of A x collider type of B
//    and the pair indices are scattered into `order`, grouped by key.  Pairs
//    of one class have similar support-sweep cost and, within a class, run
//    next to pairs of the same type combination, so the type switch of the
//    analytic supports stays wave-uniform.
// 2. one persistent kernel per class, G lanes per pair (G from the tuning
//    table; 1 for analytic and small-hull pairs, more for big hulls).  The G
//    lanes of a pair hold identical FP64 simplex state; hull sweeps stride
//    over the vertices by G and reduce inside the group.  Each group takes
//    its next pair from the class's device cursor as soon as its current
//    pair terminates, so lanes do not idle on the slowest pair of a wave.
// Counts and offsets stay on the device: no host synchronisation inside a
// batch.
#include <array>
#include <map>
#include <mutex>
#include <utility>
#include <vector>

#include "d3d_device.hpp"
#include "d3d_runtime.hpp"

namespace d3d {

constexpr int kQBlock = 256;
constexpr int kSizeClasses = 8;
constexpr int kNumClasses = D3D_NUM_CLASSES;  // size class, + 8 when both are hulls
static_assert(kNumClasses == 2 * kSizeClasses, "work classes: size x one or two hulls");
constexpr int kNumTypes = 9;
constexpr int kNumKeys = kNumClasses * kNumTypes * kNumTypes;
constexpr int kBadInput = 0x200;

// size class from the total hull vertex count of a pair
__device__ __forceinline__ int work_class(int nverts) {
    if (nverts == 0) return 0;
    if (nverts <= 32) return 1;
    if (nverts <= 128) return 2;
    if (nverts <= 512) return 3;
    if (nverts <= 1024) return 4;
    if (nverts <= 2048) return 5;
    if (nverts <= 4096) return 6;
    return 7;
}

__device__ __forceinline__ int pair_key(const d3d_scene &s, int a, int b) {
    if (a < 0 || b < 0 || (int64_t)a >= s.n_shapes || (int64_t)b >= s.n_shapes) return 0;
    const int4 A = *reinterpret_cast<const int4 *>(s.shape_i32 + 4 * (int64_t)a);
    const int4 B = *reinterpret_cast<const int4 *>(s.shape_i32 + 4 * (int64_t)b);
    const int ta = (A.x >= 0 && A.x < kNumTypes) ? A.x : 0;
    const int tb = (B.x >= 0 && B.x < kNumTypes) ? B.x : 0;
    const int na = (ta == D3D_HULL) ? (A.y > 0 ? A.y : 0) : 0;
    const int nb = (tb == D3D_HULL) ? (B.y > 0 ? B.y : 0) : 0;
    const int cls = work_class(na + nb) + (na > 0 && nb > 0 ? kSizeClasses : 0);
    return (cls * kNumTypes + ta) * kNumTypes + tb;
}

// Workspace layout (int32): hist[K] key_off[K] key_cur[K] cls_begin[C]
// cls_count[C] cls_cursor[C] pad order[n]
struct QueueWs {
    int32_t *hist, *key_off, *key_cur, *cls_begin, *cls_count, *cls_cursor, *order;
};

__host__ __device__ constexpr int64_t ws_header_ints() {
    return ((3 * kNumKeys + 3 * kNumClasses) + 63) / 64 * 64;
}

inline QueueWs ws_view(void *base, int64_t n) {
    (void)n;
    int32_t *p = (int32_t *)base;
    QueueWs w;
    w.hist = p;
    w.key_off = p + kNumKeys;
    w.key_cur = p + 2 * kNumKeys;
    w.cls_begin = p + 3 * kNumKeys;
    w.cls_count = w.cls_begin + kNumClasses;
    w.cls_cursor = w.cls_count + kNumClasses;
    w.order = p + ws_header_ints();
    return w;
}

// Clears the header with a fill kernel (no memset node, so the sequence can
// be captured in a graph and replayed).
__global__ void queue_clear_kernel(int32_t *hdr, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) hdr[i] = 0;
}

__global__ __launch_bounds__(kQBlock) void queue_hist_kernel(d3d_scene s,
                                                             const int32_t *__restrict__ pa,
                                                             const int32_t *__restrict__ pb,
                                                             int64_t n, int32_t *__restrict__ hist) {
    __shared__ int32_t h[kNumKeys];
    for (int i = threadIdx.x; i < kNumKeys; i += kQBlock) h[i] = 0;
    __syncthreads();
    for (int64_t i = (int64_t)blockIdx.x * kQBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kQBlock)
        atomicAdd(&h[pair_key(s, pa[i], pb[i])], 1);
    __syncthreads();
    for (int i = threadIdx.x; i < kNumKeys; i += kQBlock)
        if (h[i]) atomicAdd(&hist[i], h[i]);
}

// Single-block exclusive scan over keys; class ranges follow from the key
// order (class-major).
__global__ __launch_bounds__(kQBlock) void queue_scan_kernel(int32_t *__restrict__ hist,
                                                             int32_t *__restrict__ key_off,
                                                             int32_t *__restrict__ cls_begin,
                                                             int32_t *__restrict__ cls_count) {
    __shared__ int32_t part[kQBlock];
    constexpr int per = (kNumKeys + kQBlock - 1) / kQBlock;
    const int t = threadIdx.x;
    int local = 0;
    for (int j = 0; j < per; ++j) {
        const int k = t * per + j;
        if (k < kNumKeys) local += hist[k];
    }
    part[t] = local;
    __syncthreads();
    for (int off = 1; off < kQBlock; off <<= 1) {
        const int v = (t >= off) ? part[t - off] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int run = part[t] - local;
    for (int j = 0; j < per; ++j) {
        const int k = t * per + j;
        if (k < kNumKeys) {
            key_off[k] = run;
            run += hist[k];
        }
    }
    __syncthreads();
    if (t < kNumClasses) {
        const int k0 = t * kNumTypes * kNumTypes;
        const int k1 = k0 + kNumTypes * kNumTypes;
        int c = 0;
        for (int k = k0; k < k1; ++k) c += hist[k];
        cls_begin[t] = key_off[k0];
        cls_count[t] = c;
    }
}

__global__ __launch_bounds__(kQBlock) void queue_scatter_kernel(
    d3d_scene s, const int32_t *__restrict__ pa, const int32_t *__restrict__ pb, int64_t n,
    const int32_t *__restrict__ key_off, int32_t *__restrict__ key_cur,
    int32_t *__restrict__ order) {
    __shared__ int32_t h[kNumKeys];
    __shared__ int32_t base[kNumKeys];
    for (int64_t b0 = (int64_t)blockIdx.x * kQBlock; b0 < n; b0 += (int64_t)gridDim.x * kQBlock) {
        for (int i = threadIdx.x; i < kNumKeys; i += kQBlock) h[i] = 0;
        __syncthreads();
        const int64_t i = b0 + threadIdx.x;
        int key = 0, rank = 0;
        if (i < n) {
            key = pair_key(s, pa[i], pb[i]);
            rank = atomicAdd(&h[key], 1);
        }
        __syncthreads();
        for (int k = threadIdx.x; k < kNumKeys; k += kQBlock)
            if (h[k]) base[k] = atomicAdd(&key_cur[k], h[k]);
        __syncthreads();
        if (i < n) order[key_off[key] + base[key] + rank] = (int32_t)i;
        __syncthreads();
    }
}

// ------------------------------------------------------------- GJK state

// Support points on A (P) and B (Q) of the simplex rows, in LDS: they are
// touched only on append, compaction and at the end, so keeping them out of
// VGPRs leaves the registers to the simplex solve.  Structure of arrays over
// the block's threads: component c (slot * 6 + {Px,Py,Pz,Qx,Qy,Qz}) of thread
// t at base[c * kQBlock + t], conflict-free.
// P, Q rows of the simplex in LDS, structure of arrays with stride `st`
// doubles: one slot per pair, shared by the G lanes of a group (all lanes
// read it; the group's first lane, `wr`, writes it; LDS operations of a
// wavefront complete in order, so the other lanes see the writes).
struct PQ {
    double *base;
    int st;
    bool wr;
    __device__ __forceinline__ void put(int slot, v3 p, v3 q) const {
        if (!wr) return;
        double *b = base + slot * 6 * st;
        b[0] = p.x;
        b[st] = p.y;
        b[2 * st] = p.z;
        b[3 * st] = q.x;
        b[4 * st] = q.y;
        b[5 * st] = q.z;
    }
    __device__ __forceinline__ v3 P(int slot) const {
        const double *b = base + slot * 6 * st;
        return mk(b[0], b[st], b[2 * st]);
    }
    __device__ __forceinline__ v3 Q(int slot) const {
        const double *b = base + slot * 6 * st;
        return mk(b[3 * st], b[4 * st], b[5 * st]);
    }
    __device__ __forceinline__ void zero() const {
        if (!wr) return;
#pragma unroll
        for (int c = 0; c < 24; ++c) base[c * st] = 0.0;
    }
    __device__ __forceinline__ void move(int from, int to) const {
        if (from == to || !wr) return;
        const double *a = base + from * 6 * st;
        double *b = base + to * 6 * st;
#pragma unroll
        for (int c = 0; c < 6; ++c) b[c * st] = a[c * st];
    }
    // applies a compaction map (sources are >= destinations, so moving
    // rows in increasing order reads every row before it is overwritten)
    __device__ __forceinline__ void compact(const CompactMap &c) const {
        if (!c.moved) return;
        move(c.k0, 0);
        if (c.m > 1) move(c.k1, 1);
        if (c.m > 2) move(c.k2, 2);
    }
};

struct GjkState {
    Simplex S;
    v3 d;
    double vv, prev;
    int it, state;
};

__device__ __forceinline__ void gjk_init(GjkState &g, const PQ &pq) {
    simplex_zero(g.S);
    pq.zero();
    g.d = mk(1.0, 0.0, 0.0);
    g.vv = dot(g.d, g.d);
    g.prev = kMaxFloat;
    g.it = 0;
    g.state = -1;
}

// Y rows are P - Q rows (w = p - q when a row is appended; compaction moves
// Y, P, Q rows alike; zero rows stay zero), so Y is rebuilt from the LDS rows
// at the start of each step instead of being held in registers across the
// supports.  Row for row the values are the ones the reference keeps in Y.
__device__ __forceinline__ void load_y(Simplex &S, const PQ &pq) {
    S.Y0 = pq.P(0) - pq.Q(0);
    S.Y1 = pq.P(1) - pq.Q(1);
    S.Y2 = pq.P(2) - pq.Q(2);
    S.Y3 = pq.P(3) - pq.Q(3);
}

// One pass of _distance_loop (_gjk_jolt.py:224-288) given the supports p, q.
// Returns false once the pair has terminated (g.state set).
__device__ __forceinline__ bool gjk_distance_step(GjkState &g, const PQ &pq, v3 p, v3 q,
                                                  double maxd2, double tol2) {
    ++g.it;
    load_y(g.S, pq);
    const v3 w = p - q;
    const double dw = dot(g.d, w);
    if (dw < 0.0 && dw * dw > g.vv * maxd2) {
        g.state = D3D_CLIPPED;
        return false;
    }
    pq.put(g.S.n < 3 ? g.S.n : 3, p, q);
    simplex_append(g.S, w);
    v3 nvec;
    double nvv;
    int set;
    if (closest_to_origin(g.S, g.prev, nvec, nvv, set)) {
        g.d = nvec;
        g.vv = nvv;
    } else {
        g.S.n -= 1;
        set = (1 << g.S.n) - 1;
    }
    if (set == 15) {
        g.vv = 0.0;
        g.state = D3D_INTERSECT;
        return false;
    }
    // compaction moves the P/Q rows only: the next step rebuilds Y from them,
    // and the test below reads the kept rows of Y where they are
    const double ymax = max_y_len_sq_kept(g.S, set);
    const CompactMap cm = compact_map(g.S.n, set);
    pq.compact(cm);
    g.S.n = cm.m;
    if (g.vv <= tol2 || g.vv <= kEps * ymax) {
        g.vv = 0.0;
        g.state = D3D_INTERSECT;
        return false;
    }
    g.d = -g.d;
    if (g.prev - g.vv <= kEps * g.prev) {
        g.state = D3D_SEPARATED;
        return false;
    }
    g.prev = g.vv;
    return true;
}

struct GjkOut {
    double *dist, *cpa, *cpb;
    int32_t *status, *iters;
    double *simplex;
    int32_t *npts;
    bool simplex_hits_only;  // write simplex / npts only for intersecting pairs
    // GJK -> EPA (all null without EPA): intersecting pairs are appended to
    // epa_work (epa_count counts them); the others get EPA status -1 and a
    // zero output row
    int32_t *epa_work;
    unsigned long long *epa_count;
    double *epa_out;
    int32_t *epa_status;
};

__device__ __forceinline__ void epa_route(int64_t pid, bool intersecting, const GjkOut &o) {
    if (!o.epa_work) return;
    if (intersecting) {
        o.epa_work[atomicAdd(o.epa_count, 1ull)] = (int32_t)pid;
    } else {
#pragma unroll
        for (int k = 0; k < 13; ++k) o.epa_out[13 * pid + k] = 0.0;
        o.epa_status[pid] = -1;
    }
}

// calculate_closest_points (_gjk_jolt.py:667-687)
__device__ __forceinline__ void closest_points(const Simplex &S, const PQ &pq, v3 &a, v3 &b) {
    if (S.n == 1) {
        a = pq.P(0);
        b = pq.Q(0);
    } else if (S.n == 2) {
        double u, v;
        bary_line(S.Y0, S.Y1, u, v);
        a = pq.P(0) * u + pq.P(1) * v;
        b = pq.Q(0) * u + pq.Q(1) * v;
    } else if (S.n == 3) {
        double u, v, w;
        bary_plane(S.Y0, S.Y1, S.Y2, u, v, w);
        a = (pq.P(0) * u + pq.P(1) * v) + pq.P(2) * w;
        b = (pq.Q(0) * u + pq.Q(1) * v) + pq.Q(2) * w;
    } else {
        double u, v, w, x;
        bary_tetra(S.Y0, S.Y1, S.Y2, S.Y3, u, v, w, x);
        a = ((pq.P(0) * u + pq.P(1) * v) + pq.P(2) * w) + pq.P(3) * x;
        b = ((pq.Q(0) * u + pq.Q(1) * v) + pq.Q(2) * w) + pq.Q(3) * x;
    }
}

// End of gjk_distance_jolt (_gjk_jolt.py:207-221).
__device__ __forceinline__ void gjk_write(const GjkState &g, const PQ &pq, int64_t pid,
                                          double sanity, const GjkOut &o) {
    if (o.iters) o.iters[pid] = g.it;
    epa_route(pid, g.state == D3D_INTERSECT, o);
    Simplex S = g.S;
    load_y(S, pq);
    const bool want_s = !o.simplex_hits_only || g.state == D3D_INTERSECT;
    if (o.npts && want_s) o.npts[pid] = g.S.n;
    if (o.simplex && want_s) {
        double *w = o.simplex + 36 * pid;
        const v3 y[4] = {S.Y0, S.Y1, S.Y2, S.Y3};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const v3 P = pq.P(k), Q = pq.Q(k);
            w[3 * k] = y[k].x;
            w[3 * k + 1] = y[k].y;
            w[3 * k + 2] = y[k].z;
            w[12 + 3 * k] = P.x;
            w[13 + 3 * k] = P.y;
            w[14 + 3 * k] = P.z;
            w[24 + 3 * k] = Q.x;
            w[25 + 3 * k] = Q.y;
            w[26 + 3 * k] = Q.z;
        }
    }
    if (g.state == D3D_CLIPPED) {
        o.dist[pid] = kMaxFloat;
        o.cpa[3 * pid] = o.cpa[3 * pid + 1] = o.cpa[3 * pid + 2] = NAN;
        o.cpb[3 * pid] = o.cpb[3 * pid + 1] = o.cpb[3 * pid + 2] = NAN;
        o.status[pid] = D3D_CLIPPED;
        return;
    }
    v3 a, b;
    closest_points(S, pq, a, b);
    int st = g.state;
    if (!(fabs(dot(g.d, g.d) - g.vv) < sanity)) st |= D3D_SANITY_FAILED;
    const double dd = sqrt(g.vv);
    if (dd < kEps) {
        const v3 m = (a + b) * 0.5;
        a = m;
        b = m;
    }
    o.dist[pid] = dd;
    o.cpa[3 * pid] = a.x;
    o.cpa[3 * pid + 1] = a.y;
    o.cpa[3 * pid + 2] = a.z;
    o.cpb[3 * pid] = b.x;
    o.cpb[3 * pid + 1] = b.y;
    o.cpb[3 * pid + 2] = b.z;
    o.status[pid] = st;
}

__device__ __forceinline__ void gjk_write_bad(int64_t pid, const GjkOut &o) {
    o.dist[pid] = NAN;
    o.cpa[3 * pid] = o.cpa[3 * pid + 1] = o.cpa[3 * pid + 2] = NAN;
    o.cpb[3 * pid] = o.cpb[3 * pid + 1] = o.cpb[3 * pid + 2] = NAN;
    o.status[pid] = kBadInput;
    if (o.iters) o.iters[pid] = 0;
    if (o.npts) o.npts[pid] = 0;
    epa_route(pid, false, o);
}

// ------------------------------------------------------- group supports

__device__ __forceinline__ unsigned long long lanemask_lt(int lane) {
    return lane == 0 ? 0ull : (~0ull >> (kWave - lane));
}

// ---------------------------------------------- LDS-resident hull vertices
//
// A wavefront's 12 KB LDS region holds the P/Q rows of its 64/G pair slots
// (24 doubles each, interleaved) and, for hull classes with G >= 4, one vertex
// cache per slot in the rest: the FP32 vertices of the pair's hulls in their
// original order (A's at 0, B's after them; x, y and z arrays of
// cache_verts<G>() entries), loaded once when the group takes the pair.  Every
// support sweep of the pair then reads LDS instead of the pools.  Slot
// strides are congruent to G modulo 64 words, so the 64/G groups sweeping at
// the same time hit distinct banks.  Pairs whose hulls do not fit, or whose
// vertices are not exact in FP32, sweep the pools as before.
template <int G>
__host__ __device__ constexpr int cache_stride_words() {
    // (12288 bytes - P/Q rows) / slots, rounded down to G modulo 64 words
    constexpr int room = (3072 - 48 * (kWave / G)) / (kWave / G);
    return G >= kWave ? room : ((room - G) / 64) * 64 + G;
}
// With 8 to 32 lanes per pair the last kRecWords words of a slot hold the
// pair's two collider records (the FP64 blocks of A and B, then their int32
// parts), staged once when the group takes the pair, so the supports read
// them from LDS instead of reloading them from memory at every iteration
// (a wavefront-wide group loads them into scalar registers instead).
constexpr int kRecWords = 2 * 16 * 2 + 2 * 4;
template <int G>
__host__ __device__ constexpr bool slot_recs() {
    return G >= 8 && G < kWave;
}
// words of a slot for vertices or cluster records
template <int G>
__host__ __device__ constexpr int cache_words() {
    return cache_stride_words<G>() - (slot_recs<G>() ? kRecWords : 0);
}
template <int G>
__host__ __device__ constexpr int cache_verts() {
    return cache_words<G>() / 3;
}
static_assert(cache_verts<4>() >= 32, "class 1 pairs fit a 4-lane slot");
static_assert(cache_verts<64>() >= 1000, "one slot per wavefront holds 1000 vertices");
static_assert(cache_words<8>() % 2 == 0 && cache_stride_words<8>() % 2 == 0 &&
                  cache_words<16>() % 2 == 0 && cache_stride_words<16>() % 2 == 0,
              "slot records are 8-byte aligned");

// Collider record k (0: A, 1: B) of the pair in a slot (see slot_recs).
__device__ __forceinline__ Shape slot_shape(const float *recs, int k) {
    const double *f = reinterpret_cast<const double *>(recs) + 16 * k;
    const int32_t *i = reinterpret_cast<const int32_t *>(recs + 64) + 4 * k;
    Shape h;
#pragma unroll
    for (int j = 0; j < 9; ++j) h.R[j] = f[j];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        h.t[j] = f[9 + j];
        h.p[j] = f[12 + j];
    }
    h.margin = f[15];
    h.type = i[0];
    h.nv = i[1];
    h.voff = i[2];
    h.voff64 = i[3];
    return h;
}

// The collider record k of a pair: from the slot when it holds them, else
// from the scene by index (kUniform: the same record for the whole
// wavefront, so a wavefront-wide group loads it into scalar registers).
template <int G, bool kUniform>
__device__ __forceinline__ Shape pair_shape(const d3d_scene &s, const float *recs, int k, int idx) {
    if constexpr (slot_recs<G>()) {
        (void)s;
        (void)idx;
        return slot_shape(recs, k);
    } else {
        (void)recs;
        (void)k;
        return load_shape(s, (kUniform && G == kWave) ? wave_uniform(idx) : opaque(idx));
    }
}

// Copies the FP32 vertices of hull record h (validated) into the cache at
// `at`, in original order (vorig), by the G lanes of a group.
template <int G>
__device__ __forceinline__ void cache_fill(const d3d_scene &s, const Shape &h, float *cache,
                                           int at, int sub) {
    constexpr int C = cache_verts<G>();
    fill_orig<G>(s, h.voff, h.nv, cache, cache + C, cache + 2 * C, at, sub);
}

// Supports of records ia along d and ib along -d (see support_pair) for a
// pair whose hull vertices are in the group's cache (A's first, nA of them).
// Two hulls: the lower half of the group sweeps A while the upper half sweeps
// B; one hull: the whole group sweeps it.  Same results as support_pair.
template <int G>
__device__ __forceinline__ SupportPair support_pair_cached(const d3d_scene &s, int ia, int ib, v3 d,
                                                           int sub, const float *cache, int nA,
                                                           bool two) {
    constexpr int C = cache_verts<G>();
    const float *cx = cache, *cy = cache + C, *cz = cache + 2 * C;
    const float *recs = cache + cache_words<G>();
    if (two) {
        constexpr int H = G / 2;
        const bool upper = sub >= H;
        const Shape h = pair_shape<G, false>(s, recs, upper ? 1 : 0, upper ? ib : ia);
        const v3 dir = vsel(upper, -d, d);
        const int off = upper ? nA : 0;
        const int j = lds_orig_argmax<H>(cx, cy, cz, off, h.nv, h.p[0], to_local(h, dir), sub & (H - 1));
        const v3 v = mk((double)cx[off + j], (double)cy[off + j], (double)cz[off + j]);
        const v3 out = apply_margin(h, to_world(h, v), dir);
        SupportPair r;
        halves<H>(out.x, r.a.x, r.b.x);
        halves<H>(out.y, r.a.y, r.b.y);
        halves<H>(out.z, r.a.z, r.b.z);
        return r;
    }
    v3 a = mk(0.0, 0.0, 0.0), b = a;
#pragma unroll 1
    for (int k = 0; k < 2; ++k) {
        const Shape h = pair_shape<G, false>(s, recs, k, k == 0 ? ia : ib);
        const v3 dir = vsel(k == 0, d, -d);
        v3 out;
        if (h.type == D3D_HULL) {
            const int j = lds_orig_argmax<G>(cx, cy, cz, 0, h.nv, h.p[0], to_local(h, dir), sub);
            out = to_world(h, mk((double)cx[j], (double)cy[j], (double)cz[j]));
        } else {
            out = support_analytic(h, dir);
        }
        out = apply_margin(h, out, dir);
        a = vsel(k == 0, out, a);
        b = vsel(k == 0, b, out);
    }
    return SupportPair{a, b};
}

// Copies the cluster records of hull record h (validated, clustered) into
// the slot at record `at`, by the G lanes of a group.
template <int G>
__device__ __forceinline__ void bounds_fill(const d3d_scene &s, const Shape &h, float4 *slot, int at,
                                            int sub) {
    const float4 *__restrict__ cl = reinterpret_cast<const float4 *>(s.clusters) + (int64_t)h.p[1];
    const int nc = (int)h.p[2];
#pragma unroll 4
    for (int k = sub; k < nc; k += G) slot[at + k] = cl[k];
}

// Supports of records ia along d and ib along -d (see support_pair) for a pair
// whose clustered hulls have their cluster records in the group's slot (A's
// first, ncA of them): the sweeps read the bounds from LDS and the vertices
// from the pools.  Same results as support_pair.
template <int G>
__device__ __forceinline__ SupportPair support_pair_bounds(const d3d_scene &s, int ia, int ib, v3 d,
                                                           int sub, const float4 *slot, int ncA,
                                                           int nrec) {
    v3 a = mk(0.0, 0.0, 0.0), b = a;
    if (G == kWave) d = mk(wave_uniform(d.x), wave_uniform(d.y), wave_uniform(d.z));
    const float *recs = reinterpret_cast<const float *>(slot) + cache_words<G>();
#pragma unroll 1
    for (int k = 0; k < 2; ++k) {
        const Shape h = pair_shape<G, true>(s, recs, k, k == 0 ? ia : ib);
        const v3 dir = vsel(k == 0, d, -d);
        v3 out;
        if (h.type == D3D_HULL) {
            const v3 l = to_local(h, dir);
            const HullRef H = hull_ref(h);
            // the slot's words past the cluster records hold the survivor list
            int *list = const_cast<int *>(reinterpret_cast<const int *>(slot + nrec));
            const int hi = H.nc > 0 ? hull_argmax<G, true>(s, H, l, sub, slot + (k == 0 ? 0 : ncA),
                                                           list, cache_words<G>() - 4 * nrec)
                                    : hull_argmax<G>(s, H, l, sub);
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

// One work class, G lanes per pair, by one wavefront: groups of G lanes take
// pairs of the class (chunked claims from the class cursor) until the class is
// exhausted and every group of the wavefront has finished its pair.  The
// wavefront's P/Q slots are `pq_wave` (24 x 64 doubles, slot stride 64 / G);
// with kCache (hull classes, G >= 4) the rest of the region holds each slot's
// vertex cache.
template <int G, bool kHull, bool kCache = (kHull && G >= 4)>
__device__ __forceinline__ void run_class(const d3d_scene &s, const int32_t *__restrict__ pa,
                                          const int32_t *__restrict__ pb,
                                          const int32_t *__restrict__ order, int begin, int count,
                                          int32_t *__restrict__ cursor, const d3d_gjk_params &prm,
                                          const GjkOut &o, double *pq_wave, int chunk, int rot) {
    const int lane = lane_id();
    const int sub = lane & (G - 1);
    const int leader = lane - sub;
    const double tol2 = prm.tolerance * prm.tolerance;
    const PQ pq{pq_wave + lane / G, kWave / G, sub == 0};
    float *const cache = reinterpret_cast<float *>(pq_wave + 24 * (kWave / G)) +
                         (lane / G) * cache_stride_words<G>();
    GjkState g;
    gjk_init(g, pq);
    int ia = 0, ib = 0;
    int pid = 0;
    // pairs with their hull vertices in the group's cache: hull vertices of
    // A (the offset of B's), both records hulls
    bool cached = false, two = false;
    int nA = 0;
    // otherwise, pairs with their hulls' cluster records in the slot (ncA: A's
    // record count, the offset of B's)
    bool bounded = false;
    int ncA = 0, nrec = 0;
    // analytic classes: the two record types, read once per pair
    int ta = 0, tb = 0;
    bool have = false, exhausted = false;
    // The first chunk of every wavefront is fixed by its rank (wavefront index
    // minus `rot`, modulo the wavefront count): no atomic, so thousands of
    // wavefronts do not queue on one cursor when a launch starts, and the
    // rotation hands each class's fixed chunks to the wavefronts the earlier
    // classes left without one.  Later chunks come from the cursor, past all
    // fixed chunks; when the fixed chunks cover the class there are none.
    const int n_waves = (int)gridDim.x * (kQBlock / kWave);
    const int wv = (int)blockIdx.x * (kQBlock / kWave) + (int)threadIdx.x / kWave;
    const int64_t r = (wv - rot + n_waves) % n_waves;
    const int64_t fixed_end = (int64_t)n_waves * chunk;
    const bool dynamic = fixed_end < count;
    int cb = (int)min(r * chunk, (int64_t)count), ce = cb + chunk;
    // the wavefront's unclaimed items are [cb, ce) of the class
    while (true) {
        const bool need = !have && !exhausted;
        const unsigned long long m = __ballot(need && sub == 0);
        if (m != 0ull) {
            // groups take items from the wavefront's chunk of the class range;
            // one atomic fetches the next chunk when it runs out.  Groups
            // ranked past the items at hand wait for the next round.
            const int c = __popcll(m);
            const int rank = __popcll(m & lanemask_lt(leader));
            const int avail = ce - cb;
            int nb = 0;
            if (avail < c) {
                const int first = __ffsll((long long)m) - 1;
                if (dynamic) {
                    if (lane == first) nb = (int)fixed_end + atomicAdd(cursor, chunk);
                    nb = readlane_i32(nb, first);
                } else {
                    nb = count;  // nothing past the fixed chunks
                }
            }
            const int base = rank < avail ? cb : nb - avail;
            const bool served = rank < avail + (avail < c ? chunk : 0);
            if (avail < c) {
                const int taken = c - avail < chunk ? c - avail : chunk;
                cb = nb + taken;
                ce = nb + chunk;
            } else {
                cb += c;
            }
            if (need && served) {
                const int idx = base + rank;
                if (idx < count) {
                    pid = order[begin + idx];
                    ia = pa[pid];
                    ib = pb[pid];
                    Shape A, B;
                    const bool okA = load_checked(s, ia, A);
                    const bool okB = load_checked(s, ib, B);
                    const bool ok =
                        okA && okB && (kHull || (A.type != D3D_HULL && B.type != D3D_HULL));
                    gjk_init(g, pq);
                    if (ok) {
                        have = true;
                        ta = A.type;
                        tb = B.type;
                        if constexpr (kCache) {
                            const bool ha = A.type == D3D_HULL, hb = B.type == D3D_HULL;
                            nA = ha ? A.nv : 0;
                            const int nB = hb ? B.nv : 0;
                            cached = nA + nB <= cache_verts<G>() && (!ha || A.voff64 < 0) &&
                                     (!hb || B.voff64 < 0);
                            two = ha && hb;
                            ncA = ha ? (int)A.p[2] : 0;
                            const int ncB = hb ? (int)B.p[2] : 0;
                            bounded = !cached && ncA + ncB > 0 &&
                                      4 * (ncA + ncB) <= cache_words<G>();
                            nrec = ncA + ncB;
                            if (cached) {
                                if (ha) cache_fill<G>(s, A, cache, 0, sub);
                                if (hb) cache_fill<G>(s, B, cache, nA, sub);
                            } else if (bounded) {
                                float4 *slot = reinterpret_cast<float4 *>(cache);
                                if (ncA > 0) bounds_fill<G>(s, A, slot, 0, sub);
                                if (ncB > 0) bounds_fill<G>(s, B, slot, ncA, sub);
                            }
                            if constexpr (slot_recs<G>()) {
                                double *rf = reinterpret_cast<double *>(cache + cache_words<G>());
                                int32_t *ri = reinterpret_cast<int32_t *>(rf + 32);
                                for (int k = sub; k < 40; k += G) {
                                    if (k < 32)
                                        rf[k] = s.shape_f64[16 * (int64_t)(k < 16 ? ia : ib) + (k & 15)];
                                    else
                                        ri[k - 32] = s.shape_i32[4 * (int64_t)(k < 36 ? ia : ib) + (k & 3)];
                                }
                            }
                            // the group's lanes read what the others wrote
                            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                        }
                    } else if (sub == 0) {
                        gjk_write_bad(pid, o);
                    }
                } else {
                    exhausted = true;
                }
            }
            // the cursor only grows: once any group sees an item past the end
            // the class is empty for every group of the wavefront
            if (__ballot(exhausted) != 0ull) exhausted = true;
        }
        if (__ballot(have) == 0ull) {
            if (__ballot(!exhausted) == 0ull) break;
            continue;
        }
        if (have) {
            bool running = true;
            if (g.it >= prm.max_iterations) {
                g.state = D3D_ITER_CAP;
                running = false;
            } else {
                // the records were validated when the pair was taken
                SupportPair w;
                // a wavefront in its support sweep (memory rounds) issues ahead of
                // those in their FP64 simplex step, so its loads go out sooner;
                // priority only orders issue between the SIMD's wavefronts
                __builtin_amdgcn_s_setprio(2);
                if constexpr (kCache) {
                    if (cached)
                        w = support_pair_cached<G>(s, ia, ib, g.d, sub, cache, nA, two);
                    else if (bounded)
                        w = support_pair_bounds<G>(s, ia, ib, g.d, sub,
                                                   reinterpret_cast<const float4 *>(cache), ncA, nrec);
                    else
                        w = support_pair<G, kHull>(s, ia, ib, g.d, sub);
                } else if constexpr (!kHull) {
                    w = support_pair_analytic(s, ia, ib, ta, tb, g.d);
                } else {
                    w = support_pair<G, kHull>(s, ia, ib, g.d, sub);
                }
                __builtin_amdgcn_s_setprio(0);
                running = gjk_distance_step(g, pq, w.a, w.b, prm.max_distance_squared, tol2);
            }
            if (!running) {
                if (sub == 0) gjk_write(g, pq, pid, prm.sanity_check, o);
                have = false;
            }
        }
    }
}

// All work classes in one persistent launch: every wavefront walks the classes
// from the largest hulls to the analytic pairs (size classes 7 .. 0, the pairs
// of two hulls before those of one at each size), switching its lanes per pair
// at each class, so classes overlap without a launch per class.
// lanes_per_pair holds resolved values (1..64, class 0: 1, 2 or 4).
// Occupancy 3 (<= 168 VGPRs): measured faster than occupancy 2 without the
// register cap, at the cost of a few spilled registers.
__global__ __launch_bounds__(kQBlock) __attribute__((amdgpu_waves_per_eu(3))) void gjk_all_kernel(
    d3d_scene s, const int32_t *__restrict__ pa, const int32_t *__restrict__ pb,
    const int32_t *__restrict__ order, const int32_t *__restrict__ cls_begin,
    const int32_t *__restrict__ cls_count, int32_t *__restrict__ cls_cursor, d3d_gjk_params prm,
    GjkOut o) {
    __shared__ double pq_lds[24 * kQBlock];
    double *pq_wave = pq_lds + 24 * kWave * (threadIdx.x / kWave);
    const int n_waves = (int)gridDim.x * (kQBlock / kWave);
    int rot = 0;  // wavefronts given a fixed chunk by the classes walked so far
    for (int step = 0; step < kNumClasses; ++step) {
        const int cls = (step & 1) ? kSizeClasses - 1 - step / 2 : kNumClasses - 1 - step / 2;
        const int count = cls_count[cls];
        if (count == 0) continue;
        const int begin = cls_begin[cls];
        int32_t *cur = cls_cursor + cls;
        const int G = prm.lanes_per_pair[cls];
        // items per claim: claim_factor * (64 / G), or the class's fair share per
        // wavefront when that is smaller, so a small batch spreads its pairs
        // over all wavefronts instead of a few wavefronts holding all of them
        const int full = prm.claim_factor * (kWave / G);
        const int fair = (count + n_waves - 1) / n_waves;
        const int chunk = fair < full ? fair : full;
        if (cls == 0) {
            if (G <= 1) run_class<1, false>(s, pa, pb, order, begin, count, cur, prm, o, pq_wave, chunk, rot);
            else if (G == 2) run_class<2, false>(s, pa, pb, order, begin, count, cur, prm, o, pq_wave, chunk, rot);
            else run_class<4, false>(s, pa, pb, order, begin, count, cur, prm, o, pq_wave, chunk, rot);
        } else {
            switch (G) {
            case 1: run_class<1, true>(s, pa, pb, order, begin, count, cur, prm, o, pq_wave, chunk, rot); break;
            case 2: run_class<2, true>(s, pa, pb, order, begin, count, cur, prm, o, pq_wave, chunk, rot); break;
            case 4: run_class<4, true>(s, pa, pb, order, begin, count, cur, prm, o, pq_wave, chunk, rot); break;
            case 8: run_class<8, true>(s, pa, pb, order, begin, count, cur, prm, o, pq_wave, chunk, rot); break;
            case 16: run_class<16, true>(s, pa, pb, order, begin, count, cur, prm, o, pq_wave, chunk, rot); break;
            case 32: run_class<32, true>(s, pa, pb, order, begin, count, cur, prm, o, pq_wave, chunk, rot); break;
            default: run_class<64, true>(s, pa, pb, order, begin, count, cur, prm, o, pq_wave, chunk, rot); break;
            }
        }
        const int used = (count + chunk - 1) / chunk;
        rot = (int)(((int64_t)rot + (used < n_waves ? used : n_waves)) % n_waves);
    }
}

// ---------------------------------------------------------------- host

// the tuned table (distance3d_amd/tuning.json): lanes per class for the
// 1,048,576-pair mixed workload on one MI355X, class 11 from the EPA pairs
static const int kDefaultLanes[kNumClasses] = {1, 4, 16, 8, 16, 16, 32, 64,
                                               1, 8, 16, 8, 64, 16, 64, 64};

struct WsCache {
    std::mutex mu;
    std::map<std::pair<int, void *>, std::pair<void *, size_t>> bufs;
};
static WsCache g_ws;

static int workspace_for(void *stream, int64_t n, void **out) {
    int dev = 0;
    D3D_CHECK_HIP(hipGetDevice(&dev));
    const size_t need = (size_t)(ws_header_ints() + n) * sizeof(int32_t);
    std::lock_guard<std::mutex> lk(g_ws.mu);
    auto &e = g_ws.bufs[std::make_pair(dev, stream)];
    if (e.second < need) {
        if (e.first) {
            D3D_CHECK_HIP(hipStreamSynchronize((hipStream_t)stream));
            D3D_CHECK_HIP(hipFree(e.first));
            e.first = nullptr;
            e.second = 0;
        }
        const size_t grow = need + need / 4;
        D3D_CHECK_HIP(hipMalloc(&e.first, grow));
        e.second = grow;
    }
    *out = e.first;
    return D3D_OK;
}

static int cu_count() {
    static int cus = 0;
    if (!cus) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            cus <= 0)
            cus = 256;
    }
    return cus;
}

// EPA behind the GJK classes (after the join, on the caller's stream): the
// intersecting pairs are compacted into `work` and EPA runs over them.
struct EpaHook {
    int32_t *work;
    unsigned long long *count;
    int max_iter, max_faces;
    double eps;
    double *out;
    int32_t *status;
};

// the EPA list counter and the EPA claim cursor
__global__ void zero_u64_kernel(unsigned long long *c) {
    c[0] = 0ull;
    c[1] = 0ull;
}

// ---------------------------------------------------------- stage timing

// Events of the calls recorded while stage timing is enabled: before the sort,
// before the GJK kernel, after it, after EPA.  Spent event sets are reused.
struct StageTiming {
    std::mutex mu;
    bool on = false;
    std::vector<std::array<hipEvent_t, 4>> rec, pool;
};
static StageTiming g_timing;

static bool stage_events(std::array<hipEvent_t, 4> &ev) {
    std::lock_guard<std::mutex> lk(g_timing.mu);
    if (!g_timing.on) return false;
    if (!g_timing.pool.empty()) {
        ev = g_timing.pool.back();
        g_timing.pool.pop_back();
    } else {
        for (auto &e : ev)
            if (hipEventCreate(&e) != hipSuccess) return false;
    }
    g_timing.rec.push_back(ev);
    return true;
}

static int launch_all(const d3d_scene &s, const int32_t *pa, const int32_t *pb, int64_t n,
                      const d3d_gjk_params &prm, const GjkOut &o, void *ws, hipStream_t stream,
                      const EpaHook *epa) {
    const QueueWs w = ws_view(ws, n);
    const int hdr = (int)ws_header_ints();
    std::array<hipEvent_t, 4> ev;
    const bool timed = stage_events(ev);
    if (timed) D3D_CHECK_HIP(hipEventRecord(ev[0], stream));
    hipLaunchKernelGGL(queue_clear_kernel, dim3(4), dim3(kQBlock), 0, stream, (int32_t *)ws, hdr);
    D3D_CHECK_LAUNCH("queue_clear_kernel");
    int64_t g = (n + kQBlock - 1) / kQBlock;
    if (g > 2048) g = 2048;
    hipLaunchKernelGGL(queue_hist_kernel, dim3((unsigned)g), dim3(kQBlock), 0, stream, s, pa, pb, n,
                       w.hist);
    D3D_CHECK_LAUNCH("queue_hist_kernel");
    hipLaunchKernelGGL(queue_scan_kernel, dim3(1), dim3(kQBlock), 0, stream, w.hist, w.key_off,
                       w.cls_begin, w.cls_count);
    D3D_CHECK_LAUNCH("queue_scan_kernel");
    hipLaunchKernelGGL(queue_scatter_kernel, dim3((unsigned)g), dim3(kQBlock), 0, stream, s, pa, pb,
                       n, w.key_off, w.key_cur, w.order);
    D3D_CHECK_LAUNCH("queue_scatter_kernel");
    // one persistent launch for every class; lanes and launch shape resolved
    // on the host
    d3d_gjk_params rp = prm;
    for (int c = 0; c < kNumClasses; ++c) {
        int G = prm.lanes_per_pair[c] > 0 ? prm.lanes_per_pair[c] : kDefaultLanes[c];
        if (c == 0 && G > 4) G = 4;
        rp.lanes_per_pair[c] = G;
    }
    rp.blocks_per_cu = prm.blocks_per_cu > 0 ? prm.blocks_per_cu : 3;
    rp.claim_factor = prm.claim_factor > 0 ? prm.claim_factor : 2;
    rp.epa_blocks_per_cu = prm.epa_blocks_per_cu > 0 ? prm.epa_blocks_per_cu : 8;
    if (epa) {
        hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(1), 0, stream, epa->count);
        D3D_CHECK_LAUNCH("zero_u64_kernel");
    }
    // resident at occupancy 3 (4 wavefronts per block) by default
    const int64_t grid = (int64_t)cu_count() * rp.blocks_per_cu;
    if (timed) D3D_CHECK_HIP(hipEventRecord(ev[1], stream));
    hipLaunchKernelGGL(gjk_all_kernel, dim3((unsigned)grid), dim3(kQBlock), 0, stream, s, pa, pb,
                       w.order, w.cls_begin, w.cls_count, w.cls_cursor, rp, o);
    D3D_CHECK_LAUNCH("gjk_all_kernel");
    if (timed) D3D_CHECK_HIP(hipEventRecord(ev[2], stream));
    if (epa) {
        if (int rc = launch_epa(s, pa, pb, n, (const int64_t *)epa->count, epa->work, nullptr,
                                o.simplex, o.npts, epa->max_iter, epa->max_faces, epa->eps,
                                epa->out, epa->status, nullptr, epa->count + 1, stream, nullptr,
                                nullptr, rp.epa_blocks_per_cu))
            return rc;
    }
    if (timed) D3D_CHECK_HIP(hipEventRecord(ev[3], stream));
    return D3D_OK;
}

// ------------------------------------------------- GJK -> EPA pipeline

// Per-pair collider instances: record 2i is scene record a[i] posed by
// pose_a[i] (R row-major, t: 12 doubles), record 2i+1 is b[i] posed by
// pose_b[i]; NULL poses keep the scene's.  Vertex offsets are copied, so the
// instances share the scene's resident vertex pools.
__global__ __launch_bounds__(kQBlock) void instance_kernel(
    d3d_scene s, const int32_t *__restrict__ a, const int32_t *__restrict__ b,
    const double *__restrict__ pose_a, const double *__restrict__ pose_b,
    const int32_t *__restrict__ pidx_a, const int32_t *__restrict__ pidx_b, int64_t n_poses,
    int64_t n, double *__restrict__ f64, int32_t *__restrict__ i32) {
    const int64_t t = (int64_t)blockIdx.x * kQBlock + threadIdx.x;
    if (t >= 2 * n) return;
    const int64_t pair = t >> 1;
    const int side = (int)(t & 1);
    const int k = side ? b[pair] : a[pair];
    const double *pose = side ? pose_b : pose_a;
    const int32_t *pidx = side ? pidx_b : pidx_a;
    int64_t prow = pair;
    if (pidx) prow = pidx[pair];
    double *dst = f64 + 16 * t;
    int32_t *di = i32 + 4 * t;
    if (k < 0 || (int64_t)k >= s.n_shapes || (pose && pidx && (prow < 0 || prow >= n_poses))) {
        for (int j = 0; j < 16; ++j) dst[j] = 0.0;
        di[0] = -1;  // rejected as bad input downstream
        di[1] = di[2] = 0;
        di[3] = -1;
        return;
    }
    const double *src = s.shape_f64 + 16 * (int64_t)k;
#pragma unroll
    for (int j = 0; j < 16; ++j) dst[j] = src[j];
    if (pose) {
#pragma unroll
        for (int j = 0; j < 12; ++j) dst[j] = pose[12 * prow + j];
    }
    const int32_t *si = s.shape_i32 + 4 * (int64_t)k;
#pragma unroll
    for (int j = 0; j < 4; ++j) di[j] = si[j];
}

__global__ __launch_bounds__(kQBlock) void iota_pairs_kernel(int32_t *__restrict__ a,
                                                             int32_t *__restrict__ b, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * kQBlock + threadIdx.x;
    if (i < n) {
        a[i] = (int32_t)(2 * i);
        b[i] = (int32_t)(2 * i + 1);
    }
}

}  // namespace d3d

using namespace d3d;

extern "C" int d3d_stage_timing(int enable) {
    clear_error();
    std::lock_guard<std::mutex> lk(g_timing.mu);
    g_timing.on = enable != 0;
    // events of calls not yet read are dropped unread (they may still be in
    // flight: destroyed only once they have completed)
    for (auto &ev : g_timing.rec) {
        (void)hipEventSynchronize(ev[3]);
        g_timing.pool.push_back(ev);
    }
    g_timing.rec.clear();
    return D3D_OK;
}

extern "C" int d3d_stage_timing_read(double *ms_out, int64_t *calls_out) {
    clear_error();
    D3D_CHECK_ARG(ms_out && calls_out, "NULL buffer");
    std::lock_guard<std::mutex> lk(g_timing.mu);
    double sum[3] = {0.0, 0.0, 0.0};
    int64_t calls = 0;
    for (auto &ev : g_timing.rec) {
        D3D_CHECK_HIP(hipEventSynchronize(ev[3]));
        for (int k = 0; k < 3; ++k) {
            float ms = 0.0f;
            D3D_CHECK_HIP(hipEventElapsedTime(&ms, ev[k], ev[k + 1]));
            sum[k] += ms;
        }
        ++calls;
        g_timing.pool.push_back(ev);
    }
    g_timing.rec.clear();
    for (int k = 0; k < 3; ++k) ms_out[k] = sum[k];
    *calls_out = calls;
    return D3D_OK;
}

extern "C" int64_t d3d_gjk_workspace_bytes(int64_t n) {
    if (n < 0) return -1;
    return (ws_header_ints() + n) * (int64_t)sizeof(int32_t);
}

static int gjk_distance_impl(const d3d_scene *scene, const int32_t *pair_a, const int32_t *pair_b,
                             int64_t n, const d3d_gjk_params *params, double *dist, double *cp_a,
                             double *cp_b, int32_t *status, int32_t *iterations, double *simplex,
                             int32_t *n_points, void *workspace, int64_t workspace_bytes,
                             const EpaHook *epa,
                             void *stream) {
    clear_error();
    if (int rc = check_scene(scene)) return rc;
    D3D_CHECK_ARG(n >= 0, "n must be >= 0");
    if (n == 0) return D3D_OK;
    D3D_CHECK_ARG(n < (int64_t)0x7fffffff, "n must fit in int32");
    D3D_CHECK_ARG(pair_a && pair_b && dist && cp_a && cp_b && status, "NULL buffer");
    d3d_gjk_params prm;
    if (params) {
        prm = *params;
    } else {
        prm.tolerance = 1e-10;
        prm.max_distance_squared = 100000.0;
        prm.sanity_check = 1e-8;
        prm.max_iterations = 128;
        prm.hull_serial_max = 16;
        for (int c = 0; c < kNumClasses; ++c) prm.lanes_per_pair[c] = 0;
        prm.blocks_per_cu = prm.claim_factor = prm.epa_blocks_per_cu = 0;
    }
    D3D_CHECK_ARG(prm.max_iterations > 0, "max_iterations must be > 0");
    D3D_CHECK_ARG(prm.blocks_per_cu >= 0 && prm.blocks_per_cu <= 3, "blocks_per_cu must be in [0, 3]");
    D3D_CHECK_ARG(prm.claim_factor >= 0 && prm.claim_factor <= 8, "claim_factor must be in [0, 8]");
    D3D_CHECK_ARG(prm.epa_blocks_per_cu >= 0 && prm.epa_blocks_per_cu <= 16,
                  "epa_blocks_per_cu must be in [0, 16]");
    for (int c = 0; c < kNumClasses; ++c) {
        const int G = prm.lanes_per_pair[c];
        D3D_CHECK_ARG(G == 0 || (G <= 64 && (G & (G - 1)) == 0),
                      "lanes_per_pair entries must be 0 or a power of two <= 64");
    }
    void *ws = workspace;
    if (ws) {
        D3D_CHECK_ARG(workspace_bytes >= d3d_gjk_workspace_bytes(n), "workspace too small");
    } else {
        if (int rc = workspace_for(stream, n, &ws)) return rc;
    }
    GjkOut o{dist,
             cp_a,
             cp_b,
             status,
             iterations,
             simplex,
             n_points,
             epa != nullptr,
             epa ? epa->work : nullptr,
             epa ? epa->count : nullptr,
             epa ? epa->out : nullptr,
             epa ? epa->status : nullptr};
    return launch_all(*scene, pair_a, pair_b, n, prm, o, ws, (hipStream_t)stream, epa);
}

extern "C" int d3d_gjk_distance_batch(const d3d_scene *scene, const int32_t *pair_a,
                                      const int32_t *pair_b, int64_t n,
                                      const d3d_gjk_params *params, double *dist, double *cp_a,
                                      double *cp_b, int32_t *status, int32_t *iterations,
                                      double *simplex, int32_t *n_points, void *stream) {
    return gjk_distance_impl(scene, pair_a, pair_b, n, params, dist, cp_a, cp_b, status, iterations,
                             simplex, n_points, nullptr, 0, nullptr, stream);
}

extern "C" int d3d_gjk_distance_batch_ws(const d3d_scene *scene, const int32_t *pair_a,
                                         const int32_t *pair_b, int64_t n,
                                         const d3d_gjk_params *params, double *dist, double *cp_a,
                                         double *cp_b, int32_t *status, int32_t *iterations,
                                         double *simplex, int32_t *n_points, void *workspace,
                                         int64_t workspace_bytes, void *stream) {
    clear_error();
    D3D_CHECK_ARG(workspace != nullptr, "workspace is NULL");
    return gjk_distance_impl(scene, pair_a, pair_b, n, params, dist, cp_a, cp_b, status, iterations,
                             simplex, n_points, workspace, workspace_bytes, nullptr, stream);
}

// Workspace of the fused pipeline: GJK queue | simplex (n*36 f64) |
// n_points (n) | EPA work list (n) | count (u64)
static constexpr int64_t a256(int64_t b) { return (b + 255) / 256 * 256; }
static constexpr int64_t gjk_ws_bytes(int64_t n) {
    return (ws_header_ints() + n) * (int64_t)sizeof(int32_t);
}
static constexpr int64_t gjk_epa_ws_bytes(int64_t n) {
    return a256(gjk_ws_bytes(n)) + a256(n * 36 * 8) + a256(n * 4) + a256(n * 4) + 256;
}

extern "C" int64_t d3d_gjk_epa_workspace_bytes(int64_t n) {
    if (n < 0) return -1;
    return gjk_epa_ws_bytes(n);
}

extern "C" int d3d_gjk_epa_batch_ws(const d3d_scene *scene, const int32_t *pair_a,
                                    const int32_t *pair_b, int64_t n, const d3d_gjk_params *params,
                                    int32_t epa_max_iterations, int32_t epa_max_faces,
                                    double epa_epsilon, double *dist, double *cp_a, double *cp_b,
                                    int32_t *status, int32_t *iterations, double *epa_out,
                                    int32_t *epa_status, void *workspace, int64_t workspace_bytes,
                                    void *stream) {
    clear_error();
    if (int rc = check_scene(scene)) return rc;
    D3D_CHECK_ARG(n >= 0, "n must be >= 0");
    if (n == 0) return D3D_OK;
    D3D_CHECK_ARG(workspace && epa_out && epa_status, "NULL buffer");
    D3D_CHECK_ARG(workspace_bytes >= d3d_gjk_epa_workspace_bytes(n), "workspace too small");
    D3D_CHECK_ARG(epa_max_iterations >= 1, "epa_max_iterations must be >= 1");
    D3D_CHECK_ARG(epa_max_faces >= 4 && epa_max_faces <= 128, "epa_max_faces must be in [4, 128]");
    char *w = (char *)workspace;
    const int64_t gq = a256(d3d_gjk_workspace_bytes(n));
    double *simplex = (double *)(w + gq);
    int32_t *npts = (int32_t *)(w + gq + a256(n * 36 * 8));
    int32_t *work = (int32_t *)(w + gq + a256(n * 36 * 8) + a256(n * 4));
    unsigned long long *count =
        (unsigned long long *)(w + gq + a256(n * 36 * 8) + a256(n * 4) + a256(n * 4));
    const EpaHook hook{work, count, (int)epa_max_iterations, (int)epa_max_faces, epa_epsilon,
                       epa_out, epa_status};
    return gjk_distance_impl(scene, pair_a, pair_b, n, params, dist, cp_a, cp_b, status, iterations,
                             simplex, npts, workspace, gq, &hook, stream);
}

extern "C" int d3d_instance_pairs(const d3d_scene *scene, const int32_t *pair_a,
                                  const int32_t *pair_b, const double *pose_a,
                                  const double *pose_b, const int32_t *pose_idx_a,
                                  const int32_t *pose_idx_b, int64_t n_poses, int64_t n,
                                  double *inst_f64, int32_t *inst_i32, int32_t *inst_a,
                                  int32_t *inst_b, void *stream) {
    clear_error();
    if (int rc = check_scene(scene)) return rc;
    D3D_CHECK_ARG(n >= 0 && n < (int64_t)0x3fffffff, "n must be in [0, 2^30)");
    if (n == 0) return D3D_OK;
    D3D_CHECK_ARG(pair_a && pair_b && inst_f64 && inst_i32, "NULL buffer");
    hipStream_t st = (hipStream_t)stream;
    D3D_CHECK_ARG(!(pose_idx_a || pose_idx_b) || n_poses > 0, "pose indices need n_poses > 0");
    hipLaunchKernelGGL(instance_kernel, dim3(grid_for(2 * n, kQBlock)), dim3(kQBlock), 0, st, *scene,
                       pair_a, pair_b, pose_a, pose_b, pose_idx_a, pose_idx_b, n_poses, n, inst_f64,
                       inst_i32);
    D3D_CHECK_LAUNCH("instance_kernel");
    if (inst_a && inst_b) {
        hipLaunchKernelGGL(iota_pairs_kernel, dim3(grid_for(n, kQBlock)), dim3(kQBlock), 0, st, inst_a,
                           inst_b, n);
        D3D_CHECK_LAUNCH("iota_pairs_kernel");
    }
    return D3D_OK;
}

// ------------------------------------ fused pipeline without allocation

// The fused pipeline's workspace for at most D3D_NOALLOC_MAX_PAIRS pairs, as
// a device array of the library: each GPU holds its own copy from module load
// until the process ends, and every call of d3d_gjk_epa_batch_noalloc on that
// GPU works in it (the queue header and the EPA counters are cleared by fill
// kernels at the start of each call; the other regions are written before
// they are read).  In host code `scratch` names a host-side stand-in: kernels
// get the device address that hipGetSymbolAddress returns for the current
// device, kept per device in g_scratch.
static constexpr int64_t kNoallocBytes = gjk_epa_ws_bytes(D3D_NOALLOC_MAX_PAIRS);
static __device__ __align__(256) unsigned char scratch[kNoallocBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch[kMaxDevices];
}  // namespace

namespace d3d {

// Looks up the device address of `scratch` on the current device once (the
// device must be current); d3d_init calls it, and so does the first
// d3d_gjk_epa_batch_noalloc call on a device d3d_init was not called for.
int noalloc_scratch(void **out) {
    int dev = 0;
    D3D_CHECK_HIP(hipGetDevice(&dev));
    D3D_CHECK_ARG(dev >= 0 && dev < kMaxDevices, "device index beyond the 64 the library tracks");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        D3D_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        D3D_CHECK_ARG(p != nullptr, "hipGetSymbolAddress returned NULL for the scratch array");
        g_scratch[dev] = p;
    }
    if (out) *out = g_scratch[dev];
    return D3D_OK;
}

}  // namespace d3d

extern "C" int d3d_gjk_epa_batch_noalloc(const d3d_scene *scene, const int32_t *pair_a,
                                         const int32_t *pair_b, int64_t n,
                                         const d3d_gjk_params *params, int32_t epa_max_iterations,
                                         int32_t epa_max_faces, double epa_epsilon, double *dist,
                                         double *cp_a, double *cp_b, int32_t *status,
                                         int32_t *iterations, double *epa_out,
                                         int32_t *epa_status, void *stream) {
    clear_error();
    D3D_CHECK_ARG(n >= 0, "n must be >= 0");
    D3D_CHECK_ARG(n <= D3D_NOALLOC_MAX_PAIRS,
                  "n exceeds D3D_NOALLOC_MAX_PAIRS (16384 pairs per call of "
                  "d3d_gjk_epa_batch_noalloc); split the batch or use d3d_gjk_epa_batch_ws");
    if (n == 0) return D3D_OK;
    void *ws = nullptr;
    if (int rc = noalloc_scratch(&ws)) return rc;
    return d3d_gjk_epa_batch_ws(scene, pair_a, pair_b, n, params, epa_max_iterations,
                                epa_max_faces, epa_epsilon, dist, cp_a, cp_b, status, iterations,
                                epa_out, epa_status, ws, kNoallocBytes, stream);
}
