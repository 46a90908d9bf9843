// d3d_bvh.hip -- AABB broad phase on gfx950: linear BVH build and stackless
// overlap queries.
//
// Build (all on the caller's stream, inside a caller-owned buffer):
//   1. scene bounds of the AABB centres (block reduction)
//   2. 30-bit Morton codes of the centres
//   3. radix sort of (code, index), 4 passes of 8 bits (LDS histograms; the
//      order among equal codes is arbitrary, which the hierarchy tolerates)
//   4. Karras hierarchy: internal node i covers a range of sorted leaves and
//      splits at the highest differing code bit (ties broken by position)
//   5. bottom-up refit: node boxes in FP32 rounded outward, so a node box
//      contains the exact FP64 boxes below it
//   6. ropes: every node stores the node to visit after its subtree
// Query: one lane per query AABB walks left child / rope without a stack;
// FP32 node tests are conservative, and each leaf hit is confirmed with the
// reference's closed-interval FP64 test (aabb_tree.py:503-527), so the pair
// set equals the brute-force set of all_aabbs_overlap exactly.  Hits are
// appended with one atomic per wavefront (ballot compaction).
#include "d3d_bvh_core.hpp"
#include "d3d_device.hpp"
#include "d3d_runtime.hpp"

namespace d3d {

constexpr int kBvhBlock = 256;
constexpr int kSortItems = 16;  // items per thread in the radix sort
constexpr int kSortTile = kBvhBlock * kSortItems;

struct BvhNode {
    float lo[3];
    int left;
    float hi[3];
    int escape;
};

struct BvhHeader {
    int64_t n;
    double slo[3], shi[3];
    int64_t reserved[9];
};

struct BvhView {
    BvhHeader *hdr;
    BvhNode *nodes;     // 2n-1: internal 0..n-2, leaves n-1..2n-2
    int32_t *leaf_orig; // sorted position -> original index
    int32_t *parent;    // 2n-1
    int32_t *right;     // n-1 (refit / ropes)
    int32_t *flags;     // n-1 refit arrival counters
    uint32_t *keys, *keys_alt;
    int32_t *vals_alt;
    int32_t *hist;      // tiles * 256
    double *bpart;      // bound partials 6 * kBoundBlocks
};

constexpr int kBoundBlocks = 256;

__host__ __device__ inline int64_t sort_tiles(int64_t n) { return (n + kSortTile - 1) / kSortTile; }

__host__ __device__ inline int64_t align256(int64_t b) { return (b + 255) / 256 * 256; }

__host__ __device__ inline int64_t bvh_bytes(int64_t n) {
    const int64_t m = n > 0 ? n : 1;
    int64_t b = align256(sizeof(BvhHeader));
    b += align256((2 * m - 1) * (int64_t)sizeof(BvhNode));
    b += align256(m * 4);           // leaf_orig
    b += align256((2 * m - 1) * 4); // parent
    b += align256(m * 4);           // right
    b += align256(m * 4);           // flags
    b += 2 * align256(m * 4);       // keys, keys_alt
    b += align256(m * 4);           // vals_alt
    b += align256(sort_tiles(m) * 256 * 4);
    b += align256(6 * kBoundBlocks * 8);
    return b;
}

inline BvhView bvh_view(void *base, int64_t n) {
    const int64_t m = n > 0 ? n : 1;
    char *p = (char *)base;
    BvhView v;
    v.hdr = (BvhHeader *)p;
    p += align256(sizeof(BvhHeader));
    v.nodes = (BvhNode *)p;
    p += align256((2 * m - 1) * (int64_t)sizeof(BvhNode));
    v.leaf_orig = (int32_t *)p;
    p += align256(m * 4);
    v.parent = (int32_t *)p;
    p += align256((2 * m - 1) * 4);
    v.right = (int32_t *)p;
    p += align256(m * 4);
    v.flags = (int32_t *)p;
    p += align256(m * 4);
    v.keys = (uint32_t *)p;
    p += align256(m * 4);
    v.keys_alt = (uint32_t *)p;
    p += align256(m * 4);
    v.vals_alt = (int32_t *)p;
    p += align256(m * 4);
    v.hist = (int32_t *)p;
    p += align256(sort_tiles(m) * 256 * 4);
    v.bpart = (double *)p;
    return v;
}

static int cu_count_bvh() {
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

// ------------------------------------------------------------- bounds

__global__ __launch_bounds__(kBvhBlock) void bvh_bounds_partial(const double *__restrict__ aabbs,
                                                                int64_t n, double *__restrict__ part) {
    __shared__ double red[6][kBvhBlock];
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t i = (int64_t)blockIdx.x * kBvhBlock + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kBvhBlock) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const double c = (aabbs[6 * i + 2 * a] + aabbs[6 * i + 2 * a + 1]) * 0.5;
            lo[a] = fmin(lo[a], c);
            hi[a] = fmax(hi[a], c);
        }
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        red[a][threadIdx.x] = lo[a];
        red[3 + a][threadIdx.x] = hi[a];
    }
    __syncthreads();
    for (int off = kBvhBlock / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) {
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                red[a][threadIdx.x] = fmin(red[a][threadIdx.x], red[a][threadIdx.x + off]);
                red[3 + a][threadIdx.x] = fmax(red[3 + a][threadIdx.x], red[3 + a][threadIdx.x + off]);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < 6) part[6 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

__global__ __launch_bounds__(64) void bvh_bounds_final(const double *__restrict__ part, int nparts,
                                                       BvhHeader *__restrict__ hdr, int64_t n) {
    if (threadIdx.x == 0) {
        double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int b = 0; b < nparts; ++b)
            for (int a = 0; a < 3; ++a) {
                lo[a] = fmin(lo[a], part[6 * b + a]);
                hi[a] = fmax(hi[a], part[6 * b + 3 + a]);
            }
        hdr->n = n;
        for (int a = 0; a < 3; ++a) {
            hdr->slo[a] = lo[a];
            hdr->shi[a] = hi[a];
        }
    }
}

__device__ __forceinline__ uint32_t spread10(uint32_t v) {
    v &= 0x3ffu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

__global__ __launch_bounds__(kBvhBlock) void bvh_morton_kernel(const double *__restrict__ aabbs,
                                                               int64_t n,
                                                               const BvhHeader *__restrict__ hdr,
                                                               uint32_t *__restrict__ keys,
                                                               int32_t *__restrict__ vals) {
    const int64_t i = (int64_t)blockIdx.x * kBvhBlock + threadIdx.x;
    if (i >= n) return;
    uint32_t q[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const double c = (aabbs[6 * i + 2 * a] + aabbs[6 * i + 2 * a + 1]) * 0.5;
        const double ext = hdr->shi[a] - hdr->slo[a];
        double t = ext > 0.0 ? (c - hdr->slo[a]) / ext : 0.0;
        t = fmin(fmax(t, 0.0), 1.0);
        q[a] = (uint32_t)fmin(t * 1024.0, 1023.0);
    }
    keys[i] = (spread10(q[0]) << 2) | (spread10(q[1]) << 1) | spread10(q[2]);
    vals[i] = (int32_t)i;
}

// --------------------------------------------------------- radix sort

__global__ __launch_bounds__(kBvhBlock) void rs_hist_kernel(const uint32_t *__restrict__ keys,
                                                            int64_t n, int shift,
                                                            int32_t *__restrict__ hist) {
    __shared__ int32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * kSortTile;
    for (int j = 0; j < kSortItems; ++j) {
        const int64_t i = base + (int64_t)j * kBvhBlock + threadIdx.x;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & 255u], 1);
    }
    __syncthreads();
    // digit-major layout: hist[digit * tiles + tile]
    hist[(int64_t)threadIdx.x * gridDim.x + blockIdx.x] = h[threadIdx.x];
}

// exclusive scan of hist (digit-major) in place, one block
__global__ __launch_bounds__(1024) void rs_scan_kernel(int32_t *__restrict__ hist, int64_t len) {
    __shared__ int32_t part[1024];
    const int t = threadIdx.x;
    const int64_t per = (len + 1023) / 1024;
    const int64_t b0 = t * per, b1 = (b0 + per < len) ? b0 + per : len;
    int32_t s = 0;
    for (int64_t i = b0; i < b1; ++i) s += hist[i];
    part[t] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const int32_t v = t >= off ? part[t - off] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int32_t run = part[t] - s;
    for (int64_t i = b0; i < b1; ++i) {
        const int32_t c = hist[i];
        hist[i] = run;
        run += c;
    }
}

// Stable scatter: items of a tile are ranked in index order.  Per chunk of
// 256 items, each wavefront finds the lanes holding the same digit by
// matching the 8 digit bits with ballots; a lane's position is the tile's
// running count for its digit, plus the counts of the same digit in earlier
// wavefronts of the chunk, plus its rank among the earlier lanes of its own
// wavefront.  (LSD radix sort needs every pass to be stable.)
__global__ __launch_bounds__(kBvhBlock) void rs_scatter_kernel(
    const uint32_t *__restrict__ keys, const int32_t *__restrict__ vals, int64_t n, int shift,
    const int32_t *__restrict__ offs, uint32_t *__restrict__ keys_out, int32_t *__restrict__ vals_out) {
    constexpr int kWaves = kBvhBlock / kWave;
    __shared__ int32_t run[256];
    __shared__ int32_t wcnt[kWaves][256];
    const int t = threadIdx.x;
    const int lane = t & (kWave - 1);
    const int wave = t / kWave;
    run[t] = offs[(int64_t)t * gridDim.x + blockIdx.x];
#pragma unroll
    for (int w = 0; w < kWaves; ++w) wcnt[w][t] = 0;
    __syncthreads();
    const unsigned long long lt = lane == 0 ? 0ull : (~0ull >> (kWave - lane));
    const int64_t base = (int64_t)blockIdx.x * kSortTile;
    for (int j = 0; j < kSortItems; ++j) {
        const int64_t i = base + (int64_t)j * kBvhBlock + t;
        const bool valid = i < n;
        const uint32_t k = valid ? keys[i] : 0u;
        const int dig = (int)((k >> shift) & 255u);
        unsigned long long same = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const unsigned long long m = __ballot((dig >> b) & 1);
            same &= ((dig >> b) & 1) ? m : ~m;
        }
        const int rank = __popcll(same & lt);
        if (valid && rank == 0) wcnt[wave][dig] = __popcll(same);
        __syncthreads();
        if (valid) {
            int pos = run[dig] + rank;
            for (int w = 0; w < wave; ++w) pos += wcnt[w][dig];
            keys_out[pos] = k;
            vals_out[pos] = vals[i];
        }
        __syncthreads();
        int add = 0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
            add += wcnt[w][t];
            wcnt[w][t] = 0;
        }
        run[t] += add;
        __syncthreads();
    }
}

// ---------------------------------------------------------- hierarchy

__global__ __launch_bounds__(kBvhBlock) void bvh_hierarchy_kernel(const uint32_t *__restrict__ keys,
                                                                  int64_t n,
                                                                  BvhNode *__restrict__ nodes,
                                                                  int32_t *__restrict__ parent,
                                                                  int32_t *__restrict__ right,
                                                                  int32_t *__restrict__ flags) {
    const int64_t ii = (int64_t)blockIdx.x * kBvhBlock + threadIdx.x;
    if (ii >= n - 1) return;
    const int i = (int)ii;
    int lc, rc;
    karras_children(keys, n, i, &lc, &rc);
    nodes[i].left = lc;
    right[i] = rc;
    parent[lc] = i;
    parent[rc] = i;
    flags[i] = 0;
    if (i == 0) parent[0] = -1;
}

__device__ __forceinline__ void store_box(BvhNode *nd, const float lo[3], const float hi[3]) {
    nd->lo[0] = lo[0];
    nd->lo[1] = lo[1];
    nd->lo[2] = lo[2];
    nd->hi[0] = hi[0];
    nd->hi[1] = hi[1];
    nd->hi[2] = hi[2];
}

// Refit, level by level without device-wide fences: the leaf launch writes
// the leaf boxes; each internal node keeps an arrival counter (flags) and the
// child that arrives second appends the node to the next frontier list.  A
// level launch gives every node of its frontier the union of its children's
// boxes (written by earlier launches, so visible) and builds the next
// frontier the same way.  A node splits its leaf range at the highest
// differing bit of (30-bit Morton code, leaf index), so the tree has at most
// 60 internal levels for n < 2^30 and kRefitLevels launches reach the root.
constexpr int kRefitLevels = 62;

__global__ __launch_bounds__(kBvhBlock) void bvh_refit_leaves(const double *__restrict__ aabbs,
                                                              int64_t n,
                                                              const int32_t *__restrict__ sorted_vals,
                                                              int32_t *__restrict__ leaf_orig,
                                                              BvhNode *__restrict__ nodes,
                                                              const int32_t *__restrict__ parent,
                                                              int32_t *__restrict__ flags,
                                                              int32_t *__restrict__ front,
                                                              int32_t *__restrict__ count) {
    const int64_t i = (int64_t)blockIdx.x * kBvhBlock + threadIdx.x;
    if (i >= n) return;
    const int orig = sorted_vals[i];
    leaf_orig[i] = orig;
    float lo[3], hi[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        lo[a] = __double2float_rd(aabbs[6 * (int64_t)orig + 2 * a]);
        hi[a] = __double2float_ru(aabbs[6 * (int64_t)orig + 2 * a + 1]);
    }
    const int leaf = (int)(n - 1 + i);
    store_box(&nodes[leaf], lo, hi);
    nodes[leaf].left = -1;
    if (n == 1) return;
    const int p = parent[leaf];
    if (p >= 0 && atomicAdd(&flags[p], 1) == 1) front[atomicAdd(count, 1)] = p;
}

__global__ __launch_bounds__(kBvhBlock) void bvh_refit_level(BvhNode *__restrict__ nodes,
                                                             const int32_t *__restrict__ parent,
                                                             const int32_t *__restrict__ right,
                                                             int32_t *__restrict__ flags,
                                                             const int32_t *__restrict__ front_in,
                                                             const int32_t *__restrict__ count_in,
                                                             int32_t *__restrict__ front_out,
                                                             int32_t *__restrict__ count_out) {
    const int cnt = *count_in;
    for (int i = blockIdx.x * kBvhBlock + threadIdx.x; i < cnt; i += gridDim.x * kBvhBlock) {
        const int p = front_in[i];
        const BvhNode na = nodes[nodes[p].left], nb = nodes[right[p]];
        float lo[3], hi[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            lo[k] = fminf(na.lo[k], nb.lo[k]);
            hi[k] = fmaxf(na.hi[k], nb.hi[k]);
        }
        store_box(&nodes[p], lo, hi);
        const int q = parent[p];
        if (q >= 0 && atomicAdd(&flags[q], 1) == 1) front_out[atomicAdd(count_out, 1)] = q;
    }
}

__global__ void bvh_zero_ints(int32_t *__restrict__ a, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        a[i] = 0;
}

// Boxes of every node from the leaf AABBs: leaf launch, then one launch per
// tree level (see bvh_refit_leaves).  `flags` must be zero for the internal
// nodes; leaf_orig holds the sorted permutation (read and rewritten in place
// when sorted_vals == leaf_orig).
static void refit_launches(const BvhView &v, const double *aabbs, int64_t n,
                           const int32_t *sorted_vals, hipStream_t st) {
    // frontiers ping-pong between keys_alt and vals_alt (free after the sort);
    // per-level counters live in the histogram area
    int32_t *fr[2] = {(int32_t *)v.keys_alt, v.vals_alt};
    int32_t *cnt = v.hist;
    const int nb = (int)((n + kBvhBlock - 1) / kBvhBlock);
    hipLaunchKernelGGL(bvh_zero_ints, dim3(1), dim3(kBvhBlock), 0, st, cnt, (int64_t)(kRefitLevels + 1));
    hipLaunchKernelGGL(bvh_refit_leaves, dim3(nb), dim3(kBvhBlock), 0, st, aabbs, n, sorted_vals,
                       v.leaf_orig, v.nodes, v.parent, v.flags, fr[0], cnt);
    if (n > 1) {
        int lg = (int)((n / 2 + kBvhBlock - 1) / kBvhBlock);
        const int cap = 4 * cu_count_bvh();
        if (lg > cap) lg = cap;
        if (lg < 1) lg = 1;
        for (int L = 0; L < kRefitLevels; ++L)
            hipLaunchKernelGGL(bvh_refit_level, dim3(lg), dim3(kBvhBlock), 0, st, v.nodes, v.parent,
                               v.right, v.flags, fr[L & 1], cnt + L, fr[(L + 1) & 1], cnt + L + 1);
    }
}

__global__ __launch_bounds__(kBvhBlock) void bvh_rope_kernel(int64_t n, BvhNode *__restrict__ nodes,
                                                             const int32_t *__restrict__ parent,
                                                             const int32_t *__restrict__ left,
                                                             const int32_t *__restrict__ right) {
    const int64_t k = (int64_t)blockIdx.x * kBvhBlock + threadIdx.x;
    if (k >= 2 * n - 1) return;
    nodes[k].escape = rope_of((int)k, n, parent, left, right);
}

// internal nodes' left children as a plain array for rope_of
__global__ __launch_bounds__(kBvhBlock) void bvh_left_kernel(int64_t n, const BvhNode *__restrict__ nodes,
                                                             int32_t *__restrict__ left) {
    const int64_t k = (int64_t)blockIdx.x * kBvhBlock + threadIdx.x;
    if (k < n - 1) left[k] = nodes[k].left;
}

// ------------------------------------------------------------- query

__global__ __launch_bounds__(kBvhBlock) void bvh_query_kernel(
    const BvhNode *__restrict__ nodes, const int32_t *__restrict__ leaf_orig, int64_t n,
    const double *__restrict__ aabbs, const double *__restrict__ q, int64_t m,
    int32_t *__restrict__ pairs, int64_t cap, unsigned long long *__restrict__ counter) {
    const int64_t qi = (int64_t)blockIdx.x * kBvhBlock + threadIdx.x;
    if (qi >= m) return;
    double Q[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) Q[k] = q[6 * qi + k];
    float qlo[3], qhi[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        qlo[a] = __double2float_rd(Q[2 * a]);
        qhi[a] = __double2float_ru(Q[2 * a + 1]);
    }
    const int lane = lane_id();
    const int leaf0 = (int)(n - 1);
    int node = 0;
    // a rope walk visits each node at most once; the cap only guards
    // against a corrupted tree
    for (int64_t steps = 0; node >= 0 && steps < 2 * n; ++steps) {
        const BvhNode nd = nodes[node];
        const bool ov = nd.lo[0] <= qhi[0] && nd.hi[0] >= qlo[0] && nd.lo[1] <= qhi[1] &&
                        nd.hi[1] >= qlo[1] && nd.lo[2] <= qhi[2] && nd.hi[2] >= qlo[2];
        bool hit = false;
        int orig = 0;
        if (ov && node >= leaf0) {
            orig = leaf_orig[node - leaf0];
            const double *A = aabbs + 6 * (int64_t)orig;
            // aabb_overlap(tree_aabb, test_aabb), aabb_tree.py:520-527
            hit = A[0] <= Q[1] && A[1] >= Q[0] && A[2] <= Q[3] && A[3] >= Q[2] && A[4] <= Q[5] &&
                  A[5] >= Q[4];
        }
        const unsigned long long msk = __ballot(hit);
        if (hit) {
            const int first = __ffsll((long long)msk) - 1;
            unsigned long long base = 0;
            if (lane == first) base = atomicAdd(counter, (unsigned long long)__popcll(msk));
            const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)base, first);
            const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(base >> 32), first);
            const unsigned long long pos =
                (((unsigned long long)hi << 32) | lo) +
                (unsigned long long)__popcll(msk & (lane == 0 ? 0ull : (~0ull >> (64 - lane))));
            if (pos < (unsigned long long)cap) {
                pairs[2 * pos] = (int32_t)qi;
                pairs[2 * pos + 1] = orig;
            }
        }
        node = (ov && node < leaf0) ? nd.left : nd.escape;
    }
}

__global__ void bvh_zero_counter(unsigned long long *c) { *c = 0ull; }

// ------------------------------------------------------------ dense overlap
//
// all_aabbs_overlap (aabb_tree.py:465-500) by brute force, for box sets
// whose product n * m is small enough that testing every pair beats a tree
// walk (dense overlaps: a tree query with ~1000 hits is a long serial walk).
// One wavefront takes kDenseRows rows of `a` (the boxes stay in scalar
// registers) and sweeps `b` 64 boxes at a time, one per lane; a count pass,
// an exclusive scan, and a write pass that places each row's hits at the
// row's offset in b order (ballot ranks), so the pairs come out row-major
// (i outer, j inner), the reference's order, with no sort.
constexpr int kDenseRows = 4;
constexpr int kDenseBlock = 256;

__device__ __forceinline__ bool box_overlap(const double (&q)[6], const double *__restrict__ B) {
    // aabb_overlap(a_i, b_j), aabb_tree.py:520-527 (closed intervals)
    return q[0] <= B[1] && q[1] >= B[0] && q[2] <= B[3] && q[3] >= B[2] && q[4] <= B[5] &&
           q[5] >= B[4];
}

template <bool kWrite>
__global__ __launch_bounds__(kDenseBlock) void dense_overlap_kernel(
    const double *__restrict__ a, int64_t n, const double *__restrict__ b, int64_t m,
    unsigned long long *__restrict__ counts, int32_t *__restrict__ pairs, int64_t cap) {
    const int lane = lane_id();
    const int64_t wave = ((int64_t)blockIdx.x * kDenseBlock + threadIdx.x) / kWave;
    const int64_t row0 = wave * kDenseRows;
    if (row0 >= n) return;
    double q[kDenseRows][6];
    unsigned long long off[kDenseRows];
#pragma unroll
    for (int r = 0; r < kDenseRows; ++r) {
        const int64_t row = min(row0 + r, n - 1);
#pragma unroll
        for (int c = 0; c < 6; ++c) q[r][c] = wave_uniform(a[6 * row + c]);
        off[r] = kWrite ? counts[row] : 0ull;
    }
    const unsigned long long lt = lane == 0 ? 0ull : (~0ull >> (kWave - lane));
    for (int64_t j0 = 0; j0 < m; j0 += kWave) {
        const int64_t j = j0 + lane;
        double B[6];
        const bool in = j < m;
#pragma unroll
        for (int c = 0; c < 6; ++c) B[c] = in ? b[6 * j + c] : 0.0;
#pragma unroll
        for (int r = 0; r < kDenseRows; ++r) {
            const bool hit = in && row0 + r < n && box_overlap(q[r], B);
            const unsigned long long msk = __ballot(hit);
            if (kWrite && hit) {
                const unsigned long long pos = off[r] + (unsigned long long)__popcll(msk & lt);
                if (pos < (unsigned long long)cap) {
                    pairs[2 * pos] = (int32_t)(row0 + r);
                    pairs[2 * pos + 1] = (int32_t)j;
                }
            }
            off[r] += (unsigned long long)__popcll(msk);
        }
    }
    if (!kWrite && lane < kDenseRows && row0 + lane < n) {
        unsigned long long mine = off[0];
#pragma unroll
        for (int r = 1; r < kDenseRows; ++r) mine = lane == r ? off[r] : mine;
        counts[row0 + lane] = mine;
    }
}

// Exclusive scan of counts[0..n) in place by one block; the sum to *total.
__global__ __launch_bounds__(1024) void dense_scan_kernel(unsigned long long *__restrict__ counts,
                                                          int64_t n,
                                                          unsigned long long *__restrict__ total) {
    __shared__ unsigned long long part[1024];
    const int t = threadIdx.x;
    const int64_t per = (n + 1023) / 1024;
    const int64_t lo = min((int64_t)t * per, n), hi = min(lo + per, n);
    unsigned long long sum = 0;
    for (int64_t k = lo; k < hi; ++k) sum += counts[k];
    part[t] = sum;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {
        const unsigned long long v = t >= d ? part[t - d] : 0ull;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    unsigned long long run = part[t] - sum;
    for (int64_t k = lo; k < hi; ++k) {
        const unsigned long long c = counts[k];
        counts[k] = run;
        run += c;
    }
    if (t == 1023) *total = part[1023];
}

}  // namespace d3d

using namespace d3d;

extern "C" int64_t d3d_bvh_bytes(int64_t n) {
    if (n < 0) return -1;
    return bvh_bytes(n);
}

extern "C" int d3d_bvh_build(const double *aabbs, int64_t n, void *bvh, int64_t bvh_bytes_avail,
                             void *stream) {
    clear_error();
    D3D_CHECK_ARG(n >= 1 && n < (int64_t)0x40000000, "n must be in [1, 2^30)");
    D3D_CHECK_ARG(aabbs && bvh, "NULL buffer");
    D3D_CHECK_ARG(bvh_bytes_avail >= bvh_bytes(n), "bvh buffer smaller than d3d_bvh_bytes(n)");
    hipStream_t st = (hipStream_t)stream;
    BvhView v = bvh_view(bvh, n);
    int nb = (int)((n + kBvhBlock - 1) / kBvhBlock);
    const int nparts = nb < kBoundBlocks ? nb : kBoundBlocks;
    hipLaunchKernelGGL(bvh_bounds_partial, dim3(nparts), dim3(kBvhBlock), 0, st, aabbs, n, v.bpart);
    D3D_CHECK_LAUNCH("bvh_bounds_partial");
    hipLaunchKernelGGL(bvh_bounds_final, dim3(1), dim3(64), 0, st, v.bpart, nparts, v.hdr, n);
    D3D_CHECK_LAUNCH("bvh_bounds_final");
    // codes into keys_alt/vals_alt, sorted back and forth; 4 passes end in keys/leaf_orig-source
    uint32_t *kin = v.keys, *kout = v.keys_alt;
    int32_t *vin = v.leaf_orig, *vout = v.vals_alt;
    hipLaunchKernelGGL(bvh_morton_kernel, dim3(nb), dim3(kBvhBlock), 0, st, aabbs, n, v.hdr, kin, vin);
    D3D_CHECK_LAUNCH("bvh_morton_kernel");
    const int64_t tiles = sort_tiles(n);
    for (int pass = 0; pass < 4; ++pass) {
        const int shift = 8 * pass;
        hipLaunchKernelGGL(rs_hist_kernel, dim3((unsigned)tiles), dim3(kBvhBlock), 0, st, kin, n,
                           shift, v.hist);
        D3D_CHECK_LAUNCH("rs_hist_kernel");
        hipLaunchKernelGGL(rs_scan_kernel, dim3(1), dim3(1024), 0, st, v.hist, tiles * 256);
        D3D_CHECK_LAUNCH("rs_scan_kernel");
        hipLaunchKernelGGL(rs_scatter_kernel, dim3((unsigned)tiles), dim3(kBvhBlock), 0, st, kin, vin,
                           n, shift, v.hist, kout, vout);
        D3D_CHECK_LAUNCH("rs_scatter_kernel");
        uint32_t *tk = kin; kin = kout; kout = tk;
        int32_t *tv = vin; vin = vout; vout = tv;
    }
    // after 4 passes the sorted data is back in keys / leaf_orig's buffer
    if (n > 1) {
        const int ni = (int)((n - 1 + kBvhBlock - 1) / kBvhBlock);
        hipLaunchKernelGGL(bvh_hierarchy_kernel, dim3(ni), dim3(kBvhBlock), 0, st, kin, n, v.nodes,
                           v.parent, v.right, v.flags);
        D3D_CHECK_LAUNCH("bvh_hierarchy_kernel");
    }
    refit_launches(v, aabbs, n, vin, st);
    D3D_CHECK_LAUNCH("bvh_refit");
    // ropes; the refit flags (no longer needed) hold the left-child array
    if (n > 1) {
        hipLaunchKernelGGL(bvh_left_kernel, dim3((unsigned)((n - 1 + kBvhBlock - 1) / kBvhBlock)),
                           dim3(kBvhBlock), 0, st, n, v.nodes, v.flags);
        D3D_CHECK_LAUNCH("bvh_left_kernel");
    }
    const int nn = (int)((2 * n - 1 + kBvhBlock - 1) / kBvhBlock);
    hipLaunchKernelGGL(bvh_rope_kernel, dim3(nn), dim3(kBvhBlock), 0, st, n, v.nodes, v.parent,
                       v.flags, v.right);
    D3D_CHECK_LAUNCH("bvh_rope_kernel");
    return D3D_OK;
}

extern "C" int d3d_bvh_refit(const double *aabbs, int64_t n, void *bvh, int64_t bvh_bytes_avail,
                             void *stream) {
    clear_error();
    D3D_CHECK_ARG(n >= 1 && n < (int64_t)0x40000000, "n must be in [1, 2^30)");
    D3D_CHECK_ARG(aabbs && bvh, "NULL buffer");
    D3D_CHECK_ARG(bvh_bytes_avail >= bvh_bytes(n), "bvh buffer smaller than d3d_bvh_bytes(n)");
    hipStream_t st = (hipStream_t)stream;
    BvhView v = bvh_view(bvh, n);
    // the build left the ropes' left-child array in `flags`; the arrival
    // counters of the refit need it zeroed (the ropes are already in the nodes)
    if (n > 1) {
        const int64_t zb = (n - 1 + kBvhBlock - 1) / kBvhBlock;
        hipLaunchKernelGGL(bvh_zero_ints, dim3((unsigned)(zb < 1024 ? zb : 1024)), dim3(kBvhBlock), 0, st,
                           v.flags, n - 1);
    }
    refit_launches(v, aabbs, n, v.leaf_orig, st);
    D3D_CHECK_LAUNCH("bvh_refit");
    return D3D_OK;
}

extern "C" int d3d_bvh_query(const void *bvh, int64_t n, const double *aabbs, const double *queries,
                             int64_t m, int32_t *pairs, int64_t cap, uint64_t *n_pairs,
                             void *stream) {
    clear_error();
    D3D_CHECK_ARG(n >= 1 && n < (int64_t)0x40000000, "n must be in [1, 2^30)");
    D3D_CHECK_ARG(m >= 0 && m < (int64_t)0x7fffffff, "m must fit in int32");
    D3D_CHECK_ARG(bvh && aabbs && n_pairs && (m == 0 || queries), "NULL buffer");
    D3D_CHECK_ARG(cap >= 0 && (cap == 0 || pairs), "pairs buffer is NULL");
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(bvh_zero_counter, dim3(1), dim3(1), 0, st, (unsigned long long *)n_pairs);
    D3D_CHECK_LAUNCH("bvh_zero_counter");
    if (m == 0) return D3D_OK;
    BvhView v = bvh_view((void *)bvh, n);
    const int nb = (int)((m + kBvhBlock - 1) / kBvhBlock);
    hipLaunchKernelGGL(bvh_query_kernel, dim3(nb), dim3(kBvhBlock), 0, st, v.nodes, v.leaf_orig, n,
                       aabbs, queries, m, pairs, cap, (unsigned long long *)n_pairs);
    D3D_CHECK_LAUNCH("bvh_query_kernel");
    return D3D_OK;
}

extern "C" int64_t d3d_overlap_dense_workspace_bytes(int64_t n) {
    if (n < 0) return -1;
    return (n + 1) * (int64_t)sizeof(unsigned long long);
}

extern "C" int d3d_overlap_dense(const double *a, int64_t n, const double *b, int64_t m,
                                 int32_t *pairs, int64_t cap, uint64_t *n_pairs, void *workspace,
                                 int64_t workspace_bytes, void *stream) {
    clear_error();
    D3D_CHECK_ARG(n >= 0 && n < (int64_t)0x7fffffff && m >= 0 && m < (int64_t)0x7fffffff,
                  "n and m must fit in int32");
    D3D_CHECK_ARG(n_pairs && (n == 0 || (a && workspace)) && (m == 0 || b), "NULL buffer");
    D3D_CHECK_ARG(cap >= 0 && (cap == 0 || pairs), "pairs buffer is NULL");
    D3D_CHECK_ARG(workspace_bytes >= d3d_overlap_dense_workspace_bytes(n), "workspace too small");
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(bvh_zero_counter, dim3(1), dim3(1), 0, st, (unsigned long long *)n_pairs);
    D3D_CHECK_LAUNCH("bvh_zero_counter");
    if (n == 0 || m == 0) return D3D_OK;
    unsigned long long *counts = (unsigned long long *)workspace;
    const int64_t waves = (n + kDenseRows - 1) / kDenseRows;
    const unsigned blocks = (unsigned)((waves * kWave + kDenseBlock - 1) / kDenseBlock);
    hipLaunchKernelGGL(dense_overlap_kernel<false>, dim3(blocks), dim3(kDenseBlock), 0, st, a, n, b,
                       m, counts, pairs, cap);
    D3D_CHECK_LAUNCH("dense_overlap_kernel<count>");
    hipLaunchKernelGGL(dense_scan_kernel, dim3(1), dim3(1024), 0, st, counts, n,
                       (unsigned long long *)n_pairs);
    D3D_CHECK_LAUNCH("dense_scan_kernel");
    if (cap > 0) {
        hipLaunchKernelGGL(dense_overlap_kernel<true>, dim3(blocks), dim3(kDenseBlock), 0, st, a, n,
                           b, m, counts, pairs, cap);
        D3D_CHECK_LAUNCH("dense_overlap_kernel<write>");
    }
    return D3D_OK;
}
