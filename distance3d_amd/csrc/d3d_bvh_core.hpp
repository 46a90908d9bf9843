// d3d_bvh_core.hpp -- per-node logic of the linear BVH, written as plain
// __host__ __device__ functions so the same code runs in the kernels and in a
// host-side structural check (tests/bvh_host_check.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace d3d {

__host__ __device__ inline int clz32(uint32_t v) { return v ? __builtin_clz(v) : 32; }

// Length of the common prefix of sorted keys i and j; equal keys compare
// their positions so every pair of leaves differs (Karras 2012, section 4).
__host__ __device__ inline int delta(const uint32_t *k, int64_t n, int i, int j) {
    if (j < 0 || (int64_t)j >= n) return -1;
    const uint32_t a = k[i], b = k[j];
    if (a == b) return 32 + clz32((uint32_t)(i ^ j));
    return clz32(a ^ b);
}

// Children of internal node i (0 <= i < n-1) over n sorted keys.  Internal
// nodes are 0..n-2, leaves are n-1 .. 2n-2 (leaf of sorted position p is
// n-1+p).  Internal node i covers the sorted range between i and j and splits
// it at the highest bit in which its keys differ.
__host__ __device__ inline void karras_children(const uint32_t *keys, int64_t n, int i, int *left,
                                                int *right) {
    const int d = (delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
    const int dmin = delta(keys, n, i, i - d);
    int lmax = 2;
    while (delta(keys, n, i, i + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
    const int j = i + l * d;
    const int dnode = delta(keys, n, i, j);
    int s = 0, t = l;
    do {
        t = (t + 1) >> 1;
        if (delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    const int gamma = i + s * d + (d < 0 ? -1 : 0);
    const int lo = i < j ? i : j, hi = i < j ? j : i;
    const int leaf0 = (int)(n - 1);
    *left = (lo == gamma) ? leaf0 + gamma : gamma;
    *right = (hi == gamma + 1) ? leaf0 + gamma + 1 : gamma + 1;
}

// Rope of node k: the node a depth-first walk visits after k's subtree (the
// right sibling of the nearest ancestor-or-self that is a left child), -1
// after the last subtree.  The walk is capped at n steps so a malformed tree
// cannot loop.
__host__ __device__ inline int rope_of(int k, int64_t n, const int32_t *parent, const int32_t *left,
                                       const int32_t *right) {
    int cur = k;
    for (int64_t steps = 0; cur != 0 && steps < n; ++steps) {
        const int p = parent[cur];
        if (p < 0) return -1;
        if (left[p] == cur) return right[p];
        cur = p;
    }
    return -1;
}

}  // namespace d3d
