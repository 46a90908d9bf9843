// Host-side structural check of the LBVH node logic (d3d_bvh_core.hpp), run
// on the CPU by tests/test_bvh_host.py: for many sorted key sets the Karras
// hierarchy must be a tree (one parent per node, no cycle, leaves in sorted
// order left to right) and the ropes must give a complete preorder walk.
#include <algorithm>
#include <cstdio>
#include <random>
#include <vector>

#include "d3d_bvh_core.hpp"

using namespace d3d;

static int fail(const char *what, int64_t n) {
    std::printf("FAIL %s (n=%lld)\n", what, (long long)n);
    return 1;
}

static void leaves_dfs(int node, int64_t n, const std::vector<int32_t> &left,
                       const std::vector<int32_t> &right, std::vector<int> &out, int depth, bool &bad) {
    if (depth > n + 1) { bad = true; return; }
    if (node >= n - 1) { out.push_back(node - (int)(n - 1)); return; }
    leaves_dfs(left[node], n, left, right, out, depth + 1, bad);
    leaves_dfs(right[node], n, left, right, out, depth + 1, bad);
}

static int check(std::vector<uint32_t> keys) {
    std::sort(keys.begin(), keys.end());
    const int64_t n = (int64_t)keys.size();
    if (n < 2) return 0;
    std::vector<int32_t> parent(2 * n - 1, -2), left(n - 1), right(n - 1);
    for (int i = 0; i < n - 1; ++i) {
        int lc, rc;
        karras_children(keys.data(), n, i, &lc, &rc);
        if (lc < 1 || rc < 1 || lc > 2 * n - 2 || rc > 2 * n - 2 || lc == rc)
            return fail("child index", n);
        if (parent[lc] != -2 || parent[rc] != -2) return fail("two parents", n);
        parent[lc] = i;
        parent[rc] = i;
        left[i] = lc;
        right[i] = rc;
    }
    if (parent[0] != -2) return fail("root has a parent", n);
    parent[0] = -1;
    for (int64_t k = 1; k < 2 * n - 1; ++k) {
        if (parent[k] < 0) return fail("orphan node", n);
        int cur = (int)k;
        int64_t steps = 0;
        while (cur != 0 && steps <= n) { cur = parent[cur]; ++steps; }
        if (cur != 0) return fail("cycle", n);
    }
    std::vector<int> order;
    bool bad = false;
    leaves_dfs(0, n, left, right, order, 0, bad);
    if (bad || (int64_t)order.size() != n) return fail("leaf count", n);
    for (int64_t p = 0; p < n; ++p)
        if (order[p] != p) return fail("leaf order", n);
    // rope walk that always descends must visit every node once, in preorder
    std::vector<int> seen(2 * n - 1, 0);
    int node = 0;
    int64_t visits = 0;
    while (node >= 0 && visits <= 2 * n) {
        if (seen[node]++) return fail("node visited twice", n);
        ++visits;
        node = node < n - 1 ? left[node] : rope_of(node, n, parent.data(), left.data(), right.data());
    }
    if (visits != 2 * n - 1) return fail("rope walk incomplete", n);
    // ropes of internal nodes too: skipping a subtree lands after its leaves
    for (int i = 0; i < n - 1; ++i) {
        const int r = rope_of(i, n, parent.data(), left.data(), right.data());
        std::vector<int> sub;
        leaves_dfs(i, n, left, right, sub, 0, bad);
        const int last = *std::max_element(sub.begin(), sub.end());
        if (last == n - 1) {
            if (r != -1) return fail("rope after last subtree", n);
        } else {
            std::vector<int> nxt;
            leaves_dfs(r, n, left, right, nxt, 0, bad);
            if (*std::min_element(nxt.begin(), nxt.end()) != last + 1) return fail("rope target", n);
        }
    }
    return 0;
}

int main() {
    std::mt19937 rng(7);
    int bad = 0;
    for (int64_t n : {2, 3, 4, 5, 17, 1000, 5000, 65537}) {
        std::vector<uint32_t> k(n);
        for (auto &v : k) v = rng() & 0x3fffffffu;
        bad += check(k);
        for (auto &v : k) v = rng() % 50u;  // heavy duplication
        bad += check(k);
        for (auto &v : k) v = 12345u;       // all equal
        bad += check(k);
        for (int64_t i = 0; i < n; ++i) k[i] = (uint32_t)i;
        bad += check(k);
    }
    std::printf(bad ? "FAILED %d\n" : "OK\n", bad);
    return bad ? 1 : 0;
}
