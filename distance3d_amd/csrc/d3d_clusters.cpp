// d3d_clusters.cpp -- host-side scene preparation for the pruned support
// sweep: every large hull's vertices are reordered into spatial clusters of
// D3D_CLUSTER_SIZE vertices (median splits along the longest box axis), and
// each cluster gets a bounding sphere inflated by the FP32 evaluation error of
// the device bound c.d + r|d|.  The device argmax (hull_argmax in
// d3d_device.hpp) skips clusters whose bound is below its running maximum and
// breaks exact ties by the original vertex index kept in `vorig`, so supports
// are the same as a full sweep in the original order.
//
// Plain host C++ (no GPU calls); multithreaded over hull vertex ranges.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <new>
#include <numeric>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "d3d.h"

namespace d3d {
void set_error(const std::string &msg);
void clear_error();
}  // namespace d3d

namespace {

constexpr int64_t K = D3D_CLUSTER_SIZE;
constexpr double kU = 5.9604644775390625e-08;  // 2^-24, FP32 unit roundoff

// One distinct hull vertex range of the input and its place in the output.
struct Range {
    int64_t voff, nv, voff64;
    int64_t out_off, out_off64, coff;
    bool clustered;
};

struct Plan {
    std::vector<Range> ranges;
    std::vector<int64_t> shape_range;  // per record: index into ranges, -1 if not a hull
    int64_t n_vertices = 0, n_vertices64 = 0, n_clusters = 0;
};

int make_plan(const int32_t *i32, int64_t n, int32_t min_vertices, Plan &P) {
    if (!i32 || n <= 0) {
        d3d::set_error("empty collider table");
        return D3D_E_INVALID;
    }
    std::map<std::tuple<int64_t, int64_t, int64_t>, int64_t> seen;
    P.shape_range.assign((size_t)n, -1);
    int64_t cur = 0, cur64 = 0, cc = 0;
    for (int64_t k = 0; k < n; ++k) {
        const int32_t type = i32[4 * k], nv = i32[4 * k + 1], voff = i32[4 * k + 2],
                      voff64 = i32[4 * k + 3];
        if (type < D3D_HULL || type > D3D_ELLIPSE) {
            d3d::set_error("unknown collider type in record " + std::to_string(k));
            return D3D_E_INVALID;
        }
        if (type != D3D_HULL) continue;
        if (nv < 1 || voff < 0) {
            d3d::set_error("hull record " + std::to_string(k) + " has no vertex range");
            return D3D_E_INVALID;
        }
        const auto key = std::make_tuple((int64_t)voff, (int64_t)nv, (int64_t)(voff64 < 0 ? -1 : voff64));
        const auto it = seen.find(key);
        if (it != seen.end()) {
            P.shape_range[(size_t)k] = it->second;
            continue;
        }
        Range r{voff, nv, voff64 < 0 ? -1 : voff64, 0, -1, -1, min_vertices > 0 && nv >= min_vertices};
        if (r.clustered) cur = (cur + K - 1) / K * K;
        r.out_off = cur;
        cur += nv;
        if (r.voff64 >= 0) {
            r.out_off64 = cur64;
            cur64 += nv;
        }
        if (r.clustered) {
            r.coff = cc;
            cc += (nv + K - 1) / K;
        }
        seen.emplace(key, (int64_t)P.ranges.size());
        P.shape_range[(size_t)k] = (int64_t)P.ranges.size();
        P.ranges.push_back(r);
    }
    if (cur > INT32_MAX || cur64 > INT32_MAX) {
        d3d::set_error("clustered vertex pool exceeds 2^31 vertices");
        return D3D_E_INVALID;
    }
    P.n_vertices = cur;
    P.n_vertices64 = cur64;
    P.n_clusters = cc;
    return D3D_OK;
}

// Median split of idx[lo, hi) along the longest axis of its bounding box so
// that the left part holds a multiple of K points; recursion leaves every
// cluster full except the last one of the hull.
void kd_split(int32_t *idx, int64_t lo, int64_t hi, const double *P) {
    if (hi - lo <= K) return;
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t i = lo; i < hi; ++i) {
        const double *p = P + 3 * (int64_t)idx[i];
        for (int c = 0; c < 3; ++c) {
            mn[c] = std::min(mn[c], p[c]);
            mx[c] = std::max(mx[c], p[c]);
        }
    }
    int ax = 0;
    for (int c = 1; c < 3; ++c)
        if (mx[c] - mn[c] > mx[ax] - mn[ax]) ax = c;
    const int64_t mid = lo + K * ((hi - lo + 2 * K - 1) / (2 * K));
    std::nth_element(idx + lo, idx + mid, idx + hi, [P, ax](int32_t a, int32_t b) {
        const double pa = P[3 * (int64_t)a + ax], pb = P[3 * (int64_t)b + ax];
        return pa < pb || (pa == pb && a < b);
    });
    kd_split(idx, lo, mid, P);
    kd_split(idx, mid, hi, P);
}

struct Io {
    const float *vx, *vy, *vz;
    const double *v64;
    float *ovx, *ovy, *ovz;
    double *ov64;
    int32_t *vorig;
    float *clusters;
};

void build_range(const Range &r, const Io &io, std::vector<double> &P, std::vector<int32_t> &idx) {
    const int64_t nv = r.nv;
    P.resize((size_t)(3 * nv));
    idx.resize((size_t)nv);
    bool finite = true;
    for (int64_t j = 0; j < nv; ++j) {
        double *p = P.data() + 3 * j;
        if (r.voff64 >= 0) {
            const double *q = io.v64 + 3 * (r.voff64 + j);
            p[0] = q[0];
            p[1] = q[1];
            p[2] = q[2];
        } else {
            p[0] = io.vx[r.voff + j];
            p[1] = io.vy[r.voff + j];
            p[2] = io.vz[r.voff + j];
        }
        finite = finite && std::isfinite(p[0]) && std::isfinite(p[1]) && std::isfinite(p[2]);
    }
    std::iota(idx.begin(), idx.end(), 0);
    // non-finite coordinates keep the original order and unbounded clusters
    if (r.clustered && finite) kd_split(idx.data(), 0, nv, P.data());
    for (int64_t j = 0; j < nv; ++j) {
        const int64_t o = idx[(size_t)j];
        io.ovx[r.out_off + j] = io.vx[r.voff + o];
        io.ovy[r.out_off + j] = io.vy[r.voff + o];
        io.ovz[r.out_off + j] = io.vz[r.voff + o];
        io.vorig[r.out_off + j] = (int32_t)o;
        if (r.voff64 >= 0)
            for (int c = 0; c < 3; ++c) io.ov64[3 * (r.out_off64 + j) + c] = io.v64[3 * (r.voff64 + o) + c];
    }
    if (!r.clustered) return;
    const int64_t nc = (nv + K - 1) / K;
    for (int64_t q = 0; q < nc; ++q) {
        float *rec = io.clusters + 4 * (r.coff + q);
        if (!finite) {
            rec[0] = rec[1] = rec[2] = 0.0f;
            rec[3] = INFINITY;
            continue;
        }
        const int64_t a = q * K, b = std::min(nv, a + K);
        double c[3] = {0.0, 0.0, 0.0};
        for (int64_t j = a; j < b; ++j)
            for (int t = 0; t < 3; ++t) c[t] += P[3 * (size_t)idx[(size_t)j] + t];
        float c32[3];
        for (int t = 0; t < 3; ++t) c32[t] = (float)(c[t] / (double)(b - a));
        double rad = 0.0;
        for (int64_t j = a; j < b; ++j) {
            const double *p = P.data() + 3 * (size_t)idx[(size_t)j];
            const double ex = p[0] - (double)c32[0], ey = p[1] - (double)c32[1], ez = p[2] - (double)c32[2];
            rad = std::max(rad, std::sqrt(ex * ex + ey * ey + ez * ez));
        }
        // covers the FP32 rounding of c.d (direction and three FMA steps) and of
        // the bound's final FMA, with margin; then rounded up to FP32
        const double l1 = std::fabs((double)c32[0]) + std::fabs((double)c32[1]) + std::fabs((double)c32[2]);
        const double inflated = (rad + 8.0 * kU * (l1 + rad)) * (1.0 + 0x1p-20);
        float r32 = (float)inflated;
        if ((double)r32 < inflated) r32 = std::nextafter(r32, INFINITY);
        rec[0] = c32[0];
        rec[1] = c32[1];
        rec[2] = c32[2];
        rec[3] = r32;
    }
}

}  // namespace

extern "C" int d3d_cluster_plan(const int32_t *shape_i32, int64_t n_shapes, int32_t min_vertices,
                                int64_t *n_vertices, int64_t *n_vertices64, int64_t *n_clusters) {
    d3d::clear_error();
    try {
        Plan P;
        if (int rc = make_plan(shape_i32, n_shapes, min_vertices, P)) return rc;
        if (n_vertices) *n_vertices = P.n_vertices;
        if (n_vertices64) *n_vertices64 = P.n_vertices64;
        if (n_clusters) *n_clusters = P.n_clusters;
    } catch (const std::bad_alloc &) {
        d3d::set_error("d3d_cluster_plan: out of host memory");
        return D3D_E_INVALID;
    }
    return D3D_OK;
}

extern "C" int d3d_cluster_scene(const double *shape_f64, const int32_t *shape_i32, int64_t n_shapes,
                                 const float *vx, const float *vy, const float *vz, const double *v64,
                                 int64_t n_vertices, int64_t n_vertices64, int32_t min_vertices,
                                 int32_t n_threads, double *out_f64, int32_t *out_i32, float *ovx,
                                 float *ovy, float *ovz, double *ov64, int32_t *vorig, float *clusters) {
    d3d::clear_error();
    try {
        Plan P;
        if (int rc = make_plan(shape_i32, n_shapes, min_vertices, P)) return rc;
        if (!shape_f64 || !out_f64 || !out_i32) {
            d3d::set_error("NULL collider table");
            return D3D_E_INVALID;
        }
        for (const Range &r : P.ranges) {
            if (r.voff + r.nv > n_vertices || (r.voff64 >= 0 && r.voff64 + r.nv > n_vertices64)) {
                d3d::set_error("hull vertex range outside the pools");
                return D3D_E_INVALID;
            }
        }
        if (P.n_vertices > 0 && (!vx || !vy || !vz || !ovx || !ovy || !ovz || !vorig)) {
            d3d::set_error("NULL vertex pool");
            return D3D_E_INVALID;
        }
        if ((P.n_vertices64 > 0 && (!v64 || !ov64)) || (P.n_clusters > 0 && !clusters)) {
            d3d::set_error("NULL FP64 pool or cluster array");
            return D3D_E_INVALID;
        }
        if (P.n_vertices > 0) {
            std::memset(ovx, 0, sizeof(float) * (size_t)P.n_vertices);
            std::memset(ovy, 0, sizeof(float) * (size_t)P.n_vertices);
            std::memset(ovz, 0, sizeof(float) * (size_t)P.n_vertices);
            std::memset(vorig, 0, sizeof(int32_t) * (size_t)P.n_vertices);
        }
        std::memcpy(out_f64, shape_f64, sizeof(double) * 16 * (size_t)n_shapes);
        std::memcpy(out_i32, shape_i32, sizeof(int32_t) * 4 * (size_t)n_shapes);
        for (int64_t k = 0; k < n_shapes; ++k) {
            const int64_t ri = P.shape_range[(size_t)k];
            if (ri < 0) continue;
            const Range &r = P.ranges[(size_t)ri];
            out_i32[4 * k + 2] = (int32_t)r.out_off;
            out_i32[4 * k + 3] = (int32_t)r.out_off64;
            out_f64[16 * k + 13] = r.clustered ? (double)r.coff : 0.0;
            out_f64[16 * k + 14] = r.clustered ? (double)((r.nv + K - 1) / K) : 0.0;
        }
        const Io io{vx, vy, vz, v64, ovx, ovy, ovz, ov64, vorig, clusters};
        int nt = n_threads > 0 ? n_threads : (int)std::thread::hardware_concurrency();
        nt = std::max(1, std::min<int>(nt, (int)std::max<size_t>(1, P.ranges.size() / 16)));
        std::atomic<int64_t> next(0);
        std::atomic<bool> failed(false);
        auto work = [&]() {
            try {
                std::vector<double> buf;
                std::vector<int32_t> idx;
                for (;;) {
                    const int64_t i = next.fetch_add(1);
                    if (i >= (int64_t)P.ranges.size()) break;
                    build_range(P.ranges[(size_t)i], io, buf, idx);
                }
            } catch (const std::bad_alloc &) {
                failed = true;
            }
        };
        std::vector<std::thread> pool;
        for (int t = 1; t < nt; ++t) pool.emplace_back(work);
        work();
        for (auto &th : pool) th.join();
        if (failed) {
            d3d::set_error("d3d_cluster_scene: out of host memory");
            return D3D_E_INVALID;
        }
    } catch (const std::bad_alloc &) {
        d3d::set_error("d3d_cluster_scene: out of host memory");
        return D3D_E_INVALID;
    }
    return D3D_OK;
}
