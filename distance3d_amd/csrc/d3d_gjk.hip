// d3d_gjk.hip -- batched GJK distance / intersection, support points and
// AABBs for gfx950.
//
// (The batched GJK distance pipeline is in d3d_gjk_queue.hip.)
//
// Kernel shape: one lane per pair, 64 pairs per wavefront.  The FP64 simplex
// (Y, P, Q of up to four points) lives in VGPRs of its lane; support points of
// analytic colliders and of small hulls are computed by the lane itself, while
// large hulls are swept by the whole wavefront (coalesced FP32 SoA rows, see
// hull_argmax), one requesting lane at a time.  The loop runs until every
// lane of the wavefront has terminated; finished lanes keep helping with the
// sweeps of the others.
#include "d3d_device.hpp"
#include "d3d_runtime.hpp"

namespace d3d {

constexpr int kBlock = 256;
constexpr int kStatusBadInput = 0x200;

// gjk_intersection_jolt (_gjk_jolt.py:29-80) and _intersection_loop (:83-135).
__global__ __launch_bounds__(kBlock) void gjk_intersection_kernel(
    d3d_scene s, const int32_t *__restrict__ pa, const int32_t *__restrict__ pb, int64_t n,
    double tolerance, int max_iterations, int serial_max, uint8_t *__restrict__ hit,
    int32_t *__restrict__ status) {
    const int64_t pid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = pid < n;
    Shape A = empty_shape(), B = empty_shape();
    bool ok = false;
    if (valid) {
        const bool okA = load_checked(s, pa[pid], A);
        const bool okB = load_checked(s, pb[pid], B);
        ok = okA && okB;
    }
    Simplex S;
    simplex_zero(S);
    v3 d = mk(1.0, 0.0, 0.0);
    double prev = kMaxFloat;
    const double tol2 = tolerance * tolerance;
    int state = -1, it = 0;
    bool active = ok;

    while (__ballot(active) != 0ull) {
        if (active && it >= max_iterations) {
            state = D3D_ITER_CAP;
            active = false;
        }
        const v3 p = support_wave(s, A, d, active, serial_max);
        const v3 q = support_wave(s, B, -d, active, serial_max);
        if (!active) continue;
        ++it;
        const v3 w = p - q;
        if (dot(d, w) < -kEps) {
            state = D3D_SEPARATED;
            active = false;
            continue;
        }
        simplex_append(S, w);
        v3 v;
        double vv;
        int set;
        if (!closest_to_origin(S, prev, v, vv, set)) {
            state = D3D_SEPARATED;
            active = false;
            continue;
        }
        if (set == 15 || vv <= tol2 || vv <= kEps * max_y_len_sq(S)) {
            state = D3D_INTERSECT;
            active = false;
            continue;
        }
        d = -v;
        if (prev - vv <= kEps * prev) {
            state = D3D_SEPARATED;
            active = false;
            continue;
        }
        prev = vv;
        simplex_compact(S, set);
    }
    if (!valid) return;
    if (!ok) {
        hit[pid] = 0;
        if (status) status[pid] = kStatusBadInput;
        return;
    }
    hit[pid] = state == D3D_INTERSECT ? 1 : 0;
    if (status) status[pid] = state;
}

// collider.support_function(d) for a batch of (collider, direction) queries.
__global__ __launch_bounds__(kBlock) void support_kernel(d3d_scene s,
                                                         const int32_t *__restrict__ shape,
                                                         const double *__restrict__ dirs, int64_t n,
                                                         int serial_max, double *__restrict__ out) {
    const int64_t pid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = pid < n;
    Shape h = empty_shape();
    bool ok = false;
    v3 d = mk(0.0, 0.0, 0.0);
    if (valid) {
        ok = load_checked(s, shape[pid], h);
        d = mk(dirs[3 * pid], dirs[3 * pid + 1], dirs[3 * pid + 2]);
    }
    const v3 p = support_wave(s, h, d, ok, serial_max);
    if (!valid) return;
    out[3 * pid] = ok ? p.x : NAN;
    out[3 * pid + 1] = ok ? p.y : NAN;
    out[3 * pid + 2] = ok ? p.z : NAN;
}

// collider.aabb() (containment.py:6-229), lane per collider.
__global__ __launch_bounds__(kBlock) void aabb_kernel(d3d_scene s, const int32_t *__restrict__ shape,
                                                      int64_t n, double *__restrict__ out) {
    const int64_t pid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (pid >= n) return;
    Shape h;
    double mn[3], mx[3];
    if (!load_checked(s, shape[pid], h)) {
#pragma unroll
        for (int a = 0; a < 6; ++a) out[6 * pid + a] = NAN;
        return;
    }
    const double *R = h.R, *t = h.t;
    switch (h.type) {
    case D3D_HULL: {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            mn[a] = INFINITY;
            mx[a] = -INFINITY;
        }
        for (int j = 0; j < h.nv; ++j) {
            const v3 w = to_world(h, hull_vertex(s, h.voff, h.voff64, j));
            mn[0] = w.x < mn[0] ? w.x : mn[0];
            mx[0] = w.x > mx[0] ? w.x : mx[0];
            mn[1] = w.y < mn[1] ? w.y : mn[1];
            mx[1] = w.y > mx[1] ? w.y : mx[1];
            mn[2] = w.z < mn[2] ? w.z : mn[2];
            mx[2] = w.z > mx[2] ? w.z : mx[2];
        }
        break;
    }
    case D3D_BOX: {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            mn[a] = INFINITY;
            mx[a] = -INFINITY;
        }
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const double l0 = (((c >> 2) & 1) ? 0.5 : -0.5) * h.p[0];
            const double l1 = (((c >> 1) & 1) ? 0.5 : -0.5) * h.p[1];
            const double l2 = ((c & 1) ? 0.5 : -0.5) * h.p[2];
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const double v = t[a] + (l0 * R[3 * a] + l1 * R[3 * a + 1] + l2 * R[3 * a + 2]);
                mn[a] = v < mn[a] ? v : mn[a];
                mx[a] = v > mx[a] ? v : mx[a];
            }
        }
        break;
    }
    case D3D_SPHERE:
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            mn[a] = t[a] - h.p[0];
            mx[a] = t[a] + h.p[0];
        }
        break;
    case D3D_CAPSULE:
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const double e = 0.5 * h.p[1] * fabs(R[3 * a + 2]) + h.p[0];
            mn[a] = t[a] - e;
            mx[a] = t[a] + e;
        }
        break;
    case D3D_CYLINDER:
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const double ax = R[3 * a + 2];
            const double e = 0.5 * h.p[1] * fabs(ax) + h.p[0] * sqrt(1.0 - ax * ax);
            mn[a] = t[a] - e;
            mx[a] = t[a] + e;
        }
        break;
    case D3D_ELLIPSOID: {
        double E[3][3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) E[i][j] = R[3 * i + j] * h.p[j];
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const double cn = sqrt(E[0][j] * E[0][j] + E[1][j] * E[1][j] + E[2][j] * E[2][j]);
#pragma unroll
            for (int i = 0; i < 3; ++i) E[i][j] = E[i][j] / cn;
        }
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) E[i][j] = E[i][j] * h.p[j];
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            double m = -INFINITY;
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                const double v = R[3 * i] * E[k][0] + R[3 * i + 1] * E[k][1] + R[3 * i + 2] * E[k][2];
                m = v > m ? v : m;
            }
            mn[k] = t[k] - m;
            mx[k] = t[k] + m;
        }
        break;
    }
    case D3D_CONE:
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const double ht = h.p[1], r = h.p[0];
            const double pa_ = t[a], pb_ = t[a] + ht * R[3 * a + 2];
            const double ad = pb_ - pa_;
            const double e = sqrt(1.0 - ad * ad / (ht * ht));
            const double lo = pa_ - e * r, hi = pa_ + e * r;
            mn[a] = lo < pb_ ? lo : pb_;
            mx[a] = hi > pb_ ? hi : pb_;
        }
        break;
    case D3D_DISK:
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const double nr = R[3 * a + 2];
            const double e = h.p[0] * sqrt(1.0 - nr * nr);
            mn[a] = t[a] - e;
            mx[a] = t[a] + e;
        }
        break;
    default:  // D3D_ELLIPSE
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const double u = h.p[0] * R[3 * a], v = h.p[1] * R[3 * a + 1];
            const double e = sqrt(u * u + v * v);
            mn[a] = t[a] - e;
            mx[a] = t[a] + e;
        }
        break;
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (h.margin != 0.0) {
            mn[a] = mn[a] - h.margin;
            mx[a] = mx[a] + h.margin;
        }
        out[6 * pid + 2 * a] = mn[a];
        out[6 * pid + 2 * a + 1] = mx[a];
    }
}

}  // namespace d3d

using namespace d3d;

static d3d_gjk_params default_params() {
    d3d_gjk_params p;
    p.tolerance = 1e-10;
    p.max_distance_squared = 100000.0;
    p.sanity_check = 1e-8;
    p.max_iterations = 128;
    p.hull_serial_max = 16;
    for (int c = 0; c < D3D_NUM_CLASSES; ++c) p.lanes_per_pair[c] = 0;
    p.blocks_per_cu = 0;
    p.claim_factor = 0;
    p.epa_blocks_per_cu = 0;
    return p;
}

extern "C" int d3d_support_batch(const d3d_scene *scene, const int32_t *shape, const double *dirs,
                                 int64_t n, double *out, void *stream) {
    clear_error();
    if (int rc = check_scene(scene)) return rc;
    D3D_CHECK_ARG(n >= 0, "n must be >= 0");
    if (n == 0) return D3D_OK;
    D3D_CHECK_ARG(shape && dirs && out, "NULL buffer");
    hipLaunchKernelGGL(support_kernel, dim3(grid_for(n, kBlock)), dim3(kBlock), 0,
                       (hipStream_t)stream, *scene, shape, dirs, n, default_params().hull_serial_max,
                       out);
    D3D_CHECK_LAUNCH("support_kernel");
    return D3D_OK;
}

extern "C" int d3d_aabb_batch(const d3d_scene *scene, const int32_t *shape, int64_t n, double *out,
                              void *stream) {
    clear_error();
    if (int rc = check_scene(scene)) return rc;
    D3D_CHECK_ARG(n >= 0, "n must be >= 0");
    if (n == 0) return D3D_OK;
    D3D_CHECK_ARG(shape && out, "NULL buffer");
    hipLaunchKernelGGL(aabb_kernel, dim3(grid_for(n, kBlock)), dim3(kBlock), 0, (hipStream_t)stream,
                       *scene, shape, n, out);
    D3D_CHECK_LAUNCH("aabb_kernel");
    return D3D_OK;
}

extern "C" int d3d_gjk_intersection_batch(const d3d_scene *scene, const int32_t *pair_a,
                                          const int32_t *pair_b, int64_t n, double tolerance,
                                          int32_t max_iterations, uint8_t *hit, int32_t *status,
                                          void *stream) {
    clear_error();
    if (int rc = check_scene(scene)) return rc;
    D3D_CHECK_ARG(n >= 0, "n must be >= 0");
    if (n == 0) return D3D_OK;
    D3D_CHECK_ARG(pair_a && pair_b && hit, "NULL buffer");
    D3D_CHECK_ARG(max_iterations > 0, "max_iterations must be > 0");
    hipLaunchKernelGGL(gjk_intersection_kernel, dim3(grid_for(n, kBlock)), dim3(kBlock), 0,
                       (hipStream_t)stream, *scene, pair_a, pair_b, n, tolerance, max_iterations,
                       default_params().hull_serial_max, hit, status);
    D3D_CHECK_LAUNCH("gjk_intersection_kernel");
    return D3D_OK;
}
