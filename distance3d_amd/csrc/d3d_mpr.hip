// Minkowski portal refinement (mpr.py:20-393, after libccd): intersection
// and penetration depth, direction and contact position for a batch of pairs.
//
// One lane per pair, FP64 throughout, the oracle's expressions in its order.
// The algorithm calls the pair's support function from several places
// (portal discovery, refinement, penetration search); each lane runs it as a
// state machine whose phases end at a support call, and every loop round all
// lanes of the wavefront evaluate their supports together (support_wave
// sweeps large hulls with the whole wavefront).  The portal (4 rows of the
// Minkowski point and both support points) stays in registers: rows are only
// addressed with compile-time indices, runtime row choices become selects.
#include "d3d_device.hpp"
#include "d3d_runtime.hpp"

namespace d3d {

constexpr int kMprBlock = 256;
// hulls with more vertices than this are swept by the whole wavefront (the
// value the other lane-per-pair kernels use)
constexpr int kMprSerialMax = 16;

struct MprPortal {
    v3 v[4], a[4], b[4];
};

// row r (runtime) = (v, a, b)
__device__ __forceinline__ void mpr_set(MprPortal &P, int r, v3 v, v3 a, v3 b) {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        P.v[k] = vsel(k == r, v, P.v[k]);
        P.a[k] = vsel(k == r, a, P.a[k]);
        P.b[k] = vsel(k == r, b, P.b[k]);
    }
}

// _portal_direction (mpr.py:286-293)
__device__ __forceinline__ v3 mpr_portal_dir(const MprPortal &P) {
    return norm_vector(cross(P.v[2] - P.v[1], P.v[3] - P.v[1]));
}

// _encapsulates_origin (mpr.py:296-299)
__device__ __forceinline__ bool mpr_encapsulates(v3 v, v3 d) { return dot(v, d) > -10.0 * kEps; }

// _portal_reach_tolerance (mpr.py:302-311)
__device__ __forceinline__ bool mpr_reach_tol(const MprPortal &P, v3 v4, v3 d, double tol) {
    const double p = dot(v4, d);
    double m = p - dot(P.v[1], d);
    const double t2 = p - dot(P.v[2], d);
    if (t2 < m) m = t2;
    const double t3 = p - dot(P.v[3], d);
    if (t3 < m) m = t3;
    return m < tol + kEps;
}

// _expand_portal (mpr.py:314-329)
__device__ __forceinline__ void mpr_expand(MprPortal &P, v3 v4, v3 a4, v3 b4) {
    const v3 x = cross(v4, P.v[0]);
    int r;
    if (dot(P.v[1], x) > 0.0) r = dot(P.v[2], x) > 0.0 ? 1 : 3;
    else r = dot(P.v[3], x) > 0.0 ? 2 : 1;
    mpr_set(P, r, v4, a4, b4);
}

__device__ __forceinline__ double vnorm(v3 v) { return sqrt(dot(v, v)); }

// point_to_triangle(zeros, (A, B, C)) (distance/_triangle.py:12-89)
__device__ __forceinline__ double origin_to_triangle(v3 A, v3 B, v3 C, v3 &cp) {
    const v3 o = mk(0.0, 0.0, 0.0);
    const v3 ab = B - A, ac = C - A;
    const v3 ap = o - A;
    const double d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0.0 && d2 <= 0.0) {
        cp = A;
        return vnorm(o - cp);
    }
    const v3 bp = o - B;
    const double d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0.0 && d4 <= d3) {
        cp = B;
        return vnorm(o - cp);
    }
    const double vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0 && 0.0 <= d1 && d3 <= 0.0) {
        const double v = d1 / (d1 - d3);
        cp = A + ab * v;
        return vnorm(o - cp);
    }
    const v3 cpt = o - C;
    const double d5 = dot(ab, cpt), d6 = dot(ac, cpt);
    if (d6 >= 0.0 && d5 <= d6) {
        cp = C;
        return vnorm(o - cp);
    }
    const double vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0 && 0.0 <= d2 && d6 <= 0.0) {
        const double w = d2 / (d2 - d6);
        cp = A + ac * w;
        return vnorm(o - cp);
    }
    const double va = d3 * d6 - d5 * d4;
    if (va <= 0.0 && 0.0 <= d4 - d3 && d5 - d6 >= 0.0) {
        const double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
        cp = B + (C - B) * w;
        return vnorm(o - cp);
    }
    const double denom = 1.0 / (va + vb + vc);
    const double v = vb * denom, w = vc * denom;
    cp = (A + ab * v) + ac * w;
    return vnorm(o - cp);
}

// _contact_position (mpr.py:378-393)
__device__ __forceinline__ v3 mpr_contact(const MprPortal &P, v3 d) {
    double b0 = dot(cross(P.v[1], P.v[2]), P.v[3]);
    double b1 = dot(cross(P.v[3], P.v[2]), P.v[0]);
    double b2 = dot(cross(P.v[0], P.v[1]), P.v[3]);
    double b3 = dot(cross(P.v[2], P.v[1]), P.v[0]);
    double sum = ((b0 + b1) + b2) + b3;
    if (sum < kEps) {
        b0 = 0.0;
        b1 = dot(cross(P.v[2], P.v[3]), d);
        b2 = dot(cross(P.v[3], P.v[1]), d);
        b3 = dot(cross(P.v[1], P.v[2]), d);
        sum = ((b0 + b1) + b2) + b3;
    }
    b0 = b0 / sum;
    b1 = b1 / sum;
    b2 = b2 / sum;
    b3 = b3 / sum;
    const v3 p1 = ((P.a[0] * b0 + P.a[1] * b1) + P.a[2] * b2) + P.a[3] * b3;
    const v3 p2 = ((P.b[0] * b0 + P.b[1] * b1) + P.b[2] * b2) + P.b[3] * b3;
    return (p1 + p2) * 0.5;
}

// phases: each ends at a support call along d
enum : int {
    kMprDone = 0,
    kMprV1 = 1,      // _find_support_in_direction_of_origin_ray (mpr.py:180-188)
    kMprV2 = 2,      // _find_support_perpendicular_to_plane_containing_origin_v01 (:191-205)
    kMprV3 = 3,      // the discovery loop of _discover_portal (:140-160, 217-235)
    kMprRefine = 4,  // _refine_portal (:253-283)
    kMprPen = 5,     // _find_penetration_info (:332-346)
};
constexpr int kMprRefineCap = 0x100;
constexpr int kMprBadInput = 0x200;

__global__ __launch_bounds__(kMprBlock) void mpr_kernel(
    d3d_scene s, const int32_t *__restrict__ pa, const int32_t *__restrict__ pb,
    const double *__restrict__ centers, int64_t n, double tol, int max_iter, int penetration,
    int serial_max, int32_t *__restrict__ status_out, double *__restrict__ out) {
    const int64_t pid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = pid < n;
    Shape A = empty_shape(), B = empty_shape();
    bool ok = false;
    int ia = 0, ib = 0;
    if (valid) {
        ia = pa[pid];
        ib = pb[pid];
        const bool okA = load_checked(s, ia, A);
        const bool okB = load_checked(s, ib, B);
        ok = okA && okB;
    }
    MprPortal P;
#pragma unroll
    for (int k = 0; k < 4; ++k) P.v[k] = P.a[k] = P.b[k] = mk(0.0, 0.0, 0.0);
    double res[7];
#pragma unroll
    for (int k = 0; k < 7; ++k) res[k] = NAN;
    int status = 0, phase = kMprDone, it = 0;
    v3 d = mk(0.0, 0.0, 0.0);
    if (ok) {
        // _find_origin_ray (mpr.py:163-177)
        const v3 ca = mk(centers[3 * (int64_t)ia], centers[3 * (int64_t)ia + 1], centers[3 * (int64_t)ia + 2]);
        const v3 cb = mk(centers[3 * (int64_t)ib], centers[3 * (int64_t)ib + 1], centers[3 * (int64_t)ib + 2]);
        P.v[0] = ca - cb;
        P.a[0] = ca;
        P.b[0] = cb;
        if (P.v[0].x == 0.0 && P.v[0].y == 0.0 && P.v[0].z == 0.0) P.v[0].x += kEps * 10.0;
        d = norm_vector(-P.v[0]);
        phase = kMprV1;
    }
    // the refinement has succeeded: penetration info, or done
    auto intersecting = [&]() {
        if (penetration) {
            phase = kMprPen;
            it = 0;
            d = mpr_portal_dir(P);
        } else {
            status = 1;
            phase = kMprDone;
        }
    };
    // set up a refinement round along the current portal
    auto refine_round = [&]() {
        d = mpr_portal_dir(P);
        if (mpr_encapsulates(P.v[1], d)) {
            intersecting();
        } else if (it >= max_iter) {
            // the reference refines without bound; here it gives up
            status = kMprRefineCap;
            phase = kMprDone;
        } else {
            phase = kMprRefine;
        }
    };
    while (__ballot(phase != kMprDone) != 0ull) {
        const bool want = phase != kMprDone;
        const v3 sa = support_wave(s, A, d, want, serial_max);
        const v3 sb = support_wave(s, B, -d, want, serial_max);
        if (!want) continue;
        const v3 w = sa - sb;
        if (phase == kMprV1) {
            mpr_set(P, 1, w, sa, sb);
            const bool v1_zero = w.x == 0.0 && w.y == 0.0 && w.z == 0.0;
            if (!v1_zero && dot(w, d) < kEps) {
                phase = kMprDone;  // origin outside
                continue;
            }
            d = cross(P.v[0], P.v[1]);
            if (dot(d, d) < kEps) {
                status = 1;
                phase = kMprDone;
                const v3 pos = (P.a[1] + P.b[1]) * 0.5;
                if (v1_zero) {
                    // _find_penetration_touch (mpr.py:361-367)
                    res[0] = 0.0;
                    res[1] = 0.0; res[2] = 0.0; res[3] = 0.0;
                } else {
                    // _find_penetration_segment (mpr.py:370-378)
                    const v3 dir = norm_vector(P.v[1]);
                    res[0] = vnorm(P.v[1]);
                    res[1] = dir.x; res[2] = dir.y; res[3] = dir.z;
                }
                res[4] = pos.x; res[5] = pos.y; res[6] = pos.z;
                continue;
            }
            d = norm_vector(d);
            phase = kMprV2;
        } else if (phase == kMprV2) {
            mpr_set(P, 2, w, sa, sb);
            if (dot(w, d) < kEps) {
                phase = kMprDone;
                continue;
            }
            // _search_direction_perpendicular_to_plane_containing_v012 (mpr.py:208-214)
            d = norm_vector(cross(P.v[1] - P.v[0], P.v[2] - P.v[0]));
            if (dot(d, P.v[0]) > 0.0) {
                const v3 tv = P.v[1], ta = P.a[1], tb = P.b[1];
                P.v[1] = P.v[2]; P.a[1] = P.a[2]; P.b[1] = P.b[2];
                P.v[2] = tv; P.a[2] = ta; P.b[2] = tb;
                d = d * -1.0;
            }
            it = 0;
            phase = kMprV3;
        } else if (phase == kMprV3) {
            mpr_set(P, 3, w, sa, sb);
            if (dot(w, d) < kEps) {
                phase = kMprDone;
                continue;
            }
            // _iterate_discover_portal (mpr.py:217-235)
            bool cont = false;
            if (dot(cross(P.v[1], P.v[3]), P.v[0]) < kEps) {
                P.v[2] = P.v[3]; P.a[2] = P.a[3]; P.b[2] = P.b[3];
                cont = true;
            }
            if (!cont && dot(cross(P.v[3], P.v[2]), P.v[0]) < kEps) {
                P.v[1] = P.v[3]; P.a[1] = P.a[3]; P.b[1] = P.b[3];
                cont = true;
            }
            if (cont) d = norm_vector(cross(P.v[1] - P.v[0], P.v[2] - P.v[0]));
            ++it;
            if (!cont || it >= max_iter) {
                it = 0;
                refine_round();
            }
        } else if (phase == kMprRefine) {
            if (!mpr_encapsulates(w, d) || mpr_reach_tol(P, w, d, tol)) {
                phase = kMprDone;  // separated
                continue;
            }
            mpr_expand(P, w, sa, sb);
            ++it;
            refine_round();
        } else {  // kMprPen
            if (mpr_reach_tol(P, w, d, tol) || it > max_iter) {
                // _penetration_info (mpr.py:349-358)
                v3 cp;
                const double depth = origin_to_triangle(P.v[1], P.v[2], P.v[3], cp);
                v3 dir = vsel(fabs(depth) < kEps, mk(0.0, 0.0, 0.0), cp);
                dir = norm_vector(dir);
                const v3 pos = mpr_contact(P, mpr_portal_dir(P));
                res[0] = depth;
                res[1] = dir.x; res[2] = dir.y; res[3] = dir.z;
                res[4] = pos.x; res[5] = pos.y; res[6] = pos.z;
                status = 1;
                phase = kMprDone;
                continue;
            }
            mpr_expand(P, w, sa, sb);
            ++it;
            d = mpr_portal_dir(P);
        }
    }
    if (!valid) return;
    if (!ok) status = kMprBadInput;
    status_out[pid] = status;
    if (out) {
        double *o = out + 7 * pid;
#pragma unroll
        for (int k = 0; k < 7; ++k) o[k] = res[k];
    }
}

}  // namespace d3d

using namespace d3d;

extern "C" int d3d_mpr_batch(const d3d_scene *scene, const int32_t *pair_a, const int32_t *pair_b,
                             const double *centers, int64_t n, double tolerance,
                             int32_t max_iterations, int32_t penetration, int32_t *status,
                             double *out, void *stream) {
    clear_error();
    if (int rc = check_scene(scene)) return rc;
    D3D_CHECK_ARG(n >= 0, "n must be >= 0");
    if (n == 0) return D3D_OK;
    D3D_CHECK_ARG(pair_a && pair_b && centers && status, "NULL buffer");
    D3D_CHECK_ARG(penetration == 0 || out, "out is NULL but penetration was requested");
    D3D_CHECK_ARG(max_iterations > 0, "max_iterations must be > 0");
    hipLaunchKernelGGL(mpr_kernel, dim3(grid_for(n, kMprBlock)), dim3(kMprBlock), 0,
                       (hipStream_t)stream, *scene, pair_a, pair_b, centers, n, tolerance,
                       max_iterations, penetration, kMprSerialMax, status, out);
    D3D_CHECK_LAUNCH("mpr_kernel");
    return D3D_OK;
}
