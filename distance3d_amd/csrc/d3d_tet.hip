// Hydroelastic contact of tetrahedron pairs (hydroelastic_contact/
// _tetrahedron_intersection.py:7-423, _halfplanes.py:8-64): for each pair the
// contact plane (where the Young's-modulus-weighted pressure fields agree) and
// the contact polygon (that plane clipped by both tetrahedra).
//
// One lane per pair, FP64 throughout, in the order of the reference's numpy
// expressions.  The eight half-planes stay in registers (every loop over them
// is unrolled with compile-time indices); the up to 23 polygon points of a
// pair go to a per-pair workspace row, their angles and the sort order to
// LDS, so nothing is indexed dynamically in registers.
#include "d3d_atan2.hpp"
#include "d3d_device.hpp"
#include "d3d_runtime.hpp"

namespace d3d {

constexpr int kTetBlock = 64;
constexpr int kTetMaxPoly = 24;  // 3 x 8 half-planes, as the reference reserves
constexpr int kTetSame = 0x10;
constexpr int kTetCuts = 0x20;  // the plane cuts both tetrahedra
constexpr int kTetOverflow = 0x100;

// np.mean of a[0], a[st], ... as order_points computes it: the function is
// compiled by numba (_tetrahedron_intersection.py:299-318), whose np.mean adds
// the elements one after another from 0.0 and divides by the count
__device__ __forceinline__ double seq_mean(const double *a, int n, int st) {
    double r = 0.0;
    for (int i = 0; i < n; ++i) r += a[i * st];
    return r / n;
}

// plane_basis_from_normal (utils.py:78-122)
__device__ __forceinline__ void plane_basis(v3 n, v3 &x, v3 &y) {
    if (fabs(n.x) >= fabs(n.y)) {
        const double length = sqrt(n.x * n.x + n.z * n.z);
        x = mk(-n.z / length, 0.0, n.x / length);
        y = mk(n.y * x.z, n.z * x.x - n.x * x.z, -n.y * x.x);
    } else {
        const double length = sqrt(n.y * n.y + n.z * n.z);
        x = mk(0.0, n.z / length, -n.y / length);
        y = mk(n.y * x.z - n.z * x.y, -n.x * x.z, n.x * x.y);
    }
}

__global__ __launch_bounds__(kTetBlock) void tet_pair_kernel(
    const double *__restrict__ T1, const double *__restrict__ e1, const double *__restrict__ X1,
    const double *__restrict__ T2, const double *__restrict__ e2, const double *__restrict__ X2,
    const int32_t *__restrict__ pairs, int64_t m, double E1, double E2,
    int32_t *__restrict__ status_out, double *__restrict__ planes, int32_t *__restrict__ nv_out,
    double *__restrict__ polys, double *__restrict__ work) {
    __shared__ double ang_s[kTetMaxPoly][kTetBlock];
    __shared__ unsigned char ord_s[kTetMaxPoly][kTetBlock];
    const int lane = threadIdx.x;
    // grid-stride over the pairs; the 2-D polygon points of the lane's
    // current pair live in its own workspace row
    double *pts = work + 2 * kTetMaxPoly * ((int64_t)blockIdx.x * kTetBlock + lane);
    for (int64_t k = (int64_t)blockIdx.x * kTetBlock + lane; k < m;
         k += (int64_t)gridDim.x * kTetBlock) {
    const int64_t i = pairs[2 * k], j = pairs[2 * k + 1];
    const double *t1 = T1 + 12 * i, *t2 = T2 + 12 * j;
    const double *x1 = X1 + 16 * i, *x2 = X2 + 16 * j;
    const double *p1 = e1 + 4 * i, *p2 = e2 + 4 * j;
    // planes and polygons are optional outputs (NULL: status and vertex
    // count only)
    double *poly = polys ? polys + 3 * kTetMaxPoly * k : nullptr;
    int status = 0, nv = 0;

    // contact_plane (_tetrahedron_intersection.py:176-229)
    double w1[4], w2[4], h[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        w1[r] = p1[r] * E1;
        w2[r] = p2[r] * E2;
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const double a = ((w1[0] * x1[c] + w1[1] * x1[4 + c]) + w1[2] * x1[8 + c]) + w1[3] * x1[12 + c];
        const double b = ((w2[0] * x2[c] + w2[1] * x2[4 + c]) + w2[2] * x2[8 + c]) + w2[3] * x2[12 + c];
        h[c] = a - b;
    }
    const double norm = sqrt(h[0] * h[0] + h[1] * h[1] + h[2] * h[2]);
    bool same = norm == 0.0;
    if (!same) {
#pragma unroll
        for (int c = 0; c < 4; ++c) h[c] = h[c] / norm;
        h[3] = h[3] * -1.0;
        same = fabs(h[3]) < 10.0 * kEps;
    }
    if (same) {
        // _handle_same_tetrahedron (:142-159)
        double sum = 0.0;
#pragma unroll
        for (int r = 0; r < 4; ++r) sum += p2[r];
        double wt[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) wt[r] = p2[r] / sum;
        const v3 pp = mk(((wt[0] * t2[0] + wt[1] * t2[3]) + wt[2] * t2[6]) + wt[3] * t2[9],
                         ((wt[0] * t2[1] + wt[1] * t2[4]) + wt[2] * t2[7]) + wt[3] * t2[10],
                         ((wt[0] * t2[2] + wt[1] * t2[5]) + wt[2] * t2[8]) + wt[3] * t2[11]);
        const double d = sqrt(dot(pp, pp));
        const v3 n = d > 0.0 ? mk(pp.x / d, pp.y / d, pp.z / d) : mk(0.0, 0.0, 1.0);
        h[0] = n.x; h[1] = n.y; h[2] = n.z; h[3] = d;
        if (poly)
            for (int q = 0; q < 3; ++q) {
                poly[3 * q] = pp.x; poly[3 * q + 1] = pp.y; poly[3 * q + 2] = pp.z;
            }
        nv = 3;
        status = 1 | kTetSame;
    } else {
        const v3 n = mk(h[0], h[1], h[2]);
        const double d = h[3];
        // check_tetrahedra_intersect_contact_plane (:232-258), tolerance 1e-6
        bool crosses = true;
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            const double *T = t == 0 ? t1 : t2;
            double mn = 0.0, mx = 0.0;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const double dist = (T[3 * r] * n.x + T[3 * r + 1] * n.y + T[3 * r + 2] * n.z) - d;
                if (r == 0 || dist < mn) mn = dist;
                if (r == 0 || dist > mx) mx = dist;
            }
            if (!(mn < -1e-6 && mx > 1e-6)) crosses = false;
        }
        if (crosses) {
            status = kTetCuts;
            // compute_contact_polygon (:363-423), make_halfplanes (:261-297);
            // faces whose projection vanishes are dropped
            const v3 pp = n * d;
            v3 ax, ay;
            plane_basis(n, ax, ay);
            double hp[8][4];
#pragma unroll
            for (int g = 0; g < 8; ++g) hp[g][0] = hp[g][1] = hp[g][2] = hp[g][3] = 0.0;
            int nh = 0;
#pragma unroll
            for (int f = 0; f < 8; ++f) {
                const double *X = f < 4 ? x1 + 4 * f : x2 + 4 * (f - 4);
                const v3 fn = mk(X[0], X[1], X[2]);
                const double n0 = dot(fn, ax), n1 = dot(fn, ay);
                const double ds = -X[3] - dot(fn, pp);
                const double nn = sqrt(n0 * n0 + n1 * n1);
                const bool keep = nn > kEps;
                // row nh of the compacted list: written to every slot f' >= nh
                // it could land in, by selects (f' is a constant)
#pragma unroll
                for (int g = 0; g <= f; ++g) {
                    const bool w = keep && g == nh;
                    hp[g][0] = w ? n0 * ds / (nn * nn) : hp[g][0];
                    hp[g][1] = w ? n1 * ds / (nn * nn) : hp[g][1];
                    hp[g][2] = w ? n1 : hp[g][2];
                    hp[g][3] = w ? -n0 : hp[g][3];
                }
                nh += keep ? 1 : 0;
            }
            // intersect_halfplanes (_halfplanes.py:29-64)
            int cnt = 0;
            bool overflow = false;
#pragma unroll
            for (int a = 0; a < 8; ++a) {
#pragma unroll
                for (int b = a + 1; b < 8; ++b) {
                    if (b >= nh) continue;
                    const double denom = hp[a][2] * hp[b][3] - hp[a][3] * hp[b][2];
                    if (fabs(denom) < kEps) continue;
                    const double q0 = hp[b][0] - hp[a][0], q1 = hp[b][1] - hp[a][1];
                    const double t = (q0 * hp[b][3] - q1 * hp[b][2]) / denom;
                    const double px = hp[a][0] + hp[a][2] * t, py = hp[a][1] + hp[a][3] * t;
                    bool valid = true;
#pragma unroll
                    for (int c = 0; c < 8; ++c) {
                        if (c == a || c == b || c >= nh || !valid) continue;
                        const double r0 = px - hp[c][0], r1 = py - hp[c][1];
                        if (hp[c][2] * r1 - hp[c][3] * r0 < -kEps) valid = false;
                    }
                    if (valid) {
                        if (cnt + 1 >= 3 * nh) {
                            overflow = true;
                        } else {
                            pts[2 * cnt] = px;
                            pts[2 * cnt + 1] = py;
                            ++cnt;
                        }
                    }
                }
            }
            if (overflow) {
                status = kTetOverflow | kTetCuts;
            } else if (cnt >= 3) {
                // order_points (:300-318): by the angle about the mean point
                // (atan2_cr, correctly rounded), stable insertion sort
                const double cx = seq_mean(pts, cnt, 2), cy = seq_mean(pts + 1, cnt, 2);
                for (int q = 0; q < cnt; ++q) {
                    ang_s[q][lane] = atan2_cr(pts[2 * q + 1] - cy, pts[2 * q] - cx);
                    ord_s[q][lane] = (unsigned char)q;
                }
                for (int q = 1; q < cnt; ++q) {
                    const int o = ord_s[q][lane];
                    const double ao = ang_s[o][lane];
                    int r = q - 1;
                    while (r >= 0 && ang_s[ord_s[r][lane]][lane] > ao) {
                        ord_s[r + 1][lane] = ord_s[r][lane];
                        --r;
                    }
                    ord_s[r + 1][lane] = (unsigned char)o;
                }
                // filter_unique_points (:321-340): dropped within 10 eps of the
                // previous point in angular order; project_polygon_to_3d (:343-360)
                int nu = 0;
                for (int q = 0; q < cnt; ++q) {
                    const int o = ord_s[q][lane];
                    const double ux = pts[2 * o], uy = pts[2 * o + 1];
                    if (q > 0) {
                        const int pr = ord_s[q - 1][lane];
                        const double dx = ux - pts[2 * pr], dy = uy - pts[2 * pr + 1];
                        if (!(sqrt(dx * dx + dy * dy) > 10.0 * kEps)) continue;
                    }
                    if (poly) {
                        poly[3 * nu] = (ux * ax.x + uy * ay.x) + pp.x;
                        poly[3 * nu + 1] = (ux * ax.y + uy * ay.y) + pp.y;
                        poly[3 * nu + 2] = (ux * ax.z + uy * ay.z) + pp.z;
                    }
                    ++nu;
                }
                if (nu >= 3) {
                    nv = nu;
                    status = 1 | kTetCuts;
                }
            }
        }
    }
    if (planes) {
        double *plane = planes + 4 * k;
        plane[0] = h[0]; plane[1] = h[1]; plane[2] = h[2]; plane[3] = h[3];
    }
    if (poly)
        for (int q = 3 * nv; q < 3 * kTetMaxPoly; ++q) poly[q] = NAN;
    status_out[k] = status;
    nv_out[k] = nv;
    }
}

// LU factorisation with partial pivoting of a 4x4 matrix in the order of
// LAPACK's unblocked getf2 (np.linalg.solve reaches it through gesv): per
// column the first row of largest |a|, the row swap (by selects, so every
// index stays a compile-time constant and the matrix stays in registers),
// multipliers by the pivot's reciprocal (by division below DBL_MIN), the
// rank-1 update of the trailing block.  False when a pivot is zero.
__device__ __forceinline__ bool lu4(double (&A)[4][4], int (&piv)[4]) {
    bool ok = true;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        int p = j;
        double best = fabs(A[j][j]);
#pragma unroll
        for (int i = j + 1; i < 4; ++i) {
            const double v = fabs(A[i][j]);
            const bool t = v > best;
            best = t ? v : best;
            p = t ? i : p;
        }
        piv[j] = p;
#pragma unroll
        for (int i = j + 1; i < 4; ++i) {
            const bool sw = p == i;
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const double t = A[j][c];
                A[j][c] = sw ? A[i][c] : t;
                A[i][c] = sw ? t : A[i][c];
            }
        }
        const double d = A[j][j];
        if (d != 0.0) {
            const bool recip = fabs(d) >= 2.2250738585072014e-308;
            const double r = 1.0 / d;
#pragma unroll
            for (int i = j + 1; i < 4; ++i) A[i][j] = recip ? A[i][j] * r : A[i][j] / d;
        } else {
            ok = false;
        }
#pragma unroll
        for (int i = j + 1; i < 4; ++i)
#pragma unroll
            for (int c = j + 1; c < 4; ++c) A[i][c] = A[i][c] - A[i][j] * A[j][c];
    }
    return ok;
}

// getrs for one right-hand side: the row swaps in order, the unit lower and
// the upper triangular solves (column-oriented; a zero entry skips its
// column, as the reference triangular solve does).
__device__ __forceinline__ void lu4_solve(const double (&A)[4][4], const int (&piv)[4],
                                          double (&b)[4]) {
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int i = j + 1; i < 4; ++i) {
            const bool sw = piv[j] == i;
            const double t = b[j];
            b[j] = sw ? b[i] : t;
            b[i] = sw ? t : b[i];
        }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const bool nz = b[k] != 0.0;
#pragma unroll
        for (int i = k + 1; i < 4; ++i) b[i] = nz ? b[i] - b[k] * A[i][k] : b[i];
    }
#pragma unroll
    for (int k = 3; k >= 0; --k) {
        const bool nz = b[k] != 0.0;
        b[k] = nz ? b[k] / A[k][k] : b[k];
#pragma unroll
        for (int i = 0; i < k; ++i) b[i] = nz ? b[i] - b[k] * A[i][k] : b[i];
    }
}

// [T^T; 1 1 1 1] of a tetrahedron (T: 4 points, row-major): its inverse maps
// (x, 1) to barycentric coordinates (_barycentric_transform.py:4-9,
// _forces.py:94).
__device__ __forceinline__ void tet_matrix(const double *__restrict__ T, double (&A)[4][4]) {
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        A[0][c] = T[3 * c];
        A[1][c] = T[3 * c + 1];
        A[2][c] = T[3 * c + 2];
        A[3][c] = 1.0;
    }
}

// Per tetrahedron: AABB [axis][min, max] (_mesh_processing.py:4-20, NaN
// propagates as np.min/np.max do) and the barycentric transform, the inverse
// of [T^T; 1] (the reference's pseudo-inverse equals it for a non-degenerate
// tetrahedron); flags = 1 for a zero pivot.
__global__ __launch_bounds__(kTetBlock) void tet_prep_kernel(const double *__restrict__ T, int64_t n,
                                                             double *__restrict__ aabbs,
                                                             double *__restrict__ X,
                                                             int32_t *__restrict__ flags) {
    const int64_t k = (int64_t)blockIdx.x * kTetBlock + threadIdx.x;
    if (k >= n) return;
    const double *t = T + 12 * k;
    if (aabbs) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            double mn = t[a], mx = t[a];
#pragma unroll
            for (int r = 1; r < 4; ++r) {
                const double v = t[3 * r + a];
                mn = (v < mn || v != v) ? v : mn;
                mx = (v > mx || v != v) ? v : mx;
            }
            aabbs[6 * k + 2 * a] = mn;
            aabbs[6 * k + 2 * a + 1] = mx;
        }
    }
    if (X || flags) {
        double A[4][4];
        int piv[4];
        tet_matrix(t, A);
        const bool ok = lu4(A, piv);
        if (flags) flags[k] = ok ? 0 : 1;
        if (X) {
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                double b[4] = {0.0, 0.0, 0.0, 0.0};
                b[c] = 1.0;
                lu4_solve(A, piv, b);
#pragma unroll
                for (int r = 0; r < 4; ++r) X[16 * k + 4 * r + c] = b[r];
            }
        }
    }
}

// compute_contact_force (_forces.py:52-111) for each pair with status bit 0
// (zero rows for the others), from the first body's tetrahedron and
// potentials: the polygon as the fan (0, q, q + 1), at most the reference's
// six triangles (its TRIANGLES table is the fan of an 8-gon, _forces.py:
// 47-48); per triangle the pressure at the centroid (barycentric coordinates
// by the LU solve, weighted by potentials x Young's modulus) times the area.
// out row: centre (3), force (3), area.
__global__ __launch_bounds__(kTetBlock) void tet_force_kernel(
    const double *__restrict__ T1, const double *__restrict__ e1, const int32_t *__restrict__ pairs,
    const int32_t *__restrict__ status, const double *__restrict__ planes,
    const int32_t *__restrict__ n_vertices, const double *__restrict__ polys, int64_t m, double E1,
    double *__restrict__ out) {
    const int64_t k = (int64_t)blockIdx.x * kTetBlock + threadIdx.x;
    if (k >= m) return;
    double *o = out + 7 * k;
    if (!(status[k] & 1)) {
#pragma unroll
        for (int c = 0; c < 7; ++c) o[c] = 0.0;
        return;
    }
    const int64_t i = pairs[2 * k];
    const double *p1 = e1 + 4 * i;
    const double *poly = polys + 3 * kTetMaxPoly * k;
    double A[4][4];
    int piv[4];
    tet_matrix(T1 + 12 * i, A);
    lu4(A, piv);
    double we[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) we[r] = p1[r] * E1;
    const int ntri = min(n_vertices[k] - 2, 6);
    const v3 v0 = mk(poly[0], poly[1], poly[2]);
    double tf = 0.0, ta = 0.0;
    v3 com = mk(0.0, 0.0, 0.0);
    for (int q = 1; q <= ntri; ++q) {
        const v3 v1 = mk(poly[3 * q], poly[3 * q + 1], poly[3 * q + 2]);
        const v3 v2 = mk(poly[3 * q + 3], poly[3 * q + 4], poly[3 * q + 5]);
        const v3 c3 = mk(((v0.x + v1.x) + v2.x) / 3.0, ((v0.y + v1.y) + v2.y) / 3.0,
                         ((v0.z + v1.z) + v2.z) / 3.0);
        double b[4] = {c3.x, c3.y, c3.z, 1.0};
        lu4_solve(A, piv, b);
        double pr = 0.0;
#pragma unroll
        for (int r = 0; r < 4; ++r) pr += b[r] * we[r];
        const v3 cr = cross(v1 - v0, v2 - v0);
        const double area = 0.5 * sqrt(cr.x * cr.x + cr.y * cr.y + cr.z * cr.z);
        tf += pr * area;
        ta += area;
        com = mk(com.x + area * c3.x, com.y + area * c3.y, com.z + area * c3.z);
    }
    if (ta > 0.0) {
        o[0] = com.x / ta; o[1] = com.y / ta; o[2] = com.z / ta;
    } else {
        o[0] = v0.x; o[1] = v0.y; o[2] = v0.z;
    }
    const double *pl = planes + 4 * k;
    o[3] = tf * pl[0]; o[4] = tf * pl[1]; o[5] = tf * pl[2];
    o[6] = ta;
}

// intersect_halfplanes (_halfplanes.py:29-64) for one set of n half-planes
// (rows p, pq) by one wavefront: pairs (i, j > i) in the reference's loop
// order, lanes taking consecutive j; valid points land in that order (ballot
// ranks).  count = the number of points, or -1 where the reference asserts
// (a valid point past the 3n - 1 it has room for).
__global__ __launch_bounds__(kWave) void halfplane_kernel(const double *__restrict__ hp, int n,
                                                         double *__restrict__ pts,
                                                         int32_t *__restrict__ count) {
    const int lane = threadIdx.x;
    const unsigned long long lt = lane == 0 ? 0ull : (~0ull >> (kWave - lane));
    int cnt = 0;
    bool overflow = false;
    for (int i = 0; i < n && !overflow; ++i) {
        const double *a = hp + 4 * i;
        for (int j0 = i + 1; j0 < n && !overflow; j0 += kWave) {
            const int j = j0 + lane;
            bool valid = false;
            double px = 0.0, py = 0.0;
            if (j < n) {
                const double *b = hp + 4 * j;
                const double denom = a[2] * b[3] - a[3] * b[2];
                if (!(fabs(denom) < kEps)) {
                    const double q0 = b[0] - a[0], q1 = b[1] - a[1];
                    const double t = (q0 * b[3] - q1 * b[2]) / denom;
                    px = a[0] + a[2] * t;
                    py = a[1] + a[3] * t;
                    valid = true;
                    for (int k = 0; k < n && valid; ++k) {
                        if (k == i || k == j) continue;
                        const double *c = hp + 4 * k;
                        const double r0 = px - c[0], r1 = py - c[1];
                        if (c[2] * r1 - c[3] * r0 < -kEps) valid = false;
                    }
                }
            }
            const unsigned long long m = __ballot(valid);
            const int pos = cnt + __popcll(m & lt);
            if (valid && pos + 1 < 3 * n) {
                pts[2 * pos] = px;
                pts[2 * pos + 1] = py;
            }
            overflow = __ballot(valid && pos + 1 >= 3 * n) != 0ull;
            cnt += __popcll(m);
        }
    }
    if (lane == 0) *count = overflow ? -1 : cnt;
}

__global__ __launch_bounds__(kTetBlock) void angle_kernel(const double *__restrict__ y,
                                                          const double *__restrict__ x, int64_t n,
                                                          double *__restrict__ out) {
    const int64_t k = (int64_t)blockIdx.x * kTetBlock + threadIdx.x;
    if (k < n) out[k] = atan2_cr(y[k], x[k]);
}

}  // namespace d3d

using namespace d3d;

// blocks of the grid-stride pair kernel (4096 x 64 lanes: more than the
// chip holds at the kernel's occupancy); the workspace has one row per lane
constexpr int64_t kTetMaxBlocks = 4096;

extern "C" int64_t d3d_tet_pairs_workspace_bytes(int64_t m) {
    if (m < 0) return -1;
    const int64_t lanes = min((int64_t)grid_for(m, kTetBlock), kTetMaxBlocks) * kTetBlock;
    return lanes * 2 * kTetMaxPoly * (int64_t)sizeof(double);
}

extern "C" int d3d_tet_pairs_batch(const double *T1, const double *e1, const double *X1,
                                   const double *T2, const double *e2, const double *X2,
                                   const int32_t *pairs, int64_t m, double youngs_modulus1,
                                   double youngs_modulus2, int32_t *status, double *planes,
                                   int32_t *n_vertices, double *polygons, void *workspace,
                                   int64_t workspace_bytes, void *stream) {
    clear_error();
    D3D_CHECK_ARG(m >= 0, "m must be >= 0");
    if (m == 0) return D3D_OK;
    D3D_CHECK_ARG(T1 && e1 && X1 && T2 && e2 && X2 && pairs && status && n_vertices && workspace,
                  "NULL buffer");
    D3D_CHECK_ARG(workspace_bytes >= d3d_tet_pairs_workspace_bytes(m), "workspace too small");
    const unsigned blocks = (unsigned)min((int64_t)grid_for(m, kTetBlock), kTetMaxBlocks);
    hipLaunchKernelGGL(tet_pair_kernel, dim3(blocks), dim3(kTetBlock), 0,
                       (hipStream_t)stream, T1, e1, X1, T2, e2, X2, pairs, m, youngs_modulus1,
                       youngs_modulus2, status, planes, n_vertices, polygons, (double *)workspace);
    D3D_CHECK_LAUNCH("tet_pair_kernel");
    return D3D_OK;
}

extern "C" int d3d_tet_mesh_prep(const double *T, int64_t n, double *aabbs, double *X,
                                 int32_t *flags, void *stream) {
    clear_error();
    D3D_CHECK_ARG(n >= 0, "n must be >= 0");
    if (n == 0) return D3D_OK;
    D3D_CHECK_ARG(T && (aabbs || X || flags), "NULL buffer");
    hipLaunchKernelGGL(tet_prep_kernel, dim3(grid_for(n, kTetBlock)), dim3(kTetBlock), 0,
                       (hipStream_t)stream, T, n, aabbs, X, flags);
    D3D_CHECK_LAUNCH("tet_prep_kernel");
    return D3D_OK;
}

extern "C" int d3d_tet_contact_forces(const double *T1, const double *e1, const int32_t *pairs,
                                      const int32_t *status, const double *planes,
                                      const int32_t *n_vertices, const double *polygons,
                                      int64_t m, double youngs_modulus1, double *out,
                                      void *stream) {
    clear_error();
    D3D_CHECK_ARG(m >= 0, "m must be >= 0");
    if (m == 0) return D3D_OK;
    D3D_CHECK_ARG(T1 && e1 && pairs && status && planes && n_vertices && polygons && out,
                  "NULL buffer");
    hipLaunchKernelGGL(tet_force_kernel, dim3(grid_for(m, kTetBlock)), dim3(kTetBlock), 0,
                       (hipStream_t)stream, T1, e1, pairs, status, planes, n_vertices, polygons, m,
                       youngs_modulus1, out);
    D3D_CHECK_LAUNCH("tet_force_kernel");
    return D3D_OK;
}

extern "C" int d3d_polygon_angles(const double *y, const double *x, int64_t n, double *out,
                                  void *stream) {
    clear_error();
    D3D_CHECK_ARG(n >= 0, "n must be >= 0");
    if (n == 0) return D3D_OK;
    D3D_CHECK_ARG(y && x && out, "NULL buffer");
    hipLaunchKernelGGL(angle_kernel, dim3(grid_for(n, kTetBlock)), dim3(kTetBlock), 0,
                       (hipStream_t)stream, y, x, n, out);
    D3D_CHECK_LAUNCH("angle_kernel");
    return D3D_OK;
}

extern "C" int d3d_intersect_halfplanes(const double *halfplanes, int32_t n, double *points,
                                        int32_t *count, void *stream) {
    clear_error();
    // one wavefront does O(n^3) work (every pair against every half-plane);
    // the reference calls it with at most 8 half-planes
    D3D_CHECK_ARG(n >= 0 && n <= D3D_MAX_HALFPLANES, "n must be in [0, D3D_MAX_HALFPLANES = 1024]");
    D3D_CHECK_ARG(count && (n == 0 || (halfplanes && points)), "NULL buffer");
    hipLaunchKernelGGL(halfplane_kernel, dim3(1), dim3(kWave), 0, (hipStream_t)stream, halfplanes,
                       n, points, count);
    D3D_CHECK_LAUNCH("halfplane_kernel");
    return D3D_OK;
}
