// d3d_runtime.hpp -- error plumbing shared by the C-ABI entry points.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdio>
#include <string>

#include "d3d.h"

namespace d3d {

void set_error(const std::string &msg);
void clear_error();

#define D3D_CHECK_ARG(cond, msg)            \
    do {                                    \
        if (!(cond)) {                      \
            ::d3d::set_error(msg);          \
            return D3D_E_INVALID;           \
        }                                   \
    } while (0)

#define D3D_CHECK_HIP(expr)                                                          \
    do {                                                                             \
        hipError_t e_ = (expr);                                                      \
        if (e_ != hipSuccess) {                                                      \
            ::d3d::set_error(std::string(#expr) + ": " + hipGetErrorString(e_));      \
            return D3D_E_HIP;                                                        \
        }                                                                            \
    } while (0)

// Checks a kernel launch (launch errors only; execution is asynchronous).
#define D3D_CHECK_LAUNCH(name)                                                       \
    do {                                                                             \
        hipError_t e_ = hipGetLastError();                                           \
        if (e_ != hipSuccess) {                                                      \
            ::d3d::set_error(std::string(name) + " launch: " + hipGetErrorString(e_)); \
            return D3D_E_HIP;                                                        \
        }                                                                            \
    } while (0)

inline int check_scene(const d3d_scene *s) {
    if (!s) {
        set_error("scene is NULL");
        return D3D_E_INVALID;
    }
    if (!s->shape_f64 || !s->shape_i32 || s->n_shapes <= 0) {
        set_error("scene has no collider table");
        return D3D_E_INVALID;
    }
    if (s->n_vertices > 0 && (!s->vx || !s->vy || !s->vz)) {
        set_error("scene has vertices but a NULL FP32 pool pointer");
        return D3D_E_INVALID;
    }
    if (s->n_vertices64 > 0 && !s->v64) {
        set_error("scene has FP64 vertices but a NULL v64 pointer");
        return D3D_E_INVALID;
    }
    if (s->n_clusters < 0 || (s->n_clusters > 0 && !s->clusters)) {
        set_error("scene has clusters but a NULL cluster pointer");
        return D3D_E_INVALID;
    }
    // the kernels read the int32 table, the FP32 pools, the original indices
    // and the cluster records with 16-byte loads
    const uintptr_t a16 = (uintptr_t)s->shape_i32 | (uintptr_t)s->vx | (uintptr_t)s->vy |
                          (uintptr_t)s->vz | (uintptr_t)s->vorig | (uintptr_t)s->clusters;
    const uintptr_t a8 = (uintptr_t)s->shape_f64 | (uintptr_t)s->v64;
    if ((a16 & 15) != 0 || (a8 & 7) != 0) {
        set_error("scene arrays must be 16-byte aligned (shape_f64 and v64: 8-byte)");
        return D3D_E_INVALID;
    }
    return D3D_OK;
}

inline unsigned grid_for(int64_t n, int block) { return (unsigned)((n + block - 1) / block); }

// EPA launcher shared by d3d_epa_batch and the fused GJK+EPA pipeline: with
// n_dev/work set, a persistent grid takes pair indices work[0 .. *n_dev).
// EPA over the pairs work[*work_off + i], i < *n_dev (device values; work_off
// may be NULL), or over pairs 0..n_max-1 when work is NULL.  With `next` (a
// device counter, zero at launch) wavefronts claim items from it one at a
// time; without it each takes a fixed stride of items.  faces / n_faces
// (optional) receive the final polytope of each pair.
int launch_epa(const d3d_scene &s, const int32_t *pa, const int32_t *pb, int64_t n_max,
               const int64_t *n_dev, const int32_t *work, const int32_t *work_off,
               const double *simplex, const int32_t *npts, int max_iter, int max_faces,
               double eps, double *out, int32_t *status, int32_t *iters,
               unsigned long long *next, hipStream_t stream, double *faces = nullptr,
               int32_t *n_faces = nullptr, int blocks_per_cu = 8);

// Device address of d3d_gjk_epa_batch_noalloc's scratch array on the current
// device (looked up once per device with hipGetSymbolAddress); out may be NULL.
int noalloc_scratch(void **out);

}  // namespace d3d
