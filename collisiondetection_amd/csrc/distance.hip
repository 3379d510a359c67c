// Proximity on the GPU: the Distance statics in batches (include/Distance.h:14-174) and
// Distance::meshSelfDistance (src/Distance.cpp:12-66).
//
// meshSelfDistance runs in three steps, as the reference does:
//   1. the smallest squared distance between a vertex i and a vertex m of a face not containing i
//      (Distance.cpp:21-34).  The reference loops over vertices x faces (O(V F)).  Here: an upper bound
//      r0 from each face's own vertex pairs (in a manifold mesh every edge is such a pair, so r0 is at
//      most the shortest edge); then a uniform grid of cells at least r0 wide, vertices sorted by cell,
//      and each vertex compares itself only with the vertices of its 27 surrounding cells -- any pair
//      closer than r0 lies there -- so the pass is O(V) for meshes whose vertices are not piled up.
//      The face condition (some face of m avoids i) is checked from the vertex->face table only when m
//      would improve the lane's minimum.  Squared norms are exact FP64 values in the reference's
//      order, so the minimum is the reference's bit for bit.  Meshes without such face pairs (no face
//      shares a vertex with another) keep the all-pairs sweep in LDS tiles;
//   2. AABBBroadPhase candidates at that radius (the broadphase's 3-axis mode);
//   3. the minimum, over the candidate stencils, of the norm of Distance::vertexFaceDistance /
//      edgeEdgeDistance (Distance.cpp:51-64).  Minima are reduced with 64-bit atomics on the bits of
//      non-negative doubles, whose integer order is their numeric order.
#include <cmath>
#include <cstring>
#include <vector>

#include "ccd_core.h"
#include "ccd_internal.h"

namespace {

using ccd::V3;

inline unsigned blocks_for(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

__device__ __forceinline__ V3 ld3(const double *__restrict__ p, int64_t i)
{
    return ccd::mk(p[3 * i], p[3 * i + 1], p[3 * i + 2]);
}

__global__ __launch_bounds__(256) void k_distance_primitives(int kind, const double *__restrict__ q, double eta,
                                                             int64_t n, uint8_t *__restrict__ flag,
                                                             double *__restrict__ vec, double *__restrict__ bary)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (kind == CCD_DIST_VERTEX_PLANE_LT || kind == CCD_DIST_LINE_LINE_LT) {
        const double *p = q + 12 * i;
        V3 a = ld3(p, 0), b = ld3(p, 1), c = ld3(p, 2), d = ld3(p, 3);
        bool r = kind == CCD_DIST_VERTEX_PLANE_LT ? ccd::vertex_plane_distance_less_than(a, b, c, d, eta)
                                                  : ccd::line_line_distance_less_than(a, b, c, d, eta);
        flag[i] = r ? 1 : 0;
        return;
    }
    const double *p = q + 12 * i;
    V3 a = ld3(p, 0), b = ld3(p, 1), c = ld3(p, 2), d = ld3(p, 3);
    double w[4] = {0, 0, 0, 0};
    V3 r = kind == CCD_DIST_VERTEX_FACE ? ccd::vertex_face_distance(a, b, c, d, w[0], w[1], w[2])
                                        : ccd::edge_edge_distance(a, b, c, d, w[0], w[1], w[2], w[3]);
    if (vec) {
        vec[3 * i] = r.x;
        vec[3 * i + 1] = r.y;
        vec[3 * i + 2] = r.z;
    }
    if (bary)
        for (int k = 0; k < 4; k++) bary[4 * i + k] = w[k];
    if (flag) flag[i] = 1;
}

__device__ __forceinline__ void atomic_min_nonneg(unsigned long long *bits, double v)
{
    if (!(v >= 0)) return;   // negative or NaN never occurs for squared norms and norms
    unsigned long long b = (unsigned long long)__double_as_longlong(v);
    atomicMin(bits, b);
}

// Some face of m avoids vertex i (Distance.cpp:25-26, Mesh::vertexOfFace).
__device__ __forceinline__ bool has_face_avoiding(const int32_t *__restrict__ faces, const int32_t *__restrict__ voff,
                                                  const int32_t *__restrict__ vlist, int m, int i)
{
    for (int k = voff[m]; k < voff[m + 1]; k++) {
        int g = vlist[k];
        if (faces[3 * g] != i && faces[3 * g + 1] != i && faces[3 * g + 2] != i) return true;
    }
    return false;
}

// Step 1: block (x = tile of 256 vertices i, y = slice of the vertices m).
__global__ __launch_bounds__(256) void k_closest_vertex_pair(const double *__restrict__ v, int64_t V, int64_t mslice,
                                                             const int32_t *__restrict__ faces,
                                                             const int32_t *__restrict__ voff,
                                                             const int32_t *__restrict__ vlist,
                                                             unsigned long long *best_bits)
{
    __shared__ double tile[3 * 256];
    __shared__ double red[256];
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    bool vi = i < V;
    V3 pi = vi ? ld3(v, i) : ccd::mk(0, 0, 0);
    double best = __builtin_inf();
    int64_t m0 = blockIdx.y * mslice;
    int64_t m1 = m0 + mslice < V ? m0 + mslice : V;
    for (int64_t base = m0; base < m1; base += 256) {
        int64_t m = base + threadIdx.x;
        if (m < m1) {
            tile[3 * threadIdx.x] = v[3 * m];
            tile[3 * threadIdx.x + 1] = v[3 * m + 1];
            tile[3 * threadIdx.x + 2] = v[3 * m + 2];
        }
        __syncthreads();
        int cnt = (int)(m1 - base < 256 ? m1 - base : 256);
        if (vi) {
            for (int k = 0; k < cnt; k++) {
                V3 d = pi - ccd::mk(tile[3 * k], tile[3 * k + 1], tile[3 * k + 2]);
                double dd = ccd::dot(d, d);
                if (dd < best) {
                    int mm = (int)(base + k);
                    if (mm != (int)i && has_face_avoiding(faces, voff, vlist, mm, (int)i)) best = dd;
                }
            }
        }
        __syncthreads();
    }
    red[threadIdx.x] = best;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w && red[threadIdx.x + w] < red[threadIdx.x]) red[threadIdx.x] = red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0 && red[0] < __builtin_inf()) atomic_min_nonneg(best_bits, red[0]);
}

// Block minimum of `best` into best_bits (every thread of the block calls it).
__device__ __forceinline__ void block_min_into(double best, unsigned long long *best_bits)
{
    __shared__ double red[256];
    red[threadIdx.x] = best;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w && red[threadIdx.x + w] < red[threadIdx.x]) red[threadIdx.x] = red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0 && red[0] < __builtin_inf()) atomic_min_nonneg(best_bits, red[0]);
}

// Step 1a: the pairs (i, m) of each face's own vertices for which some face of m avoids i.
__global__ __launch_bounds__(256) void k_face_pair_bound(const double *__restrict__ v, const int32_t *__restrict__ faces,
                                                         int64_t F, const int32_t *__restrict__ voff,
                                                         const int32_t *__restrict__ vlist,
                                                         unsigned long long *best_bits)
{
    double best = __builtin_inf();
    for (int64_t f = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; f < F; f += (int64_t)gridDim.x * blockDim.x) {
        const int c[3] = {faces[3 * f], faces[3 * f + 1], faces[3 * f + 2]};
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
            for (int b = 0; b < 3; b++) {
                if (a == b || c[a] == c[b]) continue;
                V3 d = ld3(v, c[a]) - ld3(v, c[b]);
                double dd = ccd::dot(d, d);
                if (dd < best && has_face_avoiding(faces, voff, vlist, c[b], c[a])) best = dd;
            }
    }
    block_min_into(best, best_bits);
}

// Step 1b: cell keys.  Cell coordinates (21 bits each, clamped) of (x - lo) * inv, packed z, y, x so that
// the three cells x-1, x, x+1 of one (y, z) row are consecutive keys.
constexpr uint64_t CELL_MAX = (1ull << 21) - 1;
__device__ __forceinline__ uint64_t cell_coord(double x, double lo, double inv)
{
    double s = (x - lo) * inv;
    if (!(s > 0.0)) return 0;
    if (s >= (double)CELL_MAX) return CELL_MAX;
    return (uint64_t)s;
}
__global__ __launch_bounds__(256) void k_cell_keys(const double *__restrict__ v, int64_t V, double lx, double ly,
                                                   double lz, double inv, uint64_t *__restrict__ key,
                                                   int32_t *__restrict__ idx)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= V) return;
    uint64_t cx = cell_coord(v[3 * i], lx, inv), cy = cell_coord(v[3 * i + 1], ly, inv),
             cz = cell_coord(v[3 * i + 2], lz, inv);
    key[i] = (cz << 42) | (cy << 21) | cx;
    idx[i] = (int32_t)i;
}

// Step 1c: each vertex against the vertices of its 27 surrounding cells (9 rows of 3 consecutive keys).
__global__ __launch_bounds__(256) void k_grid_closest(const double *__restrict__ v, int64_t V, double lx, double ly,
                                                      double lz, double inv, const uint64_t *__restrict__ key,
                                                      const int32_t *__restrict__ idx,
                                                      const int32_t *__restrict__ faces,
                                                      const int32_t *__restrict__ voff,
                                                      const int32_t *__restrict__ vlist,
                                                      unsigned long long *best_bits)
{
    double best = __builtin_inf();
    const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (j < V) {
        const int i = idx[j];   // sorted order: neighbouring lanes look at the same cells
        const V3 pi = ld3(v, i);
        const int64_t cx = (int64_t)cell_coord(pi.x, lx, inv), cy = (int64_t)cell_coord(pi.y, ly, inv),
                      cz = (int64_t)cell_coord(pi.z, lz, inv);
        for (int dz = -1; dz <= 1; dz++)
            for (int dy = -1; dy <= 1; dy++) {
                const int64_t y = cy + dy, z = cz + dz;
                if (y < 0 || z < 0 || y > (int64_t)CELL_MAX || z > (int64_t)CELL_MAX) continue;
                const uint64_t row = ((uint64_t)z << 42) | ((uint64_t)y << 21);
                const uint64_t k0 = row | (uint64_t)(cx > 0 ? cx - 1 : 0);
                const uint64_t k1 = row | (uint64_t)(cx < (int64_t)CELL_MAX ? cx + 1 : CELL_MAX);
                int64_t lo = 0, hi = V;   // first key >= k0
                while (lo < hi) {
                    const int64_t mid = lo + (hi - lo) / 2;
                    if (key[mid] < k0)
                        lo = mid + 1;
                    else
                        hi = mid;
                }
                for (int64_t t = lo; t < V && key[t] <= k1; t++) {
                    const int m = idx[t];
                    if (m == i) continue;
                    V3 d = pi - ld3(v, m);
                    double dd = ccd::dot(d, d);
                    if (dd < best && has_face_avoiding(faces, voff, vlist, m, i)) best = dd;
                }
            }
    }
    block_min_into(best, best_bits);
}

// Step 3: minimum distance over the candidate stencils; the stencil count is read on the device.
template <bool VF>
__global__ __launch_bounds__(256) void k_stencil_min_distance(const double *__restrict__ v, const int4 *__restrict__ keys,
                                                              const int64_t *__restrict__ n_dev, int64_t n_max,
                                                              unsigned long long *best_bits)
{
    __shared__ double red[256];
    int64_t n = *n_dev < n_max ? *n_dev : n_max;   // never past the stored keys
    double best = __builtin_inf();
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        int4 k = keys[i];
        double b0, b1, b2, b3;
        V3 r = VF ? ccd::vertex_face_distance(ld3(v, k.x), ld3(v, k.y), ld3(v, k.z), ld3(v, k.w), b0, b1, b2)
                  : ccd::edge_edge_distance(ld3(v, k.x), ld3(v, k.y), ld3(v, k.z), ld3(v, k.w), b0, b1, b2, b3);
        double d = ccd::norm(r);
        if (d < best) best = d;
    }
    red[threadIdx.x] = best;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w && red[threadIdx.x + w] < red[threadIdx.x]) red[threadIdx.x] = red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0 && red[0] < __builtin_inf()) atomic_min_nonneg(best_bits, red[0]);
}

__global__ void k_set_u64(unsigned long long *p, unsigned long long v) { *p = v; }

}   // namespace

namespace ccdi {

int launch_distance_primitives(ccd_ctx *ctx, int kind, const double *q, double eta, int64_t n, uint8_t *flag,
                               double *vec, double *bary)
{
    if (n == 0) return CCD_OK;
    hipLaunchKernelGGL(k_distance_primitives, dim3(blocks_for(n, 256)), dim3(256), 0, ctx->stream, kind, q, eta, n,
                       flag, vec, bary);
    CCD_LAUNCH_CHECK(ctx);
    return CCD_OK;
}

int mesh_self_distance(ccd_ctx *ctx, const double *verts, int64_t V, const int32_t *faces, int64_t F,
                       const uint8_t *fixed, double *out, int64_t *nvf, int64_t *nee)
{
    hipStream_t st = ctx->stream;
    unsigned long long *best = (unsigned long long *)buffer(ctx, "selfdist_best", 16);
    if (!best) return CCD_E_OOM;
    const unsigned long long inf_bits = 0x7ff0000000000000ull;
    int rc = build_vertex_face_table(ctx, faces, F, V);
    if (rc) return rc;
    const int32_t *voff = nullptr, *vlist = nullptr;
    vertex_face_table(ctx, &voff, &vlist);
    hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(1), 0, st, best, inf_bits);
    CCD_LAUNCH_CHECK(ctx);
    // Step 1a: the bound r0 from each face's own vertex pairs, and the vertices' bounding box.
    double r02 = __builtin_inf();
    double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    if (V > 0 && F > 0) {
        unsigned g = (unsigned)std::min<int64_t>(blocks_for(F, 256), 4096);
        hipLaunchKernelGGL(k_face_pair_bound, dim3(g), dim3(256), 0, st, verts, faces, F, voff, vlist, best);
        CCD_LAUNCH_CHECK(ctx);
        std::vector<double> hv((size_t)3 * V);
        CCD_HIP_CHECK(ctx, hipMemcpyAsync(hv.data(), verts, (size_t)V * 24, hipMemcpyDeviceToHost, st));
        unsigned long long b0 = inf_bits;
        CCD_HIP_CHECK(ctx, hipMemcpyAsync(&b0, best, 8, hipMemcpyDeviceToHost, st));
        CCD_HIP_CHECK(ctx, hipStreamSynchronize(st));
        memcpy(&r02, &b0, 8);
        for (int k = 0; k < 3; k++) lo[k] = hi[k] = hv[k];
        for (int64_t i = 1; i < V; i++)
            for (int k = 0; k < 3; k++) {
                lo[k] = std::min(lo[k], hv[3 * i + k]);
                hi[k] = std::max(hi[k], hv[3 * i + k]);
            }
    }
    // Cells at least r0 wide (with a relative margin that covers the rounding of the cell coordinates),
    // and never more than 2^20 per axis.
    const double extent = std::max(hi[0] - lo[0], std::max(hi[1] - lo[1], hi[2] - lo[2]));
    double cell = std::sqrt(r02) * (1.0 + 1e-6);
    cell = std::max(cell, extent / (double)(1 << 20));
    const bool grid = V > 0 && F > 0 && r02 > 0.0 && std::isfinite(r02) && std::isfinite(extent) && cell > 0.0;
    if (grid) {
        // Step 1b/1c: vertices sorted by cell, each against its 27 surrounding cells.
        const double inv = 1.0 / cell;
        uint64_t *key = (uint64_t *)buffer(ctx, "grid_key", (size_t)V * 8);
        uint64_t *key2 = (uint64_t *)buffer(ctx, "grid_key2", (size_t)V * 8);
        int32_t *idx = (int32_t *)buffer(ctx, "grid_idx", (size_t)V * 4);
        int32_t *idx2 = (int32_t *)buffer(ctx, "grid_idx2", (size_t)V * 4);
        void *tmp = buffer(ctx, "grid_tmp", radix_sort_temp_bytes(V));
        if (!key || !key2 || !idx || !idx2 || !tmp) return CCD_E_OOM;
        hipLaunchKernelGGL(k_cell_keys, dim3(blocks_for(V, 256)), dim3(256), 0, st, verts, V, lo[0], lo[1], lo[2], inv,
                           key, idx);
        CCD_LAUNCH_CHECK(ctx);
        bool in_alt = false;
        rc = radix_sort_pairs(ctx, key, idx, key2, idx2, V, 64, tmp, in_alt);
        if (rc) return rc;
        hipLaunchKernelGGL(k_grid_closest, dim3(blocks_for(V, 256)), dim3(256), 0, st, verts, V, lo[0], lo[1], lo[2],
                           inv, in_alt ? key2 : key, in_alt ? idx2 : idx, faces, voff, vlist, best);
        CCD_LAUNCH_CHECK(ctx);
    } else if (V > 0 && F > 0 && r02 != 0.0) {
        int64_t slices = std::max<int64_t>(1, std::min<int64_t>(64, (int64_t)8192 / std::max<int64_t>(1, blocks_for(V, 256))));
        int64_t mslice = (V + slices - 1) / slices;
        mslice = ((mslice + 255) / 256) * 256;
        slices = (V + mslice - 1) / mslice;
        hipLaunchKernelGGL(k_closest_vertex_pair, dim3(blocks_for(V, 256), (unsigned)slices), dim3(256), 0, st, verts, V,
                           mslice, faces, voff, vlist, best);
        CCD_LAUNCH_CHECK(ctx);
    }
    unsigned long long hb = inf_bits;
    CCD_HIP_CHECK(ctx, hipMemcpyAsync(&hb, best, 8, hipMemcpyDeviceToHost, st));
    CCD_HIP_CHECK(ctx, hipStreamSynchronize(st));
    double closest2;
    memcpy(&closest2, &hb, 8);
    double closest = ccd::dsqrt(closest2);   // Distance.cpp:35 (sqrt of +inf is +inf)

    Candidates c;
    rc = broadphase(ctx, faces, F, verts, verts, nullptr, nullptr, V, closest, fixed, c, true);
    if (rc) return rc;
    if (nvf) *nvf = c.nvf;
    if (nee) *nee = c.nee;
    hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(1), 0, st, best + 1, inf_bits);
    CCD_LAUNCH_CHECK(ctx);
    if (c.nvf) {
        unsigned g = (unsigned)std::min<int64_t>(blocks_for(c.nvf, 256), 4096);
        hipLaunchKernelGGL(k_stencil_min_distance<true>, dim3(g), dim3(256), 0, st, verts, (const int4 *)c.vf,
                           ctx->d_counts + C_VF, c.nvf, best + 1);
        CCD_LAUNCH_CHECK(ctx);
    }
    if (c.nee) {
        unsigned g = (unsigned)std::min<int64_t>(blocks_for(c.nee, 256), 4096);
        hipLaunchKernelGGL(k_stencil_min_distance<false>, dim3(g), dim3(256), 0, st, verts, (const int4 *)c.ee,
                           ctx->d_counts + C_EE, c.nee, best + 1);
        CCD_LAUNCH_CHECK(ctx);
    }
    CCD_HIP_CHECK(ctx, hipMemcpyAsync(&hb, best + 1, 8, hipMemcpyDeviceToHost, st));
    CCD_HIP_CHECK(ctx, hipStreamSynchronize(st));
    memcpy(out, &hb, 8);
    return CCD_OK;
}

}   // namespace ccdi
