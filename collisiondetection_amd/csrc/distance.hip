// Proximity on the GPU: the Distance statics in batches (include/Distance.h:14-174) and
// Distance::meshSelfDistance (src/Distance.cpp:12-66).
//
// meshSelfDistance runs in three steps, as the reference does:
//   1. the smallest squared distance between a vertex i and a vertex m of a face not containing i
//      (Distance.cpp:21-34).  The reference loops over vertices x faces; here one lane per vertex i
//      sweeps all vertices m in LDS tiles (the n-body pattern), and the face condition -- some face of
//      m avoids i -- is checked from the vertex->face table only when m would improve the lane's
//      minimum.  Squared norms are exact FP64 values in the reference's order, so the minimum is the
//      reference's bit for bit;
//   2. AABBBroadPhase candidates at that radius (the broadphase's 3-axis mode);
//   3. the minimum, over the candidate stencils, of the norm of Distance::vertexFaceDistance /
//      edgeEdgeDistance (Distance.cpp:51-64).  Minima are reduced with 64-bit atomics on the bits of
//      non-negative doubles, whose integer order is their numeric order.
#include <cstring>

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
    if (V > 0 && F > 0) {
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
