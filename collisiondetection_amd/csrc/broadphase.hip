// Broadphase: the candidate stencil sets of KDOPBroadPhase::findCollisionCandidates
// (src/KDOPBroadPhase.cpp:34-185), built MI355X-first.
//
// The reference builds a top-down median-split tree of 13-axis swept k-DOPs and traverses it against
// itself.  Its output does not depend on the tree's shape (interior bounds are exact unions of child
// bounds, KDOPBroadPhase.cpp:85-98, so a leaf pair is visited iff its two leaf bounds overlap), so this
// file builds a linear BVH instead:
//   1. leaf bounds: per face, 13 axes x (3 vertices x history entries), min/max of pos.axis -/+ eta in
//      FP64 exactly as KDOPBroadPhase.cpp:60-70 (bit-identical), plus an FP32 copy rounded outward;
//   2. 63-bit Morton codes of the leaf centres, radix-sorted (rocPRIM);
//   3. Karras' binary radix tree over the sorted codes (index tie-break for equal codes);
//   4. bottom-up refit of FP32 interior bounds with one atomic arrival counter per node;
//   5. escape ("rope") links: in the Karras layout the node after the subtree of a node covering
//      leaves [l, r] is internal node r+1 when that node's range starts at r+1, else leaf r+1;
//   6. stackless traversal, one lane per leaf i, starting at leaf i's escape link, which visits
//      exactly the subtrees holding leaves after i: every unordered face pair is found once.  The FP32
//      test culls; the leaf-leaf decision is the reference's FP64 13-axis test (touching overlaps)
//      followed by the shared-vertex test (Mesh.cpp:13-20).  Pairs are queued per lane in registers
//      and written in wavefront-wide flushes (one atomic each);
//   7. expansion of each face pair into its 6 vertex-face and 9 edge-edge stencils
//      (KDOPBroadPhase.cpp:150-184) with owner-computes deduplication: a stencil is written only by
//      the smallest face pair that generates it, which is decided from the face -> related faces table
//      built from the traversal's pairs (rows sorted; one wavefront per face) and the vertex->face
//      table, so the unique set comes out with no global sort.
// Meshes with degenerate faces (a repeated vertex) or duplicate faces break the owner rule's
// assumptions; for them the raw stencils are emitted and deduplicated by sorting.
#include <cstring>
#include <rocprim/device/device_merge_sort.hpp>
#include <rocprim/device/device_select.hpp>

#include "ccd_core.h"
#include "ccd_internal.h"

namespace {

constexpr int KAX = 13;
constexpr int NB = 2 * KAX;   // bounds per node: 13 mins then 13 maxs
constexpr uint32_t NIL = 0xffffffffu;

struct Axes {
    double a[KAX][3];
};

// KDOPBroadPhase.cpp:10-32: the 13 directions, each divided by its norm.
Axes make_axes()
{
    static const double raw[KAX][3] = {{1, 0, 0},  {0, 1, 0},  {0, 0, 1},  {1, 1, 0},  {1, -1, 0},
                                       {0, 1, 1},  {0, 1, -1}, {1, 0, 1},  {1, 0, -1}, {1, 1, 1},
                                       {1, 1, -1}, {1, -1, 1}, {1, -1, -1}};
    Axes ax;
    for (int i = 0; i < KAX; i++) {
        ccd::V3 v = ccd::mk(raw[i][0], raw[i][1], raw[i][2]);
        double n = ccd::norm(v);
        ax.a[i][0] = v.x / n;
        ax.a[i][1] = v.y / n;
        ax.a[i][2] = v.z / n;
    }
    return ax;
}

inline unsigned blocks_for(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// ------------------------------------------------------------------------------------------------
// 1. Leaf bounds
// ------------------------------------------------------------------------------------------------
__device__ __forceinline__ void dop_point(const Axes &ax, double x, double y, double z, double eta, double *mn,
                                          double *mx)
{
    ccd::V3 p = ccd::mk(x, y, z);
#pragma unroll
    for (int k = 0; k < KAX; k++) {
        double s = ccd::dot(p, ccd::mk(ax.a[k][0], ax.a[k][1], ax.a[k][2]));
        mn[k] = ccd::smin(s - eta, mn[k]);
        mx[k] = ccd::smax(s + eta, mx[k]);
    }
}

// AABBBroadPhase.cpp:31-36: per coordinate, min/max of pos[k] -/+ eta.
__device__ __forceinline__ void aabb_point(double x, double y, double z, double eta, double *mn, double *mx)
{
    double p[3] = {x, y, z};
#pragma unroll
    for (int k = 0; k < 3; k++) {
        mn[k] = ccd::smin(p[k] - eta, mn[k]);
        mx[k] = ccd::smax(p[k] + eta, mx[k]);
    }
}

// Two-entry histories (qs = entry at t=0, qe = entry at t=1) or CSR histories (hoff/hpos).
// aabb != 0 gives AABBBroadPhase's bounds (AABBBroadPhase.cpp:9-40): the three coordinates -/+ eta,
// with the other ten slabs unbounded so that the 13-axis overlap test reduces to the 3-axis one.
__global__ __launch_bounds__(256) void k_leaf_bounds(const int32_t *__restrict__ faces, int64_t F,
                                                     const double *__restrict__ qs, const double *__restrict__ qe,
                                                     const int64_t *__restrict__ hoff,
                                                     const double *__restrict__ hpos, double eta, Axes ax, int aabb,
                                                     double *__restrict__ dop64, double *__restrict__ centre)
{
    int64_t f = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (f >= F) return;
    double mn[KAX], mx[KAX];
#pragma unroll
    for (int k = 0; k < KAX; k++) {
        mn[k] = __builtin_inf();
        mx[k] = -__builtin_inf();
    }
    for (int j = 0; j < 3; j++) {
        int v = faces[3 * f + j];
        if (hoff) {
            for (int64_t e = hoff[v]; e < hoff[v + 1]; e++) {
                if (aabb)
                    aabb_point(hpos[3 * e], hpos[3 * e + 1], hpos[3 * e + 2], eta, mn, mx);
                else
                    dop_point(ax, hpos[3 * e], hpos[3 * e + 1], hpos[3 * e + 2], eta, mn, mx);
            }
        } else {
            const double *a = qs + 3 * (int64_t)v;
            const double *b = qe + 3 * (int64_t)v;
            if (aabb) {
                aabb_point(a[0], a[1], a[2], eta, mn, mx);
                aabb_point(b[0], b[1], b[2], eta, mn, mx);
            } else {
                dop_point(ax, a[0], a[1], a[2], eta, mn, mx);
                dop_point(ax, b[0], b[1], b[2], eta, mn, mx);
            }
        }
    }
    if (aabb) {
#pragma unroll
        for (int k = 3; k < KAX; k++) {
            mn[k] = -__builtin_inf();
            mx[k] = __builtin_inf();
        }
    }
    double *o = dop64 + NB * f;
#pragma unroll
    for (int k = 0; k < KAX; k++) {
        o[k] = mn[k];
        o[KAX + k] = mx[k];
    }
    centre[3 * f + 0] = 0.5 * (mn[0] + mx[0]);
    centre[3 * f + 1] = 0.5 * (mn[1] + mx[1]);
    centre[3 * f + 2] = 0.5 * (mn[2] + mx[2]);
}

// Scene bounds of the centres: per-block partial min/max, then one block folds the partials.
__global__ __launch_bounds__(256) void k_bounds_partial(const double *__restrict__ centre, int64_t F,
                                                        double *__restrict__ partial)
{
    __shared__ double s[6][256];
    double lo[3] = {__builtin_inf(), __builtin_inf(), __builtin_inf()};
    double hi[3] = {-__builtin_inf(), -__builtin_inf(), -__builtin_inf()};
    for (int64_t f = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; f < F; f += (int64_t)gridDim.x * blockDim.x) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            double v = centre[3 * f + c];
            lo[c] = v < lo[c] ? v : lo[c];
            hi[c] = v > hi[c] ? v : hi[c];
        }
    }
#pragma unroll
    for (int c = 0; c < 3; c++) {
        s[c][threadIdx.x] = lo[c];
        s[3 + c][threadIdx.x] = hi[c];
    }
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
#pragma unroll
            for (int c = 0; c < 3; c++) {
                double a = s[c][threadIdx.x + w];
                if (a < s[c][threadIdx.x]) s[c][threadIdx.x] = a;
                double b = s[3 + c][threadIdx.x + w];
                if (b > s[3 + c][threadIdx.x]) s[3 + c][threadIdx.x] = b;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < 6) partial[6 * blockIdx.x + threadIdx.x] = s[threadIdx.x][0];
}

__global__ __launch_bounds__(256) void k_bounds_final(double *__restrict__ partial, int nparts,
                                                      double *__restrict__ out)
{
    __shared__ double s[6][256];
    double v[6] = {__builtin_inf(), __builtin_inf(), __builtin_inf(), -__builtin_inf(), -__builtin_inf(),
                   -__builtin_inf()};
    for (int p = threadIdx.x; p < nparts; p += blockDim.x) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
            double a = partial[6 * p + c];
            double b = partial[6 * p + 3 + c];
            v[c] = a < v[c] ? a : v[c];
            v[3 + c] = b > v[3 + c] ? b : v[3 + c];
        }
    }
#pragma unroll
    for (int c = 0; c < 6; c++) s[c][threadIdx.x] = v[c];
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
#pragma unroll
            for (int c = 0; c < 3; c++) {
                double a = s[c][threadIdx.x + w];
                if (a < s[c][threadIdx.x]) s[c][threadIdx.x] = a;
                double b = s[3 + c][threadIdx.x + w];
                if (b > s[3 + c][threadIdx.x]) s[3 + c][threadIdx.x] = b;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < 6) out[threadIdx.x] = s[threadIdx.x][0];
}

// ------------------------------------------------------------------------------------------------
// 2. Morton codes (21 bits per axis)
// ------------------------------------------------------------------------------------------------
__device__ __forceinline__ uint64_t spread21(uint64_t x)
{
    x &= 0x1fffffull;
    x = (x | (x << 32)) & 0x1f00000000ffffull;
    x = (x | (x << 16)) & 0x1f0000ff0000ffull;
    x = (x | (x << 8)) & 0x100f00f00f00f00full;
    x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}

// Keys are (instance id << (63 - inst_bits)) | Morton code with (63 - inst_bits) / 3 bits per axis:
// sorting groups each instance's faces, so every instance owns one subtree of the radix tree.
__global__ __launch_bounds__(256) void k_morton(const double *__restrict__ centre, int64_t F,
                                                const double *__restrict__ box, const int32_t *__restrict__ face_inst,
                                                int inst_bits, uint64_t *__restrict__ keys, int32_t *__restrict__ vals)
{
    int64_t f = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (f >= F) return;
    const int axis_bits = (63 - inst_bits) / 3;
    const double scale = (double)((1u << axis_bits) - 1);
    uint64_t code = 0;
#pragma unroll
    for (int c = 0; c < 3; c++) {
        double lo = box[c], hi = box[3 + c];
        double ext = hi - lo;
        double u = ext > 0 ? (centre[3 * f + c] - lo) / ext : 0.0;
        u = u < 0 ? 0 : (u > 1 ? 1 : u);
        if (!(u == u)) u = 0;   // NaN positions sort first; bounds still decide every pair
        uint64_t q = (uint64_t)(u * scale);
        code |= spread21(q) << c;
    }
    if (face_inst) code |= (uint64_t)face_inst[f] << (3 * axis_bits);
    keys[f] = code;
    vals[f] = (int32_t)f;
}

// Instance of each face: the last instance whose first face is <= f.
__global__ __launch_bounds__(256) void k_face_instance(const int64_t *__restrict__ inst_off, int ninst, int64_t F,
                                                       int32_t *__restrict__ face_inst)
{
    int64_t f = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (f >= F) return;
    int lo = 0, hi = ninst - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (inst_off[mid] <= f)
            lo = mid;
        else
            hi = mid - 1;
    }
    face_inst[f] = lo;
}

// ------------------------------------------------------------------------------------------------
// 3. Karras binary radix tree.  Internal nodes 0..F-2, leaf j is node F-1+j.
// ------------------------------------------------------------------------------------------------
__device__ __forceinline__ int delta(const uint64_t *__restrict__ k, int64_t F, int64_t i, int64_t j)
{
    if (j < 0 || j >= F) return -1;
    uint64_t a = k[i], b = k[j];
    if (a == b) return 64 + __clz((uint32_t)(i ^ j));
    return __clzll(a ^ b);
}

__global__ __launch_bounds__(256) void k_karras(const uint64_t *__restrict__ keys, int64_t F,
                                                uint32_t *__restrict__ left, uint32_t *__restrict__ rng_first,
                                                uint32_t *__restrict__ rng_last, uint32_t *__restrict__ parent)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= F - 1) return;
    int d = (delta(keys, F, i, i + 1) - delta(keys, F, i, i - 1)) >= 0 ? 1 : -1;
    int dmin = delta(keys, F, i, i - d);
    int64_t lmax = 2;
    while (delta(keys, F, i, i + lmax * d) > dmin) lmax *= 2;
    int64_t l = 0;
    for (int64_t t = lmax / 2; t >= 1; t /= 2)
        if (delta(keys, F, i, i + (l + t) * d) > dmin) l += t;
    int64_t j = i + l * d;
    int dnode = delta(keys, F, i, j);
    int64_t s = 0;
    int64_t t = l;
    do {
        t = (t + 1) >> 1;
        if (delta(keys, F, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    int64_t gamma = i + s * d + (d < 0 ? -1 : 0);
    int64_t lo = i < j ? i : j, hi = i < j ? j : i;
    uint32_t lc = (lo == gamma) ? (uint32_t)(F - 1 + gamma) : (uint32_t)gamma;
    uint32_t rc = (hi == gamma + 1) ? (uint32_t)(F - 1 + gamma + 1) : (uint32_t)(gamma + 1);
    left[i] = lc;
    rng_first[i] = (uint32_t)lo;
    rng_last[i] = (uint32_t)hi;
    parent[lc] = (uint32_t)i;
    parent[rc] = (uint32_t)i;
    // The right child is found from the left: it follows the left child's subtree.
}

// Escape link of every node (internal and leaf) from the ranges.
__global__ __launch_bounds__(256) void k_escape(const uint32_t *__restrict__ rng_first,
                                                const uint32_t *__restrict__ rng_last, int64_t F,
                                                uint32_t *__restrict__ escape)
{
    int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (n >= 2 * F - 1) return;
    int64_t r = n < F - 1 ? (int64_t)rng_last[n] : n - (F - 1);
    uint32_t e;
    if (r == F - 1)
        e = NIL;
    else if (r + 1 <= F - 2 && (int64_t)rng_first[r + 1] == r + 1)
        e = (uint32_t)(r + 1);
    else
        e = (uint32_t)(F - 1 + r + 1);
    escape[n] = e;
}

// Node record used by the traversal: 26 bounds quantised to 16 bits, then the first child (internal)
// or the sorted leaf index (leaf), then the escape link: 60 bytes, one 64-byte line (112 with FP32
// bounds; the traversal is bound by fetching node records, so their size is its cost).
//
// Quantisation.  Per axis, a bound v maps to x = (v - origin) * inv, with origin the smallest finite
// lower bound of all leaves and inv = 65535 / (largest finite upper bound - origin); lower bounds take
// floor(x), upper bounds ceil(x), both clamped to [0, 65535], NaN to the full range.  Correctly rounded
// subtraction and multiplication are monotone, and so are floor, ceil and the clamps, so overlapping
// boxes keep overlapping (a.hi >= b.lo gives q(a.hi) >= q(b.lo)): the quantised test only ever keeps
// more, and the FP64 leaf test decides.  Values outside the range (a refit whose mesh moved) clamp,
// which stays conservative.
struct __align__(16) NodeRec {
    uint16_t b[NB];
    uint32_t child;
    uint32_t esc;
};
constexpr double QMAX = 65535.0;
__device__ __forceinline__ uint16_t quant_lo(double v, double o, double inv)
{
    const double x = (v - o) * inv;
    if (!(x >= 0.0)) return 0;   // below the range, -inf, NaN: the smallest
    if (x >= QMAX) return (uint16_t)QMAX;
    return (uint16_t)floor(x);
}
__device__ __forceinline__ uint16_t quant_hi(double v, double o, double inv)
{
    const double x = (v - o) * inv;
    if (x != x) return (uint16_t)QMAX;   // NaN: the largest
    if (x <= 0.0) return 0;
    if (x >= QMAX) return (uint16_t)QMAX;
    return (uint16_t)ceil(x);
}

// Quantisation parameters: per-block partial min of the finite lower and max of the finite upper
// bounds per axis (wavefront shuffles, then LDS), then one block folds the partials into origin[13]
// and inv[13] (qparam[0..12], qparam[13..25]).
__global__ __launch_bounds__(256) void k_qrange_partial(const double *__restrict__ dop64, int64_t F,
                                                        double *__restrict__ partial)
{
    __shared__ double s[4][NB];
    double lo[KAX], hi[KAX];
#pragma unroll
    for (int k = 0; k < KAX; k++) {
        lo[k] = __builtin_inf();
        hi[k] = -__builtin_inf();
    }
    for (int64_t f = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; f < F; f += (int64_t)gridDim.x * blockDim.x) {
#pragma unroll
        for (int k = 0; k < KAX; k++) {
            const double a = dop64[NB * f + k], b = dop64[NB * f + KAX + k];
            if (__builtin_isfinite(a) && a < lo[k]) lo[k] = a;
            if (__builtin_isfinite(b) && b > hi[k]) hi[k] = b;
        }
    }
#pragma unroll
    for (int k = 0; k < KAX; k++) {
#pragma unroll
        for (int d = 32; d > 0; d >>= 1) {
            const double a = __shfl_xor(lo[k], d), b = __shfl_xor(hi[k], d);
            lo[k] = a < lo[k] ? a : lo[k];
            hi[k] = b > hi[k] ? b : hi[k];
        }
    }
    const int w = threadIdx.x >> 6;
    if (__lane_id() == 0) {
#pragma unroll
        for (int k = 0; k < KAX; k++) {
            s[w][k] = lo[k];
            s[w][KAX + k] = hi[k];
        }
    }
    __syncthreads();
    if (threadIdx.x < NB) {
        const int c = threadIdx.x;
        double v = s[0][c];
        for (int ww = 1; ww < 4; ww++) v = c < KAX ? (s[ww][c] < v ? s[ww][c] : v) : (s[ww][c] > v ? s[ww][c] : v);
        partial[NB * blockIdx.x + c] = v;
    }
}

__global__ __launch_bounds__(64) void k_qrange_final(const double *__restrict__ partial, int nparts,
                                                     double *__restrict__ qparam)
{
    const int k = blockIdx.x;   // one wavefront per axis
    double lo = __builtin_inf(), hi = -__builtin_inf();
    for (int p = threadIdx.x; p < nparts; p += 64) {
        const double a = partial[NB * p + k], b = partial[NB * p + KAX + k];
        lo = a < lo ? a : lo;
        hi = b > hi ? b : hi;
    }
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) {
        const double a = __shfl_xor(lo, d), b = __shfl_xor(hi, d);
        lo = a < lo ? a : lo;
        hi = b > hi ? b : hi;
    }
    if (threadIdx.x != 0) return;
    const double ext = hi - lo;
    const bool ok = __builtin_isfinite(lo) && __builtin_isfinite(hi) && ext > 0.0 && __builtin_isfinite(QMAX / ext);
    qparam[k] = __builtin_isfinite(lo) ? lo : 0.0;
    qparam[KAX + k] = ok ? QMAX / ext : 0.0;
}

// Leaf records in sorted order, and the FP64 bounds and vertex triples in sorted order.
__global__ __launch_bounds__(256) void k_leaf_records(const int32_t *__restrict__ order,
                                                      const double *__restrict__ dop64,
                                                      const int32_t *__restrict__ faces, int64_t F,
                                                      const uint32_t *__restrict__ escape,
                                                      const double *__restrict__ qparam,
                                                      NodeRec *__restrict__ nodes, double *__restrict__ sdop64,
                                                      int4 *__restrict__ sfaces, int32_t *__restrict__ rank)
{
    const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int lane = __lane_id();
    const int64_t j0 = j - lane;   // the wavefront's first leaf
    if (j0 >= F) return;           // the whole wavefront is past the leaves
    const bool active = j < F;
    const int64_t f = active ? order[j] : 0;
    // The wavefront's 64 rows of sorted FP64 bounds are one contiguous block: written element by element
    // across the lanes (each store instruction covers consecutive doubles), not row by row per lane.
    const int nrow = F - j0 < 64 ? (int)(F - j0) : 64;
    for (int e = lane; e < 64 * NB; e += 64) {
        const int rr = e / NB, c = e - rr * NB;
        const int64_t fr = __shfl(f, rr);
        if (rr < nrow) sdop64[NB * j0 + e] = dop64[NB * fr + c];
    }
    if (!active) return;
    rank[f] = (int32_t)j;
    const double *src = dop64 + NB * f;
    NodeRec r;
#pragma unroll
    for (int k = 0; k < KAX; k++) {
        r.b[k] = quant_lo(src[k], qparam[k], qparam[KAX + k]);
        r.b[KAX + k] = quant_hi(src[KAX + k], qparam[k], qparam[KAX + k]);
    }
    r.child = (uint32_t)j;
    r.esc = escape[F - 1 + j];
    nodes[F - 1 + j] = r;
    sfaces[j] = make_int4(faces[3 * f], faces[3 * f + 1], faces[3 * f + 2], (int)f);
}

// 4. Bottom-up refit: a lane per leaf climbs while it is the second child to arrive at a node.
__global__ __launch_bounds__(256) void k_refit(int64_t F, const uint32_t *__restrict__ parent,
                                               const uint32_t *__restrict__ left, const uint32_t *__restrict__ escape,
                                               NodeRec *nodes, unsigned int *arrivals)
{
    int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (j >= F) return;
    uint32_t node = parent[F - 1 + j];
    while (true) {
        __threadfence();   // release this lane's child bounds before announcing the arrival
        unsigned int prior = atomicAdd(&arrivals[node], 1u);
        if (prior == 0) return;   // the sibling subtree is not finished yet; its lane continues
        __threadfence();          // acquire the sibling's bounds
        uint32_t lc = left[node];
        // The right child is the node after the left child's subtree, i.e. its escape link.
        uint32_t rc = escape[lc];
        NodeRec a = nodes[lc];
        NodeRec b = nodes[rc];
        NodeRec r;
#pragma unroll
        for (int k = 0; k < KAX; k++) {
            r.b[k] = a.b[k] < b.b[k] ? a.b[k] : b.b[k];
            r.b[KAX + k] = a.b[KAX + k] > b.b[KAX + k] ? a.b[KAX + k] : b.b[KAX + k];
        }
        r.child = lc;
        r.esc = escape[node];
        nodes[node] = r;
        if (node == 0) return;
        node = parent[node];
    }
}

// Level-synchronous refit for large trees.  The arrival-counter refit above needs a device-wide fence
// per node, which writes back and invalidates the L2 of the lane's XCD; with millions of nodes those
// fences serialise.  Instead each internal node's depth is found once per build by pointer jumping
// over the parent links (8 rounds: exact up to depth 256, and a Karras tree over 63-bit keys with an
// index tie-break is at most 97 deep, one level per distinct split bit), and the refit then runs one
// launch per depth, deepest first: a kernel boundary orders each level's writes before the next
// level's reads, with no fence inside.  Launches for depths past the tree's maximum return at once.
constexpr int MAX_TREE_DEPTH = 128;
constexpr int DEPTH_ROUNDS = 8;
constexpr int64_t LEVEL_REFIT_MIN_FACES = 1 << 16;

__global__ __launch_bounds__(256) void k_depth_init(int64_t F, const uint32_t *__restrict__ parent,
                                                    uint32_t *__restrict__ jump, uint32_t *__restrict__ dist,
                                                    uint32_t *__restrict__ maxd)
{
    int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (n == 0) *maxd = 0;
    if (n >= F - 1) return;
    jump[n] = n == 0 ? NIL : parent[n];
    dist[n] = n == 0 ? 0u : 1u;
}

__global__ __launch_bounds__(256) void k_depth_jump(int64_t F, const uint32_t *__restrict__ jin,
                                                    const uint32_t *__restrict__ din, uint32_t *__restrict__ jout,
                                                    uint32_t *__restrict__ dout)
{
    int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (n >= F - 1) return;
    const uint32_t j = jin[n];
    if (j == NIL) {
        jout[n] = NIL;
        dout[n] = din[n];
    } else {
        jout[n] = jin[j];
        dout[n] = din[n] + din[j];
    }
}

__global__ __launch_bounds__(256) void k_depth_max(int64_t F, const uint32_t *__restrict__ depth,
                                                   uint32_t *__restrict__ maxd)
{
    uint32_t m = 0;
    for (int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; n < F - 1; n += (int64_t)gridDim.x * blockDim.x)
        m = depth[n] > m ? depth[n] : m;
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) {
        const uint32_t o = __shfl_xor(m, d);
        m = o > m ? o : m;
    }
    if (__lane_id() == 0) atomicMax(maxd, m);
}

__global__ __launch_bounds__(256) void k_refit_level(int64_t F, uint32_t level, const uint32_t *__restrict__ depth,
                                                     const uint32_t *__restrict__ maxd,
                                                     const uint32_t *__restrict__ left,
                                                     const uint32_t *__restrict__ escape, NodeRec *__restrict__ nodes)
{
    if (level > *maxd) return;
    for (int64_t n = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; n < F - 1; n += (int64_t)gridDim.x * blockDim.x) {
        if (depth[n] != level) continue;
        const uint32_t lc = left[n];
        const uint32_t rc = escape[lc];   // the node after the left child's subtree
        const NodeRec a = nodes[lc];
        const NodeRec b = nodes[rc];
        NodeRec r;
#pragma unroll
        for (int k = 0; k < KAX; k++) {
            r.b[k] = a.b[k] < b.b[k] ? a.b[k] : b.b[k];
            r.b[KAX + k] = a.b[KAX + k] > b.b[KAX + k] ? a.b[KAX + k] : b.b[KAX + k];
        }
        r.child = lc;
        r.esc = escape[n];
        nodes[n] = r;
    }
}

// ------------------------------------------------------------------------------------------------
// 6. Traversal: unordered overlapping non-neighbouring face pairs.
// ------------------------------------------------------------------------------------------------
__device__ __forceinline__ bool overlap16(const uint16_t *q, const uint16_t *b)
{
    bool ok = true;
#pragma unroll
    for (int k = 0; k < KAX; k++) ok = ok && !(q[KAX + k] < b[k] || b[KAX + k] < q[k]);
    return ok;
}
// KDOPBroadPhase.cpp:132-137: disjoint iff some axis separates them strictly.
__device__ __forceinline__ bool overlap64(const double *__restrict__ a, const double *__restrict__ b)
{
    bool ok = true;
#pragma unroll
    for (int k = 0; k < KAX; k++) ok = ok && !(a[KAX + k] < b[k] || b[KAX + k] < a[k]);
    return ok;
}
__device__ __forceinline__ bool neighbours(int4 a, int4 b)   // Mesh.cpp:13-20
{
    return a.x == b.x || a.x == b.y || a.x == b.z || a.y == b.x || a.y == b.y || a.y == b.z || a.z == b.x ||
           a.z == b.y || a.z == b.z;
}

// inst_end: for batched instances, the sorted index one past the query's instance (its faces are the
// sorted range [inst_off[k], inst_off[k+1])), else null (one instance).  An escape link always leads
// to a node whose first leaf is its index (internal) or its leaf number: once that reaches inst_end
// the walk has left the instance's subtree and stops.
// Pairs are queued per lane (QN face ids in registers) and written when some lane's queue is full or
// the traversal ends: one atomic per wavefront and flush, positions by a wavefront prefix sum.  A
// single counter claimed by every wavefront on every iteration that emits serialises the device.
// The query's FP64 bounds are read from memory only for the leaves that pass the FP32 test, which
// keeps the lane's registers to its FP32 bounds and its queue.
constexpr int TRAVERSE_QN = 8;
__device__ __forceinline__ void traverse_flush(int me, const int (&q)[TRAVERSE_QN], int n, int2 *__restrict__ pairs,
                                               int64_t cap, unsigned long long *counter)
{
    const int lane = __lane_id();
    int incl = n;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int t = __shfl_up(incl, d);
        if (lane >= d) incl += t;
    }
    const int total = __shfl(incl, 63);
    unsigned long long base = 0;
    if (lane == 0 && total) base = atomicAdd(counter, (unsigned long long)total);
    base = __shfl(base, 0);
    const int64_t pos = (int64_t)base + (incl - n);
#pragma unroll
    for (int k = 0; k < TRAVERSE_QN; k++) {
        if (k < n && pos + k < cap) {
            const int b = q[k];
            pairs[pos + k] = me < b ? make_int2(me, b) : make_int2(b, me);
        }
    }
}

// Packet traversal: one wavefront walks the tree once for its 64 consecutive sorted leaves.  The walk
// starts at the escape link of the wavefront's first leaf i0 (the subtrees holding the leaves after
// i0, which hold every leaf any lane pairs with) and follows child links into a subtree when some
// lane's query overlaps it and the subtree holds a leaf in that lane's range (after its own leaf,
// before its instance's end), escape links otherwise, until it passes the last instance end of the
// wavefront.  The node index is the same in every lane, so a node record is one scalar load for the
// whole wavefront, where 64 lanes walking their own paths fetched up to 64 different records per
// step.  A lane tests the leaves the walk reaches that overlap its query in the quantised bounds and
// lie in its range; every pair in a lane's range whose FP64 boxes overlap is reached (the quantised
// bounds of its leaf and of that leaf's ancestors contain the FP64 ones), so each lane finds exactly
// the pairs of a walk of its own.
// The FP64 leaf test issues all its loads together and evaluates every comparison (an early-exit
// test waits for one load after another: per leaf, up to 26 memory latencies in a row).  Loading the
// child and escape records of each node ahead of the test was measured slower (the extra registers
// halve the resident wavefronts).
struct NodeWords {
    uint4 w[4];   // a NodeRec as four 16-byte words: bounds in w[0..3] (26 halves), child w[3].y, escape w[3].z
};
static_assert(sizeof(NodeRec) == 64, "node records are 64 bytes");
__device__ __forceinline__ void load_node(const uint4 *__restrict__ recs, uint32_t n, NodeWords &r)
{
    const uint4 *p = recs + 4 * (int64_t)n;
#pragma unroll
    for (int k = 0; k < 4; k++) r.w[k] = p[k];
}
// Bound k (0 <= k < 26) of a record: 16-bit half k % 2 of 32-bit word k / 2.
__device__ __forceinline__ uint32_t node_bound(const NodeWords &r, int k)
{
    const int wd = k >> 1;
    const uint4 v = r.w[wd >> 2];
    const uint32_t x = (wd & 3) == 0 ? v.x : ((wd & 3) == 1 ? v.y : ((wd & 3) == 2 ? v.z : v.w));
    return (k & 1) ? (x >> 16) : (x & 0xffffu);
}
// overlap64 with every comparison evaluated (no early exit: the loads issue together).
__device__ __forceinline__ bool overlap64_all(const double *__restrict__ a, const double *__restrict__ b)
{
    double av[NB], bv[NB];
#pragma unroll
    for (int k = 0; k < NB; k++) {
        av[k] = a[k];
        bv[k] = b[k];
    }
    bool ok = true;
#pragma unroll
    for (int k = 0; k < KAX; k++) ok = ok & !((av[KAX + k] < bv[k]) | (bv[KAX + k] < av[k]));
    return ok;
}

__global__ __launch_bounds__(256) void k_traverse(int64_t F, const NodeRec *__restrict__ nodes,
                                                  const double *__restrict__ sdop64,
                                                  const int4 *__restrict__ sfaces, const int32_t *__restrict__ face_inst,
                                                  const int64_t *__restrict__ inst_off, int2 *__restrict__ pairs,
                                                  int64_t cap, unsigned long long *counter)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t i0 = blockIdx.x * (int64_t)blockDim.x + (threadIdx.x & ~63);   // the wavefront's first leaf
    if (i0 >= F || F < 2) return;   // the whole wavefront (lane 0 included) is past the leaves
    const uint4 *__restrict__ recs = reinterpret_cast<const uint4 *>(nodes);
    const bool active = i < F;
    uint32_t q16[NB];
    const double *q64 = sdop64 + NB * (active ? i : 0);
    int4 qf = make_int4(0, 0, 0, 0);
    int64_t iend = 0;   // inactive lanes accept no leaf
    {
        NodeWords me;
        load_node(recs, (uint32_t)(F - 1 + (active ? i : i0)), me);
#pragma unroll
        for (int k = 0; k < NB; k++) q16[k] = node_bound(me, k);
    }
    if (active) {
        qf = sfaces[i];
        iend = face_inst ? inst_off[face_inst[qf.w] + 1] : F;
    }
    // The last instance end of the wavefront's lanes.
    int64_t jmax = iend;
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) {
        const int64_t o = __shfl_xor(jmax, d);
        jmax = o > jmax ? o : jmax;
    }
    const uint32_t nint = (uint32_t)(F - 1);
    // The walk: node (uniform), its record, and the first leaf of its subtree.
    uint32_t node = nodes[F - 1 + i0].esc;
    int64_t fl = i0 + 1;
    if (fl >= jmax) node = NIL;
    node = __builtin_amdgcn_readfirstlane(node);
    NodeWords cur;
    if (node != NIL) load_node(recs, node, cur);
    int q[TRAVERSE_QN] = {};
    int nq = 0;
    while (node != NIL) {
        const uint32_t child = __builtin_amdgcn_readfirstlane(cur.w[3].y);
        const uint32_t esc = __builtin_amdgcn_readfirstlane(cur.w[3].z);
        const bool internal = node < nint;
        // The subtree's leaves are [fl, r1): r1 is the first leaf of the node after it.
        const int64_t r1 = esc == NIL ? F : (esc < nint ? (int64_t)esc : (int64_t)(esc - nint));
        bool ov = true;
#pragma unroll
        for (int k = 0; k < KAX; k++)
            ov = ov & !((q16[KAX + k] < node_bound(cur, k)) | (node_bound(cur, KAX + k) < q16[k]));
        const bool lane_ok = (r1 - 1 > i) & (fl < iend) & ov;
        bool descend = false;
        if (internal) {
            descend = __ballot(lane_ok) != 0;
        } else if (lane_ok) {
            // leaf: child is its sorted index (== fl)
            const int4 of = sfaces[child];
            if (overlap64_all(q64, sdop64 + NB * (int64_t)child) && !neighbours(qf, of)) {
#pragma unroll
                for (int k = 0; k < TRAVERSE_QN; k++)
                    if (k == nq) q[k] = of.w;
                nq++;
            }
        }
        if (descend) {
            node = child;   // the left child: its subtree starts at the same leaf
            load_node(recs, node, cur);
        } else {
            node = esc;
            fl = r1;
            if (fl >= jmax) node = NIL;
            if (node != NIL) load_node(recs, node, cur);
        }
        if (__ballot(nq == TRAVERSE_QN) || node == NIL) {
            traverse_flush(qf.w, q, nq, pairs, cap, counter);
            nq = 0;
        }
    }
}

// ------------------------------------------------------------------------------------------------
// Vertex -> face table (CSR)
// ------------------------------------------------------------------------------------------------
__global__ void k_vf_count(const int32_t *__restrict__ faces, int64_t F, int32_t *__restrict__ cnt)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= 3 * F) return;
    atomicAdd(&cnt[faces[i]], 1);
}
__global__ void k_vf_fill(const int32_t *__restrict__ faces, int64_t F, const int32_t *__restrict__ off,
                          int32_t *__restrict__ cursor, int32_t *__restrict__ list)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= 3 * F) return;
    int v = faces[i];
    int pos = atomicAdd(&cursor[v], 1);
    list[off[v] + pos] = (int32_t)(i / 3);
}

// Mesh checks for the owner rule: a face with a repeated vertex, or two faces with the same vertex
// set (found as equal sorted triples among the faces of each triple's smallest vertex).
__global__ void k_mesh_check(const int32_t *__restrict__ faces, int64_t F, const int32_t *__restrict__ off,
                             const int32_t *__restrict__ list, unsigned long long *flag)
{
    int64_t f = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (f >= F) return;
    int a = faces[3 * f], b = faces[3 * f + 1], c = faces[3 * f + 2];
    bool bad = (a == b || b == c || a == c);
    if (!bad) {
        int lo = a < b ? (a < c ? a : c) : (b < c ? b : c);
        int hi = a > b ? (a > c ? a : c) : (b > c ? b : c);
        int md = a + b + c - lo - hi;
        for (int k = off[lo]; k < off[lo + 1]; k++) {
            int g = list[k];
            if (g == (int)f) continue;
            int x = faces[3 * g], y = faces[3 * g + 1], z = faces[3 * g + 2];
            int l2 = x < y ? (x < z ? x : z) : (y < z ? y : z);
            int h2 = x > y ? (x > z ? x : z) : (y > z ? y : z);
            int m2 = x + y + z - l2 - h2;
            if (l2 == lo && m2 == md && h2 == hi) bad = true;
        }
    }
    if (bad) atomicAdd(flag, 1ull);
}

// Edge -> opposite face: opp[3f + s] is the other face on edge (v[s], v[s+1]) of face f, -1 on a
// boundary edge, -2 on an edge shared by more than two faces (the expansion then scans the vertex ->
// face table for that pair, as it does for every pair without this table).
__global__ __launch_bounds__(256) void k_edge_opp(const int32_t *__restrict__ faces, int64_t F,
                                                  const int32_t *__restrict__ off, const int32_t *__restrict__ list,
                                                  int32_t *__restrict__ opp)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= 3 * F) return;
    int64_t f = i / 3;
    int s = (int)(i - 3 * f);
    int a = faces[i], b = faces[3 * f + (s + 1) % 3];
    int found = -1, n = 0;
    for (int k = off[a]; k < off[a + 1]; k++) {
        int h = list[k];
        if (h == (int)f) continue;
        if (faces[3 * h] == b || faces[3 * h + 1] == b || faces[3 * h + 2] == b) {
            found = h;
            n++;
        }
    }
    opp[i] = n > 1 ? -2 : found;
}

// ------------------------------------------------------------------------------------------------
// 7. Stencil expansion with owner-computes deduplication
// ------------------------------------------------------------------------------------------------
struct MeshView {
    const int32_t *faces;
    const int32_t *vf_off;
    const int32_t *vf_list;
    const double *dop64;   // by original face id
    const uint8_t *fixed;
    const int32_t *adj_off = nullptr;   // [F + 1] related-face table rows, or null (no table)
    const int32_t *adj = nullptr;       // [2 * pairs]
    const int32_t *opp = nullptr;       // [3F] edge -> opposite face (k_edge_opp), or null
    const int32_t *rank = nullptr;      // [F] face id -> Morton-sorted position
    const double *sdop64 = nullptr;     // FP64 bounds in Morton order
    const int4 *sfaces = nullptr;       // vertex triples and face id in Morton order
};

__global__ __launch_bounds__(256) void k_adj_count(const int2 *__restrict__ pairs,
                                                   const unsigned long long *__restrict__ np_dev, int64_t cap,
                                                   int64_t F, int32_t *__restrict__ cnt)
{
    int64_t np = (int64_t)*np_dev;
    if (np > cap) np = cap;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < np; i += (int64_t)gridDim.x * blockDim.x) {
        int2 p = pairs[i];
        atomicAdd(&cnt[p.x], 1);
        atomicAdd(&cnt[p.y], 1);
    }
}

// cnt holds the row lengths on entry and is counted back down to zero.
__global__ __launch_bounds__(256) void k_adj_fill(const int2 *__restrict__ pairs,
                                                  const unsigned long long *__restrict__ np_dev, int64_t cap, int64_t F,
                                                  const int32_t *__restrict__ off, int32_t *__restrict__ cnt,
                                                  int32_t *__restrict__ adj)
{
    int64_t np = (int64_t)*np_dev;
    if (np > cap) np = cap;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < np; i += (int64_t)gridDim.x * blockDim.x) {
        int2 p = pairs[i];
        adj[off[p.x] + atomicSub(&cnt[p.x], 1) - 1] = p.y;
        adj[off[p.y] + atomicSub(&cnt[p.y], 1) - 1] = p.x;
    }
}

__device__ __forceinline__ int4 face3(const MeshView &m, int f)
{
    return make_int4(m.faces[3 * f], m.faces[3 * f + 1], m.faces[3 * f + 2], f);
}
// Rows of the related-face table up to this length are sorted (k_row_sort); longer ones stay in fill
// order and are searched linearly.
constexpr int ROW_SORT_CAP = 2048;

// Related faces: overlapping swept k-DOPs and no shared vertex, i.e. exactly the face pairs the
// traversal found.  ROWS: b is in a's row of the related-face table (binary search in the sorted row);
// otherwise the FP64 test is repeated on the two faces' bounds (when there is no table).
template <bool ROWS>
__device__ __forceinline__ bool related(const MeshView &m, int a, int b)
{
    if (ROWS) {
        const int end = m.adj_off[a + 1];
        int lo = m.adj_off[a], hi = end;
        if (hi - lo > ROW_SORT_CAP) {
            for (int k = lo; k < hi; k++)
                if (m.adj[k] == b) return true;
            return false;
        }
        while (lo < hi) {   // first entry not below b
            const int mid = (lo + hi) >> 1;
            if (m.adj[mid] < b)
                lo = mid + 1;
            else
                hi = mid;
        }
        return lo < end && m.adj[lo] == b;
    }
    if (m.rank) {   // the Morton-ordered copies: spatially close faces share cache lines
        int ra = m.rank[a], rb = m.rank[b];
        return !neighbours(m.sfaces[ra], m.sfaces[rb]) &&
               overlap64(m.sdop64 + NB * (int64_t)ra, m.sdop64 + NB * (int64_t)rb);
    }
    return !neighbours(face3(m, a), face3(m, b)) && overlap64(m.dop64 + NB * (int64_t)a, m.dop64 + NB * (int64_t)b);
}
__device__ __forceinline__ bool has_vertex(const MeshView &m, int f, int v)
{
    return m.faces[3 * f] == v || m.faces[3 * f + 1] == v || m.faces[3 * f + 2] == v;
}
__device__ __forceinline__ bool is_fixed(const MeshView &m, int v) { return m.fixed && m.fixed[v]; }

// VF stencil (p, face h) generated by pair {a, h} for every face a containing p with a related to h;
// the owner is the smallest such a.  `mine` is the a of the current pair.
template <bool ROWS>
__device__ __forceinline__ bool vf_owner(const MeshView &m, int p, int h, int mine)
{
    for (int k = m.vf_off[p]; k < m.vf_off[p + 1]; k++) {
        int a = m.vf_list[k];
        if (a < mine && related<ROWS>(m, a, h)) return false;
    }
    return true;
}

// EE stencil (edge (u0,u1), edge (w0,w1)) generated by pairs {a, b}, a containing both u's and b both
// w's, a related to b; the owner is the smallest (min(a,b), max(a,b)).  The current pair is (f < g).
template <bool ROWS>
__device__ __forceinline__ bool ee_owner(const MeshView &m, int u0, int u1, int w0, int w1, int f, int g)
{
    for (int ka = m.vf_off[u0]; ka < m.vf_off[u0 + 1]; ka++) {
        int a = m.vf_list[ka];
        if (!has_vertex(m, a, u1)) continue;
        for (int kb = m.vf_off[w0]; kb < m.vf_off[w0 + 1]; kb++) {
            int b = m.vf_list[kb];
            if (!has_vertex(m, b, w1)) continue;
            int lo = a < b ? a : b, hi = a < b ? b : a;
            if (lo < f || (lo == f && hi < g)) {
                if (related<ROWS>(m, a, b)) return false;
            }
        }
    }
    return true;
}

// The same rule from the edge table: the faces on edge u are f and oa, those on edge w are g and ob
// (-1: none), so the other generating pairs are (f, ob), (oa, g) and (oa, ob).
__device__ __forceinline__ bool precedes(int a, int b, int f, int g)
{
    int lo = a < b ? a : b, hi = a < b ? b : a;
    return lo < f || (lo == f && hi < g);
}
template <bool ROWS>
__device__ __forceinline__ bool ee_owner_opp(const MeshView &m, int oa, int ob, int f, int g)
{
    if (ob >= 0 && precedes(f, ob, f, g) && related<ROWS>(m, f, ob)) return false;
    if (oa >= 0 && precedes(oa, g, f, g) && related<ROWS>(m, oa, g)) return false;
    if (oa >= 0 && ob >= 0 && precedes(oa, ob, f, g) && related<ROWS>(m, oa, ob)) return false;
    return true;
}

__device__ __forceinline__ int4 vf_key(int p, int a, int b, int c)   // Stencils.h:13-57
{
    int lo = a < b ? (a < c ? a : c) : (b < c ? b : c);
    int hi = a > b ? (a > c ? a : c) : (b > c ? b : c);
    int md = a + b + c - lo - hi;
    return make_int4(p, lo, md, hi);
}
__device__ __forceinline__ int4 ee_key(int a, int b, int c, int d)   // Stencils.h:86-117
{
    int p0 = a < b ? a : b, p1 = a < b ? b : a;
    int q0 = c < d ? c : d, q1 = c < d ? d : c;
    if (p0 > q0) return make_int4(q0, q1, p0, p1);
    return make_int4(p0, p1, q0, q1);
}

// Per-pair expansion, for the two cases the row expansion below does not cover: raw emission and meshes
// whose related-face table does not fit int32 offsets (rows_path false; the owner tests then repeat the
// FP64 bounds test).  raw_dev: the mesh check's count of degenerate or duplicate faces; nonzero
// switches every lane to raw emission (all 15 stencils of a pair), deduplicated afterwards by sorting.
//
// Output slots are claimed per block, not per wavefront: each lane first decides which of its 15
// stencils it emits (two bit masks), the block scans the counts through LDS and one lane claims the
// block's range with a single atomic per stencil kind.  The counters are single words that every block
// hits, so per-wavefront claims (15 per wavefront and pass) made their serialisation the bottleneck.
// The grid-stride loop bound is uniform across a block, so every lane reaches the barriers.
constexpr int EXPAND_BLOCK = 256;
__global__ __launch_bounds__(EXPAND_BLOCK) void k_expand(MeshView m, bool rows_path, const int2 *__restrict__ pairs,
                                                         const unsigned long long *__restrict__ npairs_dev,
                                                         int64_t pairs_cap, const unsigned long long *__restrict__ raw_dev,
                                                         int4 *__restrict__ vf_out, int64_t vf_cap,
                                                         unsigned long long *vf_counter, int4 *__restrict__ ee_out,
                                                         int64_t ee_cap, unsigned long long *ee_counter)
{
    constexpr int NW = EXPAND_BLOCK / 64;
    __shared__ unsigned s_off[2][NW];
    __shared__ unsigned long long s_base[2];
    const bool OWNER = *raw_dev == 0;
    if (OWNER && rows_path) return;   // the row expansion has it
    int64_t np = (int64_t)*npairs_dev;
    if (np > pairs_cap) np = pairs_cap;
    const int lane = __lane_id(), wave = threadIdx.x >> 6;
    for (int64_t base = blockIdx.x * (int64_t)EXPAND_BLOCK; base < np; base += (int64_t)gridDim.x * EXPAND_BLOCK) {
        int64_t i = base + threadIdx.x;
        bool valid = i < np;
        int2 pr = valid ? pairs[i] : make_int2(0, 0);
        int f = pr.x, g = pr.y;
        int4 F0 = valid ? face3(m, f) : make_int4(0, 0, 0, 0);
        int4 G0 = valid ? face3(m, g) : make_int4(0, 0, 0, 0);
        int fv[3] = {F0.x, F0.y, F0.z};
        int gv[3] = {G0.x, G0.y, G0.z};
        bool ffix = valid && is_fixed(m, fv[0]) && is_fixed(m, fv[1]) && is_fixed(m, fv[2]);
        bool gfix = valid && is_fixed(m, gv[0]) && is_fixed(m, gv[1]) && is_fixed(m, gv[2]);
        int fo[3] = {-2, -2, -2}, go[3] = {-2, -2, -2};
        if (valid && m.opp) {
#pragma unroll
            for (int s = 0; s < 3; s++) {
                fo[s] = m.opp[3 * (int64_t)f + s];
                go[s] = m.opp[3 * (int64_t)g + s];
            }
        }
        // 6 vertex-face stencils (KDOPBroadPhase.cpp:157-175): bit 2s for (fv[s], g), 2s+1 for (gv[s], f)
        unsigned vmask = 0, emask = 0;
#pragma unroll
        for (int s = 0; s < 3; s++) {
            int p = fv[s];
            bool want = valid && !(is_fixed(m, p) && gfix);
            if (OWNER && want) want = vf_owner<false>(m, p, g, f);
            if (want) vmask |= 1u << (2 * s);
            int q = gv[s];
            want = valid && !(is_fixed(m, q) && ffix);
            if (OWNER && want) want = vf_owner<false>(m, q, f, g);
            if (want) vmask |= 1u << (2 * s + 1);
        }
        // 9 edge-edge stencils (KDOPBroadPhase.cpp:176-183): bit 3s+r for edges (s, r)
#pragma unroll
        for (int s = 0; s < 3; s++) {
            int u0 = fv[s], u1 = fv[(s + 1) % 3];
#pragma unroll
            for (int r = 0; r < 3; r++) {
                int w0 = gv[r], w1 = gv[(r + 1) % 3];
                bool want = valid && !(is_fixed(m, u0) && is_fixed(m, u1) && is_fixed(m, w0) && is_fixed(m, w1));
                if (OWNER && want) {
                    if (fo[s] != -2 && go[r] != -2)
                        want = ee_owner_opp<false>(m, fo[s], go[r], f, g);
                    else
                        want = ee_owner<false>(m, u0, u1, w0, w1, f, g);
                }
                if (want) emask |= 1u << (3 * s + r);
            }
        }
        // Block-wide exclusive offsets: wavefront inclusive scans, then the wavefront totals in LDS.
        const unsigned nv = __popc(vmask), ne = __popc(emask);
        unsigned pv = nv, pe = ne;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            unsigned tv = __shfl_up(pv, d), te = __shfl_up(pe, d);
            if (lane >= d) {
                pv += tv;
                pe += te;
            }
        }
        if (lane == 63) {
            s_off[0][wave] = pv;
            s_off[1][wave] = pe;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned tv = 0, te = 0;
#pragma unroll
            for (int w = 0; w < NW; w++) {
                unsigned a = s_off[0][w], b = s_off[1][w];
                s_off[0][w] = tv;
                s_off[1][w] = te;
                tv += a;
                te += b;
            }
            s_base[0] = tv ? atomicAdd(vf_counter, (unsigned long long)tv) : 0ull;
            s_base[1] = te ? atomicAdd(ee_counter, (unsigned long long)te) : 0ull;
        }
        __syncthreads();
        int64_t pos_v = (int64_t)(s_base[0] + s_off[0][wave] + (pv - nv));
        int64_t pos_e = (int64_t)(s_base[1] + s_off[1][wave] + (pe - ne));
#pragma unroll
        for (int s = 0; s < 3; s++) {
            if (vmask & (1u << (2 * s))) {
                if (pos_v < vf_cap) vf_out[pos_v] = vf_key(fv[s], gv[0], gv[1], gv[2]);
                pos_v++;
            }
            if (vmask & (1u << (2 * s + 1))) {
                if (pos_v < vf_cap) vf_out[pos_v] = vf_key(gv[s], fv[0], fv[1], fv[2]);
                pos_v++;
            }
        }
#pragma unroll
        for (int s = 0; s < 3; s++) {
#pragma unroll
            for (int r = 0; r < 3; r++) {
                if (emask & (1u << (3 * s + r))) {
                    if (pos_e < ee_cap) ee_out[pos_e] = ee_key(fv[s], fv[(s + 1) % 3], gv[r], gv[(r + 1) % 3]);
                    pos_e++;
                }
            }
        }
        __syncthreads();   // s_off / s_base are rewritten by the next pass
    }
}

// ---- Row expansion -----------------------------------------------------------------------------
// The stencils of every face pair, deduplicated by the same owner rule, organised by face: one
// wavefront per face f walks f's row of the related-face table.  For each related face g it emits the
// three vertex-face stencils (f's vertices against g) that f owns, and for g > f the nine edge-edge
// stencils that the pair (f, g) owns.  Every unordered pair is visited from both sides, so each side
// emits its own vertices' stencils.  The owner tests become membership tests in sorted rows:
//   - f's own row (related(f, ob));
//   - the rows of the faces around f's vertices and across f's edges, which every lane of the
//     wavefront shares (they stay in cache).
// The per-pair form instead re-reads two 208-byte FP64 bounds per test.  Output slots come from a
// count pass, an exclusive scan over faces and a write pass that recomputes nothing: the count pass
// stores each entry's 12-bit emission mask beside the row.  No atomics; the order is deterministic
// (by face, then by related face).
//
// Rows of at most ROW_SORT_CAP entries, sorted ascending in LDS by one wavefront (bitonic network).
__global__ __launch_bounds__(64) void k_row_sort(const unsigned long long *__restrict__ raw_dev, int64_t F,
                                                 const int32_t *__restrict__ off, int32_t *__restrict__ adj)
{
    __shared__ int32_t s[ROW_SORT_CAP];
    if (*raw_dev != 0) return;
    for (int64_t a = blockIdx.x; a < F; a += gridDim.x) {
        const int lo = off[a], len = off[a + 1] - lo;
        if (len < 2 || len > ROW_SORT_CAP) continue;   // uniform across the block
        if (len <= 64) {
            // One element per lane, the same bitonic network in registers (no LDS round trips or
            // barriers): lane i and its partner i ^ j keep the smaller value in the lower lane of an
            // ascending run, the larger in a descending one.
            const int lane = threadIdx.x;
            int x = lane < len ? adj[lo + lane] : INT32_MAX;
#pragma unroll
            for (int k = 2; k <= 64; k <<= 1) {
#pragma unroll
                for (int j = k >> 1; j > 0; j >>= 1) {
                    const int y = __shfl_xor(x, j);
                    const bool lower = (lane & j) == 0, up = (lane & k) == 0;
                    x = (lower == up) ? (x < y ? x : y) : (x < y ? y : x);
                }
            }
            if (lane < len) adj[lo + lane] = x;
            continue;
        }
        int n2 = 2;
        while (n2 < len) n2 <<= 1;
        for (int i = threadIdx.x; i < n2; i += 64) s[i] = i < len ? adj[lo + i] : INT32_MAX;
        __syncthreads();
        for (int k = 2; k <= n2; k <<= 1) {
            for (int j = k >> 1; j > 0; j >>= 1) {
                for (int i = threadIdx.x; i < n2; i += 64) {
                    const int l = i ^ j;
                    if (l > i) {
                        const int x = s[i], y = s[l];
                        if ((x > y) == ((i & k) == 0)) {
                            s[i] = y;
                            s[l] = x;
                        }
                    }
                }
                __syncthreads();
            }
        }
        for (int i = threadIdx.x; i < len; i += 64) adj[lo + i] = s[i];
        __syncthreads();   // s is rewritten for the next row
    }
}

__device__ __forceinline__ int64_t wave_sum(int64_t v)
{
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d);
    return v;
}

// Write pass of the row expansion: the keys of each entry's emission mask, at the exclusive offsets the
// scan left in cnt_vf / cnt_ee.  Returns at once for raw emission.
constexpr int ROWS_BLOCK = 256;
__global__ __launch_bounds__(ROWS_BLOCK) void k_expand_rows_write(MeshView m, int64_t F,
                                                                  const unsigned long long *__restrict__ raw_dev,
                                                                  const uint16_t *__restrict__ mask,
                                                                  const int64_t *__restrict__ cnt_vf,
                                                                  const int64_t *__restrict__ cnt_ee,
                                                                  int4 *__restrict__ vf_out, int64_t vf_cap,
                                                                  int4 *__restrict__ ee_out, int64_t ee_cap)
{
    if (*raw_dev != 0) return;
    const int lane = __lane_id();
    const int64_t nwaves = (int64_t)gridDim.x * (ROWS_BLOCK / 64);
    for (int64_t fw = blockIdx.x * (int64_t)(ROWS_BLOCK / 64) + (threadIdx.x >> 6); fw < F; fw += nwaves) {
        const int f = __builtin_amdgcn_readfirstlane((int)fw);
        const int lo = m.adj_off[f], hi = m.adj_off[f + 1];
        const int4 F0 = face3(m, f);
        const int fv[3] = {F0.x, F0.y, F0.z};
        int64_t pos_v = cnt_vf[f], pos_e = cnt_ee[f];
        {
            for (int base = lo; base < hi; base += 64) {
                const int k = base + lane;
                const bool valid = k < hi;
                const unsigned msk = valid ? mask[k] : 0u;
                const int4 G0 = valid && msk ? face3(m, m.adj[k]) : make_int4(0, 0, 0, 0);
                const unsigned nv = __popc(msk & 7u), ne = __popc(msk >> 3);
                unsigned pv = nv, pe = ne;   // inclusive prefix over the wavefront
#pragma unroll
                for (int d = 1; d < 64; d <<= 1) {
                    const unsigned tv = __shfl_up(pv, d), te = __shfl_up(pe, d);
                    if (lane >= d) {
                        pv += tv;
                        pe += te;
                    }
                }
                int64_t v = pos_v + (pv - nv), e = pos_e + (pe - ne);
                const int gv[3] = {G0.x, G0.y, G0.z};
#pragma unroll
                for (int s = 0; s < 3; s++) {
                    if (msk & (1u << s)) {
                        if (v < vf_cap) vf_out[v] = vf_key(fv[s], gv[0], gv[1], gv[2]);
                        v++;
                    }
                }
#pragma unroll
                for (int s = 0; s < 3; s++) {
#pragma unroll
                    for (int r = 0; r < 3; r++) {
                        if (msk & (1u << (3 + 3 * s + r))) {
                            if (e < ee_cap) ee_out[e] = ee_key(fv[s], fv[(s + 1) % 3], gv[r], gv[(r + 1) % 3]);
                            e++;
                        }
                    }
                }
                pos_v += __shfl(pv, 63);
                pos_e += __shfl(pe, 63);
            }
        }
    }
}

// Count pass of the row expansion with the rows its owner tests read cached in LDS: f's own row and
// those of the faces around f's vertices that precede f.  Every lane of the wavefront searches the same
// few rows, so they are loaded once, coalesced, and searched in LDS instead of with dependent global
// loads.  A directory of cached rows sits one entry per lane (face, LDS offset, length); lookups of a
// face run in converged code with a ballot.  The tests that read other rows (the pair across both
// edges when f's neighbour follows f; the generic rule for non-manifold edges) and rows that do not
// fit use the table in global memory.
constexpr int ROWS_LDS_PER_WAVE = 1536;   // int32 entries per wavefront

struct RowDir {
    const int32_t *lds;
    int face = -1;   // this lane's directory entry (-1: none)
    int off = 0, len = -1;   // its row in lds (len < 0: not cached)
};

// Directory lookup of a face (wave-uniform `a`, all lanes active): the row's LDS offset and length,
// or false when the face is not cached.
__device__ __forceinline__ bool dir_find(const RowDir &d, int a, int &off, int &len)
{
    const uint64_t hit = __ballot(d.face == a);
    if (!hit) return false;
    const int j = __ffsll((unsigned long long)hit) - 1;
    off = __shfl(d.off, j);
    len = __shfl(d.len, j);
    return len >= 0;
}

// b in a sorted LDS row.
__device__ __forceinline__ bool lds_row_has(const int32_t *lds, int off, int len, int b)
{
    int lo = off, hi = off + len;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (lds[mid] < b)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo < off + len && lds[lo] == b;
}

// related(a, b) with a's row from the directory when it is cached (cached: the lookup's result).
__device__ __forceinline__ bool related_dir(const MeshView &m, const int32_t *lds, bool cached, int off, int len, int a,
                                           int b)
{
    return cached ? lds_row_has(lds, off, len, b) : related<true>(m, a, b);
}

__global__ __launch_bounds__(ROWS_BLOCK) void k_expand_rows_count(MeshView m, int64_t F,
                                                                  const unsigned long long *__restrict__ raw_dev,
                                                                  uint16_t *__restrict__ mask,
                                                                  int64_t *__restrict__ cnt_vf,
                                                                  int64_t *__restrict__ cnt_ee)
{
    __shared__ int32_t s_rows[ROWS_BLOCK / 64][ROWS_LDS_PER_WAVE];
    const bool raw = *raw_dev != 0;
    const int lane = __lane_id();
    int32_t *lds = s_rows[threadIdx.x >> 6];
    const int64_t nwaves = (int64_t)gridDim.x * (ROWS_BLOCK / 64);
    for (int64_t fw = blockIdx.x * (int64_t)(ROWS_BLOCK / 64) + (threadIdx.x >> 6); fw < F; fw += nwaves) {
        const int f = __builtin_amdgcn_readfirstlane((int)fw);
        if (raw) {
            if (lane == 0) cnt_vf[f] = cnt_ee[f] = 0;
            continue;
        }
        const int lo = m.adj_off[f], hi = m.adj_off[f + 1];
        const int4 F0 = face3(m, f);
        const int fv[3] = {F0.x, F0.y, F0.z};
        int fo[3] = {-2, -2, -2};
        if (m.opp) {
#pragma unroll
            for (int s = 0; s < 3; s++) fo[s] = m.opp[3 * (int64_t)f + s];
        }
        // Directory: f, then the faces around f's vertices that precede f, each once.
        RowDir d;
        d.lds = lds;
        int nd = 0;
        auto add = [&](int a) {
            if (nd < 64 && __ballot(d.face == a) == 0) {
                if (lane == nd) d.face = a;
                nd++;
            }
        };
        add(f);
        // The three stars are loaded lane-parallel (64 entries per load, all three in flight together)
        // and walked in registers, in the same order as one entry at a time.
        int sb[3], se[3], sa[3];
#pragma unroll
        for (int s = 0; s < 3; s++) {
            sb[s] = m.vf_off[fv[s]];
            se[s] = m.vf_off[fv[s] + 1];
        }
#pragma unroll
        for (int s = 0; s < 3; s++) sa[s] = sb[s] + lane < se[s] ? m.vf_list[sb[s] + lane] : 0x7fffffff;
#pragma unroll
        for (int s = 0; s < 3; s++) {
            for (int base = sb[s]; base < se[s]; base += 64) {
                const int a = base == sb[s] ? sa[s] : (base + lane < se[s] ? m.vf_list[base + lane] : 0x7fffffff);
                const int cnt = se[s] - base < 64 ? se[s] - base : 64;
                for (int j = 0; j < cnt; j++) {
                    const int aj = __shfl(a, j);
                    if (aj < f) add(aj);
                }
            }
        }
        // The directory holds every face around f's vertices that precedes f unless it overflowed.
        const bool complete = nd < 64;
        // Then the faces across f's edges that follow f: the edge-edge owner tests search their rows too
        // (after the completeness test above, which concerns the faces that precede f).
#pragma unroll
        for (int s = 0; s < 3; s++)
            if (fo[s] > f) add(fo[s]);
        // Per directory entry: its row in the table, and which of f's vertices it holds (bit s: fv[s];
        // only for the faces that precede f, the ones the vertex-face owner tests walk).
        int rlen = 0, src = 0;
        unsigned vbits = 0;
        if (lane < nd) {
            src = m.adj_off[d.face];
            rlen = m.adj_off[d.face + 1] - src;
            if (lane > 0 && d.face < f) {
#pragma unroll
                for (int s = 0; s < 3; s++)
                    if (has_vertex(m, d.face, fv[s])) vbits |= 1u << s;
            }
        }
        int incl = rlen;
#pragma unroll
        for (int dd = 1; dd < 64; dd <<= 1) {
            const int t = __shfl_up(incl, dd);
            if (lane >= dd) incl += t;
        }
        if (lane < nd) {
            d.off = incl - rlen;
            d.len = (incl <= ROWS_LDS_PER_WAVE && rlen <= ROW_SORT_CAP) ? rlen : -1;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();   // the previous face's reads of lds are done
        // Row copies four entries at a time: their first 64 elements load together.
        for (int j0 = 0; j0 < nd; j0 += 4) {
            int val[4], dst[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int j = (j0 + u) & 63;
                const int off = __shfl(d.off, j), len = __shfl(d.len, j), from = __shfl(src, j);
                const bool ok = j0 + u < nd && lane < len;
                val[u] = ok ? m.adj[from + lane] : 0;
                dst[u] = ok ? off + lane : -1;
            }
#pragma unroll
            for (int u = 0; u < 4; u++)
                if (dst[u] >= 0) lds[dst[u]] = val[u];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int j = (j0 + u) & 63;
                const int off = __shfl(d.off, j), len = __shfl(d.len, j), from = __shfl(src, j);
                if (j0 + u < nd)
                    for (int i = 64 + lane; i < len; i += 64) lds[off + i] = m.adj[from + i];
            }
        }
        unsigned long long star[3];   // directory entries around fv[s] (f excluded)
#pragma unroll
        for (int s = 0; s < 3; s++) star[s] = __ballot((vbits >> s) & 1u);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        int f_off = 0, f_len = 0;
        const bool f_cached = dir_find(d, f, f_off, f_len);
        int64_t nv = 0, ne = 0;
        for (int base = lo; base < hi; base += 64) {
            const int k = base + lane;
            const bool valid = k < hi;
            const int g = valid ? m.adj[k] : 0;
            const int4 G0 = valid ? face3(m, g) : make_int4(0, 0, 0, 0);
            const int gv[3] = {G0.x, G0.y, G0.z};
            const bool gfix = valid && is_fixed(m, gv[0]) && is_fixed(m, gv[1]) && is_fixed(m, gv[2]);
            // KDOPBroadPhase.cpp:157-175 (this side's three): owned unless a face around fv[s] that
            // precedes f is related to g.
            unsigned vmask = 0;
#pragma unroll
            for (int s = 0; s < 3; s++) {
                bool own = valid && !(is_fixed(m, fv[s]) && gfix);
                if (complete) {
                    for (unsigned long long st = star[s]; st; st &= st - 1) {
                        const int j = __ffsll(st) - 1;
                        const int a = __shfl(d.face, j), off = __shfl(d.off, j), len = __shfl(d.len, j);
                        if (own && related_dir(m, lds, len >= 0, off, len, a, g)) own = false;
                    }
                } else {
                    for (int kk = m.vf_off[fv[s]]; kk < m.vf_off[fv[s] + 1]; kk++) {
                        const int a = m.vf_list[kk];
                        if (a >= f) continue;
                        int off = 0, len = 0;
                        const bool cached = dir_find(d, a, off, len);
                        if (own && related_dir(m, lds, cached, off, len, a, g)) own = false;
                    }
                }
                if (own) vmask |= 1u << s;
            }
            // KDOPBroadPhase.cpp:176-183, for g > f
            unsigned emask = 0;
            const bool ee = valid && g > f;
            int go[3] = {-2, -2, -2};
            if (ee && m.opp) {
#pragma unroll
                for (int r = 0; r < 3; r++) go[r] = m.opp[3 * (int64_t)g + r];
            }
            // ee_owner_opp's first two tests depend on one edge each: (f, ob) on g's edge r, (oa, g) on f's
            // edge s.  Each is evaluated once here, not once per edge pair.
            bool t_fob[3], t_oag[3];
#pragma unroll
            for (int r = 0; r < 3; r++) {
                const int ob = go[r];
                t_fob[r] = ee && ob >= 0 && precedes(f, ob, f, g) &&
                           related_dir(m, lds, f_cached, f_off, f_len, f, ob);
            }
            int oa_off[3] = {0, 0, 0}, oa_len[3] = {0, 0, 0};
            bool oa_cached[3];
#pragma unroll
            for (int s = 0; s < 3; s++) {
                const int oa = fo[s];
                oa_cached[s] = oa >= 0 && dir_find(d, oa, oa_off[s], oa_len[s]);
                t_oag[s] = ee && oa >= 0 && precedes(oa, g, f, g) &&
                           related_dir(m, lds, oa_cached[s], oa_off[s], oa_len[s], oa, g);
            }
#pragma unroll
            for (int s = 0; s < 3; s++) {
                const int u0 = fv[s], u1 = fv[(s + 1) % 3];
                const bool ufix = is_fixed(m, u0) && is_fixed(m, u1);
                const int oa = fo[s];
#pragma unroll
                for (int r = 0; r < 3; r++) {
                    const int w0 = gv[r], w1 = gv[(r + 1) % 3];
                    if (!ee || (ufix && is_fixed(m, w0) && is_fixed(m, w1))) continue;
                    const int ob = go[r];
                    bool own;
                    if (oa != -2 && ob != -2) {
                        // ee_owner_opp with the cached rows: the other generating pairs (f, ob),
                        // (oa, g) and (oa, ob)
                        own = !t_fob[r] && !t_oag[s];
                        if (own && oa >= 0 && ob >= 0 && precedes(oa, ob, f, g) &&
                            related_dir(m, lds, oa_cached[s], oa_off[s], oa_len[s], oa, ob))
                            own = false;
                    } else {
                        own = ee_owner<true>(m, u0, u1, w0, w1, f, g);
                    }
                    if (own) emask |= 1u << (3 * s + r);
                }
            }
            const unsigned msk = vmask | (emask << 3);
            if (valid) mask[k] = (uint16_t)msk;
            nv += __popc(vmask);
            ne += __popc(emask);
        }
        nv = wave_sum(nv);
        ne = wave_sum(ne);
        if (lane == 0) {
            cnt_vf[f] = nv;
            cnt_ee[f] = ne;
        }
    }
}

// Stencil order of the reference's std::set (Stencils.h:59-77 and :119-137): lexicographic.
struct KeyLess {
    __host__ __device__ bool operator()(const int4 &a, const int4 &b) const
    {
        if (a.x != b.x) return a.x < b.x;
        if (a.y != b.y) return a.y < b.y;
        if (a.z != b.z) return a.z < b.z;
        return a.w < b.w;
    }
};
struct KeyEq {
    __host__ __device__ bool operator()(const int4 &a, const int4 &b) const
    {
        return a.x == b.x && a.y == b.y && a.z == b.z && a.w == b.w;
    }
};

__global__ void k_copy_count(const unsigned long long *src, int64_t *dst) { *dst = (int64_t)*src; }

}   // namespace

namespace ccdi {

// Sorts n stencil keys (device, 4*n int32) into the reference's set order.
int sort_stencils(ccd_ctx *ctx, int32_t *keys, int64_t n)
{
    if (n <= 1) return CCD_OK;
    size_t tmp = 0;
    int4 *sorted = (int4 *)buffer(ctx, "sort_out", (size_t)n * 16 + 16);
    if (!sorted) return CCD_E_OOM;
    CCD_HIP_CHECK(ctx, rocprim::merge_sort((void *)nullptr, tmp, (int4 *)keys, sorted, (size_t)n, KeyLess(),
                                           ctx->stream));
    void *t = buffer(ctx, "sort_tmp", tmp + 16);
    if (!t) return CCD_E_OOM;
    CCD_HIP_CHECK(ctx, rocprim::merge_sort(t, tmp, (int4 *)keys, sorted, (size_t)n, KeyLess(), ctx->stream));
    CCD_HIP_CHECK(ctx, hipMemcpyAsync(keys, sorted, (size_t)n * 16, hipMemcpyDeviceToDevice, ctx->stream));
    return CCD_OK;
}

// Sort + unique in place; the unique count goes to *count_dev.
static int sort_unique(ccd_ctx *ctx, int32_t *keys, int64_t n, const char *tag, int64_t *count_dev)
{
    int rc = sort_stencils(ctx, keys, n);
    if (rc) return rc;
    std::string name = std::string("uniq_out_") + tag;
    int4 *out = (int4 *)buffer(ctx, name.c_str(), (size_t)n * 16 + 16);
    if (!out) return CCD_E_OOM;
    size_t tmp = 0;
    CCD_HIP_CHECK(ctx, rocprim::unique((void *)nullptr, tmp, (int4 *)keys, out, (unsigned long long *)count_dev,
                                       (size_t)n, KeyEq(), ctx->stream));
    void *t = buffer(ctx, "uniq_tmp", tmp + 16);
    if (!t) return CCD_E_OOM;
    CCD_HIP_CHECK(ctx, rocprim::unique(t, tmp, (int4 *)keys, out, (unsigned long long *)count_dev, (size_t)n,
                                       KeyEq(), ctx->stream));
    CCD_HIP_CHECK(ctx, hipMemcpyAsync(keys, out, (size_t)n * 16, hipMemcpyDeviceToDevice, ctx->stream));
    return CCD_OK;
}

int build_vertex_face_table_buffers(ccd_ctx *ctx, int64_t F, int64_t V)
{
    if (!buffer(ctx, "vf_cnt", (size_t)(V + 1) * 4) || !buffer(ctx, "vf_off", (size_t)(V + 1) * 4) ||
        !buffer(ctx, "vf_list", (size_t)(3 * F + 1) * 4) || !buffer(ctx, "scan_tmp", scan_temp_bytes(V + 1)))
        return CCD_E_OOM;
    return CCD_OK;
}

int build_vertex_face_table(ccd_ctx *ctx, const int32_t *faces, int64_t F, int64_t V)
{
    hipStream_t st = ctx->stream;
    int rc = build_vertex_face_table_buffers(ctx, F, V);
    if (rc) return rc;
    int32_t *vcnt = (int32_t *)ctx->bufs["vf_cnt"].ptr;
    int32_t *voff = (int32_t *)ctx->bufs["vf_off"].ptr;
    int32_t *vlist = (int32_t *)ctx->bufs["vf_list"].ptr;
    void *scan_tmp = ctx->bufs["scan_tmp"].ptr;
    rc = fill_u32(ctx, (uint32_t *)vcnt, V + 1, 0);
    if (rc) return rc;
    if (F > 0) {
        hipLaunchKernelGGL(k_vf_count, dim3(blocks_for(3 * F, 256)), dim3(256), 0, st, faces, F, vcnt);
        CCD_LAUNCH_CHECK(ctx);
    }
    rc = exclusive_scan_i32(ctx, vcnt, voff, V + 1, scan_tmp, nullptr);
    if (rc) return rc;
    rc = fill_u32(ctx, (uint32_t *)vcnt, V + 1, 0);
    if (rc) return rc;
    if (F > 0) {
        hipLaunchKernelGGL(k_vf_fill, dim3(blocks_for(3 * F, 256)), dim3(256), 0, st, faces, F, voff, vcnt, vlist);
        CCD_LAUNCH_CHECK(ctx);
    }
    return CCD_OK;
}

void vertex_face_table(ccd_ctx *ctx, const int32_t **off, const int32_t **list)
{
    *off = (const int32_t *)ctx->bufs["vf_off"].ptr;
    *list = (const int32_t *)ctx->bufs["vf_list"].ptr;
}

// Device buffers of one broadphase, sized for F faces and V vertices and the capacities in caps.
struct BPBufs {
    double *dop64, *centre, *partial, *sdop64;
    double *qpartial, *qparam;   // node quantisation: per-block partial ranges, origin[13] + inv[13]
    uint64_t *keys, *keys2;
    int32_t *vals, *order, *face_inst;
    uint32_t *left, *rf, *rl, *parent, *escape, *arrivals;
    NodeRec *nodes;
    int4 *sfaces;
    void *sort_tmp;
    int32_t *adj_cnt = nullptr, *adj_off = nullptr, *adj = nullptr;   // related-face table (null if too large)
    void *adj_tmp = nullptr;
    uint16_t *adj_mask = nullptr;                  // per table entry: the row expansion's emission mask
    int64_t *exp_vf = nullptr, *exp_ee = nullptr;   // per face: stencil counts, then output offsets
    int2 *pairs;
    int4 *vfk, *eek;
    int32_t *opp;    // edge -> opposite face, 3F (kept with the vertex -> face table for refits)
    int32_t *rank;   // face id -> Morton-sorted position
    // level-synchronous refit (large trees): depth of every internal node, its pointer-jumping
    // scratch, and the maximum depth
    uint32_t *tdepth = nullptr, *tdepth2 = nullptr, *tjump = nullptr, *tjump2 = nullptr, *tmaxd = nullptr;
};
constexpr int NPARTS = 512;
constexpr int QPARTS = 4096;   // blocks of the quantisation range pass (enough to fill the GPU)

static int bp_buffers(ccd_ctx *ctx, int64_t F, int64_t V, bool batched, BPBufs &b)
{
    b.dop64 = (double *)buffer(ctx, "dop64", (size_t)F * NB * 8);
    b.centre = (double *)buffer(ctx, "centre", (size_t)F * 3 * 8);
    b.partial = (double *)buffer(ctx, "bpartial", (size_t)NPARTS * 6 * 8 + 64);
    b.qpartial = (double *)buffer(ctx, "qpartial", (size_t)QPARTS * NB * 8);
    b.qparam = (double *)buffer(ctx, "qparam", (size_t)NB * 8);
    b.keys = (uint64_t *)buffer(ctx, "mkeys", (size_t)F * 8);
    b.keys2 = (uint64_t *)buffer(ctx, "mkeys2", (size_t)F * 8);
    b.vals = (int32_t *)buffer(ctx, "mvals", (size_t)F * 4);
    b.order = (int32_t *)buffer(ctx, "order", (size_t)F * 4);
    b.face_inst = batched ? (int32_t *)buffer(ctx, "face_inst", (size_t)F * 4) : nullptr;
    b.left = (uint32_t *)buffer(ctx, "left", (size_t)F * 4);
    b.rf = (uint32_t *)buffer(ctx, "rfirst", (size_t)F * 4);
    b.rl = (uint32_t *)buffer(ctx, "rlast", (size_t)F * 4);
    b.parent = (uint32_t *)buffer(ctx, "parent", (size_t)(2 * F) * 4);
    b.escape = (uint32_t *)buffer(ctx, "escape", (size_t)(2 * F) * 4);
    b.arrivals = (uint32_t *)buffer(ctx, "arrivals", (size_t)F * 4);
    b.nodes = (NodeRec *)buffer(ctx, "nodes", (size_t)(2 * F) * sizeof(NodeRec));
    b.sdop64 = (double *)buffer(ctx, "sdop64", (size_t)F * NB * 8);
    b.sfaces = (int4 *)buffer(ctx, "sfaces", (size_t)F * 16);
    b.sort_tmp = buffer(ctx, "radix_tmp", radix_sort_temp_bytes(F));
    b.pairs = (int2 *)buffer(ctx, "pairs", (size_t)ctx->cap_pairs * 8);
    b.vfk = (int4 *)buffer(ctx, "vf_keys", (size_t)ctx->cap_vf * 16);
    b.eek = (int4 *)buffer(ctx, "ee_keys", (size_t)ctx->cap_ee * 16);
    b.opp = (int32_t *)buffer(ctx, "edge_opp", (size_t)(3 * F + 1) * 4);
    b.rank = (int32_t *)buffer(ctx, "face_rank", (size_t)F * 4);
    if (F >= LEVEL_REFIT_MIN_FACES) {
        b.tdepth = (uint32_t *)buffer(ctx, "tree_depth", (size_t)F * 4);
        b.tdepth2 = (uint32_t *)buffer(ctx, "tree_depth2", (size_t)F * 4);
        b.tjump = (uint32_t *)buffer(ctx, "tree_jump", (size_t)F * 4);
        b.tjump2 = (uint32_t *)buffer(ctx, "tree_jump2", (size_t)F * 4);
        b.tmaxd = (uint32_t *)buffer(ctx, "tree_maxd", 256);
        if (!b.tdepth || !b.tdepth2 || !b.tjump || !b.tjump2 || !b.tmaxd) return CCD_E_OOM;
    }
    if (2 * ctx->cap_pairs < ((int64_t)1 << 31) - 1) {   // int32 row offsets
        b.adj_cnt = (int32_t *)buffer(ctx, "adj_cnt", (size_t)(F + 1) * 4);
        b.adj_off = (int32_t *)buffer(ctx, "adj_off", (size_t)(F + 1) * 4);
        b.adj = (int32_t *)buffer(ctx, "adj", (size_t)(2 * ctx->cap_pairs + 1) * 4);
        b.adj_tmp = buffer(ctx, "adj_scan_tmp", scan_temp_bytes(F + 1));
        b.adj_mask = (uint16_t *)buffer(ctx, "adj_mask", (size_t)(2 * ctx->cap_pairs + 1) * 2);
        b.exp_vf = (int64_t *)buffer(ctx, "exp_vf", (size_t)(F + 1) * 8);
        b.exp_ee = (int64_t *)buffer(ctx, "exp_ee", (size_t)(F + 1) * 8);
        if (!b.adj_cnt || !b.adj_off || !b.adj || !b.adj_tmp || !b.adj_mask || !b.exp_vf || !b.exp_ee)
            return CCD_E_OOM;
    }
    if (!b.dop64 || !b.centre || !b.partial || !b.qpartial || !b.qparam || !b.keys || !b.keys2 || !b.vals || !b.order ||
        (batched && !b.face_inst) || !b.left || !b.rf || !b.rl || !b.parent || !b.escape || !b.arrivals ||
        !b.nodes || !b.sdop64 || !b.sfaces || !b.sort_tmp || !b.pairs || !b.vfk || !b.eek || !b.opp || !b.rank)
        return CCD_E_OOM;
    return CCD_OK;
}

// Related-face table from the traversal's pairs (device-side count), rows in fill order.
static int build_related_table(ccd_ctx *ctx, const BPBufs &b, int64_t F, MeshView &mv)
{
    if (!b.adj) return CCD_OK;
    hipStream_t st = ctx->stream;
    unsigned long long *c = (unsigned long long *)ctx->d_counts;
    int rc = fill_u32(ctx, (uint32_t *)b.adj_cnt, F + 1, 0);
    if (rc) return rc;
    unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(blocks_for(ctx->cap_pairs, 256), 8192));
    hipLaunchKernelGGL(k_adj_count, dim3(grid), dim3(256), 0, st, b.pairs, c + C_PAIRS, ctx->cap_pairs, F, b.adj_cnt);
    CCD_LAUNCH_CHECK(ctx);
    rc = exclusive_scan_i32(ctx, b.adj_cnt, b.adj_off, F + 1, b.adj_tmp, nullptr);
    if (rc) return rc;
    hipLaunchKernelGGL(k_adj_fill, dim3(grid), dim3(256), 0, st, b.pairs, c + C_PAIRS, ctx->cap_pairs, F, b.adj_off,
                       b.adj_cnt, b.adj);
    CCD_LAUNCH_CHECK(ctx);
    mv.adj_off = b.adj_off;
    mv.adj = b.adj;
    return CCD_OK;
}

// The node quantisation's range from the leaves' FP64 bounds (before the leaf records are written).
static int launch_qrange(ccd_ctx *ctx, const BPBufs &b, int64_t F)
{
    const int nb = (int)std::min<int64_t>(QPARTS, blocks_for(F, 256));
    hipLaunchKernelGGL(k_qrange_partial, dim3(nb), dim3(256), 0, ctx->stream, b.dop64, F, b.qpartial);
    CCD_LAUNCH_CHECK(ctx);
    hipLaunchKernelGGL(k_qrange_final, dim3(KAX), dim3(64), 0, ctx->stream, b.qpartial, nb, b.qparam);
    CCD_LAUNCH_CHECK(ctx);
    return CCD_OK;
}

// Interior bounds of the tree from its leaf records: the arrival-counter refit for small trees, the
// level-synchronous one for large trees (new_topology: the tree was just built, so its depths are
// computed first; a refit of the same tree keeps them).
static int refit_tree(ccd_ctx *ctx, const BPBufs &b, int64_t F, bool new_topology)
{
    hipStream_t st = ctx->stream;
    if (F < LEVEL_REFIT_MIN_FACES) {
        int rc = fill_u32(ctx, b.arrivals, F, 0);
        if (rc) return rc;
        hipLaunchKernelGGL(k_refit, dim3(blocks_for(F, 256)), dim3(256), 0, st, F, b.parent, b.left, b.escape,
                           b.nodes, b.arrivals);
        CCD_LAUNCH_CHECK(ctx);
        return CCD_OK;
    }
    const unsigned g = blocks_for(F - 1, 256);
    if (new_topology) {
        hipLaunchKernelGGL(k_depth_init, dim3(g), dim3(256), 0, st, F, b.parent, b.tjump, b.tdepth, b.tmaxd);
        CCD_LAUNCH_CHECK(ctx);
        for (int r = 0; r < DEPTH_ROUNDS; r++) {   // an even count: the result ends in (tjump, tdepth)
            const bool even = (r & 1) == 0;
            hipLaunchKernelGGL(k_depth_jump, dim3(g), dim3(256), 0, st, F, even ? b.tjump : b.tjump2,
                               even ? b.tdepth : b.tdepth2, even ? b.tjump2 : b.tjump, even ? b.tdepth2 : b.tdepth);
            CCD_LAUNCH_CHECK(ctx);
        }
        hipLaunchKernelGGL(k_depth_max, dim3(std::min<unsigned>(g, 1024)), dim3(256), 0, st, F, b.tdepth, b.tmaxd);
        CCD_LAUNCH_CHECK(ctx);
    }
    const unsigned gl = std::min<unsigned>(g, 2048);
    for (int level = MAX_TREE_DEPTH - 1; level >= 0; level--) {
        hipLaunchKernelGGL(k_refit_level, dim3(gl), dim3(256), 0, st, F, (uint32_t)level, b.tdepth, b.tmaxd, b.left,
                           b.escape, b.nodes);
        CCD_LAUNCH_CHECK(ctx);
    }
    return CCD_OK;
}

// Expansion: with a related-face table, sort its rows and run the row expansion (count, scan, write;
// the scans leave the unique stencil counts in the device counters); the per-pair kernel follows for
// raw emission, or does everything when there is no table.
static int launch_expand(ccd_ctx *ctx, const MeshView &mv, int64_t F, const BPBufs &b, unsigned grid)
{
    hipStream_t st = ctx->stream;
    unsigned long long *c = (unsigned long long *)ctx->d_counts;
    const bool rows = mv.adj != nullptr;
    if (rows) {
        const unsigned gs = (unsigned)std::max<int64_t>(1, std::min<int64_t>(F, 65536));
        hipLaunchKernelGGL(k_row_sort, dim3(gs), dim3(64), 0, st, c + C_BAD, F, mv.adj_off, b.adj);
        CCD_LAUNCH_CHECK(ctx);
        const unsigned gw = (unsigned)std::max<int64_t>(1, std::min<int64_t>(blocks_for(F, ROWS_BLOCK / 64), 65536));
        hipLaunchKernelGGL(k_expand_rows_count, dim3(gw), dim3(ROWS_BLOCK), 0, st, mv, F, c + C_BAD, b.adj_mask,
                           b.exp_vf, b.exp_ee);
        CCD_LAUNCH_CHECK(ctx);
        int rc = exclusive_scan_i64(ctx, b.exp_vf, b.exp_vf, F, b.adj_tmp, (int64_t *)(c + C_VF));
        if (!rc) rc = exclusive_scan_i64(ctx, b.exp_ee, b.exp_ee, F, b.adj_tmp, (int64_t *)(c + C_EE));
        if (rc) return rc;
        hipLaunchKernelGGL(k_expand_rows_write, dim3(gw), dim3(ROWS_BLOCK), 0, st, mv, F, c + C_BAD, b.adj_mask,
                           b.exp_vf, b.exp_ee, b.vfk, ctx->cap_vf, b.eek, ctx->cap_ee);
        CCD_LAUNCH_CHECK(ctx);
    }
    hipLaunchKernelGGL(k_expand, dim3(grid), dim3(EXPAND_BLOCK), 0, st, mv, rows, b.pairs, c + C_PAIRS, ctx->cap_pairs,
                       c + C_BAD, b.vfk, ctx->cap_vf, c + C_VF, b.eek, ctx->cap_ee, c + C_EE);
    CCD_LAUNCH_CHECK(ctx);
    return CCD_OK;
}

int broadphase_prepare(ccd_ctx *ctx, int64_t F, int64_t V, bool batched)
{
    // Capacities: grow-only, first guesses from the face count (the 256^2 cloth has 9.3 face pairs,
    // 21 vertex-face and 38 edge-edge stencils per face); an overflow is detected after the run.
    ctx->cap_pairs = std::max<int64_t>(ctx->cap_pairs, 16 * F + 1024);
    ctx->cap_vf = std::max<int64_t>(ctx->cap_vf, 24 * F + 1024);
    ctx->cap_ee = std::max<int64_t>(ctx->cap_ee, 40 * F + 1024);
    BPBufs b;
    int rc = bp_buffers(ctx, F, V, batched, b);
    if (rc) return rc;
    rc = build_vertex_face_table_buffers(ctx, F, V);
    return rc;
}

int broadphase_launch(ccd_ctx *ctx, const int32_t *faces, int64_t F, const double *qs, const double *qe,
                      const int64_t *hist_off, const double *hist_pos, int64_t V, double eta, const uint8_t *fixed,
                      bool aabb, const int64_t *inst_off, int ninst)
{
    hipStream_t st = ctx->stream;
    int64_t *cnt = ctx->d_counts;
    unsigned long long *c = (unsigned long long *)cnt;
    int rc = fill_u64(ctx, (uint64_t *)cnt, NCOUNTERS, 0);
    if (rc) return rc;
    if (F < 2) return CCD_OK;
    const bool batched = inst_off && ninst > 1;
    BPBufs b;
    rc = bp_buffers(ctx, F, V, batched, b);
    if (rc) return rc;
    static const Axes axes = make_axes();
    int inst_bits = 0;
    if (batched)
        while ((1ll << inst_bits) < ninst) inst_bits++;

    prof_begin(ctx, CCD_T_BOUNDS);
    hipLaunchKernelGGL(k_leaf_bounds, dim3(blocks_for(F, 256)), dim3(256), 0, st, faces, F, qs, qe, hist_off,
                       hist_pos, eta, axes, aabb ? 1 : 0, b.dop64, b.centre);
    CCD_LAUNCH_CHECK(ctx);
    int nb = (int)std::min<int64_t>(NPARTS, blocks_for(F, 256));
    hipLaunchKernelGGL(k_bounds_partial, dim3(nb), dim3(256), 0, st, b.centre, F, b.partial);
    CCD_LAUNCH_CHECK(ctx);
    double *box = b.partial + 6 * NPARTS;
    hipLaunchKernelGGL(k_bounds_final, dim3(1), dim3(256), 0, st, b.partial, nb, box);
    CCD_LAUNCH_CHECK(ctx);
    if (batched) {
        hipLaunchKernelGGL(k_face_instance, dim3(blocks_for(F, 256)), dim3(256), 0, st, inst_off, ninst, F,
                           b.face_inst);
        CCD_LAUNCH_CHECK(ctx);
    }
    hipLaunchKernelGGL(k_morton, dim3(blocks_for(F, 256)), dim3(256), 0, st, b.centre, F, box,
                       batched ? b.face_inst : (const int32_t *)nullptr, inst_bits, b.keys, b.vals);
    CCD_LAUNCH_CHECK(ctx);
    prof_end(ctx, CCD_T_BOUNDS);

    prof_begin(ctx, CCD_T_SORT);
    bool in_alt = false;
    rc = radix_sort_pairs(ctx, b.keys, b.vals, b.keys2, b.order, F, 63, b.sort_tmp, in_alt);
    if (rc) return rc;
    const uint64_t *skeys = in_alt ? b.keys2 : b.keys;
    const int32_t *order = in_alt ? b.order : b.vals;
    ctx->tree_order = order;
    prof_end(ctx, CCD_T_SORT);

    prof_begin(ctx, CCD_T_BUILD);
    hipLaunchKernelGGL(k_karras, dim3(blocks_for(F - 1, 256)), dim3(256), 0, st, skeys, F, b.left, b.rf, b.rl,
                       b.parent);
    CCD_LAUNCH_CHECK(ctx);
    hipLaunchKernelGGL(k_escape, dim3(blocks_for(2 * F - 1, 256)), dim3(256), 0, st, b.rf, b.rl, F, b.escape);
    CCD_LAUNCH_CHECK(ctx);
    rc = launch_qrange(ctx, b, F);
    if (rc) return rc;
    hipLaunchKernelGGL(k_leaf_records, dim3(blocks_for(F, 256)), dim3(256), 0, st, order, b.dop64, faces, F,
                       b.escape, b.qparam, b.nodes, b.sdop64, b.sfaces, b.rank);
    CCD_LAUNCH_CHECK(ctx);
    rc = refit_tree(ctx, b, F, true);
    if (rc) return rc;
    prof_end(ctx, CCD_T_BUILD);

    prof_begin(ctx, CCD_T_TABLES);
    rc = build_vertex_face_table(ctx, faces, F, V);
    if (rc) return rc;
    const int32_t *voff = nullptr, *vlist = nullptr;
    vertex_face_table(ctx, &voff, &vlist);
    hipLaunchKernelGGL(k_mesh_check, dim3(blocks_for(F, 256)), dim3(256), 0, st, faces, F, voff, vlist, c + C_BAD);
    CCD_LAUNCH_CHECK(ctx);
    hipLaunchKernelGGL(k_edge_opp, dim3(blocks_for(3 * F, 256)), dim3(256), 0, st, faces, F, voff, vlist, b.opp);
    CCD_LAUNCH_CHECK(ctx);
    prof_end(ctx, CCD_T_TABLES);

    prof_begin(ctx, CCD_T_TRAVERSE);
    hipLaunchKernelGGL(k_traverse, dim3(blocks_for(F, 256)), dim3(256), 0, st, F, b.nodes, b.sdop64, b.sfaces,
                       batched ? b.face_inst : (const int32_t *)nullptr, inst_off, b.pairs, ctx->cap_pairs,
                       c + C_PAIRS);
    CCD_LAUNCH_CHECK(ctx);
    prof_end(ctx, CCD_T_TRAVERSE);

    prof_begin(ctx, CCD_T_EXPAND);
    MeshView mv{faces, voff, vlist, b.dop64, fixed};
    mv.opp = b.opp;
    mv.rank = b.rank;
    mv.sdop64 = b.sdop64;
    mv.sfaces = b.sfaces;
    rc = build_related_table(ctx, b, F, mv);
    if (rc) return rc;
    unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(blocks_for(ctx->cap_pairs, 256), 16384));
    rc = launch_expand(ctx, mv, F, b, grid);
    if (rc) return rc;
    prof_end(ctx, CCD_T_EXPAND);
    return CCD_OK;
}

// Refit: the tree of the last full build on this context (same faces, same instance layout) with new
// positions -- new leaf bounds and records, then the bottom-up refit; the Morton order, the radix tree,
// the escape links and the vertex->face table are kept.  Candidate sets do not depend on the tree's
// shape, so this is exact; only culling efficiency depends on how far the mesh moved since the build.
int broadphase_refit_launch(ccd_ctx *ctx, const int32_t *faces, int64_t F, const double *qs, const double *qe,
                            int64_t V, double eta, const uint8_t *fixed, const int64_t *inst_off, int ninst)
{
    hipStream_t st = ctx->stream;
    int64_t *cnt = ctx->d_counts;
    unsigned long long *c = (unsigned long long *)cnt;
    int rc = fill_u64(ctx, (uint64_t *)cnt, NCOUNTERS, 0);
    if (rc) return rc;
    if (F < 2) return CCD_OK;
    const bool batched = inst_off && ninst > 1;
    BPBufs b;
    rc = bp_buffers(ctx, F, V, batched, b);
    if (rc) return rc;
    static const Axes axes = make_axes();
    prof_begin(ctx, CCD_T_BOUNDS);
    hipLaunchKernelGGL(k_leaf_bounds, dim3(blocks_for(F, 256)), dim3(256), 0, st, faces, F, qs, qe,
                       (const int64_t *)nullptr, (const double *)nullptr, eta, axes, 0, b.dop64, b.centre);
    CCD_LAUNCH_CHECK(ctx);
    prof_end(ctx, CCD_T_BOUNDS);
    prof_begin(ctx, CCD_T_BUILD);
    rc = launch_qrange(ctx, b, F);
    if (rc) return rc;
    hipLaunchKernelGGL(k_leaf_records, dim3(blocks_for(F, 256)), dim3(256), 0, st, ctx->tree_order, b.dop64, faces,
                       F, b.escape, b.qparam, b.nodes, b.sdop64, b.sfaces, b.rank);
    CCD_LAUNCH_CHECK(ctx);
    rc = refit_tree(ctx, b, F, false);
    if (rc) return rc;
    prof_end(ctx, CCD_T_BUILD);
    const int32_t *voff = nullptr, *vlist = nullptr;
    vertex_face_table(ctx, &voff, &vlist);
    prof_begin(ctx, CCD_T_TRAVERSE);
    hipLaunchKernelGGL(k_traverse, dim3(blocks_for(F, 256)), dim3(256), 0, st, F, b.nodes, b.sdop64, b.sfaces,
                       batched ? b.face_inst : (const int32_t *)nullptr, inst_off, b.pairs, ctx->cap_pairs,
                       c + C_PAIRS);
    CCD_LAUNCH_CHECK(ctx);
    prof_end(ctx, CCD_T_TRAVERSE);
    prof_begin(ctx, CCD_T_EXPAND);
    MeshView mv{faces, voff, vlist, b.dop64, fixed};
    mv.opp = b.opp;
    mv.rank = b.rank;
    mv.sdop64 = b.sdop64;
    mv.sfaces = b.sfaces;
    rc = build_related_table(ctx, b, F, mv);
    if (rc) return rc;
    unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(blocks_for(ctx->cap_pairs, 256), 16384));
    rc = launch_expand(ctx, mv, F, b, grid);
    if (rc) return rc;
    prof_end(ctx, CCD_T_EXPAND);
    return CCD_OK;
}

int broadphase(ccd_ctx *ctx, const int32_t *faces, int64_t F, const double *qs, const double *qe,
               const int64_t *hist_off, const double *hist_pos, int64_t V, double eta, const uint8_t *fixed,
               Candidates &out, bool aabb, const int64_t *inst_off, int ninst)
{
    out = Candidates();
    hipStream_t st = ctx->stream;
    int rc = broadphase_prepare(ctx, F, V, inst_off && ninst > 1);
    if (rc) return rc;
    for (int attempt = 0; attempt < 3; attempt++) {
        rc = broadphase_launch(ctx, faces, F, qs, qe, hist_off, hist_pos, V, eta, fixed, aabb, inst_off, ninst);
        if (rc) return rc;
        CCD_HIP_CHECK(ctx, hipMemcpyAsync(ctx->h_counts, ctx->d_counts, sizeof(int64_t) * NCOUNTERS,
                                          hipMemcpyDeviceToHost, st));
        CCD_HIP_CHECK(ctx, hipStreamSynchronize(st));
        int64_t P = ctx->h_counts[C_PAIRS], nv = ctx->h_counts[C_VF], ne = ctx->h_counts[C_EE];
        if (P <= ctx->cap_pairs && nv <= ctx->cap_vf && ne <= ctx->cap_ee) break;
        // Too small: grow to the counts seen (the stencil counts are exact only once the pairs fit).
        ctx->cap_pairs = std::max(ctx->cap_pairs, P + P / 8 + 1024);
        ctx->cap_vf = std::max(ctx->cap_vf, nv + nv / 8 + 1024);
        ctx->cap_ee = std::max(ctx->cap_ee, ne + ne / 8 + 1024);
        rc = broadphase_prepare(ctx, F, V, inst_off && ninst > 1);
        if (rc) return rc;
        if (attempt == 2) return set_error(ctx, CCD_E_CAPACITY, "broadphase buffers did not converge");
    }
    out.npairs = ctx->h_counts[C_PAIRS];
    out.nvf = ctx->h_counts[C_VF];
    out.nee = ctx->h_counts[C_EE];
    out.vf = (int32_t *)ctx->bufs["vf_keys"].ptr;
    out.ee = (int32_t *)ctx->bufs["ee_keys"].ptr;
    out.vf_cap = ctx->cap_vf;
    out.ee_cap = ctx->cap_ee;
    if (F >= 2 && ctx->h_counts[C_BAD] != 0) {   // raw emission: deduplicate by sorting
        int64_t *cnt = ctx->d_counts;
        rc = sort_unique(ctx, out.vf, out.nvf, "vf", cnt + C_VF);
        if (rc) return rc;
        rc = sort_unique(ctx, out.ee, out.nee, "ee", cnt + C_EE);
        if (rc) return rc;
        CCD_HIP_CHECK(ctx, hipMemcpyAsync(ctx->h_counts, cnt, sizeof(int64_t) * NCOUNTERS, hipMemcpyDeviceToHost, st));
        CCD_HIP_CHECK(ctx, hipStreamSynchronize(st));
        out.nvf = ctx->h_counts[C_VF];
        out.nee = ctx->h_counts[C_EE];
    }
    return CCD_OK;
}

}   // namespace ccdi
