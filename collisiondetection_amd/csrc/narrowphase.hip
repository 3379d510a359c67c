// Narrowphase kernels: one lane per stencil (or per primitive item), FP64, no fused multiply-add.
//
// Layout: positions are read straight from the caller's flat double[3V] arrays (24-byte vertex
// records); stencil keys are int4 (16 bytes, one dwordx4 load per lane).  Consecutive stencils share
// vertices (the broadphase emits them face pair by face pair), so the gathers hit L2 after the first
// touch.  The work per stencil is the reference's fixed sub-test order with early exit
// (CTCDNarrowPhase.cpp:43-69, :91-132), so lanes diverge; the kernel is FP64-VALU bound.
#include "ccd_core.h"
#include <algorithm>

#include "ccd_internal.h"

namespace {

using ccd::V3;

// Every kernel here that reaches the root finder asks for two wavefronts per SIMD.  The register
// allocator otherwise gives the reference-root kernels (and the out-of-line solver they share) up to
// 512 registers and one wavefront per SIMD, which leaves the FP64 pipeline idle on dependent
// instructions; at two wavefronts the extra spills cost less than the latency hiding gains (about 12%
// on the 256x256 cloth).  The attribute must be on every caller: the shared solver gets the tightest
// budget of its callers only when all of them carry it.
#define CCD_TWO_WAVES __attribute__((amdgpu_waves_per_eu(2)))

__device__ __forceinline__ V3 load3(const double *__restrict__ p, int v)
{
    const double *q = p + 3 * (int64_t)v;
    return ccd::mk(q[0], q[1], q[2]);
}

template <class M>
__global__ __launch_bounds__(256) CCD_TWO_WAVES void k_primitives(int kind, const double *__restrict__ q,
                                                    const double *__restrict__ eta, double eta_s, int64_t n,
                                                    uint8_t *__restrict__ hit, double *__restrict__ tout)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    double e = eta ? eta[i] : eta_s;
    double t = 0.0;
    bool r = false;
    if (kind == CCD_PRIM_VERTEX_FACE || kind == CCD_PRIM_EDGE_EDGE) {
        const double *p = q + 24 * i;
        V3 a[8];
#pragma unroll
        for (int k = 0; k < 8; k++) a[k] = ccd::mk(p[3 * k], p[3 * k + 1], p[3 * k + 2]);
        if (kind == CCD_PRIM_VERTEX_FACE)
            r = ccd::vertex_face(a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7], e, t, M());
        else
            r = ccd::edge_edge(a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7], e, t, M());
    } else if (kind == CCD_PRIM_VERTEX_EDGE) {
        const double *p = q + 18 * i;
        V3 a[6];
#pragma unroll
        for (int k = 0; k < 6; k++) a[k] = ccd::mk(p[3 * k], p[3 * k + 1], p[3 * k + 2]);
        r = ccd::vertex_edge(a[0], a[1], a[2], a[3], a[4], a[5], e, t, M());
    } else {
        const double *p = q + 12 * i;
        V3 a[4];
#pragma unroll
        for (int k = 0; k < 4; k++) a[k] = ccd::mk(p[3 * k], p[3 * k + 1], p[3 * k + 2]);
        r = ccd::vertex_vertex(a[0], a[1], a[2], a[3], e, t, M());
    }
    hit[i] = r ? 1 : 0;
    if (tout) tout[i] = r ? t : 0.0;
}

// Two-entry histories: window (stitch0(start), end), CTCDNarrowPhase.cpp:24-36 with History.cpp:127-133.
template <bool VF, class M>
__device__ __forceinline__ int stencil2(const double *__restrict__ qs, const double *__restrict__ qe, int4 k,
                                        double eta, double &t)
{
    int v[4] = {k.x, k.y, k.z, k.w};
    V3 s[4], e[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        V3 a = load3(qs, v[j]);
        V3 b = load3(qe, v[j]);
        s[j] = ccd::mk(ccd::stitch0(a.x, b.x), ccd::stitch0(a.y, b.y), ccd::stitch0(a.z, b.z));
        e[j] = b;
    }
    return VF ? ccd::check_vf_window(s, e, eta, t, M()) : ccd::check_ee_window(s, e, eta, t, M());
}

template <bool VF, class M>
__global__ __launch_bounds__(256) CCD_TWO_WAVES void k_narrow2(const double *__restrict__ qs, const double *__restrict__ qe,
                                                 const int4 *__restrict__ keys, const double *__restrict__ eta,
                                                 double eta_s, int64_t n, uint8_t *__restrict__ hit,
                                                 double *__restrict__ tout, uint8_t *__restrict__ sub,
                                                 const int32_t *__restrict__ order)
{
    const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int64_t i = order ? (int64_t)order[j] : j;   // lane j runs stencil order[j]
    double t = 0.0;
    int r = stencil2<VF, M>(qs, qe, keys[i], eta ? eta[i] : eta_s, t);
    hit[i] = r >= 0 ? 1 : 0;
    if (tout) tout[i] = r >= 0 ? t : 0.0;
    if (sub) sub[i] = r >= 0 ? (uint8_t)r : (uint8_t)255;
}

// Narrowphase + compaction of the colliding stencils.  The stencil count is read from the device
// (n_dev) so the launch needs no host round trip; the grid covers n_max.
template <bool VF, class M>
__global__ __launch_bounds__(256) CCD_TWO_WAVES void k_narrow2_compact(const double *__restrict__ qs, const double *__restrict__ qe,
                                                         const int4 *__restrict__ keys,
                                                         const int64_t *__restrict__ n_dev, int64_t n_max,
                                                         double eta, int4 *__restrict__ out_keys,
                                                         double *__restrict__ out_toi, int64_t cap,
                                                         unsigned long long *counter)
{
    // The device count may exceed the key buffer (an overflowing broadphase counts every stencil but
    // stores only n_max); never read past what was stored.
    int64_t n = *n_dev < n_max ? *n_dev : n_max;
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    bool hit = false;
    double t = 0.0;
    int4 k = make_int4(0, 0, 0, 0);
    if (i < n) {
        k = keys[i];
        hit = stencil2<VF, M>(qs, qe, k, eta, t) >= 0;
    }
    // Wave-aggregated append: one atomic per wavefront.
    uint64_t mask = __ballot(hit);
    if (mask == 0) return;
    int lane = __lane_id();
    int leader = __ffsll((unsigned long long)mask) - 1;
    unsigned long long base = 0;
    if (lane == leader) base = atomicAdd(counter, (unsigned long long)__popcll(mask));
    base = __shfl(base, leader);
    if (hit) {
        uint64_t below = mask & ((lane == 0) ? 0ull : (~0ull >> (64 - lane)));
        int64_t pos = (int64_t)base + __popcll(below);
        if (pos < cap) {
            out_keys[pos] = k;
            if (out_toi) out_toi[pos] = t;
        }
    }
}

// Persistent work-queue form of the compacting narrowphase: a grid sized to the machine, each
// wavefront repeatedly takes the next `tile` stencils with one atomic on *work and tests them 64 at a
// time.  Per-stencil cost varies by orders of magnitude (early exits vs. full root isolation), so
// dynamic tiles keep every SIMD busy to the end; tile and block size are tuning parameters.
// Three wavefronts per SIMD (168 registers; its spills cost less than the latency a third wavefront
// hides once the grid keeps 12 per CU resident: vertex-face 20.3 -> 19.1 ms on the batch).
template <bool VF, int BS, class M>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(3))) void k_narrow2_queue(const double *__restrict__ qs, const double *__restrict__ qe,
                                                      const int4 *__restrict__ keys, const int64_t *__restrict__ n_dev,
                                                      int64_t n_max, double eta, int4 *__restrict__ out_keys,
                                                      double *__restrict__ out_toi, int64_t cap,
                                                      unsigned long long *counter, unsigned long long *work, int tile)
{
    // Clamp to the stored keys: an overflowing broadphase counts more stencils than it stores.
    const int64_t n = *n_dev < n_max ? *n_dev : n_max;
    const int lane = __lane_id();
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    while (true) {
        unsigned long long start = 0;
        if (lane == 0) start = atomicAdd(work, (unsigned long long)tile);
        start = __shfl(start, 0);
        if ((int64_t)start >= n) break;
        int64_t end = (int64_t)start + tile < n ? (int64_t)start + tile : n;
        for (int64_t j0 = (int64_t)start; j0 < end; j0 += 64) {
            int64_t i = j0 + lane;
            bool hit = false;
            double t = 0.0;
            int4 k = make_int4(0, 0, 0, 0);
            if (i < end) {
                k = keys[i];
                hit = stencil2<VF, M>(qs, qe, k, eta, t) >= 0;
            }
            uint64_t mask = __ballot(hit);
            if (mask) {
                int leader = __ffsll((unsigned long long)mask) - 1;
                unsigned long long base = 0;
                if (lane == leader) base = atomicAdd(counter, (unsigned long long)__popcll(mask));
                base = __shfl(base, leader);
                if (hit) {
                    int64_t pos = (int64_t)base + __popcll(mask & lt);
                    if (pos < cap) {
                        out_keys[pos] = k;
                        if (out_toi) out_toi[pos] = t;
                    }
                }
            }
        }
    }
}

// Edge-edge queue kernel with the window's vertex-edge sub-tests shared within a wavefront.  After the
// edge-edge test, a stencil that it did not decide runs four vertex-edge tests, (s0; s2 s3),
// (s1; s2 s3), (s2; s0 s1), (s3; s0 s1) (CTCDNarrowPhase.cpp:104-115), then four vertex-vertex tests.
// A vertex-edge test depends only on its three vertices in call order (every stencil of a two-entry
// history has the same window), and neighbouring stencils share them: a wavefront's 64 consecutive
// stencils come from a few neighbouring face pairs, where each vertex meets the same edge in several
// stencils.  So the wavefront hashes its up to 256 tests in LDS by their vertex triples, runs each
// distinct one once (spread over the lanes), and every stencil then takes the first hit in its own
// order, then its vertex-vertex tests: the same sub-test, time and result as the sequential cascade.
// The deferred list is filled in blocks of DEF_BLOCK slots, each claimed by one wavefront with one
// atomic: an atomic per 64-stencil chunk on the list's single counter serialised the whole pass (a
// third of its time).  A wavefront marks the slots of a block it leaves unused DEF_HOLE, and the second
// pass skips them.  ccdi::ee_deferral_capacity bounds the slots this takes.
constexpr int DEF_BLOCK = 512;
constexpr uint32_t DEF_HOLE = 0xffffffffu;

constexpr int VE_TASKS = 256;   // 4 per lane
constexpr int VE_TAB = 512;     // hash slots (more than the tasks: probing always ends)
struct VeShare {
    int kv[VE_TASKS], ka[VE_TASKS], kb[VE_TASKS];   // task: vertex, edge ends (call order); kv -1: none
    int tab[VE_TAB];                                // hash table: a task index, -1 empty
    int rep[VE_TASKS];                              // each task's representative (the first with its key)
    int uniq[VE_TASKS];                             // representatives, compacted
    double t[VE_TASKS];                             // representative's time of impact
    unsigned char hit[VE_TASKS];                    // representative's result
};

// Orders a wavefront's LDS accesses (its own region; LDS instructions of one wavefront execute in
// order, so this only stops the compiler from moving them across).
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ void window_pos(const double *__restrict__ qs, const double *__restrict__ qe, int v, V3 &s,
                                           V3 &e)
{
    const V3 a = load3(qs, v), b = load3(qe, v);
    s = ccd::mk(ccd::stitch0(a.x, b.x), ccd::stitch0(a.y, b.y), ccd::stitch0(a.z, b.z));
    e = b;
}

//
// SCREEN: the first pass of the two-pass form.  The edge-edge test only screens its distance sextic
// (ccd::ScreenSextic): a stencil the screen cannot clear is appended to the deferred list, position i
// and the sub-test of its first vertex-edge / vertex-vertex hit (255: none), instead of being emitted;
// k_ee_deferred finishes it.  About a quarter of the stencils of a moving mesh are deferred, the ones
// whose sextics need the root solver, and they take most of the time; the first pass carries no solver
// for them (about half the registers, twice the resident wavefronts).
//
// The screening pass runs at four wavefronts per SIMD (128 registers, a few spills) and pays with a
// persistent grid of 16 wavefronts per CU (the 512-mesh batch's edge-edge stage: 71.0 -> 65.3 ms);
// the single-pass form stays at two.
template <int BS, class M, bool SCREEN>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(SCREEN ? 4 : 2))) void k_narrow2_queue_ee_shared(
    const double *__restrict__ qs, const double *__restrict__ qe, const int4 *__restrict__ keys,
    const int64_t *__restrict__ n_dev, int64_t n_max, double eta, int4 *__restrict__ out_keys,
    double *__restrict__ out_toi, int64_t cap, unsigned long long *counter, unsigned long long *work, int tile,
    uint32_t *__restrict__ def_idx, uint8_t *__restrict__ def_sub, unsigned long long *def_count, int64_t def_cap)
{
    __shared__ VeShare s_share[BS / 64];
    VeShare &L = s_share[threadIdx.x >> 6];
    const int64_t n = *n_dev < n_max ? *n_dev : n_max;
    const int lane = __lane_id();
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    // This wavefront's block of the deferred list (SCREEN): its first slot, and the slots used.
    int64_t blk_base = -1;
    int blk_used = DEF_BLOCK;
    while (true) {
        unsigned long long start = 0;
        if (lane == 0) start = atomicAdd(work, (unsigned long long)tile);
        start = __shfl(start, 0);
        if ((int64_t)start >= n) break;
        int64_t end = (int64_t)start + tile < n ? (int64_t)start + tile : n;
        for (int64_t j0 = (int64_t)start; j0 < end; j0 += 64) {
            const int64_t i = j0 + lane;
            const bool valid = i < end;
            int4 k = make_int4(0, 0, 0, 0);
            V3 s[4], e[4];
            double t = 0.0;
            int sub = -1;
            ccd::ScreenSextic ec;   // used when SCREEN
            if (valid) {
                k = keys[i];
                const int v[4] = {k.x, k.y, k.z, k.w};
#pragma unroll
                for (int j = 0; j < 4; j++) window_pos(qs, qe, v[j], s[j], e[j]);
                if constexpr (SCREEN)
                    ccd::edge_edge(s[0], s[1], s[2], s[3], e[0], e[1], e[2], e[3], eta, t, ec);
                else if (ccd::edge_edge(s[0], s[1], s[2], s[3], e[0], e[1], e[2], e[3], eta, t, M()))
                    sub = 0;
            }
            const bool deferred = SCREEN && ec.deferred;
            const bool need = valid && sub < 0;
            // The stencil's four vertex-edge tests as tasks r * 64 + lane.
            const int tv[4] = {k.x, k.y, k.z, k.w};
            for (int q = lane; q < VE_TAB; q += 64) L.tab[q] = -1;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int a = r < 2 ? 2 : 0;   // the edge: (s2, s3) for r = 0, 1; (s0, s1) for r = 2, 3
                const int j = r * 64 + lane;
                L.kv[j] = need ? tv[r] : -1;
                L.ka[j] = tv[a];
                L.kb[j] = tv[a + 1];
            }
            wave_lds_sync();
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int j = r * 64 + lane;
                int rep = -1;
                if (need) {
                    const int a = r < 2 ? 2 : 0;
                    const int kv = tv[r], ka = tv[a], kb = tv[a + 1];
                    unsigned h = ((unsigned)kv * 0x9E3779B1u) ^ ((unsigned)ka * 0x85EBCA6Bu) ^ ((unsigned)kb * 0xC2B2AE35u);
                    h = (h ^ (h >> 15)) & (VE_TAB - 1);
                    while (true) {
                        const int c = atomicCAS(&L.tab[h], -1, j);
                        if (c == -1) {
                            rep = j;
                            break;
                        }
                        if (L.kv[c] == kv && L.ka[c] == ka && L.kb[c] == kb) {
                            rep = c;
                            break;
                        }
                        h = (h + 1) & (VE_TAB - 1);
                    }
                }
                L.rep[j] = rep;
            }
            wave_lds_sync();
            int nu = 0;
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int j = r * 64 + lane;
                const bool own = need && L.rep[j] == j;
                const uint64_t m = __ballot(own);
                if (own) L.uniq[nu + __popcll(m & lt)] = j;
                nu += __popcll(m);
            }
            wave_lds_sync();
            for (int u = lane; u < nu; u += 64) {
                const int j = L.uniq[u];
                V3 ps, pe, as, ae, bs, be;
                window_pos(qs, qe, L.kv[j], ps, pe);
                window_pos(qs, qe, L.ka[j], as, ae);
                window_pos(qs, qe, L.kb[j], bs, be);
                double tt = 0.0;
                const bool h = ccd::vertex_edge(ps, as, bs, pe, ae, be, eta, tt, M());
                L.hit[j] = h ? 1 : 0;
                L.t[j] = tt;
            }
            wave_lds_sync();
            if (need) {
#pragma unroll
                for (int r = 0; r < 4; r++) {
                    const int rr = L.rep[r * 64 + lane];
                    if (sub < 0 && L.hit[rr]) {
                        sub = 1 + r;
                        t = L.t[rr];
                    }
                }
                if (sub < 0 && ccd::vertex_vertex(s[0], s[2], e[0], e[2], eta, t, M())) sub = 5;
                if (sub < 0 && ccd::vertex_vertex(s[0], s[3], e[0], e[3], eta, t, M())) sub = 6;
                if (sub < 0 && ccd::vertex_vertex(s[1], s[2], e[1], e[2], eta, t, M())) sub = 7;
                if (sub < 0 && ccd::vertex_vertex(s[1], s[3], e[1], e[3], eta, t, M())) sub = 8;
            }
            wave_lds_sync();   // the next chunk rewrites L
            const bool hit = sub >= 0 && !deferred;
            const uint64_t mask = __ballot(hit);
            if (mask) {
                const int leader = __ffsll((unsigned long long)mask) - 1;
                unsigned long long base = 0;
                if (lane == leader) base = atomicAdd(counter, (unsigned long long)__popcll(mask));
                base = __shfl(base, leader);
                if (hit) {
                    const int64_t pos = (int64_t)base + __popcll(mask & lt);
                    if (pos < cap) {
                        out_keys[pos] = k;
                        if (out_toi) out_toi[pos] = t;
                    }
                }
            }
            const uint64_t dmask = SCREEN ? __ballot(deferred) : 0ull;
            if (dmask) {
                const int nd = __popcll(dmask);
                if (blk_used + nd > DEF_BLOCK) {
                    // The rest of the current block becomes holes; a new block is claimed.
                    if (blk_base >= 0)
                        for (int q = blk_used + lane; q < DEF_BLOCK; q += 64)
                            if (blk_base + q < def_cap) def_idx[blk_base + q] = DEF_HOLE;
                    unsigned long long b = 0;
                    if (lane == 0) b = atomicAdd(def_count, (unsigned long long)DEF_BLOCK);
                    blk_base = (int64_t)__shfl(b, 0);
                    blk_used = 0;
                }
                if (deferred) {
                    // The host sizes the list for every stored stencil and the holes, so the guard never
                    // drops one.
                    const int64_t pos = blk_base + blk_used + __popcll(dmask & lt);
                    if (pos < def_cap) {
                        def_idx[pos] = (uint32_t)i;
                        def_sub[pos] = sub < 0 ? (uint8_t)255 : (uint8_t)sub;
                    }
                }
                blk_used += nd;
            }
        }
    }
    if (SCREEN && blk_base >= 0)
        for (int q = blk_used + lane; q < DEF_BLOCK; q += 64)
            if (blk_base + q < def_cap) def_idx[blk_base + q] = DEF_HOLE;
}

// Second pass of the two-pass edge-edge narrowphase: the deferred stencils' edge-edge tests in full
// (NoCount, the recursion included), on a persistent queue over the list.  A stencil whose edge-edge
// test hits takes its time (sub-test 0, first in CTCDNarrowPhase.cpp:91-132); otherwise the first
// pass's vertex-edge / vertex-vertex result stands, and its time comes from running that one sub-test
// again on the same positions (the same call as in the first pass, so the same value).
// Three wavefronts per SIMD (168 registers, a few spills): its root solves are chains of dependent FP64
// operations, and a third wavefront hides more of their latency than the spills cost (42.6 -> 36.4 ms
// on the batch; four wavefronts spill 636 bytes per lane and take 63.9 ms).
template <int BS>
__global__ __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(3))) void k_ee_deferred(
    const double *__restrict__ qs, const double *__restrict__ qe, const int4 *__restrict__ keys,
    const uint32_t *__restrict__ def_idx, const uint8_t *__restrict__ def_sub,
    const unsigned long long *__restrict__ def_count, int64_t def_cap, double eta, int4 *__restrict__ out_keys,
    double *__restrict__ out_toi, int64_t cap, unsigned long long *counter, unsigned long long *work, int tile)
{
    const int64_t n = (int64_t)*def_count < def_cap ? (int64_t)*def_count : def_cap;
    const int lane = __lane_id();
    const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    while (true) {
        unsigned long long start = 0;
        if (lane == 0) start = atomicAdd(work, (unsigned long long)tile);
        start = __shfl(start, 0);
        if ((int64_t)start >= n) break;
        const int64_t end = (int64_t)start + tile < n ? (int64_t)start + tile : n;
        for (int64_t j0 = (int64_t)start; j0 < end; j0 += 64) {
            const int64_t i = j0 + lane;
            bool hit = false;
            double t = 0.0;
            int4 k = make_int4(0, 0, 0, 0);
            const uint32_t di = i < end ? def_idx[i] : DEF_HOLE;
            if (di != DEF_HOLE) {
                k = keys[di];
                const int v[4] = {k.x, k.y, k.z, k.w};
                V3 s[4], e[4];
#pragma unroll
                for (int j = 0; j < 4; j++) window_pos(qs, qe, v[j], s[j], e[j]);
                hit = ccd::edge_edge(s[0], s[1], s[2], s[3], e[0], e[1], e[2], e[3], eta, t, ccd::NoCount());
                const int sub = def_sub[i];
                // The sub-test's vertices are read again by id: selecting them from s[] / e[] at a runtime
                // index would keep those arrays in scratch memory for every stencil.
                auto vid = [&](int j) { return j == 0 ? k.x : (j == 1 ? k.y : (j == 2 ? k.z : k.w)); };
                if (!hit && sub >= 1 && sub <= 4) {
                    // vertex-edge (v[r]; v[a], v[a + 1]), r = sub - 1, a = 2 for r < 2 else 0
                    const int r = sub - 1, a = r < 2 ? 2 : 0;
                    V3 ps, pe, as, ae, bs, be;
                    window_pos(qs, qe, vid(r), ps, pe);
                    window_pos(qs, qe, vid(a), as, ae);
                    window_pos(qs, qe, vid(a + 1), bs, be);
                    hit = ccd::vertex_edge(ps, as, bs, pe, ae, be, eta, t, ccd::NoCount());
                } else if (!hit && sub >= 5 && sub <= 8) {
                    // vertex-vertex (0, 2), (0, 3), (1, 2), (1, 3)
                    const int x = sub < 7 ? 0 : 1, y = (sub == 5 || sub == 7) ? 2 : 3;
                    V3 xs, xe, ys, ye;
                    window_pos(qs, qe, vid(x), xs, xe);
                    window_pos(qs, qe, vid(y), ys, ye);
                    hit = ccd::vertex_vertex(xs, ys, xe, ye, eta, t, ccd::NoCount());
                }
            }
            const uint64_t mask = __ballot(hit);
            if (mask) {
                const int leader = __ffsll((unsigned long long)mask) - 1;
                unsigned long long base = 0;
                if (lane == leader) base = atomicAdd(counter, (unsigned long long)__popcll(mask));
                base = __shfl(base, leader);
                if (hit) {
                    const int64_t pos = (int64_t)base + __popcll(mask & lt);
                    if (pos < cap) {
                        out_keys[pos] = k;
                        if (out_toi) out_toi[pos] = t;
                    }
                }
            }
        }
    }
}

// Multi-entry histories: stitch on the fly (History.cpp:98-140) and test window by window.
template <bool VF, class M>
__global__ __launch_bounds__(256) CCD_TWO_WAVES void k_narrow_hist(const int64_t *__restrict__ off, const double *__restrict__ ht,
                                                     const double *__restrict__ hp, const int4 *__restrict__ keys,
                                                     const double *__restrict__ eta, double eta_s, int64_t n,
                                                     uint8_t *__restrict__ hit, double *__restrict__ tout,
                                                     uint8_t *__restrict__ sub, int32_t *__restrict__ win)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    int4 k = keys[i];
    int v[4] = {k.x, k.y, k.z, k.w};
    double e = eta ? eta[i] : eta_s;
    int64_t its[4], ends[4];
    int64_t total = 0;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        its[j] = off[v[j]];
        ends[j] = off[v[j] + 1];
        total += ends[j] - its[j];
    }
    V3 prev[4], cur[4];
    double curtime = 0.0;
    int w = -1;
    int r = -1;
    double t = 0.0;
    // Each pass advances curtime strictly (entries at or before it are consumed), so the number of
    // stitched entries is at most the total entry count; the bound also guards malformed input.
    for (int64_t guard = 0; guard <= total && curtime <= 1.0; guard++) {
        double newtime = __builtin_inf();
#pragma unroll
        for (int j = 0; j < 4; j++) {
            int64_t next = its[j] + 1;
            while (next != ends[j] && ht[next] <= curtime) {
                ++next;
                ++its[j];
            }
            V3 oldp = ccd::mk(hp[3 * its[j]], hp[3 * its[j] + 1], hp[3 * its[j] + 2]);
            if (next == ends[j]) {
                cur[j] = oldp;
            } else {
                V3 newp = ccd::mk(hp[3 * next], hp[3 * next + 1], hp[3 * next + 2]);
                double dt = ht[next] - ht[its[j]];
                double a = curtime - ht[its[j]];
                double b = ht[next] - curtime;
                cur[j] = ccd::mk((b * oldp.x + a * newp.x) / dt, (b * oldp.y + a * newp.y) / dt,
                                 (b * oldp.z + a * newp.z) / dt);
                newtime = ccd::smin(newtime, ht[next]);
            }
        }
        if (w >= 0) {
            r = VF ? ccd::check_vf_window(prev, cur, e, t, M()) : ccd::check_ee_window(prev, cur, e, t, M());
            if (r >= 0) break;
        }
#pragma unroll
        for (int j = 0; j < 4; j++) prev[j] = cur[j];
        w++;
        curtime = newtime;
    }
    hit[i] = r >= 0 ? 1 : 0;
    if (tout) tout[i] = r >= 0 ? t : 0.0;
    if (sub) sub[i] = r >= 0 ? (uint8_t)r : (uint8_t)255;
    if (win) win[i] = r >= 0 ? w : -1;
}

__global__ void k_validate_keys(const int32_t *__restrict__ keys, int64_t n4, int64_t nverts,
                                unsigned long long *bad)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    bool b = false;
    if (i < n4) {
        int32_t v = keys[i];
        b = v < 0 || v >= nverts;
    }
    uint64_t m = __ballot(b);
    if (m && __lane_id() == __ffsll((unsigned long long)m) - 1) atomicAdd(bad, (unsigned long long)__popcll(m));
}

// A history is valid when off[] is nondecreasing within [0, nentries], each vertex has at least
// one entry, its times start at 0, end at 1 and never decrease.
__global__ void k_validate_history(const int64_t *__restrict__ off, const double *__restrict__ ht, int64_t nverts,
                                   int64_t nentries, unsigned long long *bad)
{
    int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    bool b = false;
    if (v < nverts) {
        int64_t a = off[v], z = off[v + 1];
        if (a < 0 || z > nentries || z <= a) {
            b = true;
        } else {
            if (ht[a] != 0.0 || ht[z - 1] != 1.0) b = true;
            for (int64_t e = a + 1; e < z && !b; e++)
                if (!(ht[e] >= ht[e - 1])) b = true;
        }
    }
    uint64_t m = __ballot(b);
    if (m && __lane_id() == __ffsll((unsigned long long)m) - 1) atomicAdd(bad, (unsigned long long)__popcll(m));
}

// Grouping of stencils by which of their first polynomials reach the root solver
// (ccd::first_phase_signature, the major key) and by the vertex-edge quartics their windows will solve
// (ccd::ve_signature): the narrowphase then runs on the keys in signature order (stable, so each group
// keeps the broadphase's spatial order), and a wavefront's lanes mostly take the same solver calls.
template <bool VF>
__global__ __launch_bounds__(256) void k_stencil_signature(const double *__restrict__ qs, const double *__restrict__ qe,
                                                           const int4 *__restrict__ keys, int64_t n, double eta,
                                                           const double *__restrict__ eta_arr,
                                                           const int64_t *__restrict__ n_dev,
                                                           uint64_t *__restrict__ sig, int32_t *__restrict__ idx)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (n_dev && i >= *n_dev) {   // beyond the stored stencils (count on the device): last, keys not read
        sig[i] = 0xff;             // above every signature (vertex-face <= 0xf7, edge-edge <= 0x9f)
        idx[i] = (int32_t)i;
        return;
    }
    int4 k = keys[i];
    int v[4] = {k.x, k.y, k.z, k.w};
    V3 s[4], e[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        V3 a = load3(qs, v[j]);
        V3 b = load3(qe, v[j]);
        s[j] = ccd::mk(ccd::stitch0(a.x, b.x), ccd::stitch0(a.y, b.y), ccd::stitch0(a.z, b.z));
        e[j] = b;
    }
    sig[i] = ccd::ve_signature(VF, s, e) | (ccd::first_phase_signature(VF, s, e, eta_arr ? eta_arr[i] : eta) << 4);
    idx[i] = (int32_t)i;
}
__global__ __launch_bounds__(256) void k_gather_keys(const int4 *__restrict__ keys, const int32_t *__restrict__ perm,
                                                     int64_t n, int4 *__restrict__ out)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) out[i] = keys[perm[i]];
}

// History::computeMinimumGap (History.cpp:78-93): min(1, every gap between consecutive entry times),
// reduced with an atomic min on the bit pattern (the gaps of a validated history are non-negative, so
// their bits order as the values do).  *gap_bits must hold the bits of 1.0 on entry.
__global__ __launch_bounds__(256) void k_min_gap(const int64_t *__restrict__ off, const double *__restrict__ ht,
                                                 int64_t nverts, unsigned long long *gap_bits)
{
    double g = 1.0;
    for (int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; v < nverts;
         v += (int64_t)gridDim.x * blockDim.x)
        for (int64_t j = off[v] + 1; j < off[v + 1]; j++) g = ccd::smin(ht[j] - ht[j - 1], g);
    for (int o = 32; o > 0; o >>= 1) g = ccd::smin(__shfl_xor(g, o), g);
    if (__lane_id() == 0) atomicMin(gap_bits, (unsigned long long)ccd::double_to_bits(g));
}
__global__ void k_set_bits(unsigned long long *p, unsigned long long v) { *p = v; }

// SeparatingPlaneNarrowPhase (SeparatingPlaneNarrowPhase.cpp:9-278): one lane per stencil, status 1
// (collision), 0 (none) or 2 (undecided: the reference's recursion would not end; counted in *undecided).
template <bool VF, class M>
__global__ __launch_bounds__(256) CCD_TWO_WAVES void k_separating_plane(
    const int64_t *__restrict__ off, const double *__restrict__ ht, const double *__restrict__ hp,
    const int4 *__restrict__ keys, const double *__restrict__ eta, double eta_s, int64_t n,
    const unsigned long long *__restrict__ gap_bits, uint8_t *__restrict__ status, unsigned long long *undecided,
    const int32_t *__restrict__ order)
{
    const int64_t j = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    int r = 0;
    if (j < n) {
        const int64_t i = order ? (int64_t)order[j] : j;   // lane j runs stencil order[j]
        const double eps = ccd::bits_to_double((long long)*gap_bits) / 4.0;
        int4 k = keys[i];
        int v[4] = {k.x, k.y, k.z, k.w};
        ccd::HistCSR h{off, ht, hp};
        r = ccd::separating_plane_check<VF>(h, v, eta ? eta[i] : eta_s, eps, M());
        status[i] = (uint8_t)r;
    }
    uint64_t m = __ballot(r == 2);
    if (m && __lane_id() == __ffsll((unsigned long long)m) - 1) atomicAdd(undecided, (unsigned long long)__popcll(m));
}

// Sort key for the separating-plane narrowphase: 1 if a stencil's root interval [0, 1] does not decide
// it (its recursion goes on; with two-entry histories that means one or two CTCD window tests), else 0.
// Most stencils are decided at the root; grouping the others keeps a wavefront's lanes on the same path.
template <bool VF>
__global__ __launch_bounds__(256) void k_sp_order_key(const int64_t *__restrict__ off, const double *__restrict__ ht,
                                                      const double *__restrict__ hp, const int4 *__restrict__ keys,
                                                      const double *__restrict__ eta, double eta_s, int64_t n,
                                                      const unsigned long long *__restrict__ gap_bits,
                                                      uint64_t *__restrict__ key, int32_t *__restrict__ idx)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double eps = ccd::bits_to_double((long long)*gap_bits) / 4.0;
    const int4 k = keys[i];
    const int v[4] = {k.x, k.y, k.z, k.w};
    ccd::HistCSR h{off, ht, hp};
    const int r = ccd::separating_plane_check<VF>(h, v, eta ? eta[i] : eta_s, eps, ccd::FullCascade(), ccd::SeqScan(), 1);
    key[i] = r == 3 ? 1u : 0u;
    idx[i] = (int32_t)i;
}

// The narrowphase over stitched history windows (k_narrow_hist) with one wavefront per stencil, for
// long histories.  The stitching (History.cpp:98-140) is a merge of the four vertices' timelines; every
// lane runs the same merge (uniform control flow, identical values) 65 steps at a time and keeps the
// stitched positions of two consecutive steps, so lane l holds window W0 + l of a 64-window chunk and
// tests it.  The first window in order that collides (the lowest set bit of a ballot, in the first
// chunk with any) gives the answer, as the sequential loop's early exit does.
template <bool VF, class M>
__global__ __launch_bounds__(64) CCD_TWO_WAVES void k_narrow_hist_wave(
    const int64_t *__restrict__ off, const double *__restrict__ ht, const double *__restrict__ hp,
    const int4 *__restrict__ keys, const double *__restrict__ eta, double eta_s, int64_t n, uint8_t *__restrict__ hit,
    double *__restrict__ tout, uint8_t *__restrict__ sub, int32_t *__restrict__ win)
{
    using ccd::V3;
    const int64_t i = blockIdx.x;
    if (i >= n) return;
    const int lane = (int)__lane_id();
    const int4 k = keys[i];
    const int v[4] = {k.x, k.y, k.z, k.w};
    const double e = eta ? eta[i] : eta_s;
    int64_t its[4], ends[4];
    int64_t total = 0;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        its[j] = off[v[j]];
        ends[j] = off[v[j] + 1];
        total += ends[j] - its[j];
    }
    double curtime = 0.0;
    int64_t g = 0;   // step index of the sequential loop (its guard); cur is evaluated at curtime
    int r = -1, w = -1;
    double t = 0.0;
    bool alive = true;
    while (alive) {
        const int64_t w0 = g;   // this chunk: windows w0 .. w0 + 63, window w = (step w, step w + 1)
        V3 pa[4], pb[4];
        bool va = false, vb = false;
        for (int s = 0; s <= 64; s++) {
            alive = g <= total && curtime <= 1.0;
            if (!alive) break;
            V3 cur[4];
            double newtime = __builtin_inf();
#pragma unroll
            for (int j = 0; j < 4; j++) {
                int64_t next = its[j] + 1;
                while (next != ends[j] && ht[next] <= curtime) {
                    ++next;
                    ++its[j];
                }
                V3 oldp = ccd::mk(hp[3 * its[j]], hp[3 * its[j] + 1], hp[3 * its[j] + 2]);
                if (next == ends[j]) {
                    cur[j] = oldp;
                } else {
                    V3 newp = ccd::mk(hp[3 * next], hp[3 * next + 1], hp[3 * next + 2]);
                    double dt = ht[next] - ht[its[j]];
                    double a = curtime - ht[its[j]];
                    double b = ht[next] - curtime;
                    cur[j] = ccd::mk((b * oldp.x + a * newp.x) / dt, (b * oldp.y + a * newp.y) / dt,
                                     (b * oldp.z + a * newp.z) / dt);
                    newtime = ccd::smin(newtime, ht[next]);
                }
            }
            if (s == lane) {
#pragma unroll
                for (int j = 0; j < 4; j++) pa[j] = cur[j];
                va = true;
            }
            if (s == lane + 1) {
#pragma unroll
                for (int j = 0; j < 4; j++) pb[j] = cur[j];
                vb = true;
            }
            if (s == 64) break;   // the next chunk starts by evaluating this step again (idempotent)
            curtime = newtime;
            g++;
        }
        bool h = false;
        int rl = -1;
        double tl = 0.0;
        if (va && vb) {
            rl = VF ? ccd::check_vf_window(pa, pb, e, tl, M()) : ccd::check_ee_window(pa, pb, e, tl, M());
            h = rl >= 0;
        }
        const unsigned long long m = __ballot(h);
        if (m) {
            const int f = __ffsll(m) - 1;
            r = __shfl(rl, f);
            t = __shfl(tl, f);
            w = (int)(w0 + f);
            break;
        }
    }
    if (lane == 0) {
        hit[i] = r >= 0 ? 1 : 0;
        if (tout) tout[i] = r >= 0 ? t : 0.0;
        if (sub) sub[i] = r >= 0 ? (uint8_t)r : (uint8_t)255;
        if (win) win[i] = r >= 0 ? w : -1;
    }
}

// The trajectory scan of planeTrajectoryIntersect (SeparatingPlaneNarrowPhase.cpp:216-265) split over
// the 64 lanes of a wavefront that all run the same stencil: each lane tests one history segment of a
// 64-segment chunk, and the first segment in scan order that meets the plane (the lowest lane whose
// test passes, in the first chunk with any) gives the result, computed by that lane with the
// sequential scan's own expressions.  Needs the whole wavefront active and on the same call.
struct WaveScan {
    __device__ double operator()(const ccd::HistCSR &h, int v, int start, bool forward, ccd::V3 ppos, ccd::V3 pvel,
                                 ccd::V3 nrm, ccd::V3 ptstart, double tstart, double eta) const
    {
        using ccd::V3;
        const double inf = ccd::bits_to_double(0x7ff0000000000000ll);
        const int64_t b = h.off[v];
        const int n = (int)(h.off[v + 1] - b);
        const double dt = forward ? h.t[b + start] - tstart : tstart - h.t[b + start];
        const double t = ccd::plane_intersect(ppos, (forward ? 1.0 : -1.0) * pvel, nrm, ptstart,
                                              ccd::hist_pos(h, b + start), dt, eta);
        if (t >= 0 && t <= dt) return t;
        const int lane = (int)__lane_id();
        if (forward) {
            // segments (i, i + 1) for i = start .. n - 2, in increasing i
            for (int base = start; base < n - 1; base += 64) {
                const int i = base + lane;
                bool hit = false;
                double r = 0.0;
                if (i < n - 1) {
                    V3 npp = ppos + (h.t[b + i] - tstart) * pvel;
                    double d = h.t[b + i + 1] - h.t[b + i];
                    double u = ccd::plane_intersect(npp, pvel, nrm, ccd::hist_pos(h, b + i), ccd::hist_pos(h, b + i + 1),
                                                    d, eta);
                    hit = u >= 0 && u <= d;
                    r = h.t[b + i] + u - tstart;
                }
                const unsigned long long m = __ballot(hit);
                if (m) return __shfl(r, __ffsll(m) - 1);
            }
        } else {
            // segments (i, i - 1) for i = start .. 1, in decreasing i
            for (int base = start; base >= 1; base -= 64) {
                const int i = base - lane;
                bool hit = false;
                double r = 0.0;
                if (i >= 1) {
                    V3 npp = ppos - (tstart - h.t[b + i]) * pvel;
                    double d = h.t[b + i] - h.t[b + i - 1];
                    double u = ccd::plane_intersect(npp, -1.0 * pvel, nrm, ccd::hist_pos(h, b + i),
                                                    ccd::hist_pos(h, b + i - 1), d, eta);
                    hit = u >= 0 && u <= d;
                    r = tstart - (h.t[b + i] - u);
                }
                const unsigned long long m = __ballot(hit);
                if (m) return __shfl(r, __ffsll(m) - 1);
            }
        }
        return inf;
    }
};

// SeparatingPlaneNarrowPhase with one wavefront per stencil, for long histories: the lanes share the
// trajectory scans (WaveScan) and otherwise compute the same values, so control flow stays uniform.
template <bool VF, class M>
__global__ __launch_bounds__(64) CCD_TWO_WAVES void k_separating_plane_wave(
    const int64_t *__restrict__ off, const double *__restrict__ ht, const double *__restrict__ hp,
    const int4 *__restrict__ keys, const double *__restrict__ eta, double eta_s, int64_t n,
    const unsigned long long *__restrict__ gap_bits, uint8_t *__restrict__ status, unsigned long long *undecided)
{
    const int64_t i = blockIdx.x;
    if (i >= n) return;
    const double eps = ccd::bits_to_double((long long)*gap_bits) / 4.0;
    const int4 k = keys[i];
    const int v[4] = {k.x, k.y, k.z, k.w};
    ccd::HistCSR h{off, ht, hp};
    const int r = ccd::separating_plane_check<VF>(h, v, eta ? eta[i] : eta_s, eps, M(), WaveScan());
    if (__lane_id() == 0) {
        status[i] = (uint8_t)r;
        if (r == 2) atomicAdd(undecided, 1ull);
    }
}

inline unsigned blocks_for(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// One wavefront per block for the kernels that give each lane one stencil of very uneven cost (the
// separating-plane recursion, the plain narrowphases): a wavefront's slot is refilled as soon as it
// finishes instead of when the slowest of its block's wavefronts does.  Separating plane on the 256^2
// cloth candidates: 92 ms with 256-thread blocks, 73 ms with 64.
constexpr int LANE_BLOCK = 64;

}   // namespace

namespace ccdi {

// Every narrowphase launch instantiates its kernel with the product's counter type: no work
// accounting, cascades pruned once no interval combination can overlap.
#define CCD_PRODUCT_LAUNCH(LAUNCH)                                                                                  \
    do {                                                                                                            \
        using M = ccd::NoCount;                                                                                     \
        LAUNCH;                                                                                                     \
    } while (0)

int launch_primitives(ccd_ctx *ctx, int kind, const double *q, const double *eta, double eta_s, int64_t n,
                      uint8_t *hit, double *t)
{
    if (n == 0) return CCD_OK;
    CCD_PRODUCT_LAUNCH(hipLaunchKernelGGL(k_primitives<M>, dim3(blocks_for(n, LANE_BLOCK)), dim3(LANE_BLOCK), 0, ctx->stream,
                                              kind, q, eta, eta_s, n, hit, t));
    CCD_LAUNCH_CHECK(ctx);
    return CCD_OK;
}

int launch_narrow2(ccd_ctx *ctx, bool vf, const double *qs, const double *qe, const int32_t *keys,
                   const double *eta, double eta_s, int64_t n, uint8_t *hit, double *t, uint8_t *sub,
                   const OrderScratch *order_scratch)
{
    if (n == 0) return CCD_OK;
    // Run the stencils in the grouping order of the detection pipeline (k_stencil_signature); each
    // result lands at its stencil's own index, so only the speed depends on the order.  The ordering
    // buffers are the caller's (order_scratch) or the context's own.
    const int32_t *order = nullptr;
    // (the order holds int32 stencil indices: launches past 2^31 - 1 stencils run in the given order)
    if (n >= 4096 && n <= 0x7fffffffll && ctx->group_stencils && (vf ? ctx->vf_shape : ctx->ee_shape).group) {
        OrderScratch os;
        if (order_scratch) {
            os = *order_scratch;
        } else {
            os.sig = (uint64_t *)buffer(ctx, "n2o_sig", (size_t)n * 8);
            os.sig2 = (uint64_t *)buffer(ctx, "n2o_sig2", (size_t)n * 8);
            os.idx = (int32_t *)buffer(ctx, "n2o_idx", (size_t)n * 4);
            os.idx2 = (int32_t *)buffer(ctx, "n2o_idx2", (size_t)n * 4);
            os.tmp = buffer(ctx, "n2o_tmp", radix_sort_temp_bytes(n));
        }
        uint64_t *sig = os.sig, *sig2 = os.sig2;
        int32_t *idx = os.idx, *idx2 = os.idx2;
        void *tmp = os.tmp;
        if (!sig || !sig2 || !idx || !idx2 || !tmp) return CCD_E_OOM;
        if (vf)
            hipLaunchKernelGGL(k_stencil_signature<true>, dim3(blocks_for(n, 256)), dim3(256), 0, ctx->stream, qs, qe,
                               (const int4 *)keys, n, eta_s, eta, nullptr, sig, idx);
        else
            hipLaunchKernelGGL(k_stencil_signature<false>, dim3(blocks_for(n, 256)), dim3(256), 0, ctx->stream, qs, qe,
                               (const int4 *)keys, n, eta_s, eta, nullptr, sig, idx);
        CCD_LAUNCH_CHECK(ctx);
        bool in_alt = false;
        int rc = radix_sort_pairs(ctx, sig, idx, sig2, idx2, n, 8, tmp, in_alt);   // one stable 8-bit pass
        if (rc) return rc;
        order = in_alt ? idx2 : idx;
    }
    if (vf)
        CCD_PRODUCT_LAUNCH(hipLaunchKernelGGL((k_narrow2<true, M>), dim3(blocks_for(n, LANE_BLOCK)), dim3(LANE_BLOCK), 0,
                                                  ctx->stream, qs, qe, (const int4 *)keys, eta, eta_s, n, hit, t, sub,
                                                  order));
    else
        CCD_PRODUCT_LAUNCH(hipLaunchKernelGGL((k_narrow2<false, M>), dim3(blocks_for(n, LANE_BLOCK)), dim3(LANE_BLOCK), 0,
                                                  ctx->stream, qs, qe, (const int4 *)keys, eta, eta_s, n, hit, t, sub,
                                                  order));
    CCD_LAUNCH_CHECK(ctx);
    return CCD_OK;
}

int launch_narrow2_compact(ccd_ctx *ctx, bool vf, const double *qs, const double *qe, const int32_t *keys,
                           const int64_t *n_dev, int64_t n_max, double eta, int32_t *out_keys, double *out_toi,
                           int64_t cap, int64_t *counter)
{
    if (n_max == 0) return CCD_OK;
    if (vf)
        CCD_PRODUCT_LAUNCH(hipLaunchKernelGGL((k_narrow2_compact<true, M>), dim3(blocks_for(n_max, 256)),
                                                  dim3(256), 0, ctx->stream, qs, qe, (const int4 *)keys, n_dev, n_max,
                                                  eta, (int4 *)out_keys, out_toi, cap, (unsigned long long *)counter));
    else
        CCD_PRODUCT_LAUNCH(hipLaunchKernelGGL((k_narrow2_compact<false, M>), dim3(blocks_for(n_max, 256)),
                                                  dim3(256), 0, ctx->stream, qs, qe, (const int4 *)keys, n_dev, n_max,
                                                  eta, (int4 *)out_keys, out_toi, cap, (unsigned long long *)counter));
    CCD_LAUNCH_CHECK(ctx);
    return CCD_OK;
}

template <bool VF, class M>
static void launch_queue_bs(ccd_ctx *ctx, int bs, unsigned grid, const double *qs, const double *qe,
                            const int32_t *keys, const int64_t *n_dev, int64_t n_max, double eta, int32_t *out_keys,
                            double *out_toi, int64_t cap, int64_t *counter, unsigned long long *work, int tile,
                            const EeDeferral *def)
{
    // Edge-edge stencils share their vertex-edge tests within a wavefront (k_narrow2_queue_ee_shared).
#define CCD_QUEUE_LAUNCH(B)                                                                                         \
    if (VF)                                                                                                         \
        hipLaunchKernelGGL((k_narrow2_queue<true, B, M>), dim3(grid), dim3(B), 0, ctx->stream, qs, qe,           \
                           (const int4 *)keys, n_dev, n_max, eta, (int4 *)out_keys, out_toi, cap,                  \
                           (unsigned long long *)counter, work, tile);                                            \
    else if (def)                                                                                                   \
        hipLaunchKernelGGL((k_narrow2_queue_ee_shared<B, M, true>), dim3(grid), dim3(B), 0, ctx->stream, qs, qe, \
                           (const int4 *)keys, n_dev, n_max, eta, (int4 *)out_keys, out_toi, cap,                  \
                           (unsigned long long *)counter, work, tile, def->idx, def->sub, def->count, def->cap);   \
    else                                                                                                            \
        hipLaunchKernelGGL((k_narrow2_queue_ee_shared<B, M, false>), dim3(grid), dim3(B), 0, ctx->stream, qs, qe,\
                           (const int4 *)keys, n_dev, n_max, eta, (int4 *)out_keys, out_toi, cap,                  \
                           (unsigned long long *)counter, work, tile, nullptr, nullptr, nullptr, 0)
    if (bs == 64)
        CCD_QUEUE_LAUNCH(64);
    else if (bs == 128)
        CCD_QUEUE_LAUNCH(128);
    else
        CCD_QUEUE_LAUNCH(256);
#undef CCD_QUEUE_LAUNCH
}

int launch_narrow2_queue(ccd_ctx *ctx, bool vf, const double *qs, const double *qe, const int32_t *keys,
                         const int64_t *n_dev, int64_t n_max, double eta, int32_t *out_keys, double *out_toi,
                         int64_t cap, int64_t *counter, unsigned long long *work, const QueueShape &shape,
                         const EeDeferral *def)
{
    if (n_max == 0) return CCD_OK;
    int bs = shape.block == 64 || shape.block == 128 ? shape.block : 256;
    int64_t waves = (int64_t)ctx->num_cus * shape.waves_per_cu;
    int64_t need = (n_max + shape.tile - 1) / shape.tile;   // no more waves than tiles
    waves = std::max<int64_t>(1, std::min(waves, need));
    unsigned grid = (unsigned)((waves * 64 + bs - 1) / bs);
    // Two passes only with a list that holds every stored stencil (int32 positions).
    const bool split = !vf && def && def->idx && def->sub && def->count && def->work &&
                       ee_deferral_capacity(ctx, n_max) <= def->cap && n_max <= 0x7fffffffll;
    if (vf)
        CCD_PRODUCT_LAUNCH((launch_queue_bs<true, M>(ctx, bs, grid, qs, qe, keys, n_dev, n_max, eta, out_keys,
                                                         out_toi, cap, counter, work, shape.tile, nullptr)));
    else
        CCD_PRODUCT_LAUNCH((launch_queue_bs<false, M>(ctx, bs, grid, qs, qe, keys, n_dev, n_max, eta, out_keys,
                                                          out_toi, cap, counter, work, shape.tile,
                                                          split ? def : nullptr)));
    CCD_LAUNCH_CHECK(ctx);
    if (split) {
        // The deferred stencils cost about ten times the others: small tiles keep the tail short, and
        // 256 keeps the claims on the queue counter few.
        // Three wavefronts per SIMD, what the kernel's registers are set for.
        const int64_t w2 = (int64_t)ctx->num_cus * 12;
        hipLaunchKernelGGL(k_ee_deferred<64>, dim3((unsigned)std::max<int64_t>(1, w2)), dim3(64), 0, ctx->stream, qs,
                           qe, (const int4 *)keys, def->idx, def->sub, def->count, def->cap, eta, (int4 *)out_keys,
                           out_toi, cap, (unsigned long long *)counter, def->work, 256);
        CCD_LAUNCH_CHECK(ctx);
    }
    return CCD_OK;
}

int launch_narrow_hist(ccd_ctx *ctx, bool vf, const int64_t *off, const double *ht, const double *hp,
                       const int32_t *keys, const double *eta, double eta_s, int64_t n, uint8_t *hit, double *t,
                       uint8_t *sub, int32_t *win, bool wave)
{
    if (n == 0) return CCD_OK;
    if (wave) {
        if (n > 0x7fffffffll) return set_error(ctx, CCD_E_INVALID, "too many stencils for one launch");
        if (vf)
            CCD_PRODUCT_LAUNCH(hipLaunchKernelGGL((k_narrow_hist_wave<true, M>), dim3((unsigned)n), dim3(64), 0,
                                                      ctx->stream, off, ht, hp, (const int4 *)keys, eta, eta_s, n, hit,
                                                      t, sub, win));
        else
            CCD_PRODUCT_LAUNCH(hipLaunchKernelGGL((k_narrow_hist_wave<false, M>), dim3((unsigned)n), dim3(64), 0,
                                                      ctx->stream, off, ht, hp, (const int4 *)keys, eta, eta_s, n, hit,
                                                      t, sub, win));
        CCD_LAUNCH_CHECK(ctx);
        return CCD_OK;
    }
    if (vf)
        CCD_PRODUCT_LAUNCH(hipLaunchKernelGGL((k_narrow_hist<true, M>), dim3(blocks_for(n, LANE_BLOCK)), dim3(LANE_BLOCK), 0,
                                                  ctx->stream, off, ht, hp, (const int4 *)keys, eta, eta_s, n, hit, t,
                                                  sub, win));
    else
        CCD_PRODUCT_LAUNCH(hipLaunchKernelGGL((k_narrow_hist<false, M>), dim3(blocks_for(n, LANE_BLOCK)), dim3(LANE_BLOCK), 0,
                                                  ctx->stream, off, ht, hp, (const int4 *)keys, eta, eta_s, n, hit, t,
                                                  sub, win));
    CCD_LAUNCH_CHECK(ctx);
    return CCD_OK;
}

int64_t ee_deferral_capacity(const ccd_ctx *ctx, int64_t n)
{
    // Every deferred stencil, the holes a wavefront leaves when a chunk does not fit its block (at most
    // 63 per claimed block, and a block holds at least DEF_BLOCK - 63 entries when it is left, so fewer
    // than n / 4 in all), and the unused end of each wavefront's last block (at most 32 wavefronts per
    // compute unit).
    return n + n / 4 + (int64_t)(ctx->num_cus * 32 + 1) * DEF_BLOCK;
}

EeDeferral ee_deferral(ccd_ctx *ctx, int64_t n, unsigned long long *count_and_work)
{
    EeDeferral d;
    if (!ctx->ee_split || n <= 0 || n > 0x7fffffffll || !count_and_work) return d;
    const int64_t cap = ee_deferral_capacity(ctx, n);
    d.idx = (uint32_t *)buffer(ctx, "ee_def_idx", (size_t)cap * 4);
    d.sub = (uint8_t *)buffer(ctx, "ee_def_sub", (size_t)cap);
    if (!d.idx || !d.sub) return EeDeferral();   // one pass instead (the same results)
    d.count = count_and_work;
    d.work = count_and_work + 1;
    d.cap = cap;
    return d;
}

// Sizes the grouping buffers for up to nvf / nee stencils (before a capture, which cannot grow them).
int reserve_grouping(ccd_ctx *ctx, int64_t nvf, int64_t nee)
{
    const int64_t n = std::max(nvf, nee);
    if (n < 2 || n > 0x7fffffffll) return CCD_OK;
    if (!buffer(ctx, "grp_sig", (size_t)n * 8) || !buffer(ctx, "grp_sig2", (size_t)n * 8) ||
        !buffer(ctx, "grp_idx", (size_t)n * 4) || !buffer(ctx, "grp_idx2", (size_t)n * 4) ||
        !buffer(ctx, "grp_keys_vf", (size_t)std::max<int64_t>(nvf, 1) * 16) ||
        !buffer(ctx, "grp_keys_ee", (size_t)std::max<int64_t>(nee, 1) * 16) ||
        !buffer(ctx, "grp_tmp", radix_sort_temp_bytes(n)))
        return CCD_E_OOM;
    return CCD_OK;
}

int group_by_signature(ccd_ctx *ctx, bool vf, const double *qs, const double *qe, const int32_t *keys, int64_t n,
                       double eta, const int64_t *n_dev, const int32_t **grouped)
{
    *grouped = keys;
    if (n < 2 || n > 0x7fffffffll) return CCD_OK;   // int32 indices: larger launches keep the given order
    uint64_t *sig = (uint64_t *)buffer(ctx, "grp_sig", (size_t)n * 8);
    uint64_t *sig2 = (uint64_t *)buffer(ctx, "grp_sig2", (size_t)n * 8);
    int32_t *idx = (int32_t *)buffer(ctx, "grp_idx", (size_t)n * 4);
    int32_t *idx2 = (int32_t *)buffer(ctx, "grp_idx2", (size_t)n * 4);
    int4 *out = (int4 *)buffer(ctx, vf ? "grp_keys_vf" : "grp_keys_ee", (size_t)n * 16);
    void *tmp = buffer(ctx, "grp_tmp", radix_sort_temp_bytes(n));
    if (!sig || !sig2 || !idx || !idx2 || !out || !tmp) return CCD_E_OOM;
    if (vf)
        hipLaunchKernelGGL(k_stencil_signature<true>, dim3(blocks_for(n, 256)), dim3(256), 0, ctx->stream, qs, qe,
                           (const int4 *)keys, n, eta, nullptr, n_dev, sig, idx);
    else
        hipLaunchKernelGGL(k_stencil_signature<false>, dim3(blocks_for(n, 256)), dim3(256), 0, ctx->stream, qs, qe,
                           (const int4 *)keys, n, eta, nullptr, n_dev, sig, idx);
    CCD_LAUNCH_CHECK(ctx);
    bool in_alt = false;
    int rc = radix_sort_pairs(ctx, sig, idx, sig2, idx2, n, 8, tmp, in_alt);   // one stable 8-bit pass
    if (rc) return rc;
    hipLaunchKernelGGL(k_gather_keys, dim3(blocks_for(n, 256)), dim3(256), 0, ctx->stream, (const int4 *)keys,
                       in_alt ? idx2 : idx, n, out);
    CCD_LAUNCH_CHECK(ctx);
    *grouped = (const int32_t *)out;
    return CCD_OK;
}

int launch_min_gap(ccd_ctx *ctx, const int64_t *off, const double *ht, int64_t nverts, uint64_t *gap_bits)
{
    hipLaunchKernelGGL(k_set_bits, dim3(1), dim3(1), 0, ctx->stream, (unsigned long long *)gap_bits,
                       (unsigned long long)ccd::double_to_bits(1.0));
    CCD_LAUNCH_CHECK(ctx);
    if (nverts == 0) return CCD_OK;
    unsigned grid = (unsigned)std::min<int64_t>(blocks_for(nverts, 256), 4096);
    hipLaunchKernelGGL(k_min_gap, dim3(grid), dim3(256), 0, ctx->stream, off, ht, nverts,
                       (unsigned long long *)gap_bits);
    CCD_LAUNCH_CHECK(ctx);
    return CCD_OK;
}

int launch_separating_plane(ccd_ctx *ctx, bool vf, const int64_t *off, const double *ht, const double *hp,
                            const int32_t *keys, const double *eta, double eta_s, int64_t n, const uint64_t *gap_bits,
                            uint8_t *status, int64_t *undecided, bool wave)
{
    if (n == 0) return CCD_OK;
    auto gb = (const unsigned long long *)gap_bits;
    auto und = (unsigned long long *)undecided;
    if (wave) {
        if (n > 0x7fffffffll) return set_error(ctx, CCD_E_INVALID, "too many stencils for one launch");
        if (vf)
            CCD_PRODUCT_LAUNCH(hipLaunchKernelGGL((k_separating_plane_wave<true, M>), dim3((unsigned)n), dim3(64),
                                                      0, ctx->stream, off, ht, hp, (const int4 *)keys, eta, eta_s, n,
                                                      gb, status, und));
        else
            CCD_PRODUCT_LAUNCH(hipLaunchKernelGGL((k_separating_plane_wave<false, M>), dim3((unsigned)n),
                                                      dim3(64), 0, ctx->stream, off, ht, hp, (const int4 *)keys, eta,
                                                      eta_s, n, gb, status, und));
        CCD_LAUNCH_CHECK(ctx);
        return CCD_OK;
    }
    // Order the stencils so that those decided at the root run together (results are per stencil and
    // land at the stencil's own index, so the order changes nothing else).
    const int32_t *order = nullptr;
    if (n >= 4096 && n <= 0x7fffffffll) {   // int32 order indices
        uint64_t *key = (uint64_t *)buffer(ctx, "spo_key", (size_t)n * 8);
        uint64_t *key2 = (uint64_t *)buffer(ctx, "spo_key2", (size_t)n * 8);
        int32_t *idx = (int32_t *)buffer(ctx, "spo_idx", (size_t)n * 4);
        int32_t *idx2 = (int32_t *)buffer(ctx, "spo_idx2", (size_t)n * 4);
        void *tmp = buffer(ctx, "spo_tmp", radix_sort_temp_bytes(n));
        if (!key || !key2 || !idx || !idx2 || !tmp) return CCD_E_OOM;
        if (vf)
            hipLaunchKernelGGL(k_sp_order_key<true>, dim3(blocks_for(n, 256)), dim3(256), 0, ctx->stream, off, ht, hp,
                               (const int4 *)keys, eta, eta_s, n, gb, key, idx);
        else
            hipLaunchKernelGGL(k_sp_order_key<false>, dim3(blocks_for(n, 256)), dim3(256), 0, ctx->stream, off, ht, hp,
                               (const int4 *)keys, eta, eta_s, n, gb, key, idx);
        CCD_LAUNCH_CHECK(ctx);
        bool in_alt = false;
        int rc = radix_sort_pairs(ctx, key, idx, key2, idx2, n, 1, tmp, in_alt);   // one stable pass
        if (rc) return rc;
        order = in_alt ? idx2 : idx;
    }
    if (vf)
        CCD_PRODUCT_LAUNCH(hipLaunchKernelGGL((k_separating_plane<true, M>), dim3(blocks_for(n, LANE_BLOCK)), dim3(LANE_BLOCK),
                                                  0, ctx->stream, off, ht, hp, (const int4 *)keys, eta, eta_s, n, gb,
                                                  status, und, order));
    else
        CCD_PRODUCT_LAUNCH(hipLaunchKernelGGL((k_separating_plane<false, M>), dim3(blocks_for(n, LANE_BLOCK)), dim3(LANE_BLOCK),
                                                  0, ctx->stream, off, ht, hp, (const int4 *)keys, eta, eta_s, n, gb,
                                                  status, und, order));
    CCD_LAUNCH_CHECK(ctx);
    return CCD_OK;
}

int launch_validate_keys(ccd_ctx *ctx, const int32_t *keys, int64_t n4, int64_t nverts, int64_t *bad)
{
    if (n4 == 0) return CCD_OK;
    hipLaunchKernelGGL(k_validate_keys, dim3(blocks_for(n4, 256)), dim3(256), 0, ctx->stream, keys, n4, nverts,
                       (unsigned long long *)bad);
    CCD_LAUNCH_CHECK(ctx);
    return CCD_OK;
}

int launch_validate_history(ccd_ctx *ctx, const int64_t *off, const double *ht, int64_t nverts, int64_t nentries,
                            int64_t *bad)
{
    if (nverts == 0) return CCD_OK;
    hipLaunchKernelGGL(k_validate_history, dim3(blocks_for(nverts, 256)), dim3(256), 0, ctx->stream, off, ht,
                       nverts, nentries, (unsigned long long *)bad);
    CCD_LAUNCH_CHECK(ctx);
    return CCD_OK;
}

}   // namespace ccdi
