// Internal declarations shared by the HIP translation units of libccd_amd.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <unordered_map>

#include "../../include/ccd_amd.h"

// Launch shape of the persistent work-queue narrowphase (tuned by ccd_autotune).
struct QueueShape {
    int block = 256;       // threads per block: 64, 128 or 256
    int tile = 64;         // stencils a wavefront takes per atomic
    int waves_per_cu = 12; // resident wavefronts per compute unit in the persistent grid (three per SIMD)
    int group = 1;         // order the stencils by their sub-test signature first (1) or not (0)
};

struct ccd_ctx {
    int device = 0;
    int num_cus = 256;
    bool group_stencils = true;   // group stencils by vertex-edge signature before the narrowphase
    bool ee_split = true;         // edge-edge narrowphase in two passes (deferred sextic recursions)
    QueueShape vf_shape;
    QueueShape ee_shape{256, 64, 16, 1};   // the screening pass holds four wavefronts per SIMD
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    std::string err;
    struct Buf {
        void *ptr = nullptr;
        size_t bytes = 0;
        bool external = false;   // memory owned by someone else (a server's pool): never freed here
    };
    std::unordered_map<std::string, Buf> bufs;   // named grow-only device buffers
    int64_t device_bytes = 0;
    int64_t *h_counts = nullptr;                  // pinned host mailbox for device counters
    int64_t *d_counts = nullptr;                  // device counters (CCD_NCOUNTERS int64)
    // Stage timing with HIP events on the context's stream (ccd_set_profiling).
    bool profiling = false;
    hipEvent_t ev[2 * CCD_NTIMERS] = {};
    bool pending[CCD_NTIMERS] = {};
    double stage_ms[CCD_NTIMERS] = {};
    int64_t stage_calls[CCD_NTIMERS] = {};
    // Grow-only capacities of the broadphase outputs (face pairs, stencils).
    int64_t cap_pairs = 0, cap_vf = 0, cap_ee = 0;
    // Sorted face order of the last full broadphase build (kept for refits).
    const int32_t *tree_order = nullptr;
    // Captured detection pipeline (ccd_batch_detect with CCD_GRAPH): replayed while its arguments match.
    bool capturing = false;
    bool graph_valid = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    unsigned char graph_key[256] = {};
    int64_t graph_replays = 0;
    // The candidates of the last broadphase (ccd_last_candidates).
    const int32_t *last_vf = nullptr, *last_ee = nullptr;
    int64_t last_nvf = 0, last_nee = 0;
};

namespace ccdi {

constexpr int NCOUNTERS = 16;

// Counter slots in ctx->d_counts.
enum Counter {
    C_PAIRS = 0,    // overlapping non-neighbouring face pairs found by the traversal
    C_VF = 1,       // unique vertex-face stencils
    C_EE = 2,       // unique edge-edge stencils
    C_VF_HITS = 3,  // colliding vertex-face stencils
    C_EE_HITS = 4,  // colliding edge-edge stencils
    C_BAD = 5,      // validation failures
};

int set_error(ccd_ctx *ctx, int code, const char *fmt, ...);
// Stage timers: no-ops unless profiling is on.  prof_collect needs the stream to be synchronised.
void prof_begin(ccd_ctx *ctx, int timer);
void prof_end(ccd_ctx *ctx, int timer);
void prof_collect(ccd_ctx *ctx);
// Grow-only named device buffer.  Returns nullptr (and sets the context error) on failure.
void *buffer(ccd_ctx *ctx, const char *name, size_t bytes);

}   // namespace ccdi

#define CCD_HIP_CHECK(ctx, call)                                                                          \
    do {                                                                                                  \
        hipError_t _e = (call);                                                                           \
        if (_e != hipSuccess)                                                                             \
            return ccdi::set_error((ctx), CCD_E_HIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(_e), \
                                   __FILE__, __LINE__);                                                   \
    } while (0)

#define CCD_LAUNCH_CHECK(ctx) CCD_HIP_CHECK(ctx, hipGetLastError())

// ----------------------------------------------------------------------------------------------
// Narrowphase launchers (narrowphase.hip)
// ----------------------------------------------------------------------------------------------
namespace ccdi {
int launch_primitives(ccd_ctx *ctx, int kind, const double *q, const double *eta, double eta_s, int64_t n,
                      uint8_t *hit, double *t);
// Buffers for ordering n stencils by signature before the narrowphase: sig, sig2 (n uint64), idx, idx2
// (n int32) and tmp (radix_sort_temp_bytes(n) bytes).
struct OrderScratch {
    uint64_t *sig = nullptr, *sig2 = nullptr;
    int32_t *idx = nullptr, *idx2 = nullptr;
    void *tmp = nullptr;
};
int launch_narrow2(ccd_ctx *ctx, bool vf, const double *qs, const double *qe, const int32_t *keys,
                   const double *eta, double eta_s, int64_t n, uint8_t *hit, double *t, uint8_t *sub,
                   const OrderScratch *order_scratch = nullptr);
// wave: one wavefront per stencil sharing the stitched windows (for long histories), else one lane.
int launch_narrow_hist(ccd_ctx *ctx, bool vf, const int64_t *off, const double *ht, const double *hp,
                       const int32_t *keys, const double *eta, double eta_s, int64_t n, uint8_t *hit, double *t,
                       uint8_t *sub, int32_t *win, bool wave);
// Narrowphase over a stencil list, compacting the colliding stencils (keys + toi) into out_keys /
// out_toi with a device counter at counters[slot]; at most cap are written.
int launch_narrow2_compact(ccd_ctx *ctx, bool vf, const double *qs, const double *qe, const int32_t *keys,
                           const int64_t *n_dev, int64_t n_max, double eta, int32_t *out_keys, double *out_toi,
                           int64_t cap, int64_t *counter);
// The deferred-stencil list of the two-pass edge-edge narrowphase: positions (idx) and first-pass
// vertex-edge / vertex-vertex results (sub) of up to cap stencils, the list's device counter and the
// second pass's queue counter (both zero before the launch).
struct EeDeferral {
    uint32_t *idx = nullptr;
    uint8_t *sub = nullptr;
    unsigned long long *count = nullptr;
    unsigned long long *work = nullptr;
    int64_t cap = 0;
};
// The context's deferral list for up to n stencils (grow-only buffers "ee_def_idx", "ee_def_sub"; counters
// at count_and_work[0..1]), or an empty one when the context runs edge-edge in one pass.
EeDeferral ee_deferral(ccd_ctx *ctx, int64_t n, unsigned long long *count_and_work);
// The list slots the two-pass narrowphase needs for n stored stencils (the list is filled in blocks).
int64_t ee_deferral_capacity(const ccd_ctx *ctx, int64_t n);
// Persistent work-queue narrowphase; *work must be zero before the launch.  Edge-edge runs in two
// passes when def holds a list for n_max stencils (n_max < 2^31), else in one.
int launch_narrow2_queue(ccd_ctx *ctx, bool vf, const double *qs, const double *qe, const int32_t *keys,
                         const int64_t *n_dev, int64_t n_max, double eta, int32_t *out_keys, double *out_toi,
                         int64_t cap, int64_t *counter, unsigned long long *work, const QueueShape &shape,
                         const EeDeferral *def = nullptr);
// Work accounting: per-stage execution counts (ccd::Stage order) summed over the stencils into counts.
int launch_narrow2_count(ccd_ctx *ctx, bool vf, const double *qs, const double *qe, const int32_t *keys,
                         int64_t n, double eta, unsigned long long *counts);
// Validation: counts keys outside [0, nverts) into *bad.
// The n stencils (device keys, n exact on the host) reordered by their vertex-edge signature
// (ccd::ve_signature; stable) into a context buffer; *grouped receives it (or keys when n < 2).  With
// n_dev (the stored count on the device, for captured pipelines) the n slots past it sort last.
int reserve_grouping(ccd_ctx *ctx, int64_t nvf, int64_t nee);
int group_by_signature(ccd_ctx *ctx, bool vf, const double *qs, const double *qe, const int32_t *keys, int64_t n,
                       double eta, const int64_t *n_dev,
                       const int32_t **grouped);
// Separating-plane narrowphase over a CSR history: launch_min_gap writes History::computeMinimumGap's
// value (as bits) to *gap_bits on the device; launch_separating_plane reads it (eps = gap / 4).
int launch_min_gap(ccd_ctx *ctx, const int64_t *off, const double *ht, int64_t nverts, uint64_t *gap_bits);
// wave: one wavefront per stencil sharing the trajectory scans (for long histories), else one lane.
int launch_separating_plane(ccd_ctx *ctx, bool vf, const int64_t *off, const double *ht, const double *hp,
                            const int32_t *keys, const double *eta, double eta_s, int64_t n, const uint64_t *gap_bits,
                            uint8_t *status, int64_t *undecided, bool wave);
int launch_validate_keys(ccd_ctx *ctx, const int32_t *keys, int64_t n4, int64_t nverts, int64_t *bad);
int launch_validate_history(ccd_ctx *ctx, const int64_t *off, const double *ht, int64_t nverts, int64_t nentries,
                            int64_t *bad);

// ----------------------------------------------------------------------------------------------
// Broadphase (broadphase.hip).  Produces the unique candidate stencils in device buffers owned by
// the context; *vf / *ee receive their device addresses and the counts are left in ctx->d_counts.
// ----------------------------------------------------------------------------------------------
struct Candidates {
    int32_t *vf = nullptr;
    int32_t *ee = nullptr;
    int64_t nvf = 0, nee = 0, npairs = 0;
    int64_t vf_cap = 0, ee_cap = 0;
};
// aabb selects AABBBroadPhase's 3-axis bounds (used by the mesh self-distance), else the 13-axis k-DOP.
// inst_off (device, ninst+1 face offsets) runs independent instances in one pass: no pair crosses
// instances.  Synchronous: checks capacities and reruns after growing them.
int broadphase(ccd_ctx *ctx, const int32_t *faces, int64_t F, const double *qs, const double *qe,
               const int64_t *hist_off, const double *hist_pos, int64_t V, double eta, const uint8_t *fixed,
               Candidates &out, bool aabb = false, const int64_t *inst_off = nullptr, int ninst = 1);
// The same pipeline as launches only (no host synchronisation, no allocation once prepared): it can be
// captured in a HIP graph.  Counts land in ctx->d_counts; buffers must come from broadphase_prepare.
int broadphase_prepare(ccd_ctx *ctx, int64_t F, int64_t V, bool batched);
int broadphase_launch(ccd_ctx *ctx, const int32_t *faces, int64_t F, const double *qs, const double *qe,
                      const int64_t *hist_off, const double *hist_pos, int64_t V, double eta, const uint8_t *fixed,
                      bool aabb, const int64_t *inst_off, int ninst);
int build_vertex_face_table_buffers(ccd_ctx *ctx, int64_t F, int64_t V);
// Refit of the last full build's tree with new positions (launches only; see broadphase.hip).
int broadphase_refit_launch(ccd_ctx *ctx, const int32_t *faces, int64_t F, const double *qs, const double *qe,
                            int64_t V, double eta, const uint8_t *fixed, const int64_t *inst_off, int ninst);
// Vertex -> face CSR table (device buffers "vf_off" / "vf_list"), and its current addresses.
int build_vertex_face_table(ccd_ctx *ctx, const int32_t *faces, int64_t F, int64_t V);
void vertex_face_table(ccd_ctx *ctx, const int32_t **off, const int32_t **list);
int sort_stencils(ccd_ctx *ctx, int32_t *keys, int64_t n);

// Device-wide primitives (devprims.hip): plain launches, safe inside a captured graph.
int fill_u32(ccd_ctx *ctx, uint32_t *p, int64_t n, uint32_t v);
int fill_u64(ccd_ctx *ctx, uint64_t *p, int64_t n, uint64_t v);
size_t scan_temp_bytes(int64_t n);
int exclusive_scan_i32(ccd_ctx *ctx, const int32_t *in, int32_t *out, int64_t n, void *temp, int64_t *total);
int exclusive_scan_i64(ccd_ctx *ctx, const int64_t *in, int64_t *out, int64_t n, void *temp, int64_t *total);
size_t radix_sort_temp_bytes(int64_t n);
// Sorts by the low `bits` bits of the keys; the result is in (keys_alt, vals_alt) when result_in_alt.
int radix_sort_pairs(ccd_ctx *ctx, uint64_t *keys, int32_t *vals, uint64_t *keys_alt, int32_t *vals_alt, int64_t n,
                     int bits, void *temp, bool &result_in_alt);

// Proximity (distance.hip)
int launch_distance_primitives(ccd_ctx *ctx, int kind, const double *q, double eta, int64_t n, uint8_t *flag,
                               double *vec, double *bary);
int mesh_self_distance(ccd_ctx *ctx, const double *verts, int64_t V, const int32_t *faces, int64_t F,
                       const uint8_t *fixed, double *out, int64_t *nvf, int64_t *nee);
}   // namespace ccdi
