/*
 * ccd_amd.h -- C ABI of the MI355X (gfx950) swept-primitive continuous collision detection engine.
 *
 * The entry points below replace, one for one, the reference (evouga/collisiondetection) interfaces
 * named in each comment.  Pointers marked "device" are HIP device pointers (HBM-resident data, e.g.
 * a torch CUDA tensor's data_ptr()); sizes are element counts.  Positions are FP64 flat arrays of 3V
 * values (the layout of an Eigen::VectorXd of stacked vertices), faces are int32 flat arrays of 3F
 * values (the column-major storage of an Eigen::Matrix3Xi, one column per triangle), stencil keys are
 * int32 quadruples in the reference's canonical order (VertexFaceStencil: p, q0 <= q1 <= q2;
 * EdgeEdgeStencil: p0 <= p1, q0 <= q1, p0 <= q0).
 *
 * Every call is asynchronous on the context's stream unless stated otherwise, returns a status code
 * (CCD_OK = 0) and never throws.  A context is not thread-safe: use one per host thread.
 */
#ifndef CCD_AMD_H
#define CCD_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. */
#define CCD_OK 0
#define CCD_E_INVALID 1  /* bad argument (null pointer, negative size, index out of range, ...) */
#define CCD_E_CAPACITY 2 /* an output buffer is too small; the required count was written */
#define CCD_E_OOM 3      /* device allocation failed */
#define CCD_E_HIP 4      /* a HIP runtime call failed */
#define CCD_E_BUDGET 5   /* the request does not fit the context's HBM budget */
#define CCD_E_NODEVICE 6 /* no usable gfx950 device */

/* Primitive kinds for ccd_ctcd_primitives. */
#define CCD_PRIM_VERTEX_FACE 0   /* CTCD::vertexFaceCTCD   (include/CTCD.h:49-58)  24 doubles per item */
#define CCD_PRIM_EDGE_EDGE 1     /* CTCD::edgeEdgeCTCD     (include/CTCD.h:36-44)  24 doubles per item */
#define CCD_PRIM_VERTEX_EDGE 2   /* CTCD::vertexEdgeCTCD   (include/CTCD.h:63-70)  18 doubles per item */
#define CCD_PRIM_VERTEX_VERTEX 3 /* CTCD::vertexVertexCTCD (include/CTCD.h:75-79)  12 doubles per item */

/* Proximity kinds for ccd_distance_primitives (include/Distance.h of the reference); 12 doubles per item. */
#define CCD_DIST_VERTEX_PLANE_LT 0 /* Distance::vertexPlaneDistanceLessThan(p, q0, q1, q2, eta)  (Distance.h:14-19) */
#define CCD_DIST_LINE_LINE_LT 1    /* Distance::lineLineDistanceLessThan(p0, p1, q0, q1, eta)     (Distance.h:23-28) */
#define CCD_DIST_VERTEX_FACE 2     /* Distance::vertexFaceDistance(p, q0, q1, q2, b0, b1, b2)    (Distance.h:32-115) */
#define CCD_DIST_EDGE_EDGE 3       /* Distance::edgeEdgeDistance(p0, p1, q0, q1, b0..b3)          (Distance.h:119-174) */

/* Flags for ccd_kdop_candidates and ccd_batch_detect. */
#define CCD_SORTED 1 /* return stencils in the reference's std::set order (Stencils.h:59-77, :119-137) */
#define CCD_GRAPH 2  /* ccd_batch_detect: capture the pipeline in a HIP graph and replay it while the
                        arguments (pointers, sizes, eta) stay the same; the data behind them may change */

/* Stage timers (ccd_stage_times), in pipeline order. */
#define CCD_T_BOUNDS 0    /* leaf k-DOPs, scene box, Morton codes */
#define CCD_T_SORT 1      /* Morton radix sort */
#define CCD_T_BUILD 2     /* radix tree, escape links, leaf records, bottom-up refit */
#define CCD_T_TABLES 3    /* vertex -> face table and mesh checks */
#define CCD_T_TRAVERSE 4  /* face-pair traversal */
#define CCD_T_EXPAND 5    /* stencil expansion with deduplication */
#define CCD_T_NARROW_VF 6 /* vertex-face narrowphase */
#define CCD_T_NARROW_EE 7 /* edge-edge narrowphase */
#define CCD_T_PIPELINE 8  /* a whole captured pipeline (graph replays) */
#define CCD_NTIMERS 9

/* Narrowphase work stages (ccd_narrowphase_work), in the order of ccd::Stage in ccd_core.h. */
#define CCD_NSTAGES 12

typedef struct ccd_ctx ccd_ctx;

/* Context: owns a HIP stream and a grow-only device arena.  No counterpart in the reference, whose
 * objects allocate per call (KDOPBroadPhase.cpp:38-40). */
int ccd_create(ccd_ctx **out, int device);
int ccd_destroy(ccd_ctx *ctx);
/* Run subsequent work on an external stream (a hipStream_t); NULL restores the context's own. */
int ccd_set_stream(ccd_ctx *ctx, void *hip_stream);
int ccd_synchronize(ccd_ctx *ctx);
/* Human-readable description of the last error on this context (never NULL). */
const char *ccd_last_error(ccd_ctx *ctx);
/* Bytes of device memory the context currently holds. */
int64_t ccd_device_bytes(ccd_ctx *ctx);

/*
 * Primitive tests on n independent items -- CTCD::{vertexFace,edgeEdge,vertexEdge,vertexVertex}CTCD
 * (include/CTCD.h:36-79, src/CTCD.cpp:259-692).  Item i's positions are q[i*S .. i*S+S) (device),
 * S = 24/24/18/12: the start positions of the primitive's vertices in the reference's argument order,
 * then their end positions.  eta (device, n values) may be NULL to use eta_scalar for all items.
 * hit[i] = 1 when the reference function returns true; t[i] is then its time of impact (t is
 * written as 0 when hit[i] = 0, where the reference leaves t untouched).
 */
int ccd_ctcd_primitives(ccd_ctx *ctx, int kind, const double *q, const double *eta, double eta_scalar, int64_t n,
                        uint8_t *hit, double *t);

/*
 * Narrowphase -- CTCDNarrowPhase::findCollisions (src/CTCDNarrowPhase.cpp:9-22) with checkVFS / checkEES
 * (:24-135) for two-entry histories (History(qstart) + finishHistory(qend), History.cpp:8-39).
 * qstart/qend: device, 3*nverts doubles.  vf_keys/ee_keys: device, 4*n int32 keys.  vf_eta/ee_eta:
 * device per-stencil eta (the double of the reference's std::pair<Stencil,double>), or NULL to use
 * eta_scalar.  Outputs (device, n each): hit (1 = the stencil is in the reference's output set), t (the
 * time of impact of the sub-test that fired, which the reference computes and discards), sub (which
 * sub-test fired: 0 = VF or EE, then VE in the reference's order, then VV; 255 = none).  Any output
 * pointer except hit may be NULL.
 */
int ccd_ctcd_narrowphase(ccd_ctx *ctx, const double *qstart, const double *qend, int64_t nverts,
                         const int32_t *vf_keys, const double *vf_eta, int64_t nvf, const int32_t *ee_keys,
                         const double *ee_eta, int64_t nee, double eta_scalar, uint8_t *vf_hit, double *vf_t,
                         uint8_t *vf_sub, uint8_t *ee_hit, double *ee_t, uint8_t *ee_sub);

/*
 * ccd_ctcd_narrowphase_noalloc: ccd_ctcd_narrowphase (same arguments, results and errors) without any
 * device allocation.  No counterpart in the reference: it is for callers that cannot allocate (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  Its working memory
 * (stencil ordering buffers, a validation counter) is one file-scope device array of 26 MiB per GPU,
 * held until the process ends, whose address ccd_create looks up once per device.  At most
 * CCD_NOALLOC_MAX_STENCILS stencils per kind and call: larger calls return CCD_E_CAPACITY.  Every call
 * on one GPU shares that GPU's array, so calls that overlap on one GPU (other threads, other streams)
 * race; serialise them.
 */
#define CCD_NOALLOC_MAX_STENCILS ((int64_t)1 << 20)
int ccd_ctcd_narrowphase_noalloc(ccd_ctx *ctx, const double *qstart, const double *qend, int64_t nverts,
                                 const int32_t *vf_keys, const double *vf_eta, int64_t nvf, const int32_t *ee_keys,
                                 const double *ee_eta, int64_t nee, double eta_scalar, uint8_t *vf_hit, double *vf_t,
                                 uint8_t *vf_sub, uint8_t *ee_hit, double *ee_t, uint8_t *ee_sub);

/*
 * Narrowphase over multi-entry histories -- the same as ccd_ctcd_narrowphase for histories with any
 * number of entries per vertex (History::addHistory, History.cpp:21-26), stitched per stencil as
 * History::stitchCommonHistory (History.cpp:98-140) and tested window by window.  The history is CSR
 * (device): vertex v owns entries hist_off[v] .. hist_off[v+1]-1 with times hist_t[] (the first 0,
 * the last 1, nondecreasing) and positions hist_pos[3*e ..].  t is local to the window that fired
 * (window index in win, may be NULL), as the reference's CTCD calls see it.
 */
int ccd_ctcd_narrowphase_history(ccd_ctx *ctx, const int64_t *hist_off, const double *hist_t,
                                 const double *hist_pos, int64_t nverts, int64_t nentries, const int32_t *vf_keys,
                                 const double *vf_eta, int64_t nvf, const int32_t *ee_keys, const double *ee_eta,
                                 int64_t nee, double eta_scalar, uint8_t *vf_hit, double *vf_t, uint8_t *vf_sub,
                                 int32_t *vf_win, uint8_t *ee_hit, double *ee_t, uint8_t *ee_sub, int32_t *ee_win);

/*
 * Broadphase -- KDOPBroadPhase::findCollisionCandidates (src/KDOPBroadPhase.cpp:34-41, leaf bounds
 * :43-74, traversal :130-185) for two-entry histories.  faces (device, 3*nfaces int32), qstart/qend
 * (device, 3*nverts doubles), outer_eta, fixed (device, nverts bytes, nonzero = the vertex is in the
 * reference's fixedVerts set; may be NULL).  Writes the unique candidate stencils into vf_out / ee_out
 * (device, 4*cap int32) and their counts into *vf_count / *ee_count (host).  When a capacity is too
 * small the call returns CCD_E_CAPACITY with the required counts written and the buffers partially
 * filled.  flags: CCD_SORTED for the reference's set order.  Synchronous (the counts are on the host).
 */
int ccd_kdop_candidates(ccd_ctx *ctx, const int32_t *faces, int64_t nfaces, const double *qstart, const double *qend,
                        int64_t nverts, double outer_eta, const uint8_t *fixed, int32_t *vf_out, int64_t vf_cap,
                        int64_t *vf_count, int32_t *ee_out, int64_t ee_cap, int64_t *ee_count, int flags);

/*
 * SeparatingPlaneNarrowPhase::findCollisions (SeparatingPlaneNarrowPhase.h:7-22,
 * SeparatingPlaneNarrowPhase.cpp:9-278): the narrowphase that VelocityFilter uses.  Per stencil,
 * recursive time bisection with moving separating planes over the history (layout as in
 * ccd_ctcd_narrowphase_history), eps = History::computeMinimumGap / 4, and the CTCD primitives on
 * windows narrower than 3 eps.  Writes a status per stencil (device, uint8): 1 collision, 0 none, 2
 * undecided -- the reference's recursion would not end on this stencil (a sub-interval equal to its
 * parent; the reference recurses until its stack overflows) or would go more than 160 levels deep.
 * *undecided (host, may be NULL) receives the number of undecided stencils.  Root mode as for the
 * other narrowphases.  Synchronous.
 */
int ccd_separating_plane_narrowphase(ccd_ctx *ctx, const int64_t *hist_off, const double *hist_t,
                                     const double *hist_pos, int64_t nverts, int64_t nentries, const int32_t *vf_keys,
                                     const double *vf_eta, int64_t nvf, const int32_t *ee_keys, const double *ee_eta,
                                     int64_t nee, double eta_scalar, uint8_t *vf_status, uint8_t *ee_status,
                                     int64_t *undecided);

/*
 * Broadphase over multi-entry histories: as ccd_kdop_candidates, the leaf bounds covering every history
 * entry of the face's vertices (KDOPBroadPhase.cpp:60-70).  History layout as in
 * ccd_ctcd_narrowphase_history.
 */
int ccd_kdop_candidates_history(ccd_ctx *ctx, const int32_t *faces, int64_t nfaces, const int64_t *hist_off,
                                const double *hist_pos, int64_t nverts, int64_t nentries, double outer_eta,
                                const uint8_t *fixed, int32_t *vf_out, int64_t vf_cap, int64_t *vf_count,
                                int32_t *ee_out, int64_t ee_cap, int64_t *ee_count, int flags);

/*
 * Detection: broadphase then narrowphase (the sequence of example/AlecTest.cpp:97-111 and
 * ActiveLayers::collisionDetection, src/ActiveLayers.cpp:188-215) with the candidate stencils kept on
 * the device.  Every candidate is tested with eta; the colliding stencils are written to vf_hits /
 * ee_hits (device, 4*cap int32) with their times of impact in vf_toi / ee_toi (device, cap doubles,
 * may be NULL).  Counts of candidates and of collisions are written to the host (any may be NULL).
 * Returns CCD_E_CAPACITY, with the required counts, when a collision buffer is too small.
 * Synchronous.
 */
int ccd_detect(ccd_ctx *ctx, const int32_t *faces, int64_t nfaces, const double *qstart, const double *qend,
               int64_t nverts, double outer_eta, double eta, const uint8_t *fixed, int32_t *vf_hits, double *vf_toi,
               int64_t vf_cap, int64_t *vf_hit_count, int32_t *ee_hits, double *ee_toi, int64_t ee_cap,
               int64_t *ee_hit_count, int64_t *vf_candidates, int64_t *ee_candidates);

/*
 * Proximity tests on n items -- the Distance statics (include/Distance.h:14-174).  Item i's points are
 * q[12*i .. 12*i+12) (device): (p, q0, q1, q2) or (p0, p1, q0, q1) in the reference's argument order.
 * For the two LessThan kinds flag[i] (device) receives the boolean; for the two closest-point kinds
 * vec (device, 3n, may be NULL) receives the returned vector and bary (device, 4n, may be NULL) the
 * barycentric outputs (3 used for vertex-face).
 */
int ccd_distance_primitives(ccd_ctx *ctx, int kind, const double *q, double eta, int64_t n, uint8_t *flag,
                            double *vec, double *bary);

/*
 * Distance::meshSelfDistance (src/Distance.cpp:12-66): the shortest distance between a vertex and a face
 * not containing it, or between two edges sharing no vertex, over the mesh (verts: device, 3*nverts;
 * faces: device, 3*nfaces; fixed: device, nverts bytes or NULL).  *distance (host) receives it
 * (+inf when no such pair exists); *nvf / *nee (host, may be NULL) the numbers of candidate stencils
 * checked, the counts the reference prints (Distance.cpp:47).  Synchronous.
 */
int ccd_mesh_self_distance(ccd_ctx *ctx, const double *verts, int64_t nverts, const int32_t *faces, int64_t nfaces,
                           const uint8_t *fixed, double *distance, int64_t *nvf, int64_t *nee);

/* Counters of one ccd_velocity_filter call (the reference prints some of them, ActiveLayers.cpp:219-246). */
typedef struct ccd_vf_stats {
    int64_t iterations;      /* outer iterations run */
    int64_t layers;          /* penalty layers at the end */
    int64_t firings;         /* penalty group firings, all iterations */
    int64_t history_entries; /* history entries of the last iteration */
    int64_t vf_candidates;   /* broadphase candidates, summed over iterations */
    int64_t ee_candidates;
    int64_t vf_collisions;   /* separating-plane collisions, summed over iterations */
    int64_t ee_collisions;
    double closest;          /* meshSelfDistance of the start state */
    double host_seconds;     /* penalty integration (host event loop) */
    double device_seconds;   /* collision queries on the device, transfers included */
} ccd_vf_stats;

/*
 * VelocityFilter::velocityFilter (include/VelocityFilter.h:7-45, src/VelocityFilter.cpp:16-87) with the
 * layered penalty response of src/ActiveLayers.cpp, src/PenaltyGroup.cpp and src/PenaltyPotential.cpp.
 * Host arrays: qstart and qend (3*nverts; qend is overwritten when the filter succeeds), faces (3*nfaces),
 * inv_masses (3*nverts, per DOF; 0 = infinite mass).  The event-driven penalty integration runs on the
 * host; meshSelfDistance, the k-DOP broadphase over the multi-entry history and the separating-plane
 * narrowphase run on the device.  *result receives the reference's return value: the number of outer
 * iterations taken, or -1 (DANGLING_VERTICES), -2 (STARTING_STATE_INTERSECTS), -3 (MAXROLLBACKS_EXCEEDED).
 * verbose = 1 prints the reference's progress lines to stdout.  stats (may be NULL) receives the counters.
 * Where the separating-plane recursion never terminates for a stencil (the reference overflows its
 * stack there) the call fails with CCD_E_INVALID.  Synchronous.
 */
int ccd_velocity_filter(ccd_ctx *ctx, const double *qstart, double *qend, int64_t nverts, const int32_t *faces,
                        int64_t nfaces, const double *inv_masses, double outer_eta, double inner_eta,
                        double base_stiffness, double base_substep, double cor, int max_rollbacks, int verbose,
                        int *result, ccd_vf_stats *stats);

/*
 * Device memory helpers for host callers that do not use the HIP API themselves (the C++ classes in
 * include/ccdamd/ use them to stage host data).  Copies are synchronous on the context's stream.
 */
int ccd_device_alloc(ccd_ctx *ctx, int64_t bytes, void **ptr);
int ccd_device_free(ccd_ctx *ctx, void *ptr);
int ccd_copy_to_device(ccd_ctx *ctx, void *dst, const void *src, int64_t bytes);
int ccd_copy_to_host(ccd_ctx *ctx, void *dst, const void *src, int64_t bytes);

/*
 * Instrumentation (no counterpart in the reference, which times nothing).
 * ccd_set_profiling: when on, every pipeline stage is bracketed by HIP events on the context's stream
 * and calls that run a narrowphase synchronise at their end; ccd_stage_times returns the summed
 * milliseconds and launch counts per CCD_T_* timer since the last reset (reset != 0 clears them).
 */
int ccd_set_profiling(ccd_ctx *ctx, int on);
int ccd_stage_times(ccd_ctx *ctx, double *ms, int64_t *calls, int reset);

/*
 * Work accounting for the narrowphase's roofline: runs the two-entry narrowphase in a counting
 * variant and writes, per stage, how many times it executed (vf_counts / ee_counts: host arrays of
 * CCD_NSTAGES), and the FP64 operations each stage performs (ccd_stage_flops, CCD_NSTAGES values).
 * Synchronous.
 */
int ccd_narrowphase_work(ccd_ctx *ctx, const double *qstart, const double *qend, int64_t nverts,
                         const int32_t *vf_keys, int64_t nvf, const int32_t *ee_keys, int64_t nee, double eta,
                         int64_t *vf_counts, int64_t *ee_counts);
int ccd_stage_flops(int64_t *flops);

/*
 * The candidate stencils of the last ccd_detect / ccd_kdop_candidates call on this context, still on
 * the device (owned by the context, valid until its next call): device addresses and counts.
 */
int ccd_last_candidates(ccd_ctx *ctx, const int32_t **vf_keys, int64_t *nvf, const int32_t **ee_keys,
                        int64_t *nee);

/*
 * Batched detection: ninst independent meshes in one pass.  Instance k owns faces
 * inst_face_off[k] .. inst_face_off[k+1]-1 (inst_face_off: device, ninst+1 int64, starting at 0 and
 * ending at nfaces); faces hold global vertex indices and no vertex is shared between instances.  The
 * result is the union of ccd_detect over the instances -- no stencil pairs primitives of two
 * instances -- computed by one pipeline: one forest of radix trees (instance id in the high bits of
 * the sort keys), one traversal, one persistent work-queue narrowphase.  flags: CCD_GRAPH replays a
 * captured HIP graph of the whole pipeline on calls with the same arguments (after the first call has
 * sized every buffer); a capacity overflow found after a replay falls back to a fresh synchronous run.
 * Output and counts as ccd_detect.  Synchronous.  No counterpart in the reference (one mesh per call).
 */
int ccd_batch_detect(ccd_ctx *ctx, int64_t ninst, const int64_t *inst_face_off, const int32_t *faces, int64_t nfaces,
                     const double *qstart, const double *qend, int64_t nverts, double outer_eta, double eta,
                     int32_t *vf_hits, double *vf_toi, int64_t vf_cap, int64_t *vf_hit_count, int32_t *ee_hits,
                     double *ee_toi, int64_t ee_cap, int64_t *ee_hit_count, int64_t *vf_candidates,
                     int64_t *ee_candidates, int flags);

/*
 * Tuning of the persistent narrowphase: times the work-queue kernel on the candidates of the last
 * detection on this context over block sizes {64, 128, 256}, tiles {64, 128, ..., 4096} and
 * {4, 8, 12, 16} resident wavefronts per CU, on the stencils grouped by sub-test signature, and then
 * whether that grouping (its own passes included) beats the stored order; keeps the fastest per stencil kind
 * and returns it in shape[0..7] = {vf block, vf tile, vf waves/CU, ee block, ee tile, ee waves/CU,
 * vf grouping, ee grouping} (host, may be NULL; grouping 1 or 0).  ccd_set_queue_shape installs such
 * a shape (e.g. one cached from an earlier tuning run).
 */
int ccd_autotune(ccd_ctx *ctx, const double *qstart, const double *qend, double eta, int *shape);
int ccd_set_queue_shape(ccd_ctx *ctx, const int *shape);

/*
 * ccd_set_edge_edge_split (no counterpart in the reference; a scheduling choice, same results): on (the
 * default), the persistent edge-edge narrowphase of ccd_detect / ccd_batch_detect / the server runs in
 * two passes: the first screens each stencil's coplanar-distance sextic and sets aside the stencils
 * whose sextic needs the root solver, the second finishes them.  The list of set-aside stencils takes
 * 4 + 1 bytes per slot, about 1.25 slots per stored edge-edge stencil plus 512 per resident wavefront.
 * Off, one pass.
 */
int ccd_set_edge_edge_split(ccd_ctx *ctx, int on);

/*
 * Serving (no counterpart in the reference): device-resident sessions and a continuous-batching loop.
 *
 * ccd_server_create: a server on `device` with an HBM pool of hbm_budget_bytes for the transient
 *   workspace of in-flight requests (face pairs, stencils, collisions) and at most max_inflight
 *   requests on the device at once.  A worker thread admits requests in arrival order when their
 *   session is idle, a slot is free and their workspace fits in the pool.
 * ccd_server_open_session: uploads a mesh (host arrays: faces 3F, qstart/qend 3V) and builds its tree
 *   (one full detection qstart -> qend).  Fails with CCD_E_BUDGET if less than 2 GiB of device
 *   memory would remain.  Synchronous; returns the session id.
 * ccd_server_submit: a request with new end positions (host, 3V, copied): the session's previous end
 *   becomes the start and its tree is refit, not rebuilt.  Asynchronous.
 * ccd_server_poll: completed requests (counts, checksums of the colliding stencils, timestamps).
 * ccd_server_drain: waits until every submitted request has completed.
 * ccd_server_stats: 10 int64 = {submitted, completed, pending, in flight, pool bytes in use, resident
 *   bytes, reruns after a workspace overflow, sessions, admissions that waited for pool memory, largest
 *   request workspace in bytes}.  A request whose workspace exceeds the whole pool completes with
 *   CCD_E_BUDGET.
 * ccd_server_keep_hits / ccd_server_fetch_hits: keep each request's colliding stencils (4 int32 keys,
 *   then the TOIs as doubles, vertex-face first) for retrieval by request id.
 */
typedef struct ccd_server ccd_server;
typedef struct {
    int64_t request_id;
    int32_t session;
    int32_t status;
    int64_t vf_hits, ee_hits, vf_candidates, ee_candidates;
    uint64_t checksum_vf, checksum_ee; /* order-independent hashes of (key, TOI bits) of the collisions */
    double t_submit, t_start, t_done;   /* seconds, steady clock */
    int32_t reruns;                     /* times the request was rerun after outgrowing its workspace */
    int32_t reserved;
    int64_t face_pairs;                 /* overlapping face pairs the broadphase found */
} ccd_request_result;

int ccd_server_create(int device, int64_t hbm_budget_bytes, int max_inflight, ccd_server **out);
int ccd_server_open_session(ccd_server *server, const int32_t *faces, int64_t nfaces, const double *qstart,
                            const double *qend, int64_t nverts, double outer_eta, double eta, int *session);
int ccd_server_submit(ccd_server *server, int session, const double *qend_next, int64_t request_id);
int ccd_server_poll(ccd_server *server, ccd_request_result *out, int max, int *n);
int ccd_server_drain(ccd_server *server);
int ccd_server_stats(ccd_server *server, int64_t *stats);
int ccd_server_keep_hits(ccd_server *server, int on);
int ccd_server_fetch_hits(ccd_server *server, int64_t request_id, int32_t *out, int64_t cap, int64_t *n);
int ccd_server_destroy(ccd_server *server);

#ifdef __cplusplus
}
#endif

#endif /* CCD_AMD_H */
