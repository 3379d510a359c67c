// Serving runtime: device-resident CCD sessions driven by a continuous-batching request loop.
//
// A session owns one mesh (or a set of meshes packed as one, e.g. a mesh pair) on the device: faces,
// two position buffers (the current end and the next end), its radix tree from a first full build, and
// a ccd_ctx with its own HIP stream.  A request supplies new end positions; the session's previous end
// becomes the start, the tree is refit (not rebuilt), and the narrowphase runs on the candidates.
//
// The worker thread admits requests in arrival order as soon as (a) their session is idle -- requests
// of one session are sequential because each starts where the last ended -- (b) an in-flight slot is
// free and (c) the request's transient workspace (face pairs, stencils, collisions: sized from the
// session's last counts) fits in the HBM pool.  Admitted requests run on their sessions' streams, so
// the work of several requests overlaps on the device; completion is detected with events and the
// next requests are admitted as slots and pool memory free up.  Each completed request reports its
// counts, an order-independent 64-bit checksum of its colliding stencils and times of impact, and
// submit / start / done timestamps for latency statistics.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <string>
#include <vector>

#include "ccd_internal.h"

namespace {

inline unsigned blocks_for(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

__device__ __forceinline__ uint64_t mix64(uint64_t x)
{
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdull;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ull;
    x ^= x >> 33;
    return x;
}

// Sum (mod 2^64) of a hash of each colliding stencil's key and TOI bits: equal sets give equal sums in
// any order.  slot 0 = vertex-face, slot 1 = edge-edge.
__global__ __launch_bounds__(256) void k_hit_checksum(const int4 *__restrict__ vk, const double *__restrict__ vt,
                                                      const int4 *__restrict__ ek, const double *__restrict__ et,
                                                      const int64_t *__restrict__ counts, int64_t vcap, int64_t ecap,
                                                      unsigned long long *__restrict__ out)
{
    int64_t nv = counts[0] < vcap ? counts[0] : vcap;
    int64_t ne = counts[1] < ecap ? counts[1] : ecap;
    uint64_t acc[2] = {0, 0};
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nv + ne; i += (int64_t)gridDim.x * blockDim.x) {
        int4 k = i < nv ? vk[i] : ek[i - nv];
        double t = i < nv ? vt[i] : et[i - nv];
        uint64_t h = mix64(((uint64_t)(uint32_t)k.x << 32) ^ (uint32_t)k.y);
        h = mix64(h ^ (((uint64_t)(uint32_t)k.z << 32) | (uint32_t)k.w));
        h = mix64(h ^ (uint64_t)__double_as_longlong(t));
        acc[i < nv ? 0 : 1] += h;
    }
    for (int s = 0; s < 2; s++) {
        uint64_t v = acc[s];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        if (__lane_id() == 0 && v) atomicAdd(&out[s], (unsigned long long)v);
    }
}

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Pool {   // first-fit allocator over one device allocation (the admission budget)
    char *base = nullptr;
    int64_t size = 0;
    std::map<int64_t, int64_t> free_;   // offset -> length
    int64_t used = 0;
    int64_t alloc(int64_t n)
    {
        n = (n + 255) & ~int64_t(255);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= n) {
                int64_t off = it->first, len = it->second;
                free_.erase(it);
                if (len > n) free_[off + n] = len - n;
                used += n;
                return off;
            }
        }
        return -1;
    }
    void release(int64_t off, int64_t n)
    {
        n = (n + 255) & ~int64_t(255);
        used -= n;
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
};

}   // namespace

struct ccd_request_rec {
    int64_t id;
    int session;
    double t_submit;
    std::vector<double> qend;   // new end positions (host copy)
    int reruns = 0;             // workspace overflows so far
};

struct Session {
    ccd_ctx *ctx = nullptr;
    int32_t *faces = nullptr;
    double *pos[2] = {nullptr, nullptr};   // pos[cur] = current end, pos[1-cur] = next end
    int cur = 0;
    int64_t F = 0, V = 0;
    double outer_eta = 0, eta = 0;
    int64_t resident = 0;
    bool busy = false;
    // workspace sizing from the last completed request
    int64_t last_pairs = 0, last_vf = 0, last_ee = 0, last_hits = 0;
    bool moved = false;   // a request has completed: the last counts come from a moving mesh
};

struct InFlight {
    ccd_request_rec req;
    int64_t pool_off = -1, pool_bytes = 0;
    int64_t cap_pairs = 0, cap_vf = 0, cap_ee = 0, cap_hits = 0;
    hipEvent_t done = nullptr;
    int64_t *h_rec = nullptr;   // pinned: NCOUNTERS counters + 2 checksums
    double t_start = 0;
    int32_t *hit_vk = nullptr, *hit_ek = nullptr;
    double *hit_vt = nullptr, *hit_et = nullptr;
};

struct ccd_server {
    int device = 0;
    int max_inflight = 8;
    Pool pool;
    std::vector<Session> sessions;
    std::deque<ccd_request_rec> pending;
    std::vector<InFlight> inflight;
    std::deque<ccd_request_result> done;
    std::map<int64_t, std::vector<int32_t>> kept;   // request id -> colliding stencils (when keeping)
    bool keep_hits = false;
    int64_t resident_bytes = 0, budget = 0;
    std::mutex mu;
    std::condition_variable cv;
    std::thread worker;
    bool stop = false;
    std::string err;
    int64_t submitted = 0, completed = 0, retries = 0, mem_waits = 0, max_request_bytes = 0;
    // One pinned counter record and one completion event per in-flight slot, made at creation:
    // allocating or freeing pinned memory inside the request loop can wait for the whole device.
    int64_t *rec_block = nullptr;
    std::vector<int64_t *> rec_free;
    std::vector<hipEvent_t> ev_free;
};

namespace {

using namespace ccdi;

int64_t align256(int64_t x) { return (x + 255) & ~255ll; }

// A request's transient device memory: face pairs, the related-face table and its emission masks
// (two entries per pair), the stencil keys, the edge-edge deferred list (cd slots) and the hit lists.
int64_t workspace_bytes(int64_t cp, int64_t cv, int64_t ce, int64_t cd, int64_t ch)
{
    return align256(cp * 8) + align256((2 * cp + 1) * 4) + align256((2 * cp + 1) * 2) + align256(cv * 16) +
           align256(ce * 16) + align256(cd * 4) + align256(cd) + 2 * (align256(ch * 16) + align256(ch * 8)) + 256;
}

// Point the session context's transient buffers at a slice of the pool.
void bind_workspace(ccd_server *S, Session &s, InFlight &f)
{
    char *p = S->pool.base + f.pool_off;
    auto put = [&](const char *name, int64_t bytes) {
        ccd_ctx::Buf &b = s.ctx->bufs[name];
        b.ptr = p;
        b.bytes = (size_t)bytes;
        b.external = true;
        p += align256(bytes);
    };
    put("pairs", f.cap_pairs * 8);
    put("adj", (2 * f.cap_pairs + 1) * 4);
    put("adj_mask", (2 * f.cap_pairs + 1) * 2);
    put("vf_keys", f.cap_vf * 16);
    put("ee_keys", f.cap_ee * 16);
    const int64_t cd = ee_deferral_capacity(s.ctx, f.cap_ee);   // the edge-edge deferred list
    put("ee_def_idx", cd * 4);
    put("ee_def_sub", cd);
    const int64_t kb = align256(f.cap_hits * 16), tb = align256(f.cap_hits * 8);
    f.hit_vk = (int32_t *)p;
    f.hit_vt = (double *)(p + kb);
    p += kb + tb;
    f.hit_ek = (int32_t *)p;
    f.hit_et = (double *)(p + kb);
    s.ctx->cap_pairs = f.cap_pairs;
    s.ctx->cap_vf = f.cap_vf;
    s.ctx->cap_ee = f.cap_ee;
}

// Everything a request does on its session's stream; returns nonzero on a launch failure.
int launch_request(ccd_server *S, Session &s, InFlight &f)
{
    ccd_ctx *ctx = s.ctx;
    hipStream_t st = ctx->stream;
    double *start = s.pos[s.cur], *end = s.pos[1 - s.cur];
    CCD_HIP_CHECK(ctx, hipMemcpyAsync(end, f.req.qend.data(), (size_t)s.V * 24, hipMemcpyHostToDevice, st));
    int rc = broadphase_refit_launch(ctx, s.faces, s.F, start, end, s.V, s.outer_eta, nullptr, nullptr, 1);
    if (rc) return rc;
    // work[0..1]: the two queues; [2..3]: the hit checksums; [4..5]: the edge-edge deferral list.
    unsigned long long *work = (unsigned long long *)buffer(ctx, "queue_work", 48);
    if (!work) return CCD_E_OOM;
    rc = fill_u64(ctx, (uint64_t *)(ctx->d_counts + C_VF_HITS), 2, 0);
    if (!rc) rc = fill_u64(ctx, (uint64_t *)work, 6, 0);
    if (rc) return rc;
    const EeDeferral def = ee_deferral(ctx, f.cap_ee, work + 4);
    rc = launch_narrow2_queue(ctx, true, start, end, (const int32_t *)ctx->bufs["vf_keys"].ptr, ctx->d_counts + C_VF,
                              f.cap_vf, s.eta, f.hit_vk, f.hit_vt, f.cap_hits, ctx->d_counts + C_VF_HITS, work,
                              ctx->vf_shape);
    if (rc) return rc;
    rc = launch_narrow2_queue(ctx, false, start, end, (const int32_t *)ctx->bufs["ee_keys"].ptr, ctx->d_counts + C_EE,
                              f.cap_ee, s.eta, f.hit_ek, f.hit_et, f.cap_hits, ctx->d_counts + C_EE_HITS, work + 1,
                              ctx->ee_shape, &def);
    if (rc) return rc;
    unsigned g = (unsigned)std::max<int64_t>(1, std::min<int64_t>(blocks_for(2 * f.cap_hits, 256), 1024));
    hipLaunchKernelGGL(k_hit_checksum, dim3(g), dim3(256), 0, st, (const int4 *)f.hit_vk, f.hit_vt,
                       (const int4 *)f.hit_ek, f.hit_et, ctx->d_counts + C_VF_HITS, f.cap_hits, f.cap_hits, work + 2);
    CCD_LAUNCH_CHECK(ctx);
    CCD_HIP_CHECK(ctx, hipMemcpyAsync(f.h_rec, ctx->d_counts, sizeof(int64_t) * NCOUNTERS, hipMemcpyDeviceToHost, st));
    CCD_HIP_CHECK(ctx, hipMemcpyAsync(f.h_rec + NCOUNTERS, work + 2, 16, hipMemcpyDeviceToHost, st));
    CCD_HIP_CHECK(ctx, hipEventRecord(f.done, st));
    return CCD_OK;
}

void worker_loop(ccd_server *S)
{
    (void)hipSetDevice(S->device);
    std::unique_lock<std::mutex> lk(S->mu);
    while (true) {
        // 1. completions
        bool progressed = false;
        for (size_t k = 0; k < S->inflight.size();) {
            InFlight &f = S->inflight[k];
            if (hipEventQuery(f.done) != hipSuccess) {
                ++k;
                continue;
            }
            Session &s = S->sessions[f.req.session];
            int64_t *h = f.h_rec;
            bool overflow = h[C_PAIRS] > f.cap_pairs || h[C_VF] > f.cap_vf || h[C_EE] > f.cap_ee ||
                            h[C_VF_HITS] > f.cap_hits || h[C_EE_HITS] > f.cap_hits;
            s.last_pairs = h[C_PAIRS];
            s.last_vf = h[C_VF];
            s.last_ee = h[C_EE];
            s.last_hits = std::max(h[C_VF_HITS], h[C_EE_HITS]);
            if (overflow) {
                // Rerun with workspace sized from these counts; positions have not advanced.
                S->retries++;
                f.req.reruns++;
                S->pending.push_front(f.req);
            } else {
                if (S->keep_hits) {
                    std::vector<int32_t> v;
                    int64_t nv = h[C_VF_HITS], ne = h[C_EE_HITS];
                    v.resize(4 * (nv + ne) + 2 * (nv + ne) * 2);
                    std::vector<double> tv(nv), te(ne);
                    (void)hipMemcpy(v.data(), f.hit_vk, nv * 16, hipMemcpyDeviceToHost);
                    (void)hipMemcpy(v.data() + 4 * nv, f.hit_ek, ne * 16, hipMemcpyDeviceToHost);
                    (void)hipMemcpy(tv.data(), f.hit_vt, nv * 8, hipMemcpyDeviceToHost);
                    (void)hipMemcpy(te.data(), f.hit_et, ne * 8, hipMemcpyDeviceToHost);
                    v.resize(4 * (nv + ne));
                    for (double t : tv) {
                        int32_t w[2];
                        memcpy(w, &t, 8);
                        v.push_back(w[0]);
                        v.push_back(w[1]);
                    }
                    for (double t : te) {
                        int32_t w[2];
                        memcpy(w, &t, 8);
                        v.push_back(w[0]);
                        v.push_back(w[1]);
                    }
                    S->kept[f.req.id] = std::move(v);
                }
                s.cur = 1 - s.cur;   // the new end becomes the next request's start
                s.moved = true;
                ccd_request_result r{};
                r.request_id = f.req.id;
                r.session = f.req.session;
                r.status = CCD_OK;
                r.vf_hits = h[C_VF_HITS];
                r.ee_hits = h[C_EE_HITS];
                r.vf_candidates = h[C_VF];
                r.ee_candidates = h[C_EE];
                r.checksum_vf = (uint64_t)h[NCOUNTERS];
                r.checksum_ee = (uint64_t)h[NCOUNTERS + 1];
                r.reruns = f.req.reruns;
                r.face_pairs = h[C_PAIRS];
                r.t_submit = f.req.t_submit;
                r.t_start = f.t_start;
                r.t_done = now_s();
                S->done.push_back(r);
                S->completed++;
            }
            s.busy = false;
            S->pool.release(f.pool_off, f.pool_bytes);
            S->ev_free.push_back(f.done);
            S->rec_free.push_back(f.h_rec);
            S->inflight.erase(S->inflight.begin() + k);
            progressed = true;
        }
        // 2. admissions, in arrival order
        for (auto it = S->pending.begin(); it != S->pending.end() && (int)S->inflight.size() < S->max_inflight;) {
            Session &s = S->sessions[it->session];
            if (s.busy) {
                ++it;
                continue;
            }
            InFlight f;
            // Workspace from the session's last counts with headroom: twice the last counts (four times
            // for a session's first request, whose last counts come from its opening build, where the
            // mesh may not have moved at all) plus a share of the face count.  Counts of a session that
            // keeps moving grow by factors from one request to the next (a drifting mesh starts to touch
            // itself), and every overflow costs a rerun of the whole request; the headroom costs only
            // pool memory, about 60 bytes per face.
            const int64_t mul = s.moved ? 2 : 4;
            f.cap_pairs = mul * s.last_pairs + s.F / 4 + 4096;
            f.cap_vf = mul * s.last_vf + s.F + 4096;
            f.cap_ee = mul * s.last_ee + 2 * s.F + 4096;
            f.cap_hits = mul * s.last_hits + s.F / 16 + 4096;
            f.pool_bytes = workspace_bytes(f.cap_pairs, f.cap_vf, f.cap_ee, ee_deferral_capacity(s.ctx, f.cap_ee),
                                           f.cap_hits);
            if (f.pool_bytes > S->pool.size) {   // can never be admitted: fail it rather than wait forever
                ccd_request_result r{};
                r.request_id = it->id;
                r.session = it->session;
                r.status = CCD_E_BUDGET;
                r.t_submit = it->t_submit;
                r.t_start = r.t_done = now_s();
                S->done.push_back(r);
                it = S->pending.erase(it);
                progressed = true;
                continue;
            }
            f.pool_off = S->pool.alloc(f.pool_bytes);
            if (f.pool_off < 0) {   // wait for memory: keep arrival order
                S->mem_waits++;
                break;
            }
            S->max_request_bytes = std::max(S->max_request_bytes, f.pool_bytes);
            f.req = std::move(*it);
            it = S->pending.erase(it);
            s.busy = true;
            f.t_start = now_s();
            f.h_rec = S->rec_free.back();
            S->rec_free.pop_back();
            f.done = S->ev_free.back();
            S->ev_free.pop_back();
            bind_workspace(S, s, f);
            int rc = launch_request(S, s, f);
            if (rc) {
                S->err = s.ctx->err;
                ccd_request_result r{};
                r.request_id = f.req.id;
                r.session = f.req.session;
                r.status = rc;
                r.t_submit = f.req.t_submit;
                r.t_start = f.t_start;
                r.t_done = now_s();
                S->done.push_back(r);
                s.busy = false;
                (void)hipStreamSynchronize(s.ctx->stream);
                S->pool.release(f.pool_off, f.pool_bytes);
                S->ev_free.push_back(f.done);
                S->rec_free.push_back(f.h_rec);
                continue;
            }
            S->inflight.push_back(std::move(f));
            progressed = true;
        }
        if (S->stop && S->pending.empty() && S->inflight.empty()) break;
        if (!progressed) {
            if (S->inflight.empty())
                S->cv.wait_for(lk, std::chrono::milliseconds(2));
            else {
                lk.unlock();
                std::this_thread::sleep_for(std::chrono::microseconds(20));
                lk.lock();
            }
        }
    }
}

}   // namespace

extern "C" {

int ccd_server_create(int device, int64_t hbm_budget_bytes, int max_inflight, ccd_server **out)
{
    if (!out || hbm_budget_bytes <= 0 || max_inflight < 1) return CCD_E_INVALID;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
        (void)hipGetLastError();
        return CCD_E_NODEVICE;
    }
    if (device < 0 || device >= n) return CCD_E_INVALID;
    ccd_server *S = new (std::nothrow) ccd_server;
    if (!S) return CCD_E_OOM;
    S->device = device;
    S->max_inflight = max_inflight;
    S->budget = hbm_budget_bytes;
    (void)hipSetDevice(device);
    void *p = nullptr;
    if (hipMalloc(&p, (size_t)hbm_budget_bytes) != hipSuccess) {
        (void)hipGetLastError();
        delete S;
        return CCD_E_OOM;
    }
    S->pool.base = (char *)p;
    S->pool.size = hbm_budget_bytes;
    S->pool.free_[0] = hbm_budget_bytes;
    const size_t stride = NCOUNTERS + 2;
    if (hipHostMalloc((void **)&S->rec_block, sizeof(int64_t) * stride * (size_t)max_inflight, hipHostMallocDefault) !=
        hipSuccess) {
        (void)hipGetLastError();
        (void)hipFree(S->pool.base);
        delete S;
        return CCD_E_OOM;
    }
    for (int k = 0; k < max_inflight; k++) {
        hipEvent_t e = nullptr;
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
            (void)hipGetLastError();
            for (hipEvent_t x : S->ev_free) (void)hipEventDestroy(x);
            (void)hipHostFree(S->rec_block);
            (void)hipFree(S->pool.base);
            delete S;
            return CCD_E_HIP;
        }
        S->ev_free.push_back(e);
        S->rec_free.push_back(S->rec_block + stride * (size_t)k);
    }
    S->worker = std::thread(worker_loop, S);
    *out = S;
    return CCD_OK;
}

int ccd_server_open_session(ccd_server *S, const int32_t *faces, int64_t nfaces, const double *qstart,
                            const double *qend, int64_t nverts, double outer_eta, double eta, int *session)
{
    if (!S || !faces || !qstart || !qend || !session || nfaces < 0 || nverts < 0) return CCD_E_INVALID;
    *session = -1;
    Session s;
    int rc = ccd_create(&s.ctx, S->device);
    if (rc) return rc;
    s.F = nfaces;
    s.V = nverts;
    s.outer_eta = outer_eta;
    s.eta = eta;
    ccd_ctx *ctx = s.ctx;
    auto fail = [&](int code) {
        if (s.faces) (void)hipFree(s.faces);
        for (int k = 0; k < 2; k++)
            if (s.pos[k]) (void)hipFree(s.pos[k]);
        ccd_destroy(ctx);
        return code;
    };
    if (hipMalloc((void **)&s.faces, (size_t)std::max<int64_t>(1, 3 * nfaces) * 4) != hipSuccess ||
        hipMalloc((void **)&s.pos[0], (size_t)std::max<int64_t>(1, 3 * nverts) * 8) != hipSuccess ||
        hipMalloc((void **)&s.pos[1], (size_t)std::max<int64_t>(1, 3 * nverts) * 8) != hipSuccess) {
        (void)hipGetLastError();
        return fail(CCD_E_OOM);
    }
    if (hipMemcpy(s.faces, faces, (size_t)nfaces * 12, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(s.pos[0], qstart, (size_t)nverts * 24, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(s.pos[1], qend, (size_t)nverts * 24, hipMemcpyHostToDevice) != hipSuccess)
        return fail(CCD_E_HIP);
    // First full build (start -> end) sizes every buffer and gives the first counts.
    int32_t *dummy = nullptr;
    (void)hipMalloc((void **)&dummy, 64);
    int64_t nvh = 0, neh = 0, nvc = 0, nec = 0;
    rc = ccd_detect(ctx, s.faces, nfaces, s.pos[0], s.pos[1], nverts, outer_eta, eta, nullptr, dummy, nullptr, 0, &nvh,
                    dummy, nullptr, 0, &neh, &nvc, &nec);
    (void)hipFree(dummy);
    if (rc && rc != CCD_E_CAPACITY) return fail(rc);
    if (ctx->h_counts[C_BAD] != 0) return fail(set_error(ctx, CCD_E_INVALID, "sessions need meshes without degenerate or duplicate faces"));
    s.cur = 1;   // the first request starts at qend
    s.last_pairs = ctx->h_counts[C_PAIRS];
    s.last_vf = nvc;
    s.last_ee = nec;
    s.last_hits = std::max(nvh, neh);
    // The transient buffers come from the pool per request: release the context's own copies.
    for (const char *name : {"pairs", "adj", "adj_mask", "vf_keys", "ee_keys", "ee_def_idx", "ee_def_sub"}) {
        ccd_ctx::Buf &b = ctx->bufs[name];
        if (b.ptr && !b.external) {
            (void)hipFree(b.ptr);
            ctx->device_bytes -= (int64_t)b.bytes;
        }
        b.ptr = nullptr;
        b.bytes = 0;
    }
    s.resident = ccd_device_bytes(ctx) + (int64_t)nfaces * 12 + 2 * (int64_t)nverts * 24;
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) == hipSuccess && free_b < ((size_t)2 << 30))
        return fail(CCD_E_BUDGET);   // admission by HBM footprint: keep 2 GiB of headroom
    std::lock_guard<std::mutex> lk(S->mu);
    S->resident_bytes += s.resident;
    S->sessions.push_back(s);
    *session = (int)S->sessions.size() - 1;
    return CCD_OK;
}

int ccd_server_submit(ccd_server *S, int session, const double *qend_next, int64_t request_id)
{
    if (!S || !qend_next) return CCD_E_INVALID;
    std::lock_guard<std::mutex> lk(S->mu);
    if (session < 0 || session >= (int)S->sessions.size()) return CCD_E_INVALID;
    ccd_request_rec r;
    r.id = request_id;
    r.session = session;
    r.t_submit = now_s();
    r.qend.assign(qend_next, qend_next + 3 * S->sessions[session].V);
    S->pending.push_back(std::move(r));
    S->submitted++;
    S->cv.notify_one();
    return CCD_OK;
}

int ccd_server_poll(ccd_server *S, ccd_request_result *out, int max, int *n)
{
    if (!S || !n || (max > 0 && !out)) return CCD_E_INVALID;
    std::lock_guard<std::mutex> lk(S->mu);
    int k = 0;
    while (k < max && !S->done.empty()) {
        out[k++] = S->done.front();
        S->done.pop_front();
    }
    *n = k;
    return CCD_OK;
}

int ccd_server_drain(ccd_server *S)
{
    if (!S) return CCD_E_INVALID;
    while (true) {
        {
            std::lock_guard<std::mutex> lk(S->mu);
            if (S->pending.empty() && S->inflight.empty()) return CCD_OK;
        }
        std::this_thread::sleep_for(std::chrono::microseconds(200));
    }
}

int ccd_server_stats(ccd_server *S, int64_t *stats)
{
    if (!S || !stats) return CCD_E_INVALID;
    std::lock_guard<std::mutex> lk(S->mu);
    stats[0] = S->submitted;
    stats[1] = S->completed;
    stats[2] = (int64_t)S->pending.size();
    stats[3] = (int64_t)S->inflight.size();
    stats[4] = S->pool.used;
    stats[5] = S->resident_bytes;
    stats[6] = S->retries;
    stats[7] = (int64_t)S->sessions.size();
    stats[8] = S->mem_waits;
    stats[9] = S->max_request_bytes;
    return CCD_OK;
}

int ccd_server_keep_hits(ccd_server *S, int on)
{
    if (!S) return CCD_E_INVALID;
    std::lock_guard<std::mutex> lk(S->mu);
    S->keep_hits = on != 0;
    return CCD_OK;
}

int ccd_server_fetch_hits(ccd_server *S, int64_t request_id, int32_t *out, int64_t cap, int64_t *n)
{
    if (!S || !n) return CCD_E_INVALID;
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->kept.find(request_id);
    if (it == S->kept.end()) return CCD_E_INVALID;
    *n = (int64_t)it->second.size();
    if (out && cap >= *n) memcpy(out, it->second.data(), it->second.size() * 4);
    return cap >= *n ? CCD_OK : CCD_E_CAPACITY;
}

int ccd_server_destroy(ccd_server *S)
{
    if (!S) return CCD_E_INVALID;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
        S->cv.notify_one();
    }
    if (S->worker.joinable()) S->worker.join();
    (void)hipSetDevice(S->device);
    for (Session &s : S->sessions) {
        for (const char *name : {"pairs", "adj", "adj_mask", "vf_keys", "ee_keys", "ee_def_idx", "ee_def_sub"}) s.ctx->bufs[name] = ccd_ctx::Buf();
        (void)hipFree(s.faces);
        (void)hipFree(s.pos[0]);
        (void)hipFree(s.pos[1]);
        ccd_destroy(s.ctx);
    }
    for (hipEvent_t e : S->ev_free) (void)hipEventDestroy(e);
    (void)hipHostFree(S->rec_block);
    (void)hipFree(S->pool.base);
    delete S;
    return CCD_OK;
}

}   // extern "C"
