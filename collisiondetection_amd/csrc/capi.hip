// C ABI of libccd_amd.so (declarations and contracts in include/ccd_amd.h).
#include <cstdarg>
#include <new>

#include "ccd_core.h"
#include "ccd_internal.h"

namespace ccdi {

int set_error(ccd_ctx *ctx, int code, const char *fmt, ...)
{
    if (ctx) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof(buf), fmt, ap);
        va_end(ap);
        ctx->err = buf;
    }
    return code;
}

void prof_begin(ccd_ctx *ctx, int timer)
{
    if (!ctx->profiling) return;
    (void)hipEventRecord(ctx->ev[2 * timer], ctx->stream);
}

void prof_end(ccd_ctx *ctx, int timer)
{
    if (!ctx->profiling) return;
    (void)hipEventRecord(ctx->ev[2 * timer + 1], ctx->stream);
    ctx->pending[timer] = true;
}

void prof_collect(ccd_ctx *ctx)
{
    if (!ctx->profiling) return;
    for (int t = 0; t < CCD_NTIMERS; t++) {
        if (!ctx->pending[t]) continue;
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, ctx->ev[2 * t], ctx->ev[2 * t + 1]) == hipSuccess) {
            ctx->stage_ms[t] += ms;
            ctx->stage_calls[t]++;
        } else {
            (void)hipGetLastError();
        }
        ctx->pending[t] = false;
    }
}

void *buffer(ccd_ctx *ctx, const char *name, size_t bytes)
{
    if (bytes == 0) bytes = 16;
    ccd_ctx::Buf &b = ctx->bufs[name];
    if (b.bytes >= bytes) return b.ptr;
    if (b.ptr) {
        // The old allocation may still be read by queued work on the stream.
        if (hipStreamSynchronize(ctx->stream) != hipSuccess) {
            set_error(ctx, CCD_E_HIP, "stream synchronisation failed before growing %s", name);
            return nullptr;
        }
        (void)hipFree(b.ptr);
        ctx->device_bytes -= (int64_t)b.bytes;
        b.ptr = nullptr;
        b.bytes = 0;
    }
    size_t want = bytes + bytes / 4;   // grow with headroom
    void *p = nullptr;
    if (hipMalloc(&p, want) != hipSuccess) {
        (void)hipGetLastError();
        if (hipMalloc(&p, bytes) != hipSuccess) {
            (void)hipGetLastError();
            set_error(ctx, CCD_E_OOM, "device allocation of %zu bytes for %s failed", bytes, name);
            return nullptr;
        }
        want = bytes;
    }
    b.ptr = p;
    b.bytes = want;
    ctx->device_bytes += (int64_t)want;
    return p;
}

}   // namespace ccdi

using namespace ccdi;

#define CCD_REQUIRE(ctx, cond, msg)                                         \
    do {                                                                    \
        if (!(cond)) return set_error((ctx), CCD_E_INVALID, "%s", (msg));   \
    } while (0)

// Runs a device-side validation whose failure count lands in d_counts[C_BAD]; synchronous.
static int check_bad(ccd_ctx *ctx, const char *what)
{
    CCD_HIP_CHECK(ctx, hipMemcpyAsync(ctx->h_counts + C_BAD, ctx->d_counts + C_BAD, 8, hipMemcpyDeviceToHost,
                                      ctx->stream));
    CCD_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    if (ctx->h_counts[C_BAD] != 0)
        return set_error(ctx, CCD_E_INVALID, "%s: %lld invalid entries", what, (long long)ctx->h_counts[C_BAD]);
    return CCD_OK;
}

static int validate_keys(ccd_ctx *ctx, const int32_t *keys, int64_t n, int64_t nverts, const char *what)
{
    if (n == 0) return CCD_OK;
    CCD_HIP_CHECK(ctx, hipMemsetAsync(ctx->d_counts + C_BAD, 0, 8, ctx->stream));
    int rc = launch_validate_keys(ctx, keys, 4 * n, nverts, ctx->d_counts + C_BAD);
    if (rc) return rc;
    return check_bad(ctx, what);
}

static int validate_faces(ccd_ctx *ctx, const int32_t *faces, int64_t F, int64_t nverts)
{
    if (F == 0) return CCD_OK;
    CCD_HIP_CHECK(ctx, hipMemsetAsync(ctx->d_counts + C_BAD, 0, 8, ctx->stream));
    int rc = launch_validate_keys(ctx, faces, 3 * F, nverts, ctx->d_counts + C_BAD);
    if (rc) return rc;
    return check_bad(ctx, "faces reference vertices outside [0, nverts)");
}

static int validate_history(ccd_ctx *ctx, const int64_t *off, const double *ht, int64_t nverts, int64_t nentries)
{
    CCD_HIP_CHECK(ctx, hipMemsetAsync(ctx->d_counts + C_BAD, 0, 8, ctx->stream));
    int rc = launch_validate_history(ctx, off, ht, nverts, nentries, ctx->d_counts + C_BAD);
    if (rc) return rc;
    return check_bad(ctx, "history: each vertex needs entries at t=0 .. t=1 with nondecreasing times");
}

extern "C" {

int ccd_create(ccd_ctx **out, int device)
{
    if (!out) return CCD_E_INVALID;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
        (void)hipGetLastError();
        return CCD_E_NODEVICE;
    }
    if (device < 0 || device >= n) return CCD_E_INVALID;
    ccd_ctx *ctx = new (std::nothrow) ccd_ctx;
    if (!ctx) return CCD_E_OOM;
    ctx->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->own_stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return CCD_E_HIP;
    }
    ctx->stream = ctx->own_stream;
    if (hipHostMalloc((void **)&ctx->h_counts, sizeof(int64_t) * NCOUNTERS, hipHostMallocDefault) != hipSuccess ||
        hipMalloc((void **)&ctx->d_counts, sizeof(int64_t) * NCOUNTERS) != hipSuccess) {
        (void)hipStreamDestroy(ctx->own_stream);
        delete ctx;
        return CCD_E_OOM;
    }
    ctx->device_bytes = sizeof(int64_t) * NCOUNTERS;
    *out = ctx;
    return CCD_OK;
}

int ccd_destroy(ccd_ctx *ctx)
{
    if (!ctx) return CCD_E_INVALID;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (auto &kv : ctx->bufs)
        if (kv.second.ptr) (void)hipFree(kv.second.ptr);
    if (ctx->d_counts) (void)hipFree(ctx->d_counts);
    for (int i = 0; i < 2 * CCD_NTIMERS; i++)
        if (ctx->ev[i]) (void)hipEventDestroy(ctx->ev[i]);
    if (ctx->h_counts) (void)hipHostFree(ctx->h_counts);
    if (ctx->own_stream) (void)hipStreamDestroy(ctx->own_stream);
    delete ctx;
    return CCD_OK;
}

int ccd_set_stream(ccd_ctx *ctx, void *hip_stream)
{
    if (!ctx) return CCD_E_INVALID;
    ctx->stream = hip_stream ? (hipStream_t)hip_stream : ctx->own_stream;
    return CCD_OK;
}

int ccd_synchronize(ccd_ctx *ctx)
{
    if (!ctx) return CCD_E_INVALID;
    CCD_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return CCD_OK;
}

const char *ccd_last_error(ccd_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int64_t ccd_device_bytes(ccd_ctx *ctx) { return ctx ? ctx->device_bytes : 0; }

int ccd_ctcd_primitives(ccd_ctx *ctx, int kind, const double *q, const double *eta, double eta_scalar, int64_t n,
                        uint8_t *hit, double *t)
{
    if (!ctx) return CCD_E_INVALID;
    CCD_REQUIRE(ctx, kind >= 0 && kind <= 3, "unknown primitive kind");
    CCD_REQUIRE(ctx, n >= 0, "negative item count");
    if (n == 0) return CCD_OK;
    CCD_REQUIRE(ctx, q && hit, "null q or hit");
    CCD_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    return launch_primitives(ctx, kind, q, eta, eta_scalar, n, hit, t);
}

int ccd_ctcd_narrowphase(ccd_ctx *ctx, const double *qstart, const double *qend, int64_t nverts,
                         const int32_t *vf_keys, const double *vf_eta, int64_t nvf, const int32_t *ee_keys,
                         const double *ee_eta, int64_t nee, double eta_scalar, uint8_t *vf_hit, double *vf_t,
                         uint8_t *vf_sub, uint8_t *ee_hit, double *ee_t, uint8_t *ee_sub)
{
    if (!ctx) return CCD_E_INVALID;
    CCD_REQUIRE(ctx, nverts >= 0 && nvf >= 0 && nee >= 0, "negative size");
    CCD_REQUIRE(ctx, (nvf == 0 && nee == 0) || (qstart && qend), "null positions");
    CCD_REQUIRE(ctx, nvf == 0 || (vf_keys && vf_hit), "null vertex-face keys or outputs");
    CCD_REQUIRE(ctx, nee == 0 || (ee_keys && ee_hit), "null edge-edge keys or outputs");
    CCD_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    int rc = validate_keys(ctx, vf_keys, nvf, nverts, "vertex-face keys reference vertices outside [0, nverts)");
    if (rc) return rc;
    rc = validate_keys(ctx, ee_keys, nee, nverts, "edge-edge keys reference vertices outside [0, nverts)");
    if (rc) return rc;
    prof_begin(ctx, CCD_T_NARROW_VF);
    rc = launch_narrow2(ctx, true, qstart, qend, vf_keys, vf_eta, eta_scalar, nvf, vf_hit, vf_t, vf_sub);
    if (rc) return rc;
    prof_end(ctx, CCD_T_NARROW_VF);
    prof_begin(ctx, CCD_T_NARROW_EE);
    rc = launch_narrow2(ctx, false, qstart, qend, ee_keys, ee_eta, eta_scalar, nee, ee_hit, ee_t, ee_sub);
    if (rc) return rc;
    prof_end(ctx, CCD_T_NARROW_EE);
    if (ctx->profiling) {
        CCD_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
        prof_collect(ctx);
    }
    return CCD_OK;
}

int ccd_ctcd_narrowphase_history(ccd_ctx *ctx, const int64_t *hist_off, const double *hist_t,
                                 const double *hist_pos, int64_t nverts, int64_t nentries, const int32_t *vf_keys,
                                 const double *vf_eta, int64_t nvf, const int32_t *ee_keys, const double *ee_eta,
                                 int64_t nee, double eta_scalar, uint8_t *vf_hit, double *vf_t, uint8_t *vf_sub,
                                 int32_t *vf_win, uint8_t *ee_hit, double *ee_t, uint8_t *ee_sub, int32_t *ee_win)
{
    if (!ctx) return CCD_E_INVALID;
    CCD_REQUIRE(ctx, nverts >= 0 && nentries >= 0 && nvf >= 0 && nee >= 0, "negative size");
    CCD_REQUIRE(ctx, (nvf == 0 && nee == 0) || (hist_off && hist_t && hist_pos), "null history");
    CCD_REQUIRE(ctx, nvf == 0 || (vf_keys && vf_hit), "null vertex-face keys or outputs");
    CCD_REQUIRE(ctx, nee == 0 || (ee_keys && ee_hit), "null edge-edge keys or outputs");
    CCD_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    if (nvf == 0 && nee == 0) return CCD_OK;
    int rc = validate_history(ctx, hist_off, hist_t, nverts, nentries);
    if (rc) return rc;
    rc = validate_keys(ctx, vf_keys, nvf, nverts, "vertex-face keys reference vertices outside [0, nverts)");
    if (rc) return rc;
    rc = validate_keys(ctx, ee_keys, nee, nverts, "edge-edge keys reference vertices outside [0, nverts)");
    if (rc) return rc;
    rc = launch_narrow_hist(ctx, true, hist_off, hist_t, hist_pos, vf_keys, vf_eta, eta_scalar, nvf, vf_hit, vf_t,
                            vf_sub, vf_win);
    if (rc) return rc;
    return launch_narrow_hist(ctx, false, hist_off, hist_t, hist_pos, ee_keys, ee_eta, eta_scalar, nee, ee_hit, ee_t,
                              ee_sub, ee_win);
}

static int candidates_common(ccd_ctx *ctx, const int32_t *faces, int64_t nfaces, const double *qs, const double *qe,
                             const int64_t *hoff, const double *hpos, int64_t nverts, double outer_eta,
                             const uint8_t *fixed, int32_t *vf_out, int64_t vf_cap, int64_t *vf_count,
                             int32_t *ee_out, int64_t ee_cap, int64_t *ee_count, int flags)
{
    Candidates c;
    int rc = broadphase(ctx, faces, nfaces, qs, qe, hoff, hpos, nverts, outer_eta, fixed, c);
    if (rc) return rc;
    ctx->last_vf = c.vf;
    ctx->last_ee = c.ee;
    ctx->last_nvf = c.nvf;
    ctx->last_nee = c.nee;
    if (flags & CCD_SORTED) {
        rc = sort_stencils(ctx, c.vf, c.nvf);
        if (rc) return rc;
        rc = sort_stencils(ctx, c.ee, c.nee);
        if (rc) return rc;
    }
    if (vf_count) *vf_count = c.nvf;
    if (ee_count) *ee_count = c.nee;
    int64_t nv = c.nvf < vf_cap ? c.nvf : vf_cap;
    int64_t ne = c.nee < ee_cap ? c.nee : ee_cap;
    if (nv > 0 && vf_out)
        CCD_HIP_CHECK(ctx, hipMemcpyAsync(vf_out, c.vf, (size_t)nv * 16, hipMemcpyDeviceToDevice, ctx->stream));
    if (ne > 0 && ee_out)
        CCD_HIP_CHECK(ctx, hipMemcpyAsync(ee_out, c.ee, (size_t)ne * 16, hipMemcpyDeviceToDevice, ctx->stream));
    CCD_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    prof_collect(ctx);
    if (c.nvf > vf_cap || c.nee > ee_cap)
        return set_error(ctx, CCD_E_CAPACITY, "candidate buffers too small: need %lld vertex-face, %lld edge-edge",
                         (long long)c.nvf, (long long)c.nee);
    return CCD_OK;
}

int ccd_kdop_candidates(ccd_ctx *ctx, const int32_t *faces, int64_t nfaces, const double *qstart, const double *qend,
                        int64_t nverts, double outer_eta, const uint8_t *fixed, int32_t *vf_out, int64_t vf_cap,
                        int64_t *vf_count, int32_t *ee_out, int64_t ee_cap, int64_t *ee_count, int flags)
{
    if (!ctx) return CCD_E_INVALID;
    CCD_REQUIRE(ctx, nfaces >= 0 && nverts >= 0 && vf_cap >= 0 && ee_cap >= 0, "negative size");
    CCD_REQUIRE(ctx, nfaces < (int64_t)1 << 30 && nverts < (int64_t)1 << 31, "mesh too large for 32-bit indices");
    CCD_REQUIRE(ctx, nfaces == 0 || (faces && qstart && qend), "null mesh arrays");
    CCD_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    int rc = validate_faces(ctx, faces, nfaces, nverts);
    if (rc) return rc;
    return candidates_common(ctx, faces, nfaces, qstart, qend, nullptr, nullptr, nverts, outer_eta, fixed, vf_out,
                             vf_cap, vf_count, ee_out, ee_cap, ee_count, flags);
}

int ccd_kdop_candidates_history(ccd_ctx *ctx, const int32_t *faces, int64_t nfaces, const int64_t *hist_off,
                                const double *hist_pos, int64_t nverts, int64_t nentries, double outer_eta,
                                const uint8_t *fixed, int32_t *vf_out, int64_t vf_cap, int64_t *vf_count,
                                int32_t *ee_out, int64_t ee_cap, int64_t *ee_count, int flags)
{
    if (!ctx) return CCD_E_INVALID;
    CCD_REQUIRE(ctx, nfaces >= 0 && nverts >= 0 && nentries >= 0 && vf_cap >= 0 && ee_cap >= 0, "negative size");
    CCD_REQUIRE(ctx, nfaces < (int64_t)1 << 30 && nverts < (int64_t)1 << 31, "mesh too large for 32-bit indices");
    CCD_REQUIRE(ctx, nfaces == 0 || (faces && hist_off && hist_pos), "null mesh arrays");
    CCD_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    int rc = validate_faces(ctx, faces, nfaces, nverts);
    if (rc) return rc;
    return candidates_common(ctx, faces, nfaces, nullptr, nullptr, hist_off, hist_pos, nverts, outer_eta, fixed,
                             vf_out, vf_cap, vf_count, ee_out, ee_cap, ee_count, flags);
}

int ccd_detect(ccd_ctx *ctx, const int32_t *faces, int64_t nfaces, const double *qstart, const double *qend,
               int64_t nverts, double outer_eta, double eta, const uint8_t *fixed, int32_t *vf_hits, double *vf_toi,
               int64_t vf_cap, int64_t *vf_hit_count, int32_t *ee_hits, double *ee_toi, int64_t ee_cap,
               int64_t *ee_hit_count, int64_t *vf_candidates, int64_t *ee_candidates)
{
    if (!ctx) return CCD_E_INVALID;
    CCD_REQUIRE(ctx, nfaces >= 0 && nverts >= 0 && vf_cap >= 0 && ee_cap >= 0, "negative size");
    CCD_REQUIRE(ctx, nfaces < (int64_t)1 << 30 && nverts < (int64_t)1 << 31, "mesh too large for 32-bit indices");
    CCD_REQUIRE(ctx, nfaces == 0 || (faces && qstart && qend), "null mesh arrays");
    CCD_REQUIRE(ctx, (vf_cap == 0 || vf_hits) && (ee_cap == 0 || ee_hits), "null collision buffers");
    CCD_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    int rc = validate_faces(ctx, faces, nfaces, nverts);
    if (rc) return rc;
    Candidates c;
    rc = broadphase(ctx, faces, nfaces, qstart, qend, nullptr, nullptr, nverts, outer_eta, fixed, c);
    if (rc) return rc;
    ctx->last_vf = c.vf;
    ctx->last_ee = c.ee;
    ctx->last_nvf = c.nvf;
    ctx->last_nee = c.nee;
    CCD_HIP_CHECK(ctx, hipMemsetAsync(ctx->d_counts + C_VF_HITS, 0, 16, ctx->stream));
    prof_begin(ctx, CCD_T_NARROW_VF);
    rc = launch_narrow2_compact(ctx, true, qstart, qend, c.vf, ctx->d_counts + C_VF, c.nvf, eta, vf_hits, vf_toi,
                                vf_cap, ctx->d_counts + C_VF_HITS);
    if (rc) return rc;
    prof_end(ctx, CCD_T_NARROW_VF);
    prof_begin(ctx, CCD_T_NARROW_EE);
    rc = launch_narrow2_compact(ctx, false, qstart, qend, c.ee, ctx->d_counts + C_EE, c.nee, eta, ee_hits, ee_toi,
                                ee_cap, ctx->d_counts + C_EE_HITS);
    if (rc) return rc;
    prof_end(ctx, CCD_T_NARROW_EE);
    CCD_HIP_CHECK(ctx, hipMemcpyAsync(ctx->h_counts, ctx->d_counts, sizeof(int64_t) * NCOUNTERS, hipMemcpyDeviceToHost,
                                      ctx->stream));
    CCD_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    prof_collect(ctx);
    int64_t nvh = ctx->h_counts[C_VF_HITS], neh = ctx->h_counts[C_EE_HITS];
    if (vf_hit_count) *vf_hit_count = nvh;
    if (ee_hit_count) *ee_hit_count = neh;
    if (vf_candidates) *vf_candidates = c.nvf;
    if (ee_candidates) *ee_candidates = c.nee;
    if (nvh > vf_cap || neh > ee_cap)
        return set_error(ctx, CCD_E_CAPACITY, "collision buffers too small: need %lld vertex-face, %lld edge-edge",
                         (long long)nvh, (long long)neh);
    return CCD_OK;
}

int ccd_distance_primitives(ccd_ctx *ctx, int kind, const double *q, double eta, int64_t n, uint8_t *flag,
                            double *vec, double *bary)
{
    if (!ctx) return CCD_E_INVALID;
    CCD_REQUIRE(ctx, kind >= 0 && kind <= 3, "unknown proximity kind");
    CCD_REQUIRE(ctx, n >= 0, "negative item count");
    if (n == 0) return CCD_OK;
    CCD_REQUIRE(ctx, q, "null q");
    CCD_REQUIRE(ctx, kind >= CCD_DIST_VERTEX_FACE || flag, "null flag output");
    CCD_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    return launch_distance_primitives(ctx, kind, q, eta, n, flag, vec, bary);
}

int ccd_mesh_self_distance(ccd_ctx *ctx, const double *verts, int64_t nverts, const int32_t *faces, int64_t nfaces,
                           const uint8_t *fixed, double *distance, int64_t *nvf, int64_t *nee)
{
    if (!ctx) return CCD_E_INVALID;
    CCD_REQUIRE(ctx, nverts >= 0 && nfaces >= 0, "negative size");
    CCD_REQUIRE(ctx, nfaces < (int64_t)1 << 30 && nverts < (int64_t)1 << 31, "mesh too large for 32-bit indices");
    CCD_REQUIRE(ctx, distance, "null distance output");
    CCD_REQUIRE(ctx, (nfaces == 0 || faces) && (nverts == 0 || verts), "null mesh arrays");
    CCD_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    int rc = validate_faces(ctx, faces, nfaces, nverts);
    if (rc) return rc;
    return mesh_self_distance(ctx, verts, nverts, faces, nfaces, fixed, distance, nvf, nee);
}

int ccd_device_alloc(ccd_ctx *ctx, int64_t bytes, void **ptr)
{
    if (!ctx || !ptr || bytes < 0) return CCD_E_INVALID;
    *ptr = nullptr;
    if (bytes == 0) return CCD_OK;
    CCD_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    if (hipMalloc(ptr, (size_t)bytes) != hipSuccess) {
        (void)hipGetLastError();
        return set_error(ctx, CCD_E_OOM, "device allocation of %lld bytes failed", (long long)bytes);
    }
    return CCD_OK;
}

int ccd_device_free(ccd_ctx *ctx, void *ptr)
{
    if (!ctx) return CCD_E_INVALID;
    if (!ptr) return CCD_OK;
    CCD_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    CCD_HIP_CHECK(ctx, hipFree(ptr));
    return CCD_OK;
}

int ccd_copy_to_device(ccd_ctx *ctx, void *dst, const void *src, int64_t bytes)
{
    if (!ctx || bytes < 0 || (bytes > 0 && (!dst || !src))) return CCD_E_INVALID;
    if (bytes == 0) return CCD_OK;
    CCD_HIP_CHECK(ctx, hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyHostToDevice, ctx->stream));
    CCD_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return CCD_OK;
}

int ccd_copy_to_host(ccd_ctx *ctx, void *dst, const void *src, int64_t bytes)
{
    if (!ctx || bytes < 0 || (bytes > 0 && (!dst || !src))) return CCD_E_INVALID;
    if (bytes == 0) return CCD_OK;
    CCD_HIP_CHECK(ctx, hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDeviceToHost, ctx->stream));
    CCD_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    return CCD_OK;
}

int ccd_set_profiling(ccd_ctx *ctx, int on)
{
    if (!ctx) return CCD_E_INVALID;
    if (on && !ctx->ev[0]) {
        CCD_HIP_CHECK(ctx, hipSetDevice(ctx->device));
        for (int i = 0; i < 2 * CCD_NTIMERS; i++) CCD_HIP_CHECK(ctx, hipEventCreate(&ctx->ev[i]));
    }
    ctx->profiling = on != 0;
    return CCD_OK;
}

int ccd_stage_times(ccd_ctx *ctx, double *ms, int64_t *calls, int reset)
{
    if (!ctx) return CCD_E_INVALID;
    for (int t = 0; t < CCD_NTIMERS; t++) {
        if (ms) ms[t] = ctx->stage_ms[t];
        if (calls) calls[t] = ctx->stage_calls[t];
        if (reset) {
            ctx->stage_ms[t] = 0;
            ctx->stage_calls[t] = 0;
        }
    }
    return CCD_OK;
}

int ccd_stage_flops(int64_t *flops)
{
    if (!flops) return CCD_E_INVALID;
    for (int s = 0; s < CCD_NSTAGES; s++) flops[s] = ccd::STAGE_FLOPS[s];
    return CCD_OK;
}

int ccd_narrowphase_work(ccd_ctx *ctx, const double *qstart, const double *qend, int64_t nverts,
                         const int32_t *vf_keys, int64_t nvf, const int32_t *ee_keys, int64_t nee, double eta,
                         int64_t *vf_counts, int64_t *ee_counts)
{
    if (!ctx) return CCD_E_INVALID;
    CCD_REQUIRE(ctx, nverts >= 0 && nvf >= 0 && nee >= 0, "negative size");
    CCD_REQUIRE(ctx, (nvf == 0 || vf_keys) && (nee == 0 || ee_keys) && ((nvf == 0 && nee == 0) || (qstart && qend)),
                "null input");
    CCD_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    int rc = validate_keys(ctx, vf_keys, nvf, nverts, "vertex-face keys reference vertices outside [0, nverts)");
    if (rc) return rc;
    rc = validate_keys(ctx, ee_keys, nee, nverts, "edge-edge keys reference vertices outside [0, nverts)");
    if (rc) return rc;
    unsigned long long *d = (unsigned long long *)buffer(ctx, "work_counts", 2 * CCD_NSTAGES * 8);
    if (!d) return CCD_E_OOM;
    CCD_HIP_CHECK(ctx, hipMemsetAsync(d, 0, 2 * CCD_NSTAGES * 8, ctx->stream));
    rc = launch_narrow2_count(ctx, true, qstart, qend, vf_keys, nvf, eta, d);
    if (rc) return rc;
    rc = launch_narrow2_count(ctx, false, qstart, qend, ee_keys, nee, eta, d + CCD_NSTAGES);
    if (rc) return rc;
    int64_t h[2 * CCD_NSTAGES];
    CCD_HIP_CHECK(ctx, hipMemcpyAsync(h, d, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
    CCD_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    for (int s = 0; s < CCD_NSTAGES; s++) {
        if (vf_counts) vf_counts[s] = h[s];
        if (ee_counts) ee_counts[s] = h[CCD_NSTAGES + s];
    }
    return CCD_OK;
}

int ccd_last_candidates(ccd_ctx *ctx, const int32_t **vf_keys, int64_t *nvf, const int32_t **ee_keys, int64_t *nee)
{
    if (!ctx) return CCD_E_INVALID;
    if (vf_keys) *vf_keys = ctx->last_vf;
    if (nvf) *nvf = ctx->last_nvf;
    if (ee_keys) *ee_keys = ctx->last_ee;
    if (nee) *nee = ctx->last_nee;
    return CCD_OK;
}

}   // extern "C"
