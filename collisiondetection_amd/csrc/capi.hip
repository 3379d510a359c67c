// C ABI of libccd_amd.so (declarations and contracts in include/ccd_amd.h).
#include <cstdarg>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

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
    if (!ctx->profiling || ctx->capturing) return;
    (void)hipEventRecord(ctx->ev[2 * timer], ctx->stream);
}

void prof_end(ccd_ctx *ctx, int timer)
{
    if (!ctx->profiling || ctx->capturing) return;
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
    if (b.external) {
        set_error(ctx, CCD_E_CAPACITY, "external buffer %s too small (%zu < %zu bytes)", name, b.bytes, bytes);
        return nullptr;
    }
    if (b.ptr) {
        // A captured pipeline holds the old address: it must not be replayed again.
        ctx->graph_valid = false;
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

// ------------------------------------------------------------------------------------------------
// Device scratch of ccd_ctcd_narrowphase_noalloc: a file-scope device array, one copy per GPU, held
// until the process ends.  Host code never passes the array's name to a kernel (on the host it is
// only a stand-in for the symbol): it looks up each device's own address once, at set-up, and passes
// that.  Layout for up to CCD_NOALLOC_MAX_STENCILS stencils (one stencil kind at a time):
//   [0, 256)    the invalid-key counter (uint64)
//   then sig, sig2 (uint64 per stencil), idx, idx2 (int32 per stencil), the radix sort's temporary.
// ------------------------------------------------------------------------------------------------
namespace {
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)26 << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mu;
unsigned char *g_scratch_ptr[MAX_DEVICES] = {};

// The current device's address of `scratch` (the current device must be `device`).
int scratch_address(int device, unsigned char **out)
{
    *out = nullptr;
    if (device < 0 || device >= MAX_DEVICES) return CCD_E_INVALID;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
            (void)hipGetLastError();
            return CCD_E_HIP;
        }
        g_scratch_ptr[device] = (unsigned char *)p;
    }
    *out = g_scratch_ptr[device];
    return CCD_OK;
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }
}   // namespace

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
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cus > 0)
        ctx->num_cus = cus;
    if (hipHostMalloc((void **)&ctx->h_counts, sizeof(int64_t) * NCOUNTERS, hipHostMallocDefault) != hipSuccess ||
        hipMalloc((void **)&ctx->d_counts, sizeof(int64_t) * NCOUNTERS) != hipSuccess) {
        (void)hipStreamDestroy(ctx->own_stream);
        delete ctx;
        return CCD_E_OOM;
    }
    ctx->device_bytes = sizeof(int64_t) * NCOUNTERS;
    // Set-up for the allocation-free entry point: this device's address of its scratch array.
    unsigned char *scr = nullptr;
    if (scratch_address(device, &scr) != CCD_OK) {
        (void)hipFree(ctx->d_counts);
        (void)hipHostFree(ctx->h_counts);
        (void)hipStreamDestroy(ctx->own_stream);
        delete ctx;
        return CCD_E_HIP;
    }
    *out = ctx;
    return CCD_OK;
}

int ccd_destroy(ccd_ctx *ctx)
{
    if (!ctx) return CCD_E_INVALID;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (auto &kv : ctx->bufs)
        if (kv.second.ptr && !kv.second.external) (void)hipFree(kv.second.ptr);
    if (ctx->d_counts) (void)hipFree(ctx->d_counts);
    for (int i = 0; i < 2 * CCD_NTIMERS; i++)
        if (ctx->ev[i]) (void)hipEventDestroy(ctx->ev[i]);
    if (ctx->h_counts) (void)hipHostFree(ctx->h_counts);
    if (ctx->graph_exec) (void)hipGraphExecDestroy(ctx->graph_exec);
    if (ctx->graph) (void)hipGraphDestroy(ctx->graph);
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

int ccd_ctcd_narrowphase_noalloc(ccd_ctx *ctx, const double *qstart, const double *qend, int64_t nverts,
                                 const int32_t *vf_keys, const double *vf_eta, int64_t nvf, const int32_t *ee_keys,
                                 const double *ee_eta, int64_t nee, double eta_scalar, uint8_t *vf_hit, double *vf_t,
                                 uint8_t *vf_sub, uint8_t *ee_hit, double *ee_t, uint8_t *ee_sub)
{
    if (!ctx) return CCD_E_INVALID;
    CCD_REQUIRE(ctx, nverts >= 0 && nvf >= 0 && nee >= 0, "negative size");
    if (nvf > CCD_NOALLOC_MAX_STENCILS || nee > CCD_NOALLOC_MAX_STENCILS)
        return set_error(ctx, CCD_E_CAPACITY,
                         "ccd_ctcd_narrowphase_noalloc takes at most %lld stencils per kind (got %lld vertex-face, "
                         "%lld edge-edge)",
                         (long long)CCD_NOALLOC_MAX_STENCILS, (long long)nvf, (long long)nee);
    CCD_REQUIRE(ctx, (nvf == 0 && nee == 0) || (qstart && qend), "null positions");
    CCD_REQUIRE(ctx, nvf == 0 || (vf_keys && vf_hit), "null vertex-face keys or outputs");
    CCD_REQUIRE(ctx, nee == 0 || (ee_keys && ee_hit), "null edge-edge keys or outputs");
    CCD_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    if (nvf == 0 && nee == 0) return CCD_OK;
    unsigned char *base = nullptr;
    if (scratch_address(ctx->device, &base) != CCD_OK)
        return set_error(ctx, CCD_E_HIP, "no address for the device scratch array");
    const size_t n = (size_t)CCD_NOALLOC_MAX_STENCILS;
    OrderScratch os;
    size_t o = 256;
    os.sig = (uint64_t *)(base + o);
    o = align256(o + 8 * n);
    os.sig2 = (uint64_t *)(base + o);
    o = align256(o + 8 * n);
    os.idx = (int32_t *)(base + o);
    o = align256(o + 4 * n);
    os.idx2 = (int32_t *)(base + o);
    o = align256(o + 4 * n);
    os.tmp = base + o;
    o += radix_sort_temp_bytes((int64_t)n);
    if (o > NOALLOC_SCRATCH_BYTES) return set_error(ctx, CCD_E_INVALID, "scratch layout exceeds the array");
    // Key validation into the scratch counter (cleared by a fill kernel; the array keeps the last
    // call's contents), read back through the context's pinned mailbox.
    int64_t *bad = (int64_t *)base;
    int rc = fill_u64(ctx, (uint64_t *)bad, 1, 0);
    if (!rc) rc = launch_validate_keys(ctx, vf_keys, 4 * nvf, nverts, bad);
    if (!rc) rc = launch_validate_keys(ctx, ee_keys, 4 * nee, nverts, bad);
    if (rc) return rc;
    CCD_HIP_CHECK(ctx, hipMemcpyAsync(ctx->h_counts + C_BAD, bad, 8, hipMemcpyDeviceToHost, ctx->stream));
    CCD_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    if (ctx->h_counts[C_BAD] != 0)
        return set_error(ctx, CCD_E_INVALID, "%lld stencil keys reference vertices outside [0, nverts)",
                         (long long)ctx->h_counts[C_BAD]);
    // Both kinds reuse the same ordering buffers: the edge-edge launches follow the vertex-face ones on
    // the stream.
    rc = launch_narrow2(ctx, true, qstart, qend, vf_keys, vf_eta, eta_scalar, nvf, vf_hit, vf_t, vf_sub, &os);
    if (rc) return rc;
    return launch_narrow2(ctx, false, qstart, qend, ee_keys, ee_eta, eta_scalar, nee, ee_hit, ee_t, ee_sub, &os);
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
    // Long histories (16 or more entries per vertex on average): one wavefront per stencil tests 64
    // stitched windows at a time; short ones: one lane per stencil.
    const bool wave = nentries >= 16 * nverts;
    rc = launch_narrow_hist(ctx, true, hist_off, hist_t, hist_pos, vf_keys, vf_eta, eta_scalar, nvf, vf_hit, vf_t,
                            vf_sub, vf_win, wave);
    if (rc) return rc;
    return launch_narrow_hist(ctx, false, hist_off, hist_t, hist_pos, ee_keys, ee_eta, eta_scalar, nee, ee_hit, ee_t,
                              ee_sub, ee_win, wave);
}

int ccd_separating_plane_narrowphase(ccd_ctx *ctx, const int64_t *hist_off, const double *hist_t,
                                     const double *hist_pos, int64_t nverts, int64_t nentries, const int32_t *vf_keys,
                                     const double *vf_eta, int64_t nvf, const int32_t *ee_keys, const double *ee_eta,
                                     int64_t nee, double eta_scalar, uint8_t *vf_status, uint8_t *ee_status,
                                     int64_t *undecided)
{
    if (!ctx) return CCD_E_INVALID;
    CCD_REQUIRE(ctx, nverts >= 0 && nentries >= 0 && nvf >= 0 && nee >= 0, "negative size");
    CCD_REQUIRE(ctx, (nvf == 0 && nee == 0) || (hist_off && hist_t && hist_pos), "null history");
    CCD_REQUIRE(ctx, nvf == 0 || (vf_keys && vf_status), "null vertex-face keys or outputs");
    CCD_REQUIRE(ctx, nee == 0 || (ee_keys && ee_status), "null edge-edge keys or outputs");
    CCD_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    if (undecided) *undecided = 0;
    if (nvf == 0 && nee == 0) return CCD_OK;
    int rc = validate_history(ctx, hist_off, hist_t, nverts, nentries);
    if (rc) return rc;
    rc = validate_keys(ctx, vf_keys, nvf, nverts, "vertex-face keys reference vertices outside [0, nverts)");
    if (rc) return rc;
    rc = validate_keys(ctx, ee_keys, nee, nverts, "edge-edge keys reference vertices outside [0, nverts)");
    if (rc) return rc;
    uint64_t *dev = (uint64_t *)buffer(ctx, "sp_scalars", 16);   // minimum gap bits, undecided count
    if (!dev) return CCD_E_OOM;
    rc = fill_u64(ctx, dev + 1, 1, 0);
    if (rc) return rc;
    rc = launch_min_gap(ctx, hist_off, hist_t, nverts, dev);
    if (rc) return rc;
    // Long histories (16 or more entries per vertex on average, as the velocity filter builds): one
    // wavefront per stencil shares each trajectory scan; short ones: one lane per stencil.
    const bool wave = nentries >= 16 * nverts;
    rc = launch_separating_plane(ctx, true, hist_off, hist_t, hist_pos, vf_keys, vf_eta, eta_scalar, nvf, dev,
                                 vf_status, (int64_t *)(dev + 1), wave);
    if (rc) return rc;
    rc = launch_separating_plane(ctx, false, hist_off, hist_t, hist_pos, ee_keys, ee_eta, eta_scalar, nee, dev,
                                 ee_status, (int64_t *)(dev + 1), wave);
    if (rc) return rc;
    int64_t und = 0;
    CCD_HIP_CHECK(ctx, hipMemcpyAsync(&und, dev + 1, 8, hipMemcpyDeviceToHost, ctx->stream));
    CCD_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    if (undecided) *undecided = und;
    return CCD_OK;
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

// ----------------------------------------------------------------------------------------------
// Detection core shared by ccd_detect and ccd_batch_detect.
// ----------------------------------------------------------------------------------------------
struct DetectArgs {
    const int64_t *inst_off;
    int64_t ninst;
    const int32_t *faces;
    int64_t F;
    const double *qs, *qe;
    int64_t V;
    double outer_eta, eta;
    const uint8_t *fixed;
    int32_t *vf_hits;
    double *vf_toi;
    int64_t vf_cap;
    int32_t *ee_hits;
    double *ee_toi;
    int64_t ee_cap;
};

// The narrowphase half: counters reset by fill kernels, then the two persistent queue kernels.  Pure
// launches (capturable).  n_max bounds the grids: the stencil counts are read on the device.
// exact: nvf_max / nee_max are this call's stencil counts (the synchronous path), so the stencils can
// be grouped by their vertex-edge signature first; the captured path passes capacities instead.
static int narrow_launches(ccd_ctx *ctx, const DetectArgs &a, int64_t nvf_max, int64_t nee_max, bool exact)
{
    // work[0..1]: the two queues; work[2..3]: the edge-edge deferral list's counter and second queue.
    unsigned long long *work = (unsigned long long *)buffer(ctx, "queue_work", 48);
    if (!work) return CCD_E_OOM;
    int rc = fill_u64(ctx, (uint64_t *)(ctx->d_counts + C_VF_HITS), 2, 0);
    if (!rc) rc = fill_u64(ctx, (uint64_t *)work, 4, 0);
    if (rc) return rc;
    const EeDeferral def = ee_deferral(ctx, nee_max, work + 2);
    const int32_t *vfk = (const int32_t *)ctx->bufs["vf_keys"].ptr;
    const int32_t *eek = (const int32_t *)ctx->bufs["ee_keys"].ptr;
    // The stage timers include the grouping.  Without exact host counts (a captured pipeline) the
    // signature pass reads the device counts and sorts the unused capacity last.
    const bool group_vf = ctx->group_stencils && ctx->vf_shape.group;
    const bool group_ee = ctx->group_stencils && ctx->ee_shape.group;
    prof_begin(ctx, CCD_T_NARROW_VF);
    if (group_vf) {
        rc = group_by_signature(ctx, true, a.qs, a.qe, vfk, nvf_max, a.eta, exact ? nullptr : ctx->d_counts + C_VF, &vfk);
        if (rc) return rc;
    }
    rc = launch_narrow2_queue(ctx, true, a.qs, a.qe, vfk, ctx->d_counts + C_VF, nvf_max, a.eta, a.vf_hits, a.vf_toi,
                              a.vf_cap, ctx->d_counts + C_VF_HITS, work, ctx->vf_shape);
    if (rc) return rc;
    prof_end(ctx, CCD_T_NARROW_VF);
    prof_begin(ctx, CCD_T_NARROW_EE);
    if (group_ee) {
        rc = group_by_signature(ctx, false, a.qs, a.qe, eek, nee_max, a.eta, exact ? nullptr : ctx->d_counts + C_EE, &eek);
        if (rc) return rc;
    }
    rc = launch_narrow2_queue(ctx, false, a.qs, a.qe, eek, ctx->d_counts + C_EE, nee_max, a.eta, a.ee_hits, a.ee_toi,
                              a.ee_cap, ctx->d_counts + C_EE_HITS, work + 1, ctx->ee_shape, &def);
    if (rc) return rc;
    prof_end(ctx, CCD_T_NARROW_EE);
    return CCD_OK;
}

// Synchronous run: broadphase with capacity checks, then the narrowphase.  Counts end in h_counts.
static int detect_sync(ccd_ctx *ctx, const DetectArgs &a)
{
    Candidates c;
    int rc = broadphase(ctx, a.faces, a.F, a.qs, a.qe, nullptr, nullptr, a.V, a.outer_eta, a.fixed, c, false,
                        a.inst_off, (int)a.ninst);
    if (rc) return rc;
    ctx->last_vf = c.vf;
    ctx->last_ee = c.ee;
    ctx->last_nvf = c.nvf;
    ctx->last_nee = c.nee;
    rc = narrow_launches(ctx, a, c.nvf, c.nee, true);
    if (rc) return rc;
    CCD_HIP_CHECK(ctx, hipMemcpyAsync(ctx->h_counts, ctx->d_counts, sizeof(int64_t) * NCOUNTERS, hipMemcpyDeviceToHost,
                                      ctx->stream));
    CCD_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    prof_collect(ctx);
    return CCD_OK;
}

static void graph_drop(ccd_ctx *ctx)
{
    if (ctx->graph_exec) (void)hipGraphExecDestroy(ctx->graph_exec);
    if (ctx->graph) (void)hipGraphDestroy(ctx->graph);
    ctx->graph_exec = nullptr;
    ctx->graph = nullptr;
    ctx->graph_valid = false;
}

// Graph path: replay when the key matches; otherwise run synchronously (which sizes every buffer and
// gives this call's result) and capture the pipeline for the next calls.
static int detect_graph(ccd_ctx *ctx, const DetectArgs &a)
{
    unsigned char key[256] = {};
    static_assert(sizeof(DetectArgs) + 3 * sizeof(int64_t) <= sizeof(key), "graph key");
    memcpy(key, &a, sizeof(DetectArgs));
    int64_t caps[3] = {ctx->cap_pairs, ctx->cap_vf, ctx->cap_ee};
    memcpy(key + sizeof(DetectArgs), caps, sizeof(caps));
    if (ctx->graph_valid && memcmp(key, ctx->graph_key, sizeof(key)) == 0) {
        if (ctx->profiling) (void)hipEventRecord(ctx->ev[2 * CCD_T_PIPELINE], ctx->stream);
        CCD_HIP_CHECK(ctx, hipGraphLaunch(ctx->graph_exec, ctx->stream));
        if (ctx->profiling) {
            (void)hipEventRecord(ctx->ev[2 * CCD_T_PIPELINE + 1], ctx->stream);
            ctx->pending[CCD_T_PIPELINE] = true;
        }
        CCD_HIP_CHECK(ctx, hipMemcpyAsync(ctx->h_counts, ctx->d_counts, sizeof(int64_t) * NCOUNTERS,
                                          hipMemcpyDeviceToHost, ctx->stream));
        CCD_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
        prof_collect(ctx);
        ctx->graph_replays++;
        int64_t *h = ctx->h_counts;
        if (h[C_PAIRS] <= ctx->cap_pairs && h[C_VF] <= ctx->cap_vf && h[C_EE] <= ctx->cap_ee && h[C_BAD] == 0) {
            ctx->last_vf = (const int32_t *)ctx->bufs["vf_keys"].ptr;
            ctx->last_ee = (const int32_t *)ctx->bufs["ee_keys"].ptr;
            ctx->last_nvf = h[C_VF];
            ctx->last_nee = h[C_EE];
            return CCD_OK;
        }
        graph_drop(ctx);   // outgrew its buffers (or the mesh needs sort-based deduplication)
    }
    graph_drop(ctx);
    int rc = detect_sync(ctx, a);
    if (rc) return rc;
    if (ctx->h_counts[C_BAD] != 0) return CCD_OK;   // raw emission needs a host-driven sort: no graph
    int64_t saved[NCOUNTERS];
    memcpy(saved, ctx->h_counts, sizeof(saved));
    // Capture with the capacities the synchronous run settled on; buffers used inside the capture are
    // sized for them first (nothing can be allocated while capturing).
    if (ctx->group_stencils && (ctx->vf_shape.group || ctx->ee_shape.group)) {
        rc = reserve_grouping(ctx, ctx->cap_vf, ctx->cap_ee);
        if (rc) return rc;
    }
    if (!buffer(ctx, "queue_work", 48)) return CCD_E_OOM;
    if (ctx->ee_split && ctx->cap_ee <= 0x7fffffffll) {
        const int64_t cd = ee_deferral_capacity(ctx, ctx->cap_ee);
        if (!buffer(ctx, "ee_def_idx", (size_t)cd * 4) || !buffer(ctx, "ee_def_sub", (size_t)cd)) return CCD_E_OOM;
    }
    int64_t caps2[3] = {ctx->cap_pairs, ctx->cap_vf, ctx->cap_ee};
    memcpy(key + sizeof(DetectArgs), caps2, sizeof(caps2));
    ctx->capturing = true;
    hipError_t e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        ctx->capturing = false;
        return set_error(ctx, CCD_E_HIP, "hipStreamBeginCapture failed: %s", hipGetErrorString(e));
    }
    rc = broadphase_launch(ctx, a.faces, a.F, a.qs, a.qe, nullptr, nullptr, a.V, a.outer_eta, a.fixed, false,
                           a.inst_off, (int)a.ninst);
    if (!rc) rc = narrow_launches(ctx, a, ctx->cap_vf, ctx->cap_ee, false);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(ctx->stream, &g);
    ctx->capturing = false;
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess || !g) return set_error(ctx, CCD_E_HIP, "hipStreamEndCapture failed: %s", hipGetErrorString(e));
    hipGraphExec_t ge = nullptr;
    e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        return set_error(ctx, CCD_E_HIP, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
    }
    ctx->graph = g;
    ctx->graph_exec = ge;
    ctx->graph_valid = true;
    memcpy(ctx->graph_key, key, sizeof(key));
    memcpy(ctx->h_counts, saved, sizeof(saved));   // this call's result is the synchronous run's
    return CCD_OK;
}

static int detect_common(ccd_ctx *ctx, const DetectArgs &a, bool graph, int64_t *vf_hit_count, int64_t *ee_hit_count,
                         int64_t *vf_candidates, int64_t *ee_candidates)
{
    int rc = validate_faces(ctx, a.faces, a.F, a.V);
    if (rc) return rc;
    rc = graph ? detect_graph(ctx, a) : detect_sync(ctx, a);
    if (rc) return rc;
    int64_t nvh = ctx->h_counts[C_VF_HITS], neh = ctx->h_counts[C_EE_HITS];
    if (vf_hit_count) *vf_hit_count = nvh;
    if (ee_hit_count) *ee_hit_count = neh;
    if (vf_candidates) *vf_candidates = ctx->last_nvf;
    if (ee_candidates) *ee_candidates = ctx->last_nee;
    if (nvh > a.vf_cap || neh > a.ee_cap)
        return set_error(ctx, CCD_E_CAPACITY, "collision buffers too small: need %lld vertex-face, %lld edge-edge",
                         (long long)nvh, (long long)neh);
    return CCD_OK;
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
    DetectArgs a{nullptr, 1, faces, nfaces, qstart, qend, nverts, outer_eta, eta, fixed,
                 vf_hits, vf_toi, vf_cap, ee_hits, ee_toi, ee_cap};
    return detect_common(ctx, a, false, vf_hit_count, ee_hit_count, vf_candidates, ee_candidates);
}

int ccd_batch_detect(ccd_ctx *ctx, int64_t ninst, const int64_t *inst_face_off, const int32_t *faces, int64_t nfaces,
                     const double *qstart, const double *qend, int64_t nverts, double outer_eta, double eta,
                     int32_t *vf_hits, double *vf_toi, int64_t vf_cap, int64_t *vf_hit_count, int32_t *ee_hits,
                     double *ee_toi, int64_t ee_cap, int64_t *ee_hit_count, int64_t *vf_candidates,
                     int64_t *ee_candidates, int flags)
{
    if (!ctx) return CCD_E_INVALID;
    CCD_REQUIRE(ctx, ninst >= 1 && ninst < (1 << 20), "instance count must be in [1, 2^20)");
    CCD_REQUIRE(ctx, inst_face_off, "null instance offsets");
    CCD_REQUIRE(ctx, nfaces >= 0 && nverts >= 0 && vf_cap >= 0 && ee_cap >= 0, "negative size");
    CCD_REQUIRE(ctx, nfaces < (int64_t)1 << 30 && nverts < (int64_t)1 << 31, "mesh too large for 32-bit indices");
    CCD_REQUIRE(ctx, nfaces == 0 || (faces && qstart && qend), "null mesh arrays");
    CCD_REQUIRE(ctx, (vf_cap == 0 || vf_hits) && (ee_cap == 0 || ee_hits), "null collision buffers");
    CCD_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    // The offsets must run 0 = off[0] <= off[1] <= ... <= off[ninst] = nfaces (checked on the host copy).
    std::vector<int64_t> off(ninst + 1);
    CCD_HIP_CHECK(ctx, hipMemcpyAsync(off.data(), inst_face_off, (ninst + 1) * 8, hipMemcpyDeviceToHost, ctx->stream));
    CCD_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    bool ok = off[0] == 0 && off[ninst] == nfaces;
    for (int64_t k = 0; k < ninst && ok; k++) ok = off[k] <= off[k + 1];
    CCD_REQUIRE(ctx, ok, "instance face offsets must be nondecreasing from 0 to nfaces");
    DetectArgs a{ninst > 1 ? inst_face_off : nullptr, ninst, faces, nfaces, qstart, qend, nverts, outer_eta, eta,
                 nullptr, vf_hits, vf_toi, vf_cap, ee_hits, ee_toi, ee_cap};
    return detect_common(ctx, a, (flags & CCD_GRAPH) != 0, vf_hit_count, ee_hit_count, vf_candidates, ee_candidates);
}

int ccd_set_queue_shape(ccd_ctx *ctx, const int *shape)
{
    if (!ctx || !shape) return CCD_E_INVALID;
    QueueShape s[2];
    for (int k = 0; k < 2; k++) {
        s[k].block = shape[3 * k];
        s[k].tile = shape[3 * k + 1];
        s[k].waves_per_cu = shape[3 * k + 2];
        s[k].group = shape[6 + k];
        if ((s[k].block != 64 && s[k].block != 128 && s[k].block != 256) || s[k].tile < 64 || s[k].tile % 64 ||
            s[k].waves_per_cu < 1 || s[k].waves_per_cu > 32 || (s[k].group != 0 && s[k].group != 1))
            return set_error(ctx, CCD_E_INVALID, "invalid queue shape");
    }
    ctx->vf_shape = s[0];
    ctx->ee_shape = s[1];
    graph_drop(ctx);   // captured launches carry the old shape
    return CCD_OK;
}

int ccd_set_edge_edge_split(ccd_ctx *ctx, int on)
{
    if (!ctx) return CCD_E_INVALID;
    ctx->ee_split = on != 0;
    graph_drop(ctx);   // captured launches carry the old form
    return CCD_OK;
}

int ccd_autotune(ccd_ctx *ctx, const double *qstart, const double *qend, double eta, int *shape)
{
    if (!ctx) return CCD_E_INVALID;
    CCD_REQUIRE(ctx, qstart && qend, "null positions");
    CCD_REQUIRE(ctx, ctx->last_vf || ctx->last_ee, "no candidates yet: run a detection on this context first");
    CCD_HIP_CHECK(ctx, hipSetDevice(ctx->device));
    hipStream_t st = ctx->stream;
    const int blocks[3] = {64, 128, 256};
    const int tiles[7] = {64, 128, 256, 512, 1024, 2048, 4096};
    const int waves[4] = {4, 8, 12, 16};
    int64_t nmax[2] = {ctx->last_nvf, ctx->last_nee};
    const int32_t *keys[2] = {ctx->last_vf, ctx->last_ee};
    int32_t *sink = (int32_t *)buffer(ctx, "tune_sink", 16 * 4096 + 64);
    double *sink_t = (double *)buffer(ctx, "tune_sink_t", 8 * 4096 + 64);
    unsigned long long *work = (unsigned long long *)buffer(ctx, "tune_work", 32);   // queue; [2..3] deferral
    int64_t *cnt = (int64_t *)buffer(ctx, "tune_cnt", 16);
    int64_t *ndev = (int64_t *)buffer(ctx, "tune_n", 16);
    if (!sink || !sink_t || !work || !cnt || !ndev) return CCD_E_OOM;
    hipEvent_t e0, e1;
    CCD_HIP_CHECK(ctx, hipEventCreate(&e0));
    CCD_HIP_CHECK(ctx, hipEventCreate(&e1));
    // Milliseconds of `launch` on the context's stream: the second of two runs (the first warms caches).
    auto timed = [&](auto &&launch, float &ms) -> int {
        for (int rep = 0; rep < 2; rep++) {
            CCD_HIP_CHECK(ctx, hipEventRecord(e0, st));
            int rc = launch();
            if (rc) return rc;
            CCD_HIP_CHECK(ctx, hipEventRecord(e1, st));
            CCD_HIP_CHECK(ctx, hipEventSynchronize(e1));
            CCD_HIP_CHECK(ctx, hipEventElapsedTime(&ms, e0, e1));
        }
        return CCD_OK;
    };
    // Edge-edge shapes are timed as the pipeline runs them (two passes unless switched off).
    const EeDeferral tdef = ee_deferral(ctx, nmax[1], work + 2);
    auto queue = [&](int kind, const int32_t *k, const QueueShape &s) {
        return [&, kind, k, s]() -> int {
            int rc = fill_u64(ctx, (uint64_t *)work, 4, 0);
            if (!rc) rc = fill_u64(ctx, (uint64_t *)cnt, 1, 0);
            if (!rc)
                rc = launch_narrow2_queue(ctx, kind == 0, qstart, qend, k, ndev, nmax[kind], eta, sink, sink_t, 0, cnt,
                                          work, s, kind == 1 ? &tdef : nullptr);
            return rc;
        };
    };
    QueueShape best[2];
    for (int kind = 0; kind < 2; kind++) {
        if (nmax[kind] == 0) continue;
        CCD_HIP_CHECK(ctx, hipMemcpyAsync(ndev, &nmax[kind], 8, hipMemcpyHostToDevice, st));
        // The grouping as the pipeline runs it (signature pass, sort, gather), timed; the queue shapes
        // on the grouped order; then the best shape on the stored order.  Grouping stays on only when
        // it pays for itself.
        const int32_t *grouped = keys[kind];
        float t_group = 0.f;
        int rc = timed(
            [&]() { return group_by_signature(ctx, kind == 0, qstart, qend, keys[kind], nmax[kind], eta, nullptr, &grouped); },
            t_group);
        if (rc) return rc;
        float best_ms = 1e30f;
        for (int b : blocks)
            for (int t : tiles)
                for (int w : waves) {
                    QueueShape s;
                    s.block = b;
                    s.tile = t;
                    s.waves_per_cu = w;
                    float ms = 0.f;
                    rc = timed(queue(kind, grouped, s), ms);
                    if (rc) return rc;
                    if (ms < best_ms) {
                        best_ms = ms;
                        best[kind] = s;
                    }
                }
        float t_plain = 0.f;
        rc = timed(queue(kind, keys[kind], best[kind]), t_plain);
        if (rc) return rc;
        best[kind].group = (grouped != keys[kind] && t_group + best_ms <= t_plain) ? 1 : 0;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    ctx->vf_shape = best[0];
    ctx->ee_shape = best[1];
    graph_drop(ctx);
    if (shape) {
        for (int k = 0; k < 2; k++) {
            shape[3 * k] = best[k].block;
            shape[3 * k + 1] = best[k].tile;
            shape[3 * k + 2] = best[k].waves_per_cu;
            shape[6 + k] = best[k].group;
        }
    }
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
