// Work accounting of the narrowphase (ccd_narrowphase_work): the reference's full cascade per stencil
// with a per-lane stage counter, reduced per wavefront.  Built on its own with the cascade's functions
// left to the compiler's inlining choices (the product kernels inline all of them): this kernel's
// generated code is then the one that was checked against the oracle's counts on the device.
#define CCD_OUTLINE_CASCADE 1
#include "ccd_core.h"

#include "ccd_internal.h"

namespace {

using ccd::V3;

__device__ __forceinline__ V3 load3(const double *__restrict__ p, int v)
{
    const double *q = p + 3 * (int64_t)v;
    return ccd::mk(q[0], q[1], q[2]);
}

// Its root isolation is the plain form (no Bernstein shortcuts, plain Newton): with the product's form
// this kernel stalled once on the device, on the converging-spheres candidates, for no cause found in
// the code, its assembly or sanitized host runs of the same path.  Which stages run depends only on
// the interval structure, which both forms give alike.
struct LaneCount {
    static constexpr bool plain_roots = true;
    uint32_t n[ccd::S_COUNT];
    __device__ __forceinline__ void operator()(int s) { n[s]++; }
};

// Two wavefronts per SIMD, as the product kernels.
template <bool VF>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_narrow2_count(
    const double *__restrict__ qs, const double *__restrict__ qe, const int4 *__restrict__ keys, int64_t n, double eta,
    unsigned long long *counts)
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    LaneCount c;
#pragma unroll
    for (int k = 0; k < ccd::S_COUNT; k++) c.n[k] = 0;
    if (i < n) {
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
        double t = 0.0;
        if (VF)
            ccd::check_vf_window(s, e, eta, t, c);
        else
            ccd::check_ee_window(s, e, eta, t, c);
    }
#pragma unroll
    for (int k = 0; k < ccd::S_COUNT; k++) {
        uint32_t v = c.n[k];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        if (__lane_id() == 0 && v) atomicAdd(&counts[k], (unsigned long long)v);
    }
}

inline unsigned blocks_for(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

}   // namespace

namespace ccdi {

int launch_narrow2_count(ccd_ctx *ctx, bool vf, const double *qs, const double *qe, const int32_t *keys,
                         int64_t n, double eta, unsigned long long *counts)
{
    if (n == 0) return CCD_OK;
    if (vf)
        hipLaunchKernelGGL(k_narrow2_count<true>, dim3(blocks_for(n, 256)), dim3(256), 0, ctx->stream, qs, qe,
                           (const int4 *)keys, n, eta, counts);
    else
        hipLaunchKernelGGL(k_narrow2_count<false>, dim3(blocks_for(n, 256)), dim3(256), 0, ctx->stream, qs, qe,
                           (const int4 *)keys, n, eta, counts);
    CCD_LAUNCH_CHECK(ctx);
    return CCD_OK;
}

}   // namespace ccdi
