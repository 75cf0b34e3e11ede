// lane_check.hip -- test kernel for the wave64 lane-exchange primitives of
// include/csmv_device.hpp (DPP scans and reductions, ds_bpermute gathers).  Each wave
// takes 64 words and writes its reductions and per-lane results; the test compares them
// with numpy.
#include <hip/hip_runtime.h>

#include "csmv_device.hpp"

using namespace csmv;

namespace {

// out_red per wave: min, max, sum, or64 low word, or64 high word
constexpr uint32_t kRed = 5;

__global__ void __launch_bounds__(256) lane_kernel(const uint32_t *in, const uint64_t *in64,
                                                   const uint32_t *src, uint32_t n_waves,
                                                   uint32_t *out_scan, uint32_t *out_gather,
                                                   uint32_t *out_red)
{
    const uint32_t wave = blockIdx.x * (blockDim.x / kWave) + threadIdx.x / kWave;
    if (wave >= n_waves) return;   // uniform per wave
    const uint32_t lane = lane_id();
    const uint64_t i = (uint64_t)wave * kWave + lane;
    const uint32_t v = in[i];
    out_scan[i] = wave_incl_scan(v);
    out_gather[i] = lane_gather(v, src[i] & 63u);
    const uint32_t mn = wave_min(v), mx = wave_max(v), sm = wave_sum(v);
    const uint64_t o = wave_or64(in64[i]);
    if (lane == 0) {
        uint32_t *r = out_red + (uint64_t)wave * kRed;
        r[0] = mn;
        r[1] = mx;
        r[2] = sm;
        r[3] = (uint32_t)o;
        r[4] = (uint32_t)(o >> 32);
    }
}

template <typename T> hipError_t dev(T **p, const T *h, size_t n)
{
    hipError_t e = hipMalloc((void **)p, n * sizeof(T));
    if (e == hipSuccess && h) e = hipMemcpy(*p, h, n * sizeof(T), hipMemcpyHostToDevice);
    return e;
}

}  // namespace

// Returns 0, or the HIP error code.
extern "C" int lanechk_run(const uint32_t *in, const uint64_t *in64, const uint32_t *src,
                           uint32_t n_waves, uint32_t *out_scan, uint32_t *out_gather,
                           uint32_t *out_red)
{
    if (!in || !in64 || !src || !out_scan || !out_gather || !out_red || n_waves == 0 ||
        n_waves > (1u << 20))
        return (int)hipErrorInvalidValue;
    const size_t n = (size_t)n_waves * kWave;
    uint32_t *d_in = nullptr, *d_src = nullptr, *d_scan = nullptr, *d_gather = nullptr,
             *d_red = nullptr;
    uint64_t *d_in64 = nullptr;
    hipError_t e = dev(&d_in, in, n);
    if (e == hipSuccess) e = dev(&d_in64, in64, n);
    if (e == hipSuccess) e = dev(&d_src, src, n);
    if (e == hipSuccess) e = dev<uint32_t>(&d_scan, nullptr, n);
    if (e == hipSuccess) e = dev<uint32_t>(&d_gather, nullptr, n);
    if (e == hipSuccess) e = dev<uint32_t>(&d_red, nullptr, (size_t)n_waves * kRed);
    if (e == hipSuccess) {
        const uint32_t blocks = (n_waves + 3) / 4;
        hipLaunchKernelGGL(lane_kernel, dim3(blocks), dim3(256), 0, 0, d_in, d_in64, d_src, n_waves,
                           d_scan, d_gather, d_red);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e == hipSuccess) e = hipMemcpy(out_scan, d_scan, n * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(out_gather, d_gather, n * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(out_red, d_red, (size_t)n_waves * kRed * 4, hipMemcpyDeviceToHost);
    void *bufs[6] = {d_in, d_in64, d_src, d_scan, d_gather, d_red};
    for (void *b : bufs) {
        const hipError_t f = hipFree(b);
        if (e == hipSuccess) e = f;
    }
    return (int)e;
}
