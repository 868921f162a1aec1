// Device self-checks of the arithmetic the kernels rely on.
#include <vector>
#include "ngx_kernels.h"
#include "ngx_backend.h"

namespace ngx {
namespace {

// The three-instruction reciprocal against IEEE division over every float bit pattern: each thread
// takes 256 consecutive patterns; each workgroup writes its mismatch counts (in range, anywhere finite)
// and one mismatching pattern.
__global__ __launch_bounds__(256) void reciprocal_check_kernel(uint32_t *out) {
    __shared__ uint32_t in_range[256], anywhere[256], example[256];
    const uint32_t tid = blockIdx.x * 256u + threadIdx.x;
    uint32_t a = 0, b = 0, ex = 0;
    for (uint32_t i = 0; i < 256; ++i) {
        const uint32_t bits = (tid << 8) | i;
        const float x = __uint_as_float(bits);
        if (!isfinite(x) || x == 0.0f) continue;
        const float want = 1.0f / x;
        const float r = __builtin_amdgcn_rcpf(x);
        const float e = __builtin_fmaf(-x, r, 1.0f);
        const float got = __builtin_fmaf(e, r, r);
        if (__float_as_uint(got) != __float_as_uint(want)) {
            ++b;
            if (fabsf(x) >= 0x1p-126f && fabsf(x) < 0x1p126f) {
                ++a;
                ex = bits;
            }
        }
    }
    in_range[threadIdx.x] = a;
    anywhere[threadIdx.x] = b;
    example[threadIdx.x] = ex;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            in_range[threadIdx.x] += in_range[threadIdx.x + s];
            anywhere[threadIdx.x] += anywhere[threadIdx.x + s];
            if (example[threadIdx.x] == 0) example[threadIdx.x] = example[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[3 * blockIdx.x] = in_range[0];
        out[3 * blockIdx.x + 1] = anywhere[0];
        out[3 * blockIdx.x + 2] = example[0];
    }
}

// The walks' wave-wide inclusive prefix sum on given values, 64 a wave, four waves a workgroup.
__global__ __launch_bounds__(256) void wave_scan_kernel(const int *in, int *out, const int waves) {
    const int wave = blockIdx.x * 4 + (int)(threadIdx.x >> 6);
    if (wave >= waves) return;  // whole waves leave together, so a scanning wave has every lane active
    const size_t i = (size_t)wave * 64 + (threadIdx.x & 63);
    out[i] = wave_prefix_sum(in[i]);
}

}  // namespace
}  // namespace ngx

extern "C" ngx_status ngx_debug_check_reciprocal(int32_t device_id, uint64_t *mismatches_in_range,
                                                 uint64_t *mismatches_anywhere, uint32_t *example) {
    constexpr unsigned blocks = 65536;  // 65536 x 256 threads x 256 patterns = 2^32
    if (!mismatches_in_range || !mismatches_anywhere || !example) return NGX_INVALID_ARGUMENT;
    if (hipSetDevice(device_id) != hipSuccess) return NGX_DEVICE_ERROR;
    uint32_t *d = nullptr;
    if (hipMalloc(&d, sizeof(uint32_t) * 3 * blocks) != hipSuccess) return NGX_OUT_OF_MEMORY;
    ngx::reciprocal_check_kernel<<<blocks, 256>>>(d);
    std::vector<uint32_t> h(3 * blocks);
    const bool ok = hipGetLastError() == hipSuccess &&
                    hipMemcpy(h.data(), d, sizeof(uint32_t) * h.size(), hipMemcpyDeviceToHost) == hipSuccess;
    hipFree(d);
    if (!ok) return NGX_DEVICE_ERROR;
    uint64_t a = 0, b = 0;
    uint32_t ex = 0;
    for (unsigned k = 0; k < blocks; ++k) {
        a += h[3 * k];
        b += h[3 * k + 1];
        if (ex == 0) ex = h[3 * k + 2];
    }
    *mismatches_in_range = a;
    *mismatches_anywhere = b;
    *example = ex;
    return NGX_OK;
}

extern "C" ngx_status ngx_debug_ray_order(int32_t device_id, const float *directions, int32_t rays_per_world, int32_t worlds,
                                          int32_t *order) {
    if (rays_per_world < 0 || worlds < 0 || rays_per_world > ngx::RAY_ORDER_MAX_RAYS ||
        (long long)rays_per_world * worlds > (1ll << 26))
        return NGX_INVALID_ARGUMENT;
    const size_t n = (size_t)rays_per_world * worlds;
    if (n == 0) return NGX_OK;
    if (!directions || !order) return NGX_INVALID_ARGUMENT;
    if (hipSetDevice(device_id) != hipSuccess) return NGX_DEVICE_ERROR;
    float *d = nullptr;
    uint32_t *keys = nullptr;
    int *vals = nullptr, *out = nullptr;
    bool ok = hipMalloc(&d, 3 * n * sizeof(float)) == hipSuccess && hipMalloc(&keys, 2 * n * sizeof(uint32_t)) == hipSuccess &&
              hipMalloc(&vals, n * sizeof(int)) == hipSuccess;
    if (!ok) {
        hipFree(d);
        hipFree(keys);
        hipFree(vals);
        return NGX_OUT_OF_MEMORY;
    }
    ok = hipMemcpy(d, directions, 3 * n * sizeof(float), hipMemcpyHostToDevice) == hipSuccess &&
         ngx::launch_ray_order(d, rays_per_world, worlds, keys, vals, &out, nullptr) == hipSuccess &&
         hipMemcpy(order, out, n * sizeof(int), hipMemcpyDeviceToHost) == hipSuccess;
    hipFree(d);
    hipFree(keys);
    hipFree(vals);
    return ok ? NGX_OK : NGX_DEVICE_ERROR;
}

extern "C" ngx_status ngx_debug_wave_prefix_sum(int32_t device_id, const int32_t *values, int32_t waves, int32_t *sums) {
    if (waves < 0 || (waves > 0 && (!values || !sums)) || waves > (1 << 20)) return NGX_INVALID_ARGUMENT;
    if (waves == 0) return NGX_OK;
    if (hipSetDevice(device_id) != hipSuccess) return NGX_DEVICE_ERROR;
    const size_t n = (size_t)waves * 64;
    int *d = nullptr;
    if (hipMalloc(&d, 2 * n * sizeof(int)) != hipSuccess) return NGX_OUT_OF_MEMORY;
    bool ok = hipMemcpy(d, values, n * sizeof(int), hipMemcpyHostToDevice) == hipSuccess;
    if (ok) {
        ngx::wave_scan_kernel<<<(unsigned)((waves + 3) / 4), 256>>>(d, d + n, waves);
        ok = hipGetLastError() == hipSuccess && hipMemcpy(sums, d + n, n * sizeof(int), hipMemcpyDeviceToHost) == hipSuccess;
    }
    hipFree(d);
    return ok ? NGX_OK : NGX_DEVICE_ERROR;
}
