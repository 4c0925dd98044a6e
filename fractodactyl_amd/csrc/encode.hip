// Store::put on the device (crates/fd-kernel/src/store.rs:66-153): the per-sample outcomes
// of a frame pass quantised into the fd_samples columns where they lie in HBM, so a frame
// reads back 15 bytes per sample instead of the 53 of the raw outcomes.
//
// The formulas and their operation order are the host encoder's; log2 and atan2 are the
// device library's, whose last bits may differ from the host libm's (as the reference's own
// columns differ between platforms), so nu, de and normal agree with the host encoding to
// an ulp or a quantum, not bit for bit.  The class column is exact.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "device_math.hpp"
#include "kernels.hpp"

namespace fdg {
namespace {

constexpr uint8_t kUnset = 0xff;  // outside a sample mask

__device__ inline uint16_t quantised_angle(double x, double y) {
    double t = round(atan2(y, x) * (65536.0 / 6.283185307179586));
    if (isnan(t)) return 0;
    int64_t v = t >= 9.2233720368547758e18 ? INT64_MAX
                : t <= -9.2233720368547758e18 ? INT64_MIN : (int64_t)t;
    return (uint16_t)(uint64_t)v;
}

__global__ void __launch_bounds__(256) encode_columns(EncodeDev e) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < e.count;
         i += gridDim.x * blockDim.x) {
        uint8_t kind = e.kind[i];
        double nu = 0.0;
        float de = 0.0f;
        uint16_t normal = 0;
        uint8_t cls = 0;  // unset samples read class 0 with zero columns
        if (kind == 0) {
            double zr = e.zr[i], zi = e.zi[i];
            double z2 = zr * zr + zi * zi;
            double log2z = 0.5 * log2(z2);
            nu = (double)e.n[i] + 1.0 - log2(log2z);
            if (e.de) {
                double dr = e.dr[i], di = e.di[i];
                double az = sqrt(z2);
                double lnz = log2z * 0.693147180559945309417232121458176568;
                double base = 2.0 * az * lnz / (dm::hypot_ref(dr, di) * e.px_m);
                int64_t k = -(e.dexp[i] + e.px_e);
                k = k < -4000 ? -4000 : (k > 4000 ? 4000 : k);
                de = (float)(base * dm::exp2i(k / 2) * dm::exp2i(k - k / 2));
            }
            if (e.normal) {
                double dr = e.dr[i], di = e.di[i];
                double a = zr * dr + zi * di, b = zi * dr - zr * di;
                normal = quantised_angle(e.cos_r * a + e.sin_r * b, e.sin_r * a - e.cos_r * b);
            }
        } else if (kind != kUnset) {
            cls = kind == 2 ? 2 : 1;
        }
        e.cls[i] = cls;
        if (e.nu) e.nu[i] = nu;
        if (e.de) e.de[i] = de;
        if (e.normal) e.normal[i] = normal;
    }
}

// The outcomes of the samples list[0 .. m) packed contiguously (a masked pass's computed
// samples): the readback then moves 53 bytes per computed sample instead of per sample.
__global__ void __launch_bounds__(256) gather_outcomes(GatherDev g) {
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < g.m;
         t += gridDim.x * blockDim.x) {
        uint32_t i = g.list[t];
        g.kind_c[t] = g.kind[i];
        g.n_c[t] = g.n[i];
        g.zr_c[t] = g.zr[i];
        g.zi_c[t] = g.zi[i];
        g.dr_c[t] = g.dr[i];
        g.di_c[t] = g.di[i];
        g.dexp_c[t] = g.dexp[i];
        if (g.ops_c) g.ops_c[t] = g.ops[i];
    }
}

}  // namespace

hipError_t launch_gather_outcomes(const GatherDev& g, int cus, hipStream_t stream) {
    if (g.m == 0) return hipSuccess;
    uint64_t blocks = (g.m + 255) / 256;
    uint64_t cap = (uint64_t)cus * 16;
    hipLaunchKernelGGL(gather_outcomes, dim3((unsigned)(blocks < cap ? blocks : cap)), dim3(256),
                       0, stream, g);
    return hipGetLastError();
}

hipError_t launch_encode_columns(const EncodeDev& e, int cus, hipStream_t stream) {
    if (e.count == 0) return hipSuccess;
    uint64_t blocks = (e.count + 255) / 256;
    uint64_t cap = (uint64_t)cus * 16;
    hipLaunchKernelGGL(encode_columns, dim3((unsigned)(blocks < cap ? blocks : cap)), dim3(256),
                       0, stream, e);
    return hipGetLastError();
}

}  // namespace fdg
