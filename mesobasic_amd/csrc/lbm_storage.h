// lbm_storage.h -- population storage formats.  load() returns f in the
// compute type, store() takes f; the 16-bit formats hold the shifted value
// d = f - w*rho_shift ("DDF shifting", Lehmann et al. 2022), which keeps the
// stored magnitudes small and centred on zero.
#pragma once

#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

namespace lbm {

struct StF64 {
    using T = double;
    static __device__ __forceinline__ double load(T v, double) { return v; }
    static __device__ __forceinline__ T store(double f, double) { return f; }
};
struct StF32 {
    using T = float;
    static __device__ __forceinline__ float load(T v, float) { return v; }
    static __device__ __forceinline__ T store(float f, float) { return f; }
};
struct StF16 {
    using T = __half;
    static __device__ __forceinline__ float load(T v, float sh) { return __half2float(v) + sh; }
    static __device__ __forceinline__ T store(float f, float sh) { return __float2half_rn(f - sh); }
};
struct StBF16 {
    using T = __hip_bfloat16;
    static __device__ __forceinline__ float load(T v, float sh) { return __bfloat162float(v) + sh; }
    static __device__ __forceinline__ T store(float f, float sh) { return __float2bfloat16(f - sh); }
};

}  // namespace lbm
