// lbm_storage.h -- population storage formats.  load() returns f in the
// compute type, store() takes f; the 16-bit formats hold the shifted value
// d = f - w*rho_shift ("DDF shifting", Lehmann et al. 2022), which keeps the
// stored magnitudes small and centred on zero.  FP32 and FP64 store f
// itself (shift 0).
#pragma once

#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

namespace lbm {

// d(v) / of_d(d) convert between the stored value and the shifted value d
// itself (no shift applied), for kernels that work in the shifted frame;
// put_nt() stores with the non-temporal hint (streamed past the caches).
struct StF64 {
    using T = double;
    static __device__ __forceinline__ double load(T v, double) { return v; }
    static __device__ __forceinline__ T store(double f, double) { return f; }
    static __device__ __forceinline__ double d(T v) { return v; }
    static __device__ __forceinline__ T of_d(double d) { return d; }
    static __device__ __forceinline__ void put_nt(T* p, T v) { __builtin_nontemporal_store(v, p); }
};
struct StF32 {
    using T = float;
    static __device__ __forceinline__ float load(T v, float) { return v; }
    static __device__ __forceinline__ T store(float f, float) { return f; }
    static __device__ __forceinline__ float d(T v) { return v; }
    static __device__ __forceinline__ T of_d(float d) { return d; }
    static __device__ __forceinline__ void put_nt(T* p, T v) { __builtin_nontemporal_store(v, p); }
};
struct StF16 {
    using T = __half;
    static __device__ __forceinline__ float load(T v, float sh) { return __half2float(v) + sh; }
    static __device__ __forceinline__ T store(float f, float sh) { return __float2half_rn(f - sh); }
    static __device__ __forceinline__ float d(T v) { return __half2float(v); }
    static __device__ __forceinline__ T of_d(float d) { return __float2half_rn(d); }
    static __device__ __forceinline__ void put_nt(T* p, T v)
    {
        __builtin_nontemporal_store(__half_as_ushort(v), reinterpret_cast<unsigned short*>(p));
    }
};
struct StBF16 {
    using T = __hip_bfloat16;
    static __device__ __forceinline__ float load(T v, float sh) { return __bfloat162float(v) + sh; }
    static __device__ __forceinline__ T store(float f, float sh) { return __float2bfloat16(f - sh); }
    static __device__ __forceinline__ float d(T v) { return __bfloat162float(v); }
    static __device__ __forceinline__ T of_d(float d) { return __float2bfloat16(d); }
    static __device__ __forceinline__ void put_nt(T* p, T v)
    {
        __builtin_nontemporal_store(__builtin_bit_cast(unsigned short, v),
                                    reinterpret_cast<unsigned short*>(p));
    }
};

}  // namespace lbm
