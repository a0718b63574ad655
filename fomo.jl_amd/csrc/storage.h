// storage.h -- loads and stores of stored arrays (Float32, or Float16 / BFloat16 with a per-array
// scale: a value v is stored as RN(v * enc) and read back as float(stored) * dec); arithmetic is
// always Float32.
#pragma once
#include <hip/hip_bf16.h>
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include <type_traits>

#include "fomo_common.h"

namespace fomo {

// Element c of a global array of E: with a 32-bit index the byte offset is formed in 32 bits,
// so the access compiles to a scalar base plus a 32-bit vector offset.  NT: streaming hint
// (measured slower for every array of the sweeps, which re-read their inputs through L2).
template <class E, bool NT = false, class I>
__device__ __forceinline__ E gget(const void* p, I c) {
  gp<const E> a;
  if constexpr (sizeof(I) == 4)
    a = (gp<const E>)(G((const char*)p) + (unsigned)(c * (unsigned)sizeof(E)));
  else
    a = G((const E*)p) + c;
  if constexpr (NT)
    return __builtin_nontemporal_load(a);
  else
    return *a;
}
template <class E, bool NT = false, class I>
__device__ __forceinline__ void gput(void* p, I c, E v) {
  gp<E> a;
  if constexpr (sizeof(I) == 4)
    a = (gp<E>)(G((char*)p) + (unsigned)(c * (unsigned)sizeof(E)));
  else
    a = G((E*)p) + c;
  if constexpr (NT)
    __builtin_nontemporal_store(v, a);
  else
    *a = v;
}

template <class T>
struct Cvt;
// The value a store of v followed by a load returns (v itself for Float32).
template <class T>
__device__ __forceinline__ float round_trip(float v, float enc, float dec);
template <>
struct Cvt<float> {
  template <bool NT = false, class I>
  __device__ __forceinline__ static float ld(const void* p, I c, float) {
    return gget<float, NT>(p, c);
  }
  template <bool NT = false, class I>
  __device__ __forceinline__ static void st(void* p, I c, float v, float) {
    gput<float, NT>(p, c, v);
  }
};
template <>
struct Cvt<__half> {
  template <bool NT = false, class I>
  __device__ __forceinline__ static float ld(const void* p, I c, float dec) {
    return __half2float(__ushort_as_half(gget<unsigned short, NT>(p, c))) * dec;
  }
  template <bool NT = false, class I>
  __device__ __forceinline__ static void st(void* p, I c, float v, float enc) {
    gput<unsigned short, NT>(p, c, __half_as_ushort(__float2half_rn(v * enc)));
  }
};
template <>
struct Cvt<__hip_bfloat16> {
  template <bool NT = false, class I>
  __device__ __forceinline__ static float ld(const void* p, I c, float dec) {
    return __bfloat162float(__ushort_as_bfloat16(gget<unsigned short, NT>(p, c))) * dec;
  }
  template <bool NT = false, class I>
  __device__ __forceinline__ static void st(void* p, I c, float v, float enc) {
    gput<unsigned short, NT>(p, c, __bfloat16_as_ushort(__float2bfloat16(v * enc)));
  }
};

// Store of a wavefield value.  Float32: as is.  16-bit: as Cvt<T>::st, except that a value the
// format cannot hold (the rounded result is an infinity or a NaN: overflow of v * enc) is stored
// as the largest finite value of its sign and *flag is set to 1, which the host reports as a
// storage-overflow error.  Runs without overflow store exactly the bits of Cvt<T>::st.
template <class T, class I>
__device__ __forceinline__ void st_field(void* p, I c, float v, float enc, unsigned* flag) {
  if constexpr (std::is_same<T, float>::value) {
    (void)flag;
    gput<float>(p, c, v);
  } else {
    constexpr bool kHalf = std::is_same<T, __half>::value;
    unsigned short b = kHalf ? __half_as_ushort(__float2half_rn(v * enc)) : __bfloat16_as_ushort(__float2bfloat16(v * enc));
    constexpr unsigned short kExp = kHalf ? 0x7C00 : 0x7F80, kMaxFinite = kHalf ? 0x7BFF : 0x7F7F;
    if ((b & kExp) == kExp) {
      b = (unsigned short)((b & 0x8000) | kMaxFinite);
      if (flag) G(flag)[0] = 1u;  // every writer stores the same value
    }
    gput<unsigned short>(p, c, b);
  }
}

template <>
__device__ __forceinline__ float round_trip<float>(float v, float, float) {
  return v;
}
template <>
__device__ __forceinline__ float round_trip<__half>(float v, float enc, float dec) {
  return __half2float(__float2half_rn(v * enc)) * dec;
}
template <>
__device__ __forceinline__ float round_trip<__hip_bfloat16>(float v, float enc, float dec) {
  return __bfloat162float(__float2bfloat16(v * enc)) * dec;
}

// Base of plane k of a stored array (uniform across the block, so loads address it as a scalar
// base plus a 32-bit in-plane offset).
template <class T>
__device__ __forceinline__ const void* plane_of(const void* p, int k, size_t plane_elems) {
  return (const void*)((const T*)p + (size_t)k * plane_elems);
}
template <class T>
__device__ __forceinline__ void* plane_of(void* p, int k, size_t plane_elems) {
  return (void*)((T*)p + (size_t)k * plane_elems);
}

}  // namespace fomo
