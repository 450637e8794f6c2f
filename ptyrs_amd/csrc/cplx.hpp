// Complex arithmetic for the device kernels, in the precision of the engine (float or double).
#pragma once

#include <hip/hip_runtime.h>

namespace ptyrs {

template <typename T>
struct cplx {
  T x, y;
};

template <typename T> struct vec2;
template <> struct vec2<float> { using type = float2; };
template <> struct vec2<double> { using type = double2; };

template <typename T>
__host__ __device__ __forceinline__ cplx<T> mk(T x, T y) { return cplx<T>{x, y}; }
template <typename T>
__host__ __device__ __forceinline__ cplx<T> operator+(cplx<T> a, cplx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__host__ __device__ __forceinline__ cplx<T> operator-(cplx<T> a, cplx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__host__ __device__ __forceinline__ cplx<T> operator*(cplx<T> a, cplx<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T>
__host__ __device__ __forceinline__ cplx<T> operator*(cplx<T> a, T s) { return {a.x * s, a.y * s}; }
template <typename T>
__host__ __device__ __forceinline__ cplx<T> conj(cplx<T> a) { return {a.x, -a.y}; }
template <typename T>
__host__ __device__ __forceinline__ T norm2(cplx<T> a) { return a.x * a.x + a.y * a.y; }
// conj(a) * b
template <typename T>
__host__ __device__ __forceinline__ cplx<T> cmulc(cplx<T> a, cplx<T> b) {
  return {a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x};
}
// a * (-i)
template <typename T>
__host__ __device__ __forceinline__ cplx<T> mul_mi(cplx<T> a) { return {a.y, -a.x}; }

// 8- or 16-byte vector load/store of one complex value.
// 1/x: the hardware reciprocal for float (v_rcp_f32, within 1 ulp), an IEEE division for
// double.
template <typename T>
__device__ __forceinline__ T rcp(T x) {
  if constexpr (sizeof(T) == 4) return __builtin_amdgcn_rcpf(x);
  else return T(1) / x;
}

template <typename T>
__device__ __forceinline__ cplx<T> ld(const cplx<T>* p) {
  auto v = *reinterpret_cast<const typename vec2<T>::type*>(p);
  return {v.x, v.y};
}
template <typename T>
__device__ __forceinline__ void st(cplx<T>* p, cplx<T> a) {
  typename vec2<T>::type v;
  v.x = a.x;
  v.y = a.y;
  *reinterpret_cast<typename vec2<T>::type*>(p) = v;
}

// Buffer-resource access: a uniform base in SGPRs (the resource), a per-lane 32-bit byte
// offset and a uniform byte offset, so that a kernel's many loads of one plane share one
// address VGPR. Accesses past `bytes` read 0 and write nothing. The third resource dword is
// the raw-buffer format word of gfx9-family GPUs (CDNA).
struct Buf {
  __amdgpu_buffer_rsrc_t r;
};
__device__ __forceinline__ Buf buf(const void* base, uint32_t bytes) {
  return Buf{__builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, int(bytes),
                                               0x00020000)};
}
template <typename T>
__device__ __forceinline__ cplx<T> bld(Buf b, uint32_t voff, uint32_t soff) {
  if constexpr (sizeof(T) == 4) {
    return __builtin_bit_cast(cplx<T>, __builtin_amdgcn_raw_buffer_load_b64(b.r, voff, soff, 0));
  } else {
    return __builtin_bit_cast(cplx<T>, __builtin_amdgcn_raw_buffer_load_b128(b.r, voff, soff, 0));
  }
}
__device__ __forceinline__ float bld1(Buf b, uint32_t voff, uint32_t soff) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(b.r, voff, soff, 0));
}
template <typename T>
__device__ __forceinline__ void bst(Buf b, uint32_t voff, uint32_t soff, cplx<T> a) {
  if constexpr (sizeof(T) == 4) {
    using V = decltype(__builtin_amdgcn_raw_buffer_load_b64(b.r, 0u, 0u, 0));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(V, a), b.r, voff, soff, 0);
  } else {
    using V = decltype(__builtin_amdgcn_raw_buffer_load_b128(b.r, 0u, 0u, 0));
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(V, a), b.r, voff, soff, 0);
  }
}

// Kahan-compensated accumulator: the running sum and the lost low-order part.
template <typename T>
struct kahan {
  T s = 0, c = 0;
  __device__ __forceinline__ void add(T v) {
    T y = v - c;
    T t = s + y;
    c = (t - s) - y;
    s = t;
  }
};

}  // namespace ptyrs
