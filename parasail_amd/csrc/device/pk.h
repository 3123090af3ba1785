// pk.h -- packed int16 helpers shared by the wavefront kernels: two DP
// cells per 32-bit VGPR (half A in bits 0-15, half B in bits 16-31), with
// the saturating v_pk_add_i16 / v_pk_sub_i16 and v_pk_max_i16 of gfx950.
#pragma once

#include <hip/hip_runtime.h>

typedef short s2v __attribute__((ext_vector_type(2)));

namespace pk {

__device__ __forceinline__ int add(int a, int b)
{
    return __builtin_bit_cast(int, __builtin_elementwise_add_sat(__builtin_bit_cast(s2v, a),
                                                                __builtin_bit_cast(s2v, b)));
}
__device__ __forceinline__ int sub(int a, int b)
{
    return __builtin_bit_cast(int, __builtin_elementwise_sub_sat(__builtin_bit_cast(s2v, a),
                                                                __builtin_bit_cast(s2v, b)));
}
__device__ __forceinline__ int max(int a, int b)
{
    return __builtin_bit_cast(int, __builtin_elementwise_max(__builtin_bit_cast(s2v, a),
                                                            __builtin_bit_cast(s2v, b)));
}
__device__ __forceinline__ int min(int a, int b)
{
    return __builtin_bit_cast(int, __builtin_elementwise_min(__builtin_bit_cast(s2v, a),
                                                            __builtin_bit_cast(s2v, b)));
}
// unsigned saturating subtraction of both halves (clamps at 0)
__device__ __forceinline__ int subu(int a, int b)
{
    typedef unsigned short u2v __attribute__((ext_vector_type(2)));
    return __builtin_bit_cast(int, __builtin_elementwise_sub_sat(__builtin_bit_cast(u2v, a),
                                                                __builtin_bit_cast(u2v, b)));
}
__device__ __forceinline__ int lo(int v) { return (int)(short)(v & 0xffff); }
__device__ __forceinline__ int hi(int v) { return v >> 16; }
__host__ __device__ __forceinline__ int pack(int l, int h) { return (l & 0xffff) | (h << 16); }
// maximum of three for non-negative halves below 0x7C00: read as binary16
// such a half is +0, a subnormal or a finite positive number, and those
// order as their bit patterns do, so v_pk_maximum3_f16 gives the integer
// maximum of each half (a half at or above 0x7C00 -- +inf or NaN -- yields
// a half at or above 0x7C00)
typedef _Float16 h2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ int max3u15(int a, int b, int c)
{
    const h2v x = __builtin_bit_cast(h2v, a), y = __builtin_bit_cast(h2v, b),
              z = __builtin_bit_cast(h2v, c);
    return __builtin_bit_cast(int, __builtin_elementwise_maximum(__builtin_elementwise_maximum(x, y), z));
}
// value of lane-1 within the wave (lane 0 gets an unspecified value)
__device__ __forceinline__ int shr1(int v) { return __builtin_amdgcn_mov_dpp(v, 0x138, 0xF, 0xF, false); }

}  // namespace pk

#define PK_NEG ((int)0x80008000)
#define PK_POS ((int)0x7FFF7FFF)
// substitution value of pad rows / columns: low enough that a pad cell
// never competes with a real one, high enough that adding it to any int16
// DP value saturates instead of wrapping
#define W_NEG (-12000)
