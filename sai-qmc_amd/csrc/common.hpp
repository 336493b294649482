// Shared device helpers for the gfx950 DQMC engine.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define DQMC_WAVE 64

typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// 16x16x4 MFMA tile: lane l supplies A[l%16][l/16] and B[l/16][l%16]; its
// four accumulator values v = 0..3 hold column l%16 of D at the rows given by
// Mfma<T>::row (measured on gfx950: FP32 keeps rows 4*(l/16)+v, FP64 keeps
// rows l/16 + 4v).
template <typename T> struct Mfma;
template <> struct Mfma<double> {
    typedef f64x4 acc_t;
    static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int v) { return (lane >> 4) + 4 * v; }
};
template <> struct Mfma<float> {
    typedef f32x4 acc_t;
    static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};

// Wave-wide reductions on DPP lane moves (quad_perm swaps, half-row and row
// mirrors within the 16-lane rows, then across rows); 64-bit values move as
// two 32-bit halves.  With every row written (RMASK 0xF) and every source
// lane valid, the old value of the destination is never kept, so no zero
// needs to be moved in first; partial row masks keep 0 in the rows they
// skip.
template <int CTRL, int RMASK> __device__ __forceinline__ int dpp_move_i(int x)
{
    if constexpr (RMASK == 0xF) return __builtin_amdgcn_mov_dpp(x, CTRL, RMASK, 0xF, false);
    else return __builtin_amdgcn_update_dpp(0, x, CTRL, RMASK, 0xF, false);
}
template <int CTRL, int RMASK> __device__ __forceinline__ double dpp_move(double x)
{
    const int lo = dpp_move_i<CTRL, RMASK>(__double2loint(x));
    const int hi = dpp_move_i<CTRL, RMASK>(__double2hiint(x));
    return __hiloint2double(hi, lo);
}
template <int CTRL, int RMASK> __device__ __forceinline__ float dpp_move(float x)
{
    return __int_as_float(dpp_move_i<CTRL, RMASK>(__float_as_int(x)));
}
// lane l of every 16-lane row to the whole row (gfx950 DPP row_newbcast);
// l folds to a constant in unrolled code
__device__ __forceinline__ double row_bcast(double v, int l)
{
    switch (l & 15) {
    case 0: return dpp_move<0x150, 0xF>(v);
    case 1: return dpp_move<0x151, 0xF>(v);
    case 2: return dpp_move<0x152, 0xF>(v);
    case 3: return dpp_move<0x153, 0xF>(v);
    case 4: return dpp_move<0x154, 0xF>(v);
    case 5: return dpp_move<0x155, 0xF>(v);
    case 6: return dpp_move<0x156, 0xF>(v);
    case 7: return dpp_move<0x157, 0xF>(v);
    case 8: return dpp_move<0x158, 0xF>(v);
    case 9: return dpp_move<0x159, 0xF>(v);
    case 10: return dpp_move<0x15A, 0xF>(v);
    case 11: return dpp_move<0x15B, 0xF>(v);
    case 12: return dpp_move<0x15C, 0xF>(v);
    case 13: return dpp_move<0x15D, 0xF>(v);
    case 14: return dpp_move<0x15E, 0xF>(v);
    default: return dpp_move<0x15F, 0xF>(v);
    }
}

// the value of 16-lane row h0 (lane l + 16 h0) to lane l of every row: the
// 32-lane swap brings the half holding h0 to both halves, the 16-lane swap
// the row of h0 within it to both rows; h0 folds to a constant in unrolled
// code
__device__ __forceinline__ double row_to_all(double v, int h0)
{
    const int lo = __double2loint(v), hi = __double2hiint(v);
    const auto l32 = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
    const auto h32 = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
    const int lo1 = (h0 & 2) ? l32[1] : l32[0], hi1 = (h0 & 2) ? h32[1] : h32[0];
    const auto l16 = __builtin_amdgcn_permlane16_swap(lo1, lo1, false, false);
    const auto h16 = __builtin_amdgcn_permlane16_swap(hi1, hi1, false, false);
    return __hiloint2double((h0 & 1) ? h16[1] : h16[0], (h0 & 1) ? l16[1] : l16[0]);
}

__device__ __forceinline__ double lane63(double v)
{
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), 63);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), 63);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ float lane63(float v) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63)); }

// Exchange with the lane 16 (32) apart: the gfx950 row swaps return the
// pair (own, partner) in an order that differs between the two rows, and
// their sum is the same either way.
template <int BITS> __device__ __forceinline__ double swap_add(double v)
{
    const int lo = __double2loint(v), hi = __double2hiint(v);
    const auto l = BITS == 16 ? __builtin_amdgcn_permlane16_swap(lo, lo, false, false)
                              : __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
    const auto h = BITS == 16 ? __builtin_amdgcn_permlane16_swap(hi, hi, false, false)
                              : __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
    return __hiloint2double(h[0], l[0]) + __hiloint2double(h[1], l[1]);
}
template <int BITS> __device__ __forceinline__ float swap_add(float v)
{
    const int x = __float_as_int(v);
    const auto r = BITS == 16 ? __builtin_amdgcn_permlane16_swap(x, x, false, false)
                              : __builtin_amdgcn_permlane32_swap(x, x, false, false);
    return __int_as_float(r[0]) + __int_as_float(r[1]);
}
__device__ __forceinline__ double first_lane(double v)
{
    const int lo = __builtin_amdgcn_readfirstlane(__double2loint(v));
    const int hi = __builtin_amdgcn_readfirstlane(__double2hiint(v));
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ float first_lane(float v) { return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v))); }

// Sum over the 64 lanes: butterflies within rows of 16 (DPP), then across
// rows with the gfx950 row swaps.  Every lane ends with the same bits
// ((R0 + R1) + (R2 + R3) for row sums R); the result is returned as a
// wave-uniform value.
template <typename T> __device__ __forceinline__ T wave_sum(T v)
{
    v += dpp_move<0xB1, 0xF>(v);
    v += dpp_move<0x4E, 0xF>(v);
    v += dpp_move<0x141, 0xF>(v);
    v += dpp_move<0x140, 0xF>(v);
    v = swap_add<16>(v);
    v = swap_add<32>(v);
    return first_lane(v);
}

// Sums over the 64 lanes of four values at once, returned wave-uniform.
// The row swaps go first and halve the data at each step: the 32-lane swap
// of (d0, d2) and of (d1, d3) leaves d0's and d1's partial sums in lanes
// 0-31 and d2's and d3's in lanes 32-63, the 16-lane swap of the two
// leaves dj's in row j; then the butterflies within the rows of 16 run on
// one value instead of four (21 VALU operations against 72 for four
// wave_sum).  The order of the additions is not wave_sum's.
__device__ __forceinline__ void wave_sum4(double (&d)[4])
{
    auto swap32 = [](double u, double v) {
        const auto l = __builtin_amdgcn_permlane32_swap(__double2loint(u), __double2loint(v), false, false);
        const auto h = __builtin_amdgcn_permlane32_swap(__double2hiint(u), __double2hiint(v), false, false);
        return __hiloint2double(h[0], l[0]) + __hiloint2double(h[1], l[1]);
    };
    const double s02 = swap32(d[0], d[2]), s13 = swap32(d[1], d[3]);
    const auto l = __builtin_amdgcn_permlane16_swap(__double2loint(s02), __double2loint(s13), false, false);
    const auto h = __builtin_amdgcn_permlane16_swap(__double2hiint(s02), __double2hiint(s13), false, false);
    double t = __hiloint2double(h[0], l[0]) + __hiloint2double(h[1], l[1]);
    t += dpp_move<0xB1, 0xF>(t);
    t += dpp_move<0x4E, 0xF>(t);
    t += dpp_move<0x141, 0xF>(t);
    t += dpp_move<0x140, 0xF>(t);
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int lo = __builtin_amdgcn_readlane(__double2loint(t), 16 * j);
        const int hi = __builtin_amdgcn_readlane(__double2hiint(t), 16 * j);
        d[j] = __hiloint2double(hi, lo);
    }
}

__device__ __forceinline__ double wave_max(double v)
{
    v = fmax(v, dpp_move<0xB1, 0xF>(v));
    v = fmax(v, dpp_move<0x4E, 0xF>(v));
    v = fmax(v, dpp_move<0x141, 0xF>(v));
    v = fmax(v, dpp_move<0x140, 0xF>(v));
    v = fmax(v, dpp_move<0x142, 0xA>(v));
    v = fmax(v, dpp_move<0x143, 0xC>(v));
    return lane63(v);
}

// Block-wide sum; `scratch` holds one slot per wave.  Every thread gets the
// result.  Contains two barriers.
template <typename T>
__device__ __forceinline__ T block_sum(T v, T *scratch)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nw = (blockDim.x + 63) >> 6;
    v = wave_sum(v);
    if (lane == 0) scratch[wave] = v;
    __syncthreads();
    T s = 0;
    for (int w = 0; w < nw; w++) s += scratch[w];
    __syncthreads();
    return s;
}

// xoshiro256** step returning the top 53 bits as a double in [0, 1):
// the generator and conversion of src/rng.c:23-37 (Blackman & Vigna).
__device__ __forceinline__ double xoshiro_double(uint64_t s[4])
{
    const uint64_t x = s[1] * 5;
    const uint64_t out = ((x << 7) | (x >> 57)) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = (s[3] << 45) | (s[3] >> 19);
    return (double)(out >> 11) * (1.0 / 9007199254740992.0);
}

// Optional per-replica vector used as a diagonal scale: value 1 when p is
// null, 1/p[i] when inv is set.
template <typename T> struct Vec {
    const T *p;
    long long stride;
    int inv;
    __device__ __forceinline__ T at(int b, int i) const
    {
        if (!p) return T(1);
        const T v = p[(long long)b * stride + i];
        return inv ? T(1) / v : v;
    }
    // the stored value (1 without a vector) and its use as a factor: at()
    // split in two so the load can be issued long before the value is used
    __device__ __forceinline__ T raw(int b, int i) const { return p ? p[(long long)b * stride + i] : T(1); }
    __device__ __forceinline__ T fix(T v) const { return (p && inv) ? T(1) / v : v; }
};

template <typename T> static inline Vec<T> vnone() { return Vec<T>{nullptr, 0, 0}; }
template <typename T> static inline Vec<T> vec(const T *p, long long stride, int inv = 0)
{
    return Vec<T>{p, stride, inv};
}

// Map a linear workgroup id so that consecutive logical ids stay on one XCD
// (the dispatcher deals workgroups round-robin over the 8 XCDs).  Logical
// ids are grouped per batch entry by the caller, so one replica's tiles
// share an L2.
__device__ __forceinline__ int xcd_remap(int lin, int total)
{
    if ((total & 7) != 0) return lin;
    return (lin & 7) * (total >> 3) + (lin >> 3);
}
