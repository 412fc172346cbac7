// Device helpers shared by the kernels: the 16 x 16 x 4 MFMA wrappers and
// their accumulator layout, and wave-wide reductions on DPP lane
// permutations (no LDS round trip).
#pragma once

#include <hip/hip_runtime.h>

#include <climits>

namespace lmx {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef double d2 __attribute__((ext_vector_type(2)));
template <class T>
struct V4;
template <>
struct V4<double> {
    typedef d4 type;
};
template <>
struct V4<float> {
    typedef f4 type;
};

// c += a * b over one 16 x 16 x 4 block: lane l supplies A[l % 16][l / 16]
// and B[l / 16][l % 16].
__device__ inline d4 mfma(double a, double b, d4 c) { return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0); }
__device__ inline f4 mfma(float a, float b, f4 c) { return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0); }

// Row, inside a 16 x 16 MFMA block, of accumulator value r of lane group
// lk = lane / 16 (its column is lane % 16): lk + 4r for f64, 4lk + r for
// f32 (measured on gfx950).
template <class T>
__device__ inline int acc_row(int lk, int r) {
    return sizeof(T) == 8 ? lk + 4 * r : 4 * lk + r;
}

template <int CTRL, int RM>
__device__ inline int dpp_i(int v) {
    return __builtin_amdgcn_update_dpp(v, v, CTRL, RM, 0xF, false);
}
template <int CTRL, int RM>
__device__ inline float dpp_f(float v) {
    return __int_as_float(dpp_i<CTRL, RM>(__float_as_int(v)));
}
template <int CTRL, int RM>
__device__ inline double dpp_f(double v) {
    return __hiloint2double(dpp_i<CTRL, RM>(__double2hiint(v)), dpp_i<CTRL, RM>(__double2loint(v)));
}
// Lane k's value to every lane (k a compile-time or wave-uniform lane).
__device__ inline float bcast_lane(float v, int k) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), k));
}
__device__ inline double bcast_lane(double v, int k) {
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), k),
                            __builtin_amdgcn_readlane(__double2loint(v), k));
}
__device__ inline float lane63(float v) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63)); }
__device__ inline double lane63(double v) {
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), 63),
                            __builtin_amdgcn_readlane(__double2loint(v), 63));
}

// Wave-wide max / min / sum: quad swaps, row rotations by 4 and 8, then the
// row broadcasts of lanes 15 and 31; the full result lands in lane 63 and
// is returned to every lane.  fmax ignores NaN.
template <class T>
__device__ inline T wave_max(T v) {
    v = fmax(v, dpp_f<0xB1, 0xF>(v));   // quad_perm 1,0,3,2
    v = fmax(v, dpp_f<0x4E, 0xF>(v));   // quad_perm 2,3,0,1
    v = fmax(v, dpp_f<0x124, 0xF>(v));  // row_ror 4
    v = fmax(v, dpp_f<0x128, 0xF>(v));  // row_ror 8
    v = fmax(v, dpp_f<0x142, 0xA>(v));  // row_bcast 15
    v = fmax(v, dpp_f<0x143, 0xC>(v));  // row_bcast 31
    return lane63(v);
}
__device__ inline int wave_min(int v) {
    v = min(v, dpp_i<0xB1, 0xF>(v));
    v = min(v, dpp_i<0x4E, 0xF>(v));
    v = min(v, dpp_i<0x124, 0xF>(v));
    v = min(v, dpp_i<0x128, 0xF>(v));
    v = min(v, dpp_i<0x142, 0xA>(v));
    v = min(v, dpp_i<0x143, 0xC>(v));
    return __builtin_amdgcn_readlane(v, 63);
}
__device__ inline int wave_max_int(int v) {
    v = max(v, dpp_i<0xB1, 0xF>(v));
    v = max(v, dpp_i<0x4E, 0xF>(v));
    v = max(v, dpp_i<0x124, 0xF>(v));
    v = max(v, dpp_i<0x128, 0xF>(v));
    v = max(v, dpp_i<0x142, 0xA>(v));
    v = max(v, dpp_i<0x143, 0xC>(v));
    return __builtin_amdgcn_readlane(v, 63);
}
template <class T>
__device__ inline T wave_sum(T v) {
    v += dpp_f<0xB1, 0xF>(v);
    v += dpp_f<0x4E, 0xF>(v);
    v += dpp_f<0x124, 0xF>(v);
    v += dpp_f<0x128, 0xF>(v);
    v += dpp_f<0x142, 0xA>(v);
    v += dpp_f<0x143, 0xC>(v);
    return lane63(v);
}
// Sum over each aligned group of 16 lanes; every lane of the group gets it.
template <class T>
__device__ inline T row16_sum(T v) {
    v += dpp_f<0xB1, 0xF>(v);
    v += dpp_f<0x4E, 0xF>(v);
    v += dpp_f<0x124, 0xF>(v);
    v += dpp_f<0x128, 0xF>(v);
    return v;
}
template <class T>
__device__ inline T row16_max(T v) {
    v = fmax(v, dpp_f<0xB1, 0xF>(v));
    v = fmax(v, dpp_f<0x4E, 0xF>(v));
    v = fmax(v, dpp_f<0x124, 0xF>(v));
    v = fmax(v, dpp_f<0x128, 0xF>(v));
    return v;
}

// Raises this wave's instruction-issue priority (s_setprio 3; waves start
// at 0) for kernels on the factorisation's critical path: when they share a
// CU with the lookahead GEMM, the SIMD arbiter issues their instructions
// first.
__device__ inline void critical_path_priority() { __builtin_amdgcn_s_setprio(3); }

// Lane c's value of x, to every lane (v_readlane into a scalar register).
__device__ __forceinline__ float lane_value(float x, int c) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), c));
}
__device__ __forceinline__ double lane_value(double x, int c) {
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(x), c),
                            __builtin_amdgcn_readlane(__double2loint(x), c));
}

}  // namespace lmx
