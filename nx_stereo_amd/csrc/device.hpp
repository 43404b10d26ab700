// Device-side helpers shared by the gfx950 kernels: packed u16 arithmetic,
// DPP row exchanges and reductions over a 16-lane DPP row.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nxs {

typedef uint16_t u16x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ u16x2 as_v(uint32_t x) { return __builtin_bit_cast(u16x2, x); }
__device__ __forceinline__ uint32_t as_u(u16x2 x) { return __builtin_bit_cast(uint32_t, x); }

// v_pk_min_u16 / v_pk_add_u16 / v_pk_sub_u16 on two u16 lanes held in one VGPR.
__device__ __forceinline__ uint32_t pk_min(uint32_t a, uint32_t b) {
    return as_u(__builtin_elementwise_min(as_v(a), as_v(b)));
}
__device__ __forceinline__ uint32_t pk_add(uint32_t a, uint32_t b) { return as_u(as_v(a) + as_v(b)); }
// min(lo, hi) in both halves (v_pk_min_u16 with its second operand's halves swapped).
__device__ __forceinline__ uint32_t pk_min_halves(uint32_t a) {
    const u16x2 v = as_v(a);
    return as_u(__builtin_elementwise_min(v, __builtin_shufflevector(v, v, 1, 0)));
}
__device__ __forceinline__ uint32_t pk_sub(uint32_t a, uint32_t b) { return as_u(as_v(a) - as_v(b)); }
__device__ __forceinline__ uint32_t pk_mul(uint32_t a, uint32_t b) { return as_u(as_v(a) * as_v(b)); }
__device__ __forceinline__ uint32_t pk_max(uint32_t a, uint32_t b) {
    return as_u(__builtin_elementwise_max(as_v(a), as_v(b)));
}

// DPP controls (GFX9): row_shl:1, row_shr:1, row_ror:n. A DPP row is 16 lanes.
constexpr int DPP_ROW_SHL1 = 0x101;
constexpr int DPP_ROW_SHR1 = 0x111;
constexpr int dpp_row_ror(int n) { return 0x120 + n; }

template <int CTRL>
__device__ __forceinline__ uint32_t dpp(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xF, 0xF, false);
}

// The same move, where a lane whose source falls outside its row keeps `edge`.
template <int CTRL>
__device__ __forceinline__ uint32_t dpp_or(uint32_t edge, uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)edge, (int)v, CTRL, 0xF, 0xF, false);
}

// A wave-uniform 64-bit value as scalar registers: each half through readfirstlane. The
// builtin returns int, so each half goes back through uint32_t before it is widened; a
// widened int would be sign-extended and corrupt the high half when bit 31 is set.
__device__ __forceinline__ uint64_t uniform_u64(uint64_t v) {
    const uint64_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
    const uint64_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(v >> 32));
    return hi << 32 | lo;
}

// Raw buffer access (MUBUF) with 32-bit byte offsets: an offset at or past `bytes` loads
// zeros and drops stores, which masks an access without a branch.
typedef __amdgpu_buffer_rsrc_t Rsrc;
constexpr uint32_t MASKED = 0x80000000u;  // past every buffer made here (all under 2^31 bytes)

// `base` must be the same in every lane of the wave: it is read from the first active lane,
// so the descriptor lives in scalar registers.
__device__ __forceinline__ Rsrc buffer_of(const void *base, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(
        reinterpret_cast<void *>(uniform_u64(reinterpret_cast<uint64_t>(base))), (short)0,
        (int)__builtin_amdgcn_readfirstlane((int)bytes), 0x00020000);
}

// Minimum over the 16 lanes of a DPP row, left in every lane of the row.
__device__ __forceinline__ uint32_t row_min_u32(uint32_t v) {
    v = min(v, dpp<dpp_row_ror(8)>(v));
    v = min(v, dpp<dpp_row_ror(4)>(v));
    v = min(v, dpp<dpp_row_ror(2)>(v));
    v = min(v, dpp<dpp_row_ror(1)>(v));
    return v;
}

// Wave-wide DPP shifts by one lane (GFX9): lane i reads lane i-1 (shr) or i+1 (shl).
constexpr int DPP_WAVE_SHL1 = 0x130;
constexpr int DPP_WAVE_SHR1 = 0x138;

// v_permlane16_swap with the same value in both operands leaves rows (0, 0, 2, 2) in one
// result and rows (1, 1, 3, 3) in the other; v_permlane32_swap likewise the wave's halves.
// On row-uniform values, the minimum of the two results is the minimum over rows 2r and
// 2r+1 (over the halves), in every lane.
__device__ __forceinline__ uint32_t pair_rows_min(uint32_t v) {
    const auto s = __builtin_amdgcn_permlane16_swap(v, v, false, false);
    return min((uint32_t)s[0], (uint32_t)s[1]);
}
__device__ __forceinline__ uint32_t halves_min(uint32_t v) {
    const auto s = __builtin_amdgcn_permlane32_swap(v, v, false, false);
    return min((uint32_t)s[0], (uint32_t)s[1]);
}

__device__ __forceinline__ uint32_t pair_rows_max(uint32_t v) {
    const auto s = __builtin_amdgcn_permlane16_swap(v, v, false, false);
    return max((uint32_t)s[0], (uint32_t)s[1]);
}
__device__ __forceinline__ uint32_t halves_max(uint32_t v) {
    const auto s = __builtin_amdgcn_permlane32_swap(v, v, false, false);
    return max((uint32_t)s[0], (uint32_t)s[1]);
}

// Lane-wise u16 minimum of two halves over the 16 lanes of a DPP row.
__device__ __forceinline__ uint32_t row_min_pk(uint32_t v) {
    v = pk_min(v, dpp<dpp_row_ror(8)>(v));
    v = pk_min(v, dpp<dpp_row_ror(4)>(v));
    v = pk_min(v, dpp<dpp_row_ror(2)>(v));
    v = pk_min(v, dpp<dpp_row_ror(1)>(v));
    return v;
}

// The lines of the aggregation: orientation o is 0 rows, 1 columns, 2 x-y diagonals, 3 x+y
// diagonals. Line i's first pixel in the forward direction and its length (0 past the
// last line).
struct LineGeometry {
    int y0, x0, length;
};

// Line i of orientation o, its first pixel in the forward direction and its length.
__device__ __forceinline__ LineGeometry line_geometry(int o, int i, int W, int H) {
    LineGeometry g;
    if (o == 0) {  // row i, forward L->R
        g.y0 = i;
        g.x0 = 0;
        g.length = i < H ? W : 0;
    } else if (o == 1) {  // column i, forward down
        g.y0 = 0;
        g.x0 = i;
        g.length = i < W ? H : 0;
    } else if (o == 2) {  // x - y = i - (H-1), forward down-right
        const int c = i - (H - 1);
        g.y0 = max(0, -c);
        g.x0 = g.y0 + c;
        g.length = i < W + H - 1 ? min(H - g.y0, W - g.x0) : 0;
    } else {  // x + y = i, forward down-left
        g.y0 = max(0, i - (W - 1));
        g.x0 = i - g.y0;
        g.length = i < W + H - 1 ? min(H - g.y0, g.x0 + 1) : 0;
    }
    return g;
}

}  // namespace nxs
