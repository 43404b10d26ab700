// Device-side helpers shared by the gfx950 kernels: packed u16 arithmetic,
// DPP row exchanges and reductions over a 16-lane DPP row, and the 9x7 census.
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


// ------------------------------------------------------------------------------------------
// The 9x7 census (sparse_9x7: dy in -3..3, dx in -4..4, both non-zero, 48 bits), shared by
// the full search's census and cost kernel and the band search's.

// The image rows in LDS carry CENSUS_REACH replicated pixels on each side, so the census
// reads its 9 columns without clamping each one.
constexpr int CENSUS_REACH = 4;

// An image row in LDS: CENSUS_REACH replicated pixels, the row, the right edge replicated
// on to a whole dword past the last census group's reads (bytes x0 .. x0 + 11 of the group
// at x0 = 4 ((W + 3) / 4 - 1)).
__host__ __device__ constexpr int census_row_pitch(int W) { return (W + 2 * CENSUS_REACH + 7) & ~3; }

// The census of four neighbouring pixels x0 .. x0 + 3 of one image row, from the dwords of
// the 7 padded rows (row r = dy + 3, `pitch_dwords` apart) that hold columns x0 - 4 .. x0 + 7.
// Neighbours are compared two at a time, in the halves of a packed u16: with the
// neighbours a, b as (a, b) and the centre c as (0x7FFE + c) in both halves, bit 15 of a
// half of the difference is a + 1 < c. Comparison k (0..23: row rank 0..5 over dy = -3..3
// but 0, then the column pairs (-4, -3), (-2, -1), (1, 2), (3, 4)) leaves its two bits at
// bits k and 16 + k of the low word (k < 16), or k - 16 and k of the high word: an order of
// the 48 bits of its own, the same for both images, which the Hamming costs do not see.
// census_bits_in_order puts them in the window's row-major order.
struct CensusQuad {
    uint32_t lo[4], hi[4];
};

// Two pixels at a time, row by row, so that few registers are live (the rows' dwords are
// read once for each two).
__device__ __forceinline__ CensusQuad census_quad(const uint32_t *img, int pitch_dwords) {
    CensusQuad q;
    const uint32_t centres = img[3 * pitch_dwords + 1];
#pragma unroll
    for (int k0 = 0; k0 < 4; k0 += 2) {
        uint32_t Y[4];
#pragma unroll
        for (int k = k0; k < k0 + 2; k++) {
            const uint32_t half = ((centres >> (8 * k)) & 0xFFu) + 0x7FFEu;
            Y[k] = half | (half << 16);
            q.lo[k] = q.hi[k] = 0;
        }
#pragma unroll
        for (int rank = 0; rank < 6; rank++) {
            const int r = rank < 3 ? rank : rank + 1;
            const uint32_t w0 = img[r * pitch_dwords], w1 = img[r * pitch_dwords + 1],
                           w2 = img[r * pitch_dwords + 2];
#pragma unroll
            for (int k = k0; k < k0 + 2; k++) {
#pragma unroll
                for (int pair = 0; pair < 4; pair++) {
                    // Bytes first and first + 1 of the 12 bytes w0 w1 w2.
                    const int first = k + (pair < 2 ? 2 * pair : 2 * pair + 1);
                    const uint32_t X =
                        first + 1 < 8
                            ? __builtin_amdgcn_perm(w1, w0, (uint32_t)first | 0x0C00u |
                                                                (uint32_t)(first + 1) << 16 |
                                                                0x0C000000u)
                            : __builtin_amdgcn_perm(w2, w1, (uint32_t)(first - 4) | 0x0C00u |
                                                                (uint32_t)(first - 3) << 16 |
                                                                0x0C000000u);
                    const uint32_t z = pk_sub(Y[k], X);
                    const int c = rank * 4 + pair;
                    if (c < 16) q.lo[k] |= (z >> (15 - c)) & (0x00010001u << c);
                    else q.hi[k] |= (z >> (31 - c)) & (0x00010001u << (c - 16));
                }
            }
        }
    }
    return q;
}

// Bits 0..15 of x to the even bits 0..30.
__device__ __forceinline__ uint32_t spread_even(uint32_t x) {
    x &= 0xFFFFu;
    x = (x | x << 8) & 0x00FF00FFu;
    x = (x | x << 4) & 0x0F0F0F0Fu;
    x = (x | x << 2) & 0x33333333u;
    return (x | x << 1) & 0x55555555u;
}

// A census in census_quad's order to the window's row-major order (bit 2k and 2k + 1 are
// comparison k's two).
__device__ __forceinline__ uint64_t census_bits_in_order(uint32_t lo, uint32_t hi) {
    const uint32_t low = spread_even(lo) | spread_even(lo >> 16) << 1;
    const uint32_t high = spread_even(hi & 0xFFu) | spread_even((hi >> 16) & 0xFFu) << 1;
    return (uint64_t)low | (uint64_t)high << 32;
}

}  // namespace nxs
