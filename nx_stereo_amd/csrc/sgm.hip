// Semi-global matching on gfx950: census, Hamming costs, 8-path min-plus
// aggregation, winner selection, left-right check and the 3x3 median.
//
// Semantics follow NxStereo.sgm_n (lib/nx_stereo.ex:475-683) as its
// step-by-step transcription states it (native/stereo_match/src/reference.rs).
//
// Layout in HBM: images u8 {H,W}; census u64 {H,W}; volumes pixel-major,
// volume[(y*W + x)*D + d], so the D values of one pixel are contiguous and a
// 16-lane DPP row reads them in one coalesced access.
//
// Aggregation works per line orientation: rows (L->R, R->L), columns (down,
// up), x-y diagonals (down-right, up-left) and x+y diagonals (down-left,
// up-right). Each line is independent and owned by one 16-lane DPP row: lane
// i holds disparities [i*K, i*K+K) as K/2 packed u16 pairs, neighbours d-1
// and d+1 across lanes come from row_shr/row_shl DPP moves, and the minimum
// over D is a 4-step row_ror DPP reduction. Both directions of a line run at
// once from its two ends; each stores its byte path costs until they meet,
// then adds the other's and stores the byte sum of both (each <= p2+48 <= 127,
// so the sum fits a byte) into the orientation's volume. Selection sums the
// orientations.
//
// Every kernel takes one pair by value, or a table of pairs of one size class
// in device memory with the grid's last dimension indexing it (the batched
// form: one launch per kernel for the whole class).
#include "device.hpp"
#include "kernels.hpp"
#include "profile.hpp"

namespace nxs {

// ------------------------------------------------------------------------------------------
// Census (reference.rs:20-38; nx_stereo.ex:491-509): bit k, row-major over
// dy in (-3..3 without 0) x dx in (-4..4 without 0), is set iff the clamped
// neighbour + 1 < the centre.

// The pair a block works on: entry `index` of the batch table, or the one passed by value.
__device__ __forceinline__ PairDesc pair_of(const PairDesc &single, const PairDesc *table,
                                            int index) {
    return table ? table[index] : single;
}

// ------------------------------------------------------------------------------------------
// Census and costs (reference.rs:20-63; census.rs:18-84), one block per image row. The
// seven rows around row y of both images come into LDS (rows clamped at the top and
// bottom), and the block makes row y's 48-bit census of both there: bit k is set when
// neighbour + 1 < centre, row-major over dy in -3..3 and dx in -4..4 (0 skipped), columns
// clamped. Then C[y,x,d] = popcount(FL[y,x] ^ FR[y,x-d]), with FR = 0 left of the image
// (the right census row has D zero features in front).
//
// The costs leave straight from registers. A thread takes the 16 disparities of chunk c of
// four consecutive pixels x0..x0+3, chunks fastest, so the 16-byte stores of a wave's lanes
// with the same pixels cover one contiguous run of the cost volume. The four pixels' 64
// costs read 19 right features, FR[x0-16c-15 .. x0-16c+3], once each. The right census
// row is kept in LDS with one spare feature after every 16, so that neighbouring lanes'
// features, 16c apart, fall into different banks (17c slots apart: 34c banks, distinct
// for c < 32).

typedef uint32_t v2u __attribute__((ext_vector_type(2)));
typedef uint32_t v3u __attribute__((ext_vector_type(3)));
typedef uint32_t v4u __attribute__((ext_vector_type(4)));

constexpr int MAX_D = 256;

// The LDS slot of right feature f (f = D + x - d, 0 <= f < D + W).
__host__ __device__ constexpr int right_slot(int f) { return f + (f >> 4); }

// The census rows in LDS: the left one for every pixel of the cost groups (4 pixels each),
// the right one for features 0 .. D + W + 2 (the last group's reads past the row end: loaded,
// used by nothing), each slot placed as right_slot says.
__host__ __device__ constexpr int left_slots(int W) { return (W + 3) & ~3; }
__host__ __device__ constexpr int right_slots(int W, int D) { return right_slot(D + W + 3) + 1; }

// LDS bytes of one block: the two census rows, then 7 padded rows of each image.
__host__ __device__ constexpr size_t census_cost_lds(int W, int D) {
    return 8 * ((size_t)left_slots(W) + (size_t)right_slots(W, D)) +
           14 * (size_t)census_row_pitch(W);
}

// Row y of pair pd, by one block of COST_BLOCK threads with census_cost_lds(W, D) bytes of
// LDS at `lds`. keep_census: also store both census images.
template <int COST_BLOCK, int D>
__device__ __forceinline__ void census_cost_row(const PairDesc &pd, int W, int H, int y,
                                                int keep_census, uint8_t *lds) {
    uint64_t *cl = reinterpret_cast<uint64_t *>(lds);  // [left_slots(W)]
    uint64_t *cr = cl + left_slots(W);                 // [right_slots(W, D)]
    const int row_pitch = census_row_pitch(W);         // padded row
    uint8_t *rows = reinterpret_cast<uint8_t *>(cr + right_slots(W, D));  // [2][7][row_pitch]
    // A thread loads its byte of all 14 rows before it stores any, so that the loads are in
    // flight together (a thread past the padded row loads a byte it does not store). Buffer
    // loads: the rows' offsets are scalars, the column the one vector offset.
    const Rsrc images[2] = {buffer_of(pd.left, (uint32_t)W * H), buffer_of(pd.right, (uint32_t)W * H)};
    for (int i0 = 0; i0 < row_pitch; i0 += COST_BLOCK) {
        const int i = i0 + (int)threadIdx.x;
        const int column = min(max(i - CENSUS_REACH, 0), W - 1);
        uint32_t v[14];
#pragma unroll
        for (int k = 0; k < 14; k++)
            v[k] = __builtin_amdgcn_raw_buffer_load_b8(images[k / 7], column,
                                                       min(max(y + k % 7 - 3, 0), H - 1) * W, 0);
        if (i < row_pitch) {
#pragma unroll
            for (int k = 0; k < 14; k++) rows[k * row_pitch + i] = (uint8_t)v[k];
        }
    }
    for (int f = threadIdx.x; f < D; f += COST_BLOCK) cr[right_slot(f)] = 0;
    __syncthreads();
    // Four pixels a thread (census_quad); pixels past the row's end land in slots that only
    // pixels past it read.
    const int quads = (W + 3) >> 2;
    for (int i = threadIdx.x; i < 2 * quads; i += COST_BLOCK) {
        const int side = i >= quads ? 1 : 0, x0 = 4 * (i - side * quads);
        // Columns x0 - 4 .. x0 + 7: bytes x0 .. x0 + 11 of each padded row, dword aligned.
        const uint32_t *img = reinterpret_cast<const uint32_t *>(rows + side * 7 * row_pitch + x0);
        const CensusQuad q = census_quad(img, row_pitch >> 2);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const uint64_t f = (uint64_t)q.lo[k] | (uint64_t)q.hi[k] << 32;
            if (side) cr[right_slot(D + x0 + k)] = f;
            else cl[x0 + k] = f;
            if (keep_census && x0 + k < W)
                (side ? pd.b.census_right : pd.b.census_left)[(size_t)y * W + x0 + k] =
                    census_bits_in_order(q.lo[k], q.hi[k]);
        }
    }
    __syncthreads();
    constexpr int chunks = D >> 4;
    const int groups = (W + 3) >> 2;
    const Rsrc out = buffer_of(pd.b.costs + (size_t)y * W * D, (uint32_t)W * D);
    for (int q = threadIdx.x; q < groups * chunks; q += COST_BLOCK) {
        const int g = q / chunks, c = q - g * chunks;
        const int x0 = 4 * g;
        // Pixel x0 + k at disparity 16c + j reads right feature f0 + k - j, f0 = D + x0 - 16c.
        const int f0 = D + x0 - 16 * c;
        // r[i]: feature f0 - 15 + i (past the row's end: never used). f0 is 4 (g mod 4)
        // past a multiple of 16, so the 19 features cross one multiple of 16, at i = t: their
        // slots are right_slot(f0 - 15) + i, one more from i = t on. One base and a select
        // of two addresses per read, each with its constant offset.
        const int t = 15 - (f0 & 15);
        const uint64_t *below = cr + right_slot(f0 - 15), *above = below + 1;
        uint64_t r[19];
#pragma unroll
        for (int i = 0; i < 19; i++) r[i] = (i >= t ? above : below)[i];
        const uint64_t *ls = cl + x0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const uint64_t l = ls[k];
            uint32_t words[4];
#pragma unroll
            for (int w = 0; w < 4; w++) {
                uint32_t word = 0;
#pragma unroll
                for (int j = 0; j < 4; j++)
                    word |= (uint32_t)__builtin_popcountll(l ^ r[15 + k - 4 * w - j]) << (8 * j);
                words[w] = word;
            }
            // A pixel past the row's end stores nothing (its offset is past the buffer).
            const uint32_t at = x0 + k < W ? (uint32_t)(x0 + k) * D + 16 * c : MASKED;
            __builtin_amdgcn_raw_buffer_store_b128((v4u){words[0], words[1], words[2], words[3]},
                                                   out, at, 0, 0);
        }
    }
}

// Grid y: the row; z: the pair. COST_BLOCK threads: 512 where every row of the launch is in
// flight at once (three blocks a CU), else 1024 (launch_all).
template <int COST_BLOCK, int D>
__global__ __launch_bounds__(COST_BLOCK) void census_cost_kernel(PairDesc single,
                                                                 const PairDesc *table, int W,
                                                                 int H, int keep_census) {
    extern __shared__ __align__(16) uint8_t lds[];
    census_cost_row<COST_BLOCK, D>(pair_of(single, table, blockIdx.z), W, H, blockIdx.y,
                                   keep_census, lds);
}

// ------------------------------------------------------------------------------------------
// Path recurrence (reference.rs:65-121; paths.rs:19-42), on a 16-lane row.

template <int K>
struct Bytes;  // K bytes of one lane, loaded and stored as whole dwords (K = 2: one short)
template <>
struct Bytes<2> {
    uint32_t w[1];  // the two bytes in the low half
};
template <>
struct Bytes<4> {
    uint32_t w[1];
    __device__ void load(const uint8_t *p) { w[0] = *reinterpret_cast<const uint32_t *>(p); }
    __device__ void store(uint8_t *p) const { *reinterpret_cast<uint32_t *>(p) = w[0]; }
};
template <>
struct Bytes<8> {
    uint32_t w[2];
    __device__ void load(const uint8_t *p) {
        uint2 v = *reinterpret_cast<const uint2 *>(p);
        w[0] = v.x;
        w[1] = v.y;
    }
    __device__ void store(uint8_t *p) const { *reinterpret_cast<uint2 *>(p) = make_uint2(w[0], w[1]); }
};
template <>
struct Bytes<12> {
    uint32_t w[3];
    __device__ void load(const uint8_t *p) {
        const uint32_t *q = reinterpret_cast<const uint32_t *>(p);
        w[0] = q[0];
        w[1] = q[1];
        w[2] = q[2];
    }
    __device__ void store(uint8_t *p) const {
        uint32_t *q = reinterpret_cast<uint32_t *>(p);
        q[0] = w[0];
        q[1] = w[1];
        q[2] = w[2];
    }
};
template <>
struct Bytes<16> {
    uint32_t w[4];
    __device__ void load(const uint8_t *p) {
        uint4 v = *reinterpret_cast<const uint4 *>(p);
        w[0] = v.x;
        w[1] = v.y;
        w[2] = v.z;
        w[3] = v.w;
    }
    __device__ void store(uint8_t *p) const {
        *reinterpret_cast<uint4 *>(p) = make_uint4(w[0], w[1], w[2], w[3]);
    }
};


// Orientation o's byte volume. A choice among the four pointers as scalar values: a
// run-time index into the descriptor's array (or selects the compiler folds back into one)
// would put the descriptor in scratch memory.
__device__ __forceinline__ uint8_t *orientation_volume(const MatchBuffers &b, int o) {
    uint64_t v = 0;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        v = o == k ? uniform_u64((uint64_t)b.paired[k]) : v;
    }
    return (uint8_t *)v;
}

// K bytes of one lane at a byte offset of a buffer, as whole dwords (MUBUF loads/stores).
template <int K>
__device__ __forceinline__ Bytes<K> vload(Rsrc r, uint32_t at) {
    Bytes<K> b;
    if constexpr (K == 2) {
        b.w[0] = __builtin_amdgcn_raw_buffer_load_b16(r, at, 0, 0);
    } else if constexpr (K == 4) {
        b.w[0] = __builtin_amdgcn_raw_buffer_load_b32(r, at, 0, 0);
    } else if constexpr (K == 8) {
        const v2u v = __builtin_amdgcn_raw_buffer_load_b64(r, at, 0, 0);
        b.w[0] = v[0];
        b.w[1] = v[1];
    } else if constexpr (K == 12) {
        const v3u v = __builtin_amdgcn_raw_buffer_load_b96(r, at, 0, 0);
        b.w[0] = v[0];
        b.w[1] = v[1];
        b.w[2] = v[2];
    } else {
        const v4u v = __builtin_amdgcn_raw_buffer_load_b128(r, at, 0, 0);
        b.w[0] = v[0];
        b.w[1] = v[1];
        b.w[2] = v[2];
        b.w[3] = v[3];
    }
    return b;
}

// The same with a cache policy for the load (2: non-temporal, streaming).
template <int K, int AUX>
__device__ __forceinline__ Bytes<K> vload_policy(Rsrc r, uint32_t at) {
    static_assert(K >= 4, "whole dwords");
    Bytes<K> b;
    if constexpr (K == 4) {
        b.w[0] = __builtin_amdgcn_raw_buffer_load_b32(r, at, 0, AUX);
    } else if constexpr (K == 8) {
        const v2u v = __builtin_amdgcn_raw_buffer_load_b64(r, at, 0, AUX);
        b.w[0] = v[0];
        b.w[1] = v[1];
    } else if constexpr (K == 12) {
        const v3u v = __builtin_amdgcn_raw_buffer_load_b96(r, at, 0, AUX);
        b.w[0] = v[0];
        b.w[1] = v[1];
        b.w[2] = v[2];
    } else {
        const v4u v = __builtin_amdgcn_raw_buffer_load_b128(r, at, 0, AUX);
        b.w[0] = v[0];
        b.w[1] = v[1];
        b.w[2] = v[2];
        b.w[3] = v[3];
    }
    return b;
}

// A store with a cache policy (AUX 2: non-temporal, streaming).
template <int K, int AUX = 0>
__device__ __forceinline__ void vstore(Rsrc r, uint32_t at, const Bytes<K> &b) {
    if constexpr (K == 2) {
        __builtin_amdgcn_raw_buffer_store_b16((uint16_t)b.w[0], r, at, 0, AUX);
    } else if constexpr (K == 4) {
        __builtin_amdgcn_raw_buffer_store_b32(b.w[0], r, at, 0, AUX);
    } else if constexpr (K == 8) {
        __builtin_amdgcn_raw_buffer_store_b64((v2u){b.w[0], b.w[1]}, r, at, 0, AUX);
    } else if constexpr (K == 12) {
        __builtin_amdgcn_raw_buffer_store_b96((v3u){b.w[0], b.w[1], b.w[2]}, r, at, 0, AUX);
    } else {
        __builtin_amdgcn_raw_buffer_store_b128((v4u){b.w[0], b.w[1], b.w[2], b.w[3]}, r, at, 0, AUX);
    }
}

// Bytes b0 b1 b2 b3 of a dword as the u16 pairs (b0, b1) and (b2, b3).
// v_perm_b32 (S0, S1, sel): result byte i is byte sel[i] of {S0:S1}, 0-3 from S1, 4-7 from
// S0, and 0x0C reads zero.
__device__ __forceinline__ uint32_t lo_pair(uint32_t w) { return __builtin_amdgcn_perm(0u, w, 0x0C010C00u); }
__device__ __forceinline__ uint32_t hi_pair(uint32_t w) { return __builtin_amdgcn_perm(0u, w, 0x0C030C02u); }
// Two u16 pairs (each value < 256) as one dword of bytes.
__device__ __forceinline__ uint32_t pack_pairs(uint32_t a, uint32_t b) {
    return __builtin_amdgcn_perm(b, a, 0x06040200u);
}

// A line's D disparities spread over LANES consecutive lanes of a wave (16: a DPP row; 32:
// two rows; 64: the wave). Moves from the neighbouring lane of the same line, where the
// line's first (last) lane takes `edge`, and the minimum over the line's lanes.
template <int LANES>
struct LineLanes;

template <>
struct LineLanes<16> {
    static __device__ __forceinline__ uint32_t from_left(uint32_t edge, uint32_t v, int) {
        return dpp_or<DPP_ROW_SHR1>(edge, v);
    }
    static __device__ __forceinline__ uint32_t from_right(uint32_t edge, uint32_t v, int) {
        return dpp_or<DPP_ROW_SHL1>(edge, v);
    }
    static __device__ __forceinline__ uint32_t min(uint32_t v) { return row_min_u32(v); }
};

// Two rows: wave-wide shifts (lane 32 would read lane 31, which is another line's), and the
// row minima of rows 2r and 2r+1 combined through v_permlane16_swap.
template <>
struct LineLanes<32> {
    static __device__ __forceinline__ uint32_t from_left(uint32_t edge, uint32_t v, int lane) {
        const uint32_t t = dpp_or<DPP_WAVE_SHR1>(edge, v);
        return lane == 0 ? edge : t;
    }
    static __device__ __forceinline__ uint32_t from_right(uint32_t edge, uint32_t v, int lane) {
        const uint32_t t = dpp_or<DPP_WAVE_SHL1>(edge, v);
        return lane == 31 ? edge : t;
    }
    static __device__ __forceinline__ uint32_t min(uint32_t v) { return pair_rows_min(row_min_u32(v)); }
};

template <>
struct LineLanes<64> {
    static __device__ __forceinline__ uint32_t from_left(uint32_t edge, uint32_t v, int) {
        return dpp_or<DPP_WAVE_SHR1>(edge, v);
    }
    static __device__ __forceinline__ uint32_t from_right(uint32_t edge, uint32_t v, int) {
        return dpp_or<DPP_WAVE_SHL1>(edge, v);
    }
    static __device__ __forceinline__ uint32_t min(uint32_t v) {
        return halves_min(pair_rows_min(row_min_u32(v)));
    }
};

template <int K, int LANES>
struct PathState {
    static constexpr int R = K / 2;
    static constexpr uint32_t SENT = 0x3FFF3FFFu;
    uint32_t v[R];  // disparities (2k, 2k+1) of this lane's K in the low and high halves
    uint32_t PM;    // the minimum over the line's D values in both halves, same in all its lanes
    // The neighbouring lanes' values at this lane's two edges. The lanes at a line's two ends
    // are never written by the moves, so they keep the sentinel from reset().
    uint32_t left_in, right_in;
    // The next step's min(prev[d], min(prev[d-1], prev[d+1]) + p1), which needs no minimum.
    uint32_t part[R];

    // The path entering a line: prev = 0 everywhere.
    __device__ void reset(uint32_t p1x2, int lane) {
#pragma unroll
        for (int k = 0; k < R; k++) v[k] = 0;
        PM = 0;
        left_in = right_in = SENT;
        prepare(p1x2, lane);
    }

    // Steps compute L(p,d) = min(a[d], p2, min(a[d-1], a[d+1]) + p1) + C(p,d) with
    // a = prev - min(prev), written unnormalised: min(prev[d], pm + p2,
    // min(prev[d-1], prev[d+1]) + p1) - pm + C. Out-of-range neighbours are a
    // sentinel that only ever loses to the pm + p2 term, as libSGM's p2 - p1
    // padding does. The part without pm is made as soon as prev is known, so that it
    // overlaps the reduction that makes pm.
    __device__ __forceinline__ void prepare(uint32_t p1x2, int lane) {
        // The line's first and last lanes have no neighbour past its ends: the sentinel.
        left_in = LineLanes<LANES>::from_left(left_in, v[R - 1], lane);
        right_in = LineLanes<LANES>::from_right(right_in, v[0], lane);
        uint32_t s[R + 1];
        s[0] = (v[0] << 16) | (left_in >> 16);
#pragma unroll
        for (int k = 1; k < R; k++) s[k] = (v[k] << 16) | (v[k - 1] >> 16);
        s[R] = (right_in << 16) | (v[R - 1] >> 16);
#pragma unroll
        for (int k = 0; k < R; k++) part[k] = pk_min(pk_add(pk_min(s[k], s[k + 1]), p1x2), v[k]);
    }

    // One step: the pm + p2 term, the normalisation and this pixel's costs; then the next
    // step's part and this step's minimum over the line.
    __device__ __forceinline__ void step(const Bytes<K> &cost, uint32_t p1x2, uint32_t p2,
                                         int lane) {
        const uint32_t JUMP = pk_add(PM, p2 * 0x10001u);
        uint32_t c[R];
        if constexpr (K == 2) {
            c[0] = lo_pair(cost.w[0]);
        } else {
#pragma unroll
            for (int k = 0; k < R / 2; k++) {
                c[2 * k] = lo_pair(cost.w[k]);
                c[2 * k + 1] = hi_pair(cost.w[k]);
            }
        }
        uint32_t mn = 0xFFFFFFFFu;
#pragma unroll
        for (int k = 0; k < R; k++) {
            v[k] = pk_add(pk_sub(pk_min(part[k], JUMP), PM), c[k]);
            mn = pk_min(mn, v[k]);
        }
        prepare(p1x2, lane);
        // (m << 16 | m) orders as m does, so the u32 minimum over lanes is the packed one.
        PM = LineLanes<LANES>::min(pk_min_halves(mn));
    }

    __device__ __forceinline__ Bytes<K> bytes() const {
        Bytes<K> b;
        if constexpr (K == 2) {
            b.w[0] = __builtin_amdgcn_perm(0u, v[0], 0x0C0C0200u);
        } else {
#pragma unroll
            for (int k = 0; k < R / 2; k++) b.w[k] = pack_pairs(v[2 * k], v[2 * k + 1]);
        }
        return b;
    }

    // The path whose values bytes() gave: the values, their minimum over the line, the
    // edges and the next step's part are all functions of the values. All-zero bytes give
    // the path entering a line (reset()).
    __device__ __forceinline__ void restore(const Bytes<K> &b, uint32_t p1x2, int lane) {
        if constexpr (K == 2) {
            v[0] = lo_pair(b.w[0]);
        } else {
#pragma unroll
            for (int k = 0; k < R / 2; k++) {
                v[2 * k] = lo_pair(b.w[k]);
                v[2 * k + 1] = hi_pair(b.w[k]);
            }
        }
        uint32_t mn = 0xFFFFFFFFu;
#pragma unroll
        for (int k = 0; k < R; k++) mn = pk_min(mn, v[k]);
        PM = LineLanes<LANES>::min(pk_min_halves(mn));
        left_in = right_in = SENT;
        prepare(p1x2, lane);
    }
};

struct PathArgs {
    PairDesc single;
    const PairDesc *table;
    int W, H;
    int backward_columns;  // 0 with 3 paths: columns go down only
    // The launch covers orientations first..last; block_start[o] is the first block of o.
    int first, last;
    int block_start[4];
    // Blocks longest lines first (8 paths, all four orientations in the launch): see
    // block_lines.
    int longest_first;
};

// The orientation of block b and the first of its `per_block` lines. By default blocks go
// orientation by orientation, lines in order. Longest first: the rows and the columns (the
// longer lines first), then the two diagonal orientations alternately, each from its middle
// group of lines outward, since a diagonal's length falls with its distance from the middle
// ones. The hardware starts blocks in index order as room frees up, so the long lines, which
// bound the launch, start first and the short ones fill in behind them.
struct BlockLines {
    int o, first_line;
};

__device__ __forceinline__ BlockLines block_lines(const PathArgs &a, int b, int per_block) {
    if (!a.longest_first) {
        // The last launched orientation whose first block is at or before b. Constant indices
        // only: `a` may be a kernel argument reached through a reference (the fused batch
        // launch's), where a run-time index would load it into vector registers.
        int o = a.first, start = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const bool here = k >= a.first && k <= a.last && b >= a.block_start[k];
            o = here ? k : o;
            start = here ? a.block_start[k] : start;
        }
        return {o, (b - start) * per_block};
    }
    const int rows = a.block_start[1] - a.block_start[0];
    const int columns = a.block_start[2] - a.block_start[1];
    const int diagonals = a.block_start[3] - a.block_start[2];  // each diagonal orientation
    const bool rows_first = a.W >= a.H;
    const int first_count = rows_first ? rows : columns;
    if (b < first_count) return {rows_first ? 0 : 1, b * per_block};
    b -= first_count;
    const int second_count = rows_first ? columns : rows;
    if (b < second_count) return {rows_first ? 1 : 0, b * per_block};
    b -= second_count;
    const int r = b >> 1, center = (diagonals - 1) / 2, k = (r + 1) >> 1;
    const int group = r == 0 ? center : (r & 1 ? center + k : center - k);
    return {2 + (b & 1), group * per_block};
}

// One wave per block, 64 / LANES lines per wave.
constexpr int PATH_BLOCK = 64;

// Grid y: the pair. PREFETCH: steps of costs (and, past the middle, of the other
// direction's path costs) loaded ahead of use. LANES: lanes per line, K = D / LANES
// disparities per lane.
template <int K, int PREFETCH, int LANES>
__global__ __launch_bounds__(PATH_BLOCK) void path_kernel(PathArgs a) {
    constexpr int D = LANES * K;
    constexpr int LINES_PER_BLOCK = PATH_BLOCK / LANES;
    const int lane = threadIdx.x % LANES;  // within the line
    const int row = threadIdx.x / LANES;   // the block's line
    const int b = blockIdx.x;
    int o = a.last;
    for (int k = a.first; k < a.last; k++)
        if (b < a.block_start[k + 1]) {
            o = k;
            break;
        }
    const int first_line = (b - a.block_start[o]) * LINES_PER_BLOCK;
    const int W = a.W, H = a.H;
    const PairDesc pd = pair_of(a.single, a.table, blockIdx.y);

    const LineGeometry g = line_geometry(o, first_line + row, W, H);
    int longest = 0;
#pragma unroll
    for (int r = 0; r < LINES_PER_BLOCK; r++)
        longest = max(longest, line_geometry(o, first_line + r, W, H).length);
    if (longest == 0) return;  // wave-uniform

    const int dy = o == 0 ? 0 : 1;
    const int dx = o == 0 ? 1 : (o == 1 ? 0 : (o == 2 ? 1 : -1));
    // Byte offsets in the volumes (all under 2^31): one step's pixel to the next, and this
    // lane's bytes of the line's first and last pixels. A line past the last one (length 0)
    // computes on whatever it loads and stores nothing.
    const uint32_t stride = (uint32_t)((dy * W + dx) * D);
    const int L = g.length;
    const uint32_t first = ((uint32_t)g.y0 * W + (uint32_t)g.x0) * D + lane * K;
    const uint32_t last = first + (uint32_t)max(L - 1, 0) * stride;
    const uint32_t volume = (uint32_t)W * H * D;
    const Rsrc costs = buffer_of(pd.b.costs, volume);
    const Rsrc paired = buffer_of(orientation_volume(pd.b, o), volume);
    const uint32_t p1x2 = (uint32_t)pd.p1 * 0x10001u;
    const uint32_t p2 = (uint32_t)pd.p2;
    // Every step issues the same loads and stores, and the loops run whole groups of PREFETCH
    // steps (up to PREFETCH-1 past the longest line, storing nothing): an access that must not
    // happen takes the MASKED offset, and one past a line's end reads bytes it never uses. So
    // the compiler counts outstanding accesses exactly and waits only for the step it needs.

    if (o == 1 && !a.backward_columns) {
        // Columns with 3 paths: down only, its own path costs.
        PathState<K, LANES> s;
        s.reset(p1x2, lane);
        Bytes<K> ring[PREFETCH];
        // The prologue issues what a step issues (a masked store, then the load), so the
        // loop is entered with the same accesses outstanding as it goes round with.
        Bytes<K> none;
#pragma unroll
        for (int k = 0; k < (K + 3) / 4; k++) none.w[k] = 0;
#pragma unroll
        for (int p = 0; p < PREFETCH; p++) {
            __builtin_amdgcn_sched_barrier(0);
            vstore<K>(paired, MASKED, none);
            ring[p] = vload<K>(costs, first + p * stride);
        }
        uint32_t at = first, ahead = first + PREFETCH * stride;
        for (int t0 = 0; t0 < longest; t0 += PREFETCH) {
#pragma unroll
            for (int p = 0; p < PREFETCH; p++) {
                const int t = t0 + p;
                // Keep the steps in program order, so each waits only for its own loads.
                __builtin_amdgcn_sched_barrier(0);
                s.step(ring[p], p1x2, p2, lane);
                vstore<K>(paired, t < L ? at : MASKED, s.bytes());
                ring[p] = vload<K>(costs, ahead);
                at += stride;
                ahead += stride;
            }
        }
        return;
    }

    // Both directions at once, from the two ends of the line (two independent chains,
    // half the sequential steps). At step i the forward direction is at t = i and the
    // backward one at u = L-1-i. Before they meet, each stores its own path costs; past
    // the middle, each adds the other's (stored when that one passed) and stores the byte
    // sum (each <= p2 + 48 <= 127, so no byte carries into the next). An odd line's middle
    // pixel sums both directions in registers.
    //
    // The other's bytes for step j are loaded with its costs, PREFETCH steps ahead, when
    // the other direction stored them by then: 2j - (L-1) >= PREFETCH. Just past the middle
    // (0 < 2j - (L-1) < PREFETCH) they are younger than that, and come instead from the
    // history ring of each direction's last PREFETCH own bytes.
    PathState<K, LANES> fwd, bwd;
    fwd.reset(p1x2, lane);
    bwd.reset(p1x2, lane);
    Bytes<K> ring_f[PREFETCH], ring_b[PREFETCH];    // costs, PREFETCH steps ahead
    Bytes<K> other_f[PREFETCH], other_b[PREFETCH];  // the other direction's bytes, ditto
    Bytes<K> hist_f[PREFETCH], hist_b[PREFETCH];    // own bytes of the last PREFETCH steps
#pragma unroll
    for (int p = 0; p < PREFETCH; p++) {
        // As a step issues: two (masked) stores, then the four loads; the masked loads of
        // the other's bytes read zeros.
#pragma unroll
        for (int k = 0; k < (K + 3) / 4; k++) hist_f[p].w[k] = hist_b[p].w[k] = 0;
        __builtin_amdgcn_sched_barrier(0);
        vstore<K>(paired, MASKED, hist_f[p]);
        vstore<K>(paired, MASKED, hist_b[p]);
        ring_f[p] = vload<K>(costs, first + p * stride);
        ring_b[p] = vload<K>(costs, last - p * stride);
        other_f[p] = vload<K>(paired, MASKED);
        other_b[p] = vload<K>(paired, MASKED);
    }
    uint32_t at_f = first, at_b = last;  // this step's pixels
    uint32_t ahead_f = first + PREFETCH * stride, ahead_b = last - PREFETCH * stride;
    for (int i0 = 0; i0 < longest; i0 += PREFETCH) {
#pragma unroll
        for (int p = 0; p < PREFETCH; p++) {
            const int i = i0 + p;
            // Keep the steps in program order, so each waits only for its own loads.
            __builtin_amdgcn_sched_barrier(0);
            fwd.step(ring_f[p], p1x2, p2, lane);
            bwd.step(ring_b[p], p1x2, p2, lane);
            const Bytes<K> mf = fwd.bytes(), mb = bwd.bytes();
            const int past = 2 * i - (L - 1);  // < 0 before the middle, 0 at an odd line's middle
            Bytes<K> of = other_f[p], ob = other_b[p];
            if (past > 0 && past < PREFETCH) {
                // The other passed here at step i - past, which the history ring still holds.
                const int s = (i - past) & (PREFETCH - 1);
#pragma unroll
                for (int q = 0; q < PREFETCH; q++) {
                    if (q == p) continue;  // step i itself: never the other's
#pragma unroll
                    for (int k = 0; k < (K + 3) / 4; k++) {
                        of.w[k] = s == q ? hist_b[q].w[k] : of.w[k];
                        ob.w[k] = s == q ? hist_f[q].w[k] : ob.w[k];
                    }
                }
            }
            hist_f[p] = mf;
            hist_b[p] = mb;
            Bytes<K> sf, sb;
#pragma unroll
            for (int k = 0; k < (K + 3) / 4; k++) {
                sf.w[k] = mf.w[k] + (past < 0 ? 0u : (past == 0 ? mb.w[k] : of.w[k]));
                sb.w[k] = mb.w[k] + (past <= 0 ? 0u : ob.w[k]);
            }
            vstore<K>(paired, i < L ? at_f : MASKED, sf);
            vstore<K>(paired, i < L && past != 0 ? at_b : MASKED, sb);
            // Step i + PREFETCH: its costs, and the other's bytes when already stored (the
            // stores just made come first in program order, so the loads see them).
            ring_f[p] = vload<K>(costs, ahead_f);
            ring_b[p] = vload<K>(costs, ahead_b);
            const bool stored = 2 * (i + PREFETCH) - (L - 1) >= PREFETCH;
            other_f[p] = vload<K>(paired, stored ? ahead_f : MASKED);
            other_b[p] = vload<K>(paired, stored ? ahead_b : MASKED);
            at_f += stride;
            at_b -= stride;
            ahead_f += stride;
            ahead_b -= stride;
        }
    }
}

// ------------------------------------------------------------------------------------------
// The same aggregation with a line's two directions in two waves of one block: wave 0 walks
// forward, wave 1 backward, over the same four lines (16 lanes each). Each wave issues half
// the instructions per step of the one-wave form, so a line's sequential walk, which bounds
// the kernel, takes about half the time. The two waves meet through memory and LDS:
//
// * A wave that meets the other at s_barrier first waits until its global stores from LAG
//   steps ago have completed (and its LDS writes are done). After a barrier at step b, the
//   other's stores of steps <= b - LAG are visible.
// * Past the middle, a wave needs the other's bytes of its own pixel, which the other made
//   at step i - past (past = 2i - (L-1)). When past >= X they come from memory, loaded
//   PREFETCH steps ahead. Nearer the middle they come from an LDS ring of each wave's last
//   RING steps, which past < RING - 1 keeps from being rewritten before it is read.
// * The ring and a barrier at every step are needed only within X steps (in 2i) of the
//   block's lines' middles, the window. After it the waves meet once every PREFETCH steps.
//   X = 2 PREFETCH + LAG + 2 covers a prefetch issued up to 2 PREFETCH - 1 steps after the
//   last barrier. Before it nothing is read from the other wave, so they do not meet.
// * At an odd line's middle both waves are at the same pixel in the same step: the forward
//   wave stores the sum.
// * A step of the window or after it issues one store and two loads (a load of nothing
//   where there is nothing to read), which the count in the barrier's wait relies on. A
//   step of phase A (below) issues a store and one load; the first LAG barriers after it
//   therefore wait for less, but every read of the other's bytes from memory is prefetched
//   after a barrier with LAG three-access steps before it.
//
// The walk runs in block-uniform phases, each a whole number of PREFETCH steps:
//   A  before the window: own bytes out, costs in; nothing to mask and no barrier;
//   the general step around the window and wherever a line of the block has ended;
//   C  after the window while every line of the block goes on: pair sums out, costs and the
//      other's bytes in, a barrier every PREFETCH steps.
// A CPU model of these rules (tests/test_two_wave_rules.py: line lengths up to 260, blocks
// of unequal lines, each prefetch depth, the wait's count over the accesses actually issued)
// checks that every read is of a completed store or a visible ring slot, and that phases A
// and C need no mask. The eight-path orientations use it (rows, columns, both diagonals),
// and the rows of 3 paths; the down-only columns of 3 paths have nothing to meet, so there
// each wave walks four columns of its own.

constexpr int TWO_BLOCK = 128;

// Columns with 3 paths: down only. The block takes 2 LINES columns, each wave LINES of its
// own, walked alone (no meeting, no barrier), storing its own path costs.
template <int K, int PREFETCH, int LANES>
__device__ __forceinline__ void walk_columns_down(const PathArgs &a, const PairDesc &pd, int b,
                                                  int dir, int lane, int row) {
    constexpr int LINES = 64 / LANES;
    constexpr uint32_t D_BYTES = LANES * K;
    const int W = a.W, H = a.H, o = 1;
    const int first_column = (b - a.block_start[o]) * 2 * LINES + dir * LINES;
    const LineGeometry gc = line_geometry(1, first_column + row, W, H);
    int longest_column = 0;
#pragma unroll
    for (int r = 0; r < LINES; r++)
        longest_column = max(longest_column, line_geometry(1, first_column + r, W, H).length);
    if (longest_column == 0) return;  // wave-uniform
    const int Lc = gc.length;
    const uint32_t stride_c = (uint32_t)W * D_BYTES;
    const uint32_t volume_c = (uint32_t)W * H * D_BYTES;
    const Rsrc costs_c = buffer_of(pd.b.costs, volume_c);
    const Rsrc paired_c = buffer_of(orientation_volume(pd.b, 1), volume_c);
    const uint32_t p1x2c = (uint32_t)pd.p1 * 0x10001u, p2c = (uint32_t)pd.p2;
    const uint32_t first_c = (uint32_t)gc.x0 * D_BYTES + lane * K;
    PathState<K, LANES> sc;
    sc.reset(p1x2c, lane);
    Bytes<K> ring_c[PREFETCH];
    Bytes<K> none_c;
#pragma unroll
    for (int k = 0; k < (K + 3) / 4; k++) none_c.w[k] = 0;
    // As a step issues: a (masked) store, then the load.
#pragma unroll
    for (int p = 0; p < PREFETCH; p++) {
        __builtin_amdgcn_sched_barrier(0);
        vstore<K>(paired_c, MASKED, none_c);
        ring_c[p] = vload<K>(costs_c, first_c + p * stride_c);
    }
    uint32_t at_c = first_c, ahead_c = first_c + PREFETCH * stride_c;
    for (int t0 = 0; t0 < longest_column; t0 += PREFETCH) {
#pragma unroll
        for (int p = 0; p < PREFETCH; p++) {
            const int t = t0 + p;
            __builtin_amdgcn_sched_barrier(0);
            sc.step(ring_c[p], p1x2c, p2c, lane);
            vstore<K>(paired_c, t < Lc ? at_c : MASKED, sc.bytes());
            ring_c[p] = vload<K>(costs_c, ahead_c);
            at_c += stride_c;
            ahead_c += stride_c;
        }
    }
}


// LANES: lanes per line, 16 (four lines a wave) or 32 (two lines a wave, K = D / 32: half
// the instructions per step in each wave, twice the waves). STREAM: the launch's volumes
// overflow the 256 MB Infinity Cache; phase C's pair sums (read next by select) and its
// reads of the other's bytes (their last use) then go non-temporal, which leaves the cache
// to the cost volume and the own bytes still to be read (at 1920x1080x256 path + select
// 2262 -> 2170 us; where the volumes fit, at 728x544x128, it cost 10 us).
template <int K, int PREFETCH, int LANES, bool STREAM>
__global__ __launch_bounds__(TWO_BLOCK) void path2_kernel(PathArgs a) {
    constexpr int NT = STREAM && K >= 4 ? 2 : 0;
    // Steps after which a store has completed (the compiler's own wait for the loads of
    // PREFETCH steps ago already implies it), the window's half-width, and the ring's depth.
    constexpr int LAG = PREFETCH == 4 ? 5 : 6;
    constexpr int X = 2 * PREFETCH + LAG + 2;
    constexpr int RING = X + 1 <= 16 ? 16 : 32;
    static_assert(X <= RING - 1, "the LDS ring must hold the window's late steps");
    static_assert(LANES == 16 || LANES == 32, "a line is one or two DPP rows");
    constexpr int D = LANES * K;
    constexpr int LINES = 64 / LANES;
    // Own bytes per direction, line, step (mod RING) and lane.
    __shared__ Bytes<K> ring[2][LINES][RING][LANES];
    const int dir = threadIdx.x >> 6;  // wave-uniform: 0 forward, 1 backward
    const int lane = threadIdx.x & (LANES - 1);
    const int row = (threadIdx.x & 63) / LANES;
    const int b = blockIdx.x;
    const BlockLines bl = block_lines(a, b, LINES);
    const int o = bl.o;
    const int W = a.W, H = a.H;
    const PairDesc pd = pair_of(a.single, a.table, blockIdx.y);
    if (o == 1 && !a.backward_columns) {
        walk_columns_down<K, PREFETCH, LANES>(a, pd, b, dir, lane, row);
        return;
    }
    const int first_line = bl.first_line;
    const LineGeometry g = line_geometry(o, first_line + row, W, H);
    // Block-uniform: the longest line, and the window around the lines' middles (in 2i).
    int longest = 0, mid_lo = 1 << 30, mid_hi = -(1 << 30);
#pragma unroll
    for (int r = 0; r < LINES; r++) {
        const int len = line_geometry(o, first_line + r, W, H).length;
        longest = max(longest, len);
        mid_lo = min(mid_lo, len - 1);
        mid_hi = max(mid_hi, len - 1);
    }
    if (longest == 0) return;  // both waves leave
    const int window_lo = mid_lo - X, window_hi = mid_hi + X;

    const int dy = o == 0 ? 0 : 1;
    const int dx = o == 0 ? 1 : (o == 1 ? 0 : (o == 2 ? 1 : -1));
    const uint32_t stride = (uint32_t)((dy * W + dx) * D);
    const int L = g.length;
    const uint32_t first = ((uint32_t)g.y0 * W + (uint32_t)g.x0) * D + lane * K;
    const uint32_t last = first + (uint32_t)max(L - 1, 0) * stride;
    const uint32_t volume = (uint32_t)W * H * D;
    const Rsrc costs = buffer_of(pd.b.costs, volume);
    const Rsrc paired = buffer_of(orientation_volume(pd.b, o), volume);
    const uint32_t p1x2 = (uint32_t)pd.p1 * 0x10001u;
    const uint32_t p2 = (uint32_t)pd.p2;
    // This wave's pixel at step 0 and the offset of one step (modulo 2^32 going backward).
    const uint32_t start = dir == 0 ? first : last;
    const uint32_t step = dir == 0 ? stride : 0u - stride;

    PathState<K, LANES> s;
    s.reset(p1x2, lane);
    Bytes<K> cost[PREFETCH], other[PREFETCH];
    Bytes<K> none;
#pragma unroll
    for (int k = 0; k < (K + 3) / 4; k++) none.w[k] = 0;
    // The prologue issues what a step issues: a (masked) store, then two loads.
#pragma unroll
    for (int p = 0; p < PREFETCH; p++) {
        __builtin_amdgcn_sched_barrier(0);
        vstore<K>(paired, MASKED, none);
        __builtin_amdgcn_sched_barrier(0);
        cost[p] = vload<K>(costs, start + p * step);
        other[p] = vload<K>(paired, MASKED);
    }
    uint32_t at = start, ahead = start + PREFETCH * step;
    // The store of step i - LAG is complete once at most its two loads and the three
    // accesses of each later step are outstanding; LDS writes done.
    auto meet = []() __attribute__((always_inline)) {
        asm volatile("s_waitcnt vmcnt(%0) lgkmcnt(0)\n\ts_barrier" ::"n"(3 * LAG - 1) : "memory");
    };
    // Any step: the window's ring and barriers, the middle, lines that have ended.
    auto general = [&](int i0) __attribute__((always_inline)) {
#pragma unroll
        for (int p = 0; p < PREFETCH; p++) {
            const int i = i0 + p;
            __builtin_amdgcn_sched_barrier(0);
            s.step(cost[p], p1x2, p2, lane);
            const Bytes<K> mine = s.bytes();
            const bool window = 2 * i >= window_lo && 2 * i <= window_hi;  // block-uniform
            if (window) ring[dir][row][i & (RING - 1)][lane] = mine;
            if (window || p == 0) meet();
            const int past = 2 * i - (L - 1);  // < 0 before the middle, 0 at an odd line's middle
            Bytes<K> theirs = other[p];
            if (past >= 0 && past < X) theirs = ring[dir ^ 1][row][(i - past) & (RING - 1)][lane];
            Bytes<K> out;
#pragma unroll
            for (int k = 0; k < (K + 3) / 4; k++) out.w[k] = mine.w[k] + (past < 0 ? 0u : theirs.w[k]);
            // At the middle only the forward wave stores (the sum of both). The store goes
            // before this step's loads: the count in meet() relies on that order.
            vstore<K>(paired, i < L && !(past == 0 && dir == 1) ? at : MASKED, out);
            __builtin_amdgcn_sched_barrier(0);
            // Step i + PREFETCH: its costs, and the other's bytes when its store has completed.
            cost[p] = vload<K>(costs, ahead);
            const bool stored = 2 * (i + PREFETCH) - (L - 1) >= X;
            other[p] = vload<K>(paired, stored ? ahead : MASKED);
            at += step;
            ahead += step;
        }
    };
    // Phase A: every line of the block goes on and is before its middle.
    auto before_window = [&]() __attribute__((always_inline)) {
#pragma unroll
        for (int p = 0; p < PREFETCH; p++) {
            __builtin_amdgcn_sched_barrier(0);
            s.step(cost[p], p1x2, p2, lane);
            vstore<K>(paired, at, s.bytes());
            __builtin_amdgcn_sched_barrier(0);
            cost[p] = vload<K>(costs, ahead);
            at += step;
            ahead += step;
        }
    };
    // Phase C: every line of the block goes on, with the other's bytes in memory.
    auto after_window = [&]() __attribute__((always_inline)) {
#pragma unroll
        for (int p = 0; p < PREFETCH; p++) {
            __builtin_amdgcn_sched_barrier(0);
            s.step(cost[p], p1x2, p2, lane);
            const Bytes<K> mine = s.bytes();
            if (p == 0) meet();
            Bytes<K> out;
#pragma unroll
            for (int k = 0; k < (K + 3) / 4; k++) out.w[k] = mine.w[k] + other[p].w[k];
            vstore<K, NT>(paired, at, out);
            __builtin_amdgcn_sched_barrier(0);
            cost[p] = vload<K>(costs, ahead);
            if constexpr (NT) other[p] = vload_policy<K, NT>(paired, ahead);
            else other[p] = vload<K>(paired, ahead);
            at += step;
            ahead += step;
        }
    };
    // Block-uniform phase bounds, multiples of PREFETCH: A is [0, a_end), C is
    // [c_start, c_end) (empty when c_end <= c_start).
    const int a_end = max(0, (window_lo + 1) >> 1) & ~(PREFETCH - 1);
    const int c_start = ((window_hi >> 1) + PREFETCH) & ~(PREFETCH - 1);
    const int c_end = (mid_lo + 1) & ~(PREFETCH - 1);
    int i0 = 0;
    for (; i0 < a_end; i0 += PREFETCH) before_window();
    for (; i0 < c_start && i0 < longest; i0 += PREFETCH) general(i0);
    for (; i0 < c_end; i0 += PREFETCH) after_window();
    for (; i0 < longest; i0 += PREFETCH) general(i0);
}

// ------------------------------------------------------------------------------------------
// The two-wave form meeting by recomputation. The form above stores each direction's bytes
// of its first half and reads them back in the other's second half: 2N of the 5N an
// orientation moves. Here a direction stores its path only every S steps of its first half
// (a checkpoint: the path's D bytes are the whole of its state), and the other direction,
// walking its second half, takes the first half's pixels S at a time: it loads the checkpoint
// that enters them, walks the first direction's recurrence over those S pixels again (their
// costs are the ones it walks itself) and keeps the bytes in registers, then walks its own S
// steps, adding them. The recurrence is integer arithmetic, so the recomputed bytes are the
// ones the first direction made. An orientation then moves 2N of costs, 2N/S of checkpoints
// and N of pair sums, for about one more recurrence step per pixel of the second half.
//
// Line positions 0 .. T-1, T = S * ceil(longest / S) over the block's lines: a line of length
// L has its pixels at 0 .. L-1 and nothing past them (costs load as zeros, stores drop). A
// direction that walks zero costs from its entering path keeps that path (every value stays
// 0), so the backward wave may start at T-1. The forward wave takes the first nf = ceil(n/2)
// of the n segments of S positions, the backward wave the last n - nf: its phase 1. In its
// phase 2 each wave walks the other's segments, nearest first. Segment j of a wave's phase 1
// (its steps jS .. jS+S-1) enters with the path after its step jS-1, which the wave stores at
// that step's pixel; in the other wave's step order that pixel is the one just past the
// segment, whose pair sum the other stores only one segment later. Segment 0 enters with the
// line's entering path, which restores from zero bytes (a masked load). The phases meet at
// one barrier, after every checkpoint store has completed; the waves touch nothing of each
// other's otherwise, and no LDS.
// Block b of launch geometry `a` (TWO_BLOCK threads), on the pair pair_of(a.single, table,
// index).
template <int K, int S, int LANES, bool STREAM>
__device__ __forceinline__ void path_rc_block(const PathArgs &a, const PairDesc *table, int index,
                                              int b) {
    constexpr int NT = STREAM && K >= 4 ? 2 : 0;
    static_assert(LANES == 16 || LANES == 32, "a line is one or two DPP rows");
    static_assert(S == 2 || S == 4 || S == 8, "segments of 2, 4 or 8 steps");
    constexpr int D = LANES * K;
    constexpr int LINES = 64 / LANES;
    const int dir = threadIdx.x >> 6;  // wave-uniform: 0 forward, 1 backward
    const int lane = threadIdx.x & (LANES - 1);
    const int row = (threadIdx.x & 63) / LANES;
    const BlockLines bl = block_lines(a, b, LINES);
    const int o = bl.o;
    const int W = a.W, H = a.H;
    const PairDesc pd = pair_of(a.single, table, index);
    if (o == 1 && !a.backward_columns) {
        walk_columns_down<K, S, LANES>(a, pd, b, dir, lane, row);
        return;
    }
    const int first_line = bl.first_line;
    const LineGeometry g = line_geometry(o, first_line + row, W, H);
    int longest = 0;
#pragma unroll
    for (int r = 0; r < LINES; r++) longest = max(longest, line_geometry(o, first_line + r, W, H).length);
    if (longest == 0) return;  // both waves leave
    const int n = (longest + S - 1) / S;
    const int T = n * S;
    const int h1 = (dir == 0 ? (n + 1) / 2 : n / 2) * S;  // this wave's phase 1, in steps
    const int h2 = T - h1;

    const int dy = o == 0 ? 0 : 1;
    const int dx = o == 0 ? 1 : (o == 1 ? 0 : (o == 2 ? 1 : -1));
    const uint32_t stride = (uint32_t)((dy * W + dx) * D);
    const int L = g.length;
    const uint32_t first = ((uint32_t)g.y0 * W + (uint32_t)g.x0) * D + lane * K;
    const uint32_t volume = (uint32_t)W * H * D;
    const Rsrc costs = buffer_of(pd.b.costs, volume);
    const Rsrc paired = buffer_of(orientation_volume(pd.b, o), volume);
    const uint32_t p1x2 = (uint32_t)pd.p1 * 0x10001u;
    const uint32_t p2 = (uint32_t)pd.p2;
    // Step t of this wave is at position t (forward) or T-1-t (backward): a pixel of the line
    // when t - lo < L, with lo the first step on it.
    const int lo = dir == 0 ? 0 : T - L;
    const uint32_t base = dir == 0 ? first : first + (uint32_t)(T - 1) * stride;
    const uint32_t step = dir == 0 ? stride : 0u - stride;
    auto at = [&](int t) __attribute__((always_inline)) -> uint32_t {
        return (uint32_t)(t - lo) < (uint32_t)L ? base + (uint32_t)t * step : MASKED;
    };

    PathState<K, LANES> s;
    s.reset(p1x2, lane);
    Bytes<K> ca[S], cb[S];
#pragma unroll
    for (int p = 0; p < S; p++) ca[p] = vload<K>(costs, at(p));
    // Phase 1: the wave's own path, costs S steps ahead, a checkpoint at the end of every
    // segment but the last.
    for (int t0 = 0; t0 < h1; t0 += S) {
#pragma unroll
        for (int p = 0; p < S; p++) {
            const int t = t0 + p;
            __builtin_amdgcn_sched_barrier(0);
            s.step(ca[p], p1x2, p2, lane);
            if (p == S - 1) vstore<K>(paired, t + 1 < h1 ? at(t) : MASKED, s.bytes());
            __builtin_amdgcn_sched_barrier(0);
            ca[p] = vload<K>(costs, at(t + S));
        }
    }
    // Every checkpoint stored and the other wave's too.
    asm volatile("s_waitcnt vmcnt(0) lgkmcnt(0)\n\ts_barrier" ::: "memory");
    if (h2 == 0) return;  // wave-uniform
    PathState<K, LANES> rec;
    // Phase 2, segment k: steps h1 + kS .. h1 + kS + S-1, whose costs are in cur; the
    // next segment's in the other array; the checkpoint entering it (in the other's
    // order) at step h1 + kS + S.
    Bytes<K> entering = vload<K>(paired, S < h2 ? at(h1 + S) : MASKED);
#pragma unroll
    for (int p = 0; p < S; p++) cb[p] = vload<K>(costs, at(h1 + S + p));
    auto segment = [&](int k, Bytes<K>(&cur)[S]) __attribute__((always_inline)) {
        const int t0 = h1 + k * S;
        rec.restore(entering, p1x2, lane);
        Bytes<K> theirs[S];
#pragma unroll
        for (int r = S - 1; r >= 0; r--) {
            __builtin_amdgcn_sched_barrier(0);
            rec.step(cur[r], p1x2, p2, lane);
            theirs[r] = rec.bytes();
        }
        entering = vload<K>(paired, t0 + 2 * S < T ? at(t0 + 2 * S) : MASKED);
#pragma unroll
        for (int p = 0; p < S; p++) {
            __builtin_amdgcn_sched_barrier(0);
            s.step(cur[p], p1x2, p2, lane);
            const Bytes<K> mine = s.bytes();
            Bytes<K> out;
#pragma unroll
            for (int w = 0; w < (K + 3) / 4; w++) out.w[w] = mine.w[w] + theirs[p].w[w];
            vstore<K, NT>(paired, at(t0 + p), out);
            __builtin_amdgcn_sched_barrier(0);
            cur[p] = vload<K>(costs, at(t0 + 2 * S + p));
        }
    };
    for (int k = 0; k < h2 / S; k += 2) {
        segment(k, ca);
        if ((k + 1) * S < h2) segment(k + 1, cb);
    }
}

template <int K, int S, int LANES, bool STREAM>
__global__ __launch_bounds__(TWO_BLOCK) void path_rc_kernel(PathArgs a) {
    path_rc_block<K, S, LANES, STREAM>(a, a.table, blockIdx.y, blockIdx.x);
}

// ------------------------------------------------------------------------------------------
// The two-wave form with each direction's D disparities also split into G slices of 16
// lanes, one wave each: 2G waves per block over the same four lines, each issuing about
// 1/G of a direction's per-step work. At every step the slices of a direction exchange,
// through LDS and one s_barrier, what the next step needs from the others: each slice's
// minimum (the line's minimum is theirs), its first lane's first pair (the right neighbour
// of the slice before) and its last lane's last pair (the left neighbour of the slice
// after). The exchange is double-buffered by step parity: a wave writes step t+1's values
// only after the barrier of step t, which every reader of step t-1's values has passed.
// The two directions meet as in the two-wave form, with a barrier at every step (X below).

struct SliceShare {
    uint32_t minimum, first, last;
};

template <int K, int PREFETCH, int G>
__global__ __launch_bounds__(128 * G) void path_sliced_kernel(PathArgs a) {
    constexpr int LAG = PREFETCH == 8 ? 6 : 8;  // steps after which a store has completed
    constexpr int X = PREFETCH + LAG;            // past >= X: the other's bytes from memory
    constexpr int RING = X + 1 <= 15 ? 16 : 32;  // past < RING - 1: the ring slot still holds them
    static_assert(X <= RING - 2, "the LDS ring must hold the late steps");
    constexpr int R = K / 2;
    constexpr int D = 16 * K * G;
    constexpr int LINES = 4;
    constexpr uint32_t SENT = 0x3FFF3FFFu;
    __shared__ Bytes<K> ring[2][LINES][RING][16 * G];
    __shared__ SliceShare share[2][2][LINES][G];  // [step parity][direction][line][slice]
    const int wave = threadIdx.x >> 6;
    const int dir = wave / G, slice = wave % G;  // wave-uniform
    const int lane = threadIdx.x & 15;
    const int row = (threadIdx.x >> 4) & 3;
    const int b = blockIdx.x;
    int o = a.last;
    for (int k = a.first; k < a.last; k++)
        if (b < a.block_start[k + 1]) {
            o = k;
            break;
        }
    const int first_line = (b - a.block_start[o]) * LINES;
    const int W = a.W, H = a.H;
    const PairDesc pd = pair_of(a.single, a.table, blockIdx.y);
    const LineGeometry g = line_geometry(o, first_line + row, W, H);
    int longest = 0, mid_lo = 1 << 30, mid_hi = -(1 << 30);
#pragma unroll
    for (int r = 0; r < LINES; r++) {
        const int len = line_geometry(o, first_line + r, W, H).length;
        longest = max(longest, len);
        mid_lo = min(mid_lo, len - 1);
        mid_hi = max(mid_hi, len - 1);
    }
    if (longest == 0) return;  // every wave of the block leaves
    const int window_lo = mid_lo - X - 2, window_hi = mid_hi + X + 2;  // ring writes (in 2i)

    const int dy = o == 0 ? 0 : 1;
    const int dx = o == 0 ? 1 : (o == 1 ? 0 : (o == 2 ? 1 : -1));
    const uint32_t stride = (uint32_t)((dy * W + dx) * D);
    const int L = g.length;
    const int column = slice * 16 + lane;  // this lane's disparities: [column*K, column*K+K)
    const uint32_t first = ((uint32_t)g.y0 * W + (uint32_t)g.x0) * D + column * K;
    const uint32_t last = first + (uint32_t)max(L - 1, 0) * stride;
    const uint32_t volume = (uint32_t)W * H * D;
    const Rsrc costs = buffer_of(pd.b.costs, volume);
    const Rsrc paired = buffer_of(orientation_volume(pd.b, o), volume);
    const uint32_t p1x2 = (uint32_t)pd.p1 * 0x10001u;
    const uint32_t p2 = (uint32_t)pd.p2;
    const uint32_t start = dir == 0 ? first : last;
    const uint32_t step = dir == 0 ? stride : 0u - stride;

    // The path entering a line has prev = 0 everywhere; past the line's two ends, the
    // sentinel.
    uint32_t v[R];
#pragma unroll
    for (int k = 0; k < R; k++) v[k] = 0;
    uint32_t pm = 0;
    uint32_t left_edge = slice == 0 ? SENT : 0u, right_edge = slice == G - 1 ? SENT : 0u;

    Bytes<K> cost[PREFETCH], other[PREFETCH];
    Bytes<K> none;
#pragma unroll
    for (int k = 0; k < (K + 3) / 4; k++) none.w[k] = 0;
#pragma unroll
    for (int p = 0; p < PREFETCH; p++) {
        __builtin_amdgcn_sched_barrier(0);
        vstore<K>(paired, MASKED, none);
        __builtin_amdgcn_sched_barrier(0);
        cost[p] = vload<K>(costs, start + p * step);
        other[p] = vload<K>(paired, MASKED);
    }
    uint32_t at = start, ahead = start + PREFETCH * step;
    for (int i0 = 0; i0 < longest; i0 += PREFETCH) {
#pragma unroll
        for (int p = 0; p < PREFETCH; p++) {
            const int i = i0 + p;
            __builtin_amdgcn_sched_barrier(0);
            // The recurrence on this slice (as PathState::step), its neighbours across the
            // slice's ends being the other slices' pairs exchanged last step.
            const uint32_t left_in = dpp_or<DPP_ROW_SHR1>(left_edge, v[R - 1]);
            const uint32_t right_in = dpp_or<DPP_ROW_SHL1>(right_edge, v[0]);
            uint32_t sh[R + 1];
            sh[0] = (v[0] << 16) | (left_in >> 16);
#pragma unroll
            for (int k = 1; k < R; k++) sh[k] = (v[k] << 16) | (v[k - 1] >> 16);
            sh[R] = (right_in << 16) | (v[R - 1] >> 16);
            const uint32_t PM = pm * 0x10001u;
            const uint32_t JUMP = (pm + p2) * 0x10001u;
            uint32_t c[R];
            if constexpr (K == 2) {
                c[0] = lo_pair(cost[p].w[0]);
            } else {
#pragma unroll
                for (int k = 0; k < R / 2; k++) {
                    c[2 * k] = lo_pair(cost[p].w[k]);
                    c[2 * k + 1] = hi_pair(cost[p].w[k]);
                }
            }
            uint32_t mn = 0xFFFFFFFFu;
#pragma unroll
            for (int k = 0; k < R; k++) {
                uint32_t t = pk_min(sh[k], sh[k + 1]);
                t = pk_add(t, p1x2);
                t = pk_min(t, v[k]);
                t = pk_min(t, JUMP);
                t = pk_add(pk_sub(t, PM), c[k]);
                v[k] = t;
                mn = pk_min(mn, t);
            }
            const uint32_t m = row_min_u32(min(mn & 0xFFFFu, mn >> 16));
            Bytes<K> mine;
            if constexpr (K == 2) {
                mine.w[0] = __builtin_amdgcn_perm(0u, v[0], 0x0C0C0200u);
            } else {
#pragma unroll
                for (int k = 0; k < R / 2; k++) mine.w[k] = pack_pairs(v[2 * k], v[2 * k + 1]);
            }
            // Share this step's slice minimum and end pairs; publish own bytes near the middle.
            SliceShare &out_share = share[i & 1][dir][row][slice];
            if (lane == 0) {
                out_share.minimum = m;
                out_share.first = v[0];
            }
            if (lane == 15) out_share.last = v[R - 1];
            const bool window = 2 * i >= window_lo && 2 * i <= window_hi;  // block-uniform
            if (window) ring[dir][row][i & (RING - 1)][column] = mine;
            // The store of step i - LAG is complete once at most its two loads and the three
            // accesses of each later step are outstanding; LDS writes done.
            asm volatile("s_waitcnt vmcnt(%0) lgkmcnt(0)\n\ts_barrier" ::"n"(3 * LAG - 1) : "memory");
            uint32_t line_min = share[i & 1][dir][row][0].minimum;
#pragma unroll
            for (int q = 1; q < G; q++) line_min = min(line_min, share[i & 1][dir][row][q].minimum);
            pm = line_min;
            left_edge = slice == 0 ? SENT : share[i & 1][dir][row][slice > 0 ? slice - 1 : 0].last;
            right_edge = slice == G - 1 ? SENT : share[i & 1][dir][row][slice < G - 1 ? slice + 1 : 0].first;

            const int past = 2 * i - (L - 1);
            Bytes<K> theirs = other[p];
            if (past >= 0 && past < X) theirs = ring[dir ^ 1][row][(i - past) & (RING - 1)][column];
            Bytes<K> sum;
#pragma unroll
            for (int k = 0; k < (K + 3) / 4; k++) sum.w[k] = mine.w[k] + (past < 0 ? 0u : theirs.w[k]);
            vstore<K>(paired, i < L && !(past == 0 && dir == 1) ? at : MASKED, sum);
            __builtin_amdgcn_sched_barrier(0);
            cost[p] = vload<K>(costs, ahead);
            const bool stored = 2 * (i + PREFETCH) - (L - 1) >= X;
            other[p] = vload<K>(paired, stored ? ahead : MASKED);
            at += step;
            ahead += step;
        }
    }
}

// ------------------------------------------------------------------------------------------
// Selection (reference.rs:148-186; nx_stereo.ex:608-641), a pixel's disparities over 4, 8 or
// 16 lanes.
//
// V[d] = S[d] + 0x8000 for d > x. The winner is the lowest d of the least V; d = 0 is never
// past the left edge and S <= 1016, so the winner is never a penalised d, and (V << 4 | j)
// for the lane's j-th disparity orders the unpenalised ones in 14 bits (penalised ones take
// 0xFFFF). Sums, keys and the second-best search run on u16 pairs. The second best is the
// least V outside [below, above]: those entries are lifted to 0xFFFD..0xFFFF, above every
// V (at most 0x83F8). The winner's two neighbours come from the lanes holding them.

// A block takes SELECT_GROUPS groups of 256 / PL pixels of a row, one after another, loading the
// next group's four volumes while it selects in the current one (so that each wave keeps
// loads in flight).
constexpr int SELECT_GROUPS = 2;

// STREAM: the volumes are read with non-temporal loads, which at sizes whose volumes
// overflow the Infinity Cache keep them from displacing what later launches reuse
// (1920x1080x256: 391 -> 347 us; at 728x544x128, whose volumes it holds, plain loads).
// The minimum over the PL lanes of a pixel (PL = 4, 8 or 16 consecutive lanes of a DPP
// row), left in each of them: quad_perm [1,0,3,2] and [2,3,0,1], row_half_mirror, then the
// row's two halves.
template <int PL>
__device__ __forceinline__ uint32_t pixel_min_u32(uint32_t v) {
    if constexpr (PL == 16) {
        return row_min_u32(v);
    } else {
        v = min(v, dpp<0xB1>(v));
        v = min(v, dpp<0x4E>(v));
        if constexpr (PL == 8) v = min(v, dpp<0x141>(v));
        return v;
    }
}
template <int PL>
__device__ __forceinline__ uint32_t pixel_min_pk(uint32_t v) {
    if constexpr (PL == 16) {
        return row_min_pk(v);
    } else {
        v = pk_min(v, dpp<0xB1>(v));
        v = pk_min(v, dpp<0x4E>(v));
        if constexpr (PL == 8) v = pk_min(v, dpp<0x141>(v));
        return v;
    }
}

// K disparities a lane, PL lanes a pixel (D = PL K): 16 a lane where D allows (64, 128 and
// 256 disparities on 4, 8 and 16 lanes), so that a wave selects more pixels at once and the
// per-pixel work (the reductions over the pixel's lanes, the subpixel step) is shared by
// fewer lanes; 192 disparities take 12 a lane on 16 lanes.
// Block bx of row y of pair pd (256 threads).
template <int K, int PL, bool STREAM>
__device__ __forceinline__ void select_block(const PairDesc &pd, int orientations, int W, int H,
                                             int bx, int y) {
    static_assert(K <= 16, "a key keeps the lane's index in 4 bits");
    static_assert(PL == 4 || PL == 8 || PL == 16, "a pixel's lanes lie in one DPP row");
    constexpr int D = PL * K, R = K / 2;
    constexpr int PIXELS = 256 / PL;  // a block's pixels per group
    const uint32_t volume = (uint32_t)W * H * D;
    const Rsrc vol[4] = {buffer_of(pd.b.paired[0], volume), buffer_of(pd.b.paired[1], volume),
                         buffer_of(pd.b.paired[2], volume), buffer_of(pd.b.paired[3], volume)};
    const float uniqueness = pd.uniqueness;
    float *__restrict__ selected = pd.b.selected;
    uint32_t *__restrict__ keys = pd.b.keys;
    uint16_t *__restrict__ sums_out = pd.sums;
    const int lane = threadIdx.x & (PL - 1);  // within the pixel's lanes
    const int x_first = bx * PIXELS * SELECT_GROUPS + (int)(threadIdx.x / PL);
    const int d0 = lane * K;
    // This lane's bytes of group g's pixel (the last pixel of the row past its end).
    auto offset_of = [&](int g) {
        const uint32_t pix = (uint32_t)y * W + min(x_first + PIXELS * g, W - 1);
        return pix * D + (uint32_t)lane * K;  // the volume is under 2^31 bytes
    };
    Bytes<K> cur[4], next[4];
    auto load = [&](Bytes<K>(&b)[4], int g) __attribute__((always_inline)) {
        const uint32_t at = offset_of(g);
#pragma unroll
        for (int m = 0; m < 4; m++)
            if (m < orientations) b[m] = vload_policy<K, STREAM ? 2 : 0>(vol[m], at);
    };
    load(cur, 0);
#pragma unroll
    for (int g = 0; g < SELECT_GROUPS; g++) {
        if (g + 1 < SELECT_GROUPS) load(next, g + 1);
        const int x = x_first + PIXELS * g;
        const bool valid = x < W;
        const uint32_t at = offset_of(g), pix = (uint32_t)y * W + min(x, W - 1);
        uint32_t sum[R];
#pragma unroll
        for (int k = 0; k < K / 4; k++) {
            sum[2 * k] = lo_pair(cur[0].w[k]);
            sum[2 * k + 1] = hi_pair(cur[0].w[k]);
        }
#pragma unroll
        for (int m = 1; m < 4; m++) {
            if (m >= orientations) break;
#pragma unroll
            for (int k = 0; k < K / 4; k++) {
                sum[2 * k] = pk_add(sum[2 * k], lo_pair(cur[m].w[k]));
                sum[2 * k + 1] = pk_add(sum[2 * k + 1], hi_pair(cur[m].w[k]));
            }
        }
        if (sums_out && valid) {
            uint32_t *out = reinterpret_cast<uint32_t *>(sums_out + at);
#pragma unroll
            for (int k = 0; k < R; k++) out[k] = sum[k];
        }
        uint32_t value[R], key[R];
        // (j, j + 1) of register r, as u16 pairs.
        auto pair_index = [](int r) { return (uint32_t)(2 * r) | (uint32_t)(2 * r + 1) << 16; };
        if (__any(x < D - 1)) {  // wave-uniform: a pixel of the wave has disparities past x
            const uint32_t X1 = (uint32_t)(x + 1) * 0x10001u;
            const uint32_t D0 = (uint32_t)d0 * 0x10001u;
#pragma unroll
            for (int r = 0; r < R; r++) {
                // d - (x + 1) wraps past 0xF000 for d <= x and stays under 0x100 for d > x.
                const uint32_t q = pk_sub(pk_add(D0, pair_index(r)), X1);
                const uint32_t past = pk_sub(pk_min(q, 0x01000100u), 0x01000100u);  // 0xFF.. or 0
                value[r] = sum[r] | (past & 0x80008000u);
                key[r] = pk_max(pk_add(pk_mul(sum[r], 0x00100010u), pair_index(r)), past);
            }
        } else {
#pragma unroll
            for (int r = 0; r < R; r++) {
                value[r] = sum[r];
                key[r] = pk_add(pk_mul(sum[r], 0x00100010u), pair_index(r));
            }
        }
        uint32_t kmin = key[0];
#pragma unroll
        for (int r = 1; r < R; r++) kmin = pk_min(kmin, key[r]);
        const uint32_t k16 = pk_min_halves(kmin) & 0xFFFFu;
        const uint32_t best = pixel_min_u32<PL>((k16 >> 4) << 8 | (uint32_t)(d0 + (int)(k16 & 15u)));
        const uint32_t minimum = best >> 8;
        const int winner = (int)(best & 0xFFu);
        const int below = max(winner - 1, 0), above = min(winner + 1, D - 1);

        // Entries with below <= d0 + j <= above: (j - (below - d0)) <= span as u16.
        const uint32_t T = (uint32_t)(below - d0) & 0xFFFFu;
        const uint32_t S1 = (uint32_t)(above - below + 1);
        const uint32_t T2 = T * 0x10001u, S2 = S1 * 0x10001u;
        uint32_t smin = 0xFFFFFFFFu;
#pragma unroll
        for (int r = 0; r < R; r++) {
            const uint32_t lifted = pk_sub(pk_min(pk_sub(pair_index(r), T2), S2), S2);
            smin = pk_min(smin, pk_max(value[r], lifted));
        }
        const uint32_t second = pixel_min_u32<PL>(pk_min_halves(smin)) & 0xFFFFu;
        // The winner's two neighbours' sums, from the lanes that hold them: the owner of
        // d = below gives S[below] in the low half (others 0xFFFF), of above in the high.
        const int jb = below - d0, ja = above - d0;
        uint32_t wb = sum[0], wa = sum[0];
#pragma unroll
        for (int r = 1; r < R; r++) {
            wb = (jb >> 1) == r ? sum[r] : wb;
            wa = (ja >> 1) == r ? sum[r] : wa;
        }
        const uint32_t sb = (jb & 1 ? wb >> 16 : wb) & 0xFFFFu;
        const uint32_t sa = (ja & 1 ? wa >> 16 : wa) & 0xFFFFu;
        const uint32_t near = pixel_min_pk<PL>(((uint32_t)jb < (uint32_t)K ? sb : 0xFFFFu) |
                                         ((uint32_t)ja < (uint32_t)K ? sa : 0xFFFFu) << 16);
        if (lane == 0 && valid) {
            const uint32_t lower = (near & 0xFFFFu) + (x < below ? 0x8000u : 0u);
            const uint32_t upper = (near >> 16) + (x < above ? 0x8000u : 0u);
            const float lower_f = (float)lower;
            const float upper_f = (float)upper;
            const float minimum_f32 = (float)minimum;
            // An empty "second" set would be i32::MAX in the reference; D >= 64 never
            // leaves it empty.
            const bool accepted = (float)second > __fmul_rn(minimum_f32, uniqueness);
            const float denominator =
                __fsub_rn(__fadd_rn(lower_f, upper_f), __fmul_rn(2.0f, minimum_f32));
            const float correction =
                fminf(fmaxf(__fdiv_rn(__fsub_rn(lower_f, upper_f),
                                      __fmul_rn(2.0f, fmaxf(denominator, 1.0f))),
                            -0.5f),
                      0.5f);
            const bool refine = denominator > 0.0f && winner > 0 && winner < min(D - 1, x);
            const float disparity = __fadd_rn((float)winner, refine ? correction : 0.0f);
            selected[pix] = accepted ? disparity : -1.0f;
            keys[pix] = accepted ? ((minimum << 16) | (uint32_t)winner) : INVALID_KEY;
        }
#pragma unroll
        for (int m = 0; m < 4; m++) cur[m] = next[m];
    }
}

// Grid x: groups of pixels of a row; y: the row; z: the pair.
template <int K, int PL, bool STREAM>
__global__ __launch_bounds__(256) void select_kernel(PairDesc single, const PairDesc *table,
                                                     int orientations, int W) {
    select_block<K, PL, STREAM>(pair_of(single, table, blockIdx.z), orientations, W, gridDim.y,
                                blockIdx.x, blockIdx.y);
}

// ------------------------------------------------------------------------------------------
// Left-right check (reference.rs:188-235; check.rs:7-47), per row: each right pixel keeps
// the cheapest left match projected onto it, ties to the larger disparity, via an LDS
// atomic minimum of (cost << 16) | (65535 - d) into the row's map.

constexpr int MAX_WIDTH = MAX_IMAGE_WIDTH;

// The checked value of pixel x of a row: `sel` is the row's selected disparities and
// `right` its map, already made.
__device__ __forceinline__ float checked_value(const float *sel, const uint32_t *right, int x,
                                               int W) {
    const float value = sel[x];
    if (value < 0.0f) return -1.0f;
    const int lo = (int)floorf(value), hi = (int)ceilf(value);
    bool rejected = true;
#pragma unroll
    for (int k = 0; k < 2; k++) {
        const int d = k ? hi : lo;
        const int rx = x - d;
        bool dis = false;
        if (rx >= 0 && rx < W) {
            const uint32_t found = right[rx];
            dis = found != INVALID_KEY && abs(65535 - (int)(found & 0xFFFFu) - d) > 1;
        }
        rejected = rejected && dis;
    }
    return rejected ? -1.0f : value;
}

// ------------------------------------------------------------------------------------------
// 3x3 median with replicated borders (reference.rs:237-254; filters.rs:10-30 network),
// on i32 bit patterns, which order -1.0 and non-negative values as f32 does; then the
// Disparity stage's d > 0 ? d : -1 (stage/disparity.ex:162).
//
// One block per output row y: it checks rows y-1, y and y+1 (clamped) in LDS, each with
// its own map, then takes the median of row y from them. The checked map is stored too
// when asked for (stage-by-stage checks).

__host__ __device__ constexpr size_t check_median_lds(int W) { return 6 * 4 * (size_t)W; }

// Grid x: the row; y: the pair.
constexpr int CHECK_BLOCK = 1024;

// Row y of pair pd, by one block of CHECK_BLOCK threads with check_median_lds(W) bytes of
// LDS at `check_lds`.
__device__ __forceinline__ void check_median_row(const PairDesc &pd, int W, int H, int y,
                                                 int keep_checked, int positive,
                                                 uint32_t *check_lds) {
    uint32_t *right = check_lds;                                    // [3][W]
    float *checked = reinterpret_cast<float *>(check_lds + 3 * W);  // [3][W]
    int rows[3];
#pragma unroll
    for (int r = 0; r < 3; r++) rows[r] = min(max(y - 1 + r, 0), H - 1);
    for (int i = threadIdx.x; i < 3 * W; i += blockDim.x) right[i] = INVALID_KEY;
    __syncthreads();
    // The three rows' keys and selected values are loaded together; the selected values wait
    // in `checked`, which the check then overwrites pixel by pixel (each its own).
    for (int x = threadIdx.x; x < W; x += blockDim.x) {
        uint32_t key[3];
        float sel[3];
#pragma unroll
        for (int r = 0; r < 3; r++) {
            key[r] = pd.b.keys[(size_t)rows[r] * W + x];
            sel[r] = pd.b.selected[(size_t)rows[r] * W + x];
        }
#pragma unroll
        for (int r = 0; r < 3; r++) {
            checked[r * W + x] = sel[r];
            if (key[r] != INVALID_KEY) {
                const uint32_t w = key[r] & 0xFFFFu;
                atomicMin(&right[r * W + x - (int)w], (key[r] & 0xFFFF0000u) | (65535u - w));
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 3; r++) {
        for (int x = threadIdx.x; x < W; x += blockDim.x) {
            const float c = checked_value(checked + r * W, right + r * W, x, W);
            checked[r * W + x] = c;
            if (r == 1 && keep_checked) pd.b.checked[(size_t)y * W + x] = c;
        }
    }
    __syncthreads();
    float *__restrict__ out = pd.disparity + (size_t)y * W;
    for (int x = threadIdx.x; x < W; x += blockDim.x) {
        int v[9];
        int n = 0;
#pragma unroll
        for (int r = 0; r < 3; r++) {
#pragma unroll
            for (int dx = -1; dx <= 1; dx++)
                v[n++] = __float_as_int(checked[r * W + min(max(x + dx, 0), W - 1)]);
        }
#define CX(a, b)                        \
    {                                   \
        const int lo = min(v[a], v[b]); \
        const int hi = max(v[a], v[b]); \
        v[a] = lo;                      \
        v[b] = hi;                      \
    }
        CX(1, 2) CX(4, 5) CX(7, 8) CX(0, 1) CX(3, 4) CX(6, 7) CX(1, 2) CX(4, 5) CX(7, 8) CX(0, 3)
        CX(5, 8) CX(4, 7) CX(3, 6) CX(1, 4) CX(2, 5) CX(4, 7) CX(4, 2) CX(6, 4) CX(4, 2)
#undef CX
        const float m = __int_as_float(v[4]);
        out[x] = !positive || m > 0.0f ? m : -1.0f;
    }
}

__global__ __launch_bounds__(CHECK_BLOCK) void check_median_kernel(PairDesc single,
                                                                   const PairDesc *table, int W,
                                                                   int H, int keep_checked,
                                                                   int positive) {
    extern __shared__ __align__(16) uint32_t check_lds[];
    check_median_row(pair_of(single, table, blockIdx.y), W, H, blockIdx.x, keep_checked,
                     positive, check_lds);
}

// ------------------------------------------------------------------------------------------

static int orientations_of(int paths) { return paths == 8 ? 4 : 2; }

size_t match_bytes(int W, int H, int D, int paths) {
    const size_t P = (size_t)W * H, N = P * D;
    auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
    return up(8 * P) * 2 + up(N) * (1 + orientations_of(paths)) + up(4 * P) * 3;
}

void match_carve(MatchBuffers &b, void *base, int W, int H, int D, int paths) {
    const size_t P = (size_t)W * H, N = P * D;
    auto up = [](size_t n) { return (n + 255) & ~(size_t)255; };
    uint8_t *p = static_cast<uint8_t *>(base);
    b.census_left = reinterpret_cast<uint64_t *>(p);
    p += up(8 * P);
    b.census_right = reinterpret_cast<uint64_t *>(p);
    p += up(8 * P);
    b.costs = p;
    p += up(N);
    for (int o = 0; o < 4; o++) {
        if (o < orientations_of(paths)) {
            b.paired[o] = p;
            p += up(N);
        } else {
            b.paired[o] = nullptr;
        }
    }
    b.selected = reinterpret_cast<float *>(p);
    p += up(4 * P);
    b.keys = reinterpret_cast<uint32_t *>(p);
    p += up(4 * P);
    b.checked = reinterpret_cast<float *>(p);
}

// ------------------------------------------------------------------------------------------
// Launchers. One launch per kernel covers `count` pairs of one size class: grid y (or z)
// indexes the pair, whose descriptor comes by value (count 1) or from `table`.

template <int K, int PREFETCH, int LANES>
static void launch_paths(const PathArgs &a, int blocks, int count, hipStream_t stream) {
    hipLaunchKernelGGL((path_kernel<K, PREFETCH, LANES>), dim3(blocks, count), dim3(PATH_BLOCK), 0,
                       stream, a);
}

template <int K, int LANES>
static void launch_paths_prefetch(const PathArgs &a, int blocks, int count, int prefetch,
                                  hipStream_t stream) {
    switch (prefetch) {
        case 2: launch_paths<K, 2, LANES>(a, blocks, count, stream); break;
        case 8: launch_paths<K, 8, LANES>(a, blocks, count, stream); break;
        default: launch_paths<K, 4, LANES>(a, blocks, count, stream); break;
    }
}

bool lanes_fit(int disparities, int lanes) {
    if (lanes != 16 && lanes != 32 && lanes != 64) return false;
    const int k = disparities / lanes;
    return disparities % lanes == 0 && (k == 2 || k == 4 || k == 8 || k == 12 || k == 16);
}

// The path launch of D disparities with `lanes` lanes per line (16 where that does not fit).
template <int D>
static void launch_paths_lanes(const PathArgs &a, int blocks, int count, const Schedule &sch,
                               hipStream_t stream) {
    if constexpr (D == 128 || D == 256) {
        if (sch.lanes == 64) return launch_paths_prefetch<D / 64, 64>(a, blocks, count, sch.prefetch, stream);
    }
    if constexpr (D == 64 || D == 128 || D == 256) {
        if (sch.lanes == 32) return launch_paths_prefetch<D / 32, 32>(a, blocks, count, sch.prefetch, stream);
    }
    launch_paths_prefetch<D / 16, 16>(a, blocks, count, sch.prefetch, stream);
}

// Slices per direction that fit D disparities (D / (16 slices) an even count up to 16), or 1.
int slices_fit(int disparities, int slices) {
    if (slices != 2 && slices != 4) return 1;
    const int k = disparities / (16 * slices);
    return disparities % (16 * slices) == 0 && (k == 2 || k == 4 || k == 8) ? slices : 1;
}

template <int K, int G>
static void launch_sliced_prefetch(const PathArgs &a, int blocks, int count, int prefetch,
                                   hipStream_t stream) {
    switch (prefetch) {
        case 2: hipLaunchKernelGGL((path_sliced_kernel<K, 2, G>), dim3(blocks, count), dim3(128 * G), 0, stream, a); break;
        case 8: hipLaunchKernelGGL((path_sliced_kernel<K, 8, G>), dim3(blocks, count), dim3(128 * G), 0, stream, a); break;
        default: hipLaunchKernelGGL((path_sliced_kernel<K, 4, G>), dim3(blocks, count), dim3(128 * G), 0, stream, a); break;
    }
}

template <int D>
static void launch_sliced(const PathArgs &a, int blocks, int count, int prefetch, int slices,
                          hipStream_t stream) {
    if constexpr (D == 64) {
        launch_sliced_prefetch<2, 2>(a, blocks, count, prefetch, stream);
    } else if constexpr (D == 128 || D == 256) {
        if (slices == 4)
            launch_sliced_prefetch<D / 64, 4>(a, blocks, count, prefetch, stream);
        else
            launch_sliced_prefetch<D / 32, 2>(a, blocks, count, prefetch, stream);
    }
}

bool two_lanes_fit(int disparities) {
    return disparities == 64 || disparities == 128 || disparities == 256;
}

template <int K, int LANES, bool STREAM>
static void launch_two_as(const PathArgs &a, int blocks, int count, int prefetch, hipStream_t stream) {
    switch (prefetch) {
        case 2: hipLaunchKernelGGL((path2_kernel<K, 2, LANES, STREAM>), dim3(blocks, count), dim3(TWO_BLOCK), 0, stream, a); break;
        case 8: hipLaunchKernelGGL((path2_kernel<K, 8, LANES, STREAM>), dim3(blocks, count), dim3(TWO_BLOCK), 0, stream, a); break;
        default: hipLaunchKernelGGL((path2_kernel<K, 4, LANES, STREAM>), dim3(blocks, count), dim3(TWO_BLOCK), 0, stream, a); break;
    }
}

template <int K, int LANES, bool STREAM>
static void launch_rc_as(const PathArgs &a, int blocks, int count, int segment, hipStream_t stream) {
    switch (segment) {
        case 2: hipLaunchKernelGGL((path_rc_kernel<K, 2, LANES, STREAM>), dim3(blocks, count), dim3(TWO_BLOCK), 0, stream, a); break;
        case 8: hipLaunchKernelGGL((path_rc_kernel<K, 8, LANES, STREAM>), dim3(blocks, count), dim3(TWO_BLOCK), 0, stream, a); break;
        default: hipLaunchKernelGGL((path_rc_kernel<K, 4, LANES, STREAM>), dim3(blocks, count), dim3(TWO_BLOCK), 0, stream, a); break;
    }
}

// Two waves per line, meeting through memory (meet 0) or by recomputation from checkpoints
// every `prefetch` steps (meet 1).
template <int K, int LANES>
static void launch_two(const PathArgs &a, int blocks, int count, int prefetch, int meet,
                       bool stream_volumes, hipStream_t stream) {
    if (meet == 1) {
        if (stream_volumes)
            launch_rc_as<K, LANES, true>(a, blocks, count, prefetch, stream);
        else
            launch_rc_as<K, LANES, false>(a, blocks, count, prefetch, stream);
    } else if (stream_volumes) {
        launch_two_as<K, LANES, true>(a, blocks, count, prefetch, stream);
    } else {
        launch_two_as<K, LANES, false>(a, blocks, count, prefetch, stream);
    }
}

template <int K>
static hipError_t launch_paths_and_select(const MatchParams &p, const PairDesc &single,
                                          const PairDesc *table, int count, const Schedule &sch,
                                          hipStream_t stream) {
    const int W = p.width, H = p.height;
    const int n = orientations_of(p.paths);
    const int lines[4] = {H, W, W + H - 1, W + H - 1};
    PathArgs a;
    a.single = single;
    a.table = table;
    a.W = W;
    a.H = H;
    a.backward_columns = p.paths == 8;
    a.longest_first = 0;
    // Two waves per line at 8 paths, and at 3 paths for the rows (the down-only columns then
    // take four lines a wave, eight a block); 4 or 8 waves (slices) at 8 paths only.
    const bool two = sch.waves >= 2;
    const int slices = two && p.paths == 8 ? slices_fit(16 * K, sch.waves / 2) : 1;
    // Two waves per line take 16 lanes a line, or 32 where D / 32 is 2, 4 or 8.
    const int two_lanes = two && slices == 1 && sch.lanes == 32 && two_lanes_fit(16 * K) ? 32 : 16;
    const int per_block = two ? 64 / two_lanes
                              : PATH_BLOCK / (lanes_fit(16 * K, sch.lanes) ? sch.lanes : 16);
    // Volumes the 256 MB Infinity Cache cannot hold are streamed where read for the last time.
    const bool stream_volumes = (size_t)n * W * H * 16 * K * count > ((size_t)256 << 20);
    auto launch = [&](int first, int last) {
        a.first = first;
        a.last = last;
        // The recomputing kernel holds about two waves a SIMD at 256 disparities, so a large
        // pair's blocks do not all fit at once, and the diagonals' longest lines would start
        // last: 1920x1080x256 1426 -> 1203 us, 1280x720x192 518 -> 451 us longest first.
        // Where every block fits at once (728x544x128) the order changes nothing.
        a.longest_first = two && slices == 1 && sch.meet == 1 && p.paths == 8 && first == 0 &&
                          last == 3;
        int blocks = 0;
        for (int o = 0; o < 4; o++) {
            a.block_start[o] = blocks;
            const int lines_per_block = two && o == 1 && !a.backward_columns ? 2 * per_block : per_block;
            if (o >= first && o <= last) blocks += (lines[o] + lines_per_block - 1) / lines_per_block;
        }
        if (two && slices > 1) {
            launch_sliced<16 * K>(a, blocks, count, sch.prefetch, slices, stream);
        } else if (two && two_lanes == 32) {
            if constexpr (K % 2 == 0 && (K / 2 == 2 || K / 2 == 4 || K / 2 == 8))
                launch_two<K / 2, 32>(a, blocks, count, sch.prefetch, sch.meet, stream_volumes, stream);
        } else if (two) {
            launch_two<K, 16>(a, blocks, count, sch.prefetch, sch.meet, stream_volumes, stream);
        } else {
            launch_paths_lanes<16 * K>(a, blocks, count, sch, stream);
        }
    };
    int slot = profile_begin(PK_PATH, stream);
    if (sch.split) {
        for (int o = 0; o < n; o++) launch(o, o);
    } else {
        launch(0, n - 1);
    }
    profile_end(slot, stream);
    slot = profile_begin(PK_SELECT, stream);
    // 16 disparities a select lane where D allows (K = D / 16 lanes a pixel), else 12.
    constexpr int SK = K == 12 ? 12 : 16, SPL = 16 * K / SK;
    const int select_pixels = 256 / SPL * SELECT_GROUPS;
    const dim3 select_grid((W + select_pixels - 1) / select_pixels, H, count);
    if (stream_volumes)
        hipLaunchKernelGGL((select_kernel<SK, SPL, true>), select_grid, dim3(256), 0, stream,
                           single, table, n, W);
    else
        hipLaunchKernelGGL((select_kernel<SK, SPL, false>), select_grid, dim3(256), 0, stream,
                           single, table, n, W);
    profile_end(slot, stream);
    return hipGetLastError();
}

template <int COST_BLOCK>
static void launch_census_cost(const PairDesc &single, const PairDesc *table, int W, int H, int D,
                               int count, int keep_census, hipStream_t stream) {
    const dim3 grid(1, H, count), block(COST_BLOCK);
    const size_t lds = census_cost_lds(W, D);
    switch (D) {
        case 64: hipLaunchKernelGGL((census_cost_kernel<COST_BLOCK, 64>), grid, block, lds, stream, single, table, W, H, keep_census); break;
        case 128: hipLaunchKernelGGL((census_cost_kernel<COST_BLOCK, 128>), grid, block, lds, stream, single, table, W, H, keep_census); break;
        case 192: hipLaunchKernelGGL((census_cost_kernel<COST_BLOCK, 192>), grid, block, lds, stream, single, table, W, H, keep_census); break;
        case 256: hipLaunchKernelGGL((census_cost_kernel<COST_BLOCK, 256>), grid, block, lds, stream, single, table, W, H, keep_census); break;
        default: break;  // refused by the caller
    }
}

// The device's CUs (256 on MI355X), read once.
static int compute_unit_count() {
    static const int compute_units = [] {
        int device = 0, n = 0;
        if (hipGetDevice(&device) != hipSuccess ||
            hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
            n = 0;
        return n > 0 ? n : 256;
    }();
    return compute_units;
}

static hipError_t launch_all(const MatchParams &p, const PairDesc &single, const PairDesc *table,
                             int count, const Schedule &sch, hipStream_t stream) {
    const int W = p.width, H = p.height, D = p.disparities;
    if (W > MAX_WIDTH || W < 1 || H < 1 || count < 1) return hipErrorInvalidValue;
    // Above 64 KB of LDS a kernel must say so first (up to the CU's 160 KB).
    static const hipError_t lds_ok = [] {
        const void *kernels[] = {
            reinterpret_cast<const void *>(census_cost_kernel<512, 64>),
            reinterpret_cast<const void *>(census_cost_kernel<512, 128>),
            reinterpret_cast<const void *>(census_cost_kernel<512, 192>),
            reinterpret_cast<const void *>(census_cost_kernel<512, 256>),
            reinterpret_cast<const void *>(census_cost_kernel<1024, 64>),
            reinterpret_cast<const void *>(census_cost_kernel<1024, 128>),
            reinterpret_cast<const void *>(census_cost_kernel<1024, 192>),
            reinterpret_cast<const void *>(census_cost_kernel<1024, 256>)};
        hipError_t e = hipSuccess;
        for (const void *k : kernels)
            if (e == hipSuccess)
                e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize,
                                        (int)census_cost_lds(MAX_WIDTH, MAX_D));
        return e;
    }();
    if (lds_ok != hipSuccess) return lds_ok;
    // 512-thread blocks when every row of the launch runs at once (three blocks a CU at the
    // kernel's 74 VGPRs): the cost volume is then written in one sweep, and at 728x544x128,
    // whose volumes fill the 256 MB Infinity Cache, the path kernel after it runs 8% faster
    // (180 -> 165 us). Otherwise 1024 (at 1920x1080x256 either takes about 150 us: the
    // rows' blocks run in rounds, each round's stores draining before the next starts).
    const bool one_sweep = (size_t)H * count <= 3 * (size_t)compute_unit_count();
    int slot = profile_begin(PK_COST, stream);
    if (one_sweep)
        launch_census_cost<512>(single, table, W, H, D, count, p.keep_stages ? 1 : 0, stream);
    else
        launch_census_cost<1024>(single, table, W, H, D, count, p.keep_stages ? 1 : 0, stream);
    profile_end(slot, stream);
    hipError_t e;
    switch (D) {
        case 64: e = launch_paths_and_select<4>(p, single, table, count, sch, stream); break;
        case 128: e = launch_paths_and_select<8>(p, single, table, count, sch, stream); break;
        case 192: e = launch_paths_and_select<12>(p, single, table, count, sch, stream); break;
        case 256: e = launch_paths_and_select<16>(p, single, table, count, sch, stream); break;
        default: return hipErrorInvalidValue;
    }
    if (e != hipSuccess) return e;
    static const hipError_t check_lds_ok = hipFuncSetAttribute(
        reinterpret_cast<const void *>(check_median_kernel),
        hipFuncAttributeMaxDynamicSharedMemorySize, (int)check_median_lds(MAX_WIDTH));
    if (check_lds_ok != hipSuccess) return check_lds_ok;
    slot = profile_begin(PK_CHECK_MEDIAN, stream);
    hipLaunchKernelGGL(check_median_kernel, dim3(H, count), dim3(CHECK_BLOCK), check_median_lds(W), stream,
                       single, table, W, H, p.keep_stages ? 1 : 0, p.positive ? 1 : 0);
    profile_end(slot, stream);
    return hipGetLastError();
}

hipError_t launch_check_median(const float *selected, const uint32_t *keys, float *checked,
                               float *out, int W, int H, bool keep_checked, bool positive,
                               hipStream_t stream) {
    if (W > MAX_WIDTH || W < 1 || H < 1) return hipErrorInvalidValue;
    static const hipError_t check_lds_ok = hipFuncSetAttribute(
        reinterpret_cast<const void *>(check_median_kernel),
        hipFuncAttributeMaxDynamicSharedMemorySize, (int)check_median_lds(MAX_WIDTH));
    if (check_lds_ok != hipSuccess) return check_lds_ok;
    PairDesc single{};
    single.disparity = out;
    single.b.selected = const_cast<float *>(selected);
    single.b.keys = const_cast<uint32_t *>(keys);
    single.b.checked = checked;
    const int slot = profile_begin(PK_CHECK_MEDIAN, stream);
    hipLaunchKernelGGL(check_median_kernel, dim3(H, 1), dim3(CHECK_BLOCK), check_median_lds(W),
                       stream, single, nullptr, W, H, keep_checked ? 1 : 0, positive ? 1 : 0);
    profile_end(slot, stream);
    return hipGetLastError();
}

hipError_t launch_match(const MatchParams &p, const uint8_t *left, const uint8_t *right,
                        MatchBuffers &b, float *disparity, uint16_t *sums, hipStream_t stream) {
    PairDesc single;
    single.left = left;
    single.right = right;
    single.disparity = disparity;
    single.sums = sums;
    single.b = b;
    single.p1 = p.p1;
    single.p2 = p.p2;
    single.uniqueness = p.uniqueness;
    return launch_all(p, single, nullptr, 1, schedule_for(p.width, p.height, p.disparities, p.paths),
                      stream);
}

hipError_t launch_match_batch(const MatchParams &p, const PairDesc *table, int count,
                              const Schedule &schedule, hipStream_t stream) {
    PairDesc unused{};
    return launch_all(p, unused, table, count, schedule, stream);
}

// ------------------------------------------------------------------------------------------
// The batched form's fused launches: every size class of a batch in one launch per kernel.
// The grid of each launch is the classes' blocks one after the other (the class with the
// most work a pair first), each class its pairs in table order, each pair its blocks as the
// class's own launch orders them. A block finds its class from the classes' first blocks,
// then runs the very body of the per-class kernel at the class's disparity count: results
// are the per-class launches' bit for bit. One kernel holds every body up to the batch's
// largest disparity count, so its registers are that body's: the classes share one
// occupancy, the largest class's.

struct MixedClass {
    int W, H, D, orientations;
    int first;     // the class's first table entry
    int start[4];  // its first block in the census, path, select and check launches
    int path_blocks, select_x;  // a pair's blocks in the path launch; select blocks a row
    PathArgs path;  // one pair's path launch geometry (its pair fields unused)
};

struct MixedArgs {
    const PairDesc *table;
    int classes;
    int keep_stages, positive;
    MixedClass c[MAX_MIXED_CLASSES];
};
static_assert(sizeof(MixedArgs) <= 4096, "a kernel's arguments take at most 4 KB");

enum { MIX_COST = 0, MIX_PATH = 1, MIX_SELECT = 2, MIX_CHECK = 3 };

// The class of block b of launch `which`: the last whose first block is at or before b.
__device__ __forceinline__ int mixed_class(const MixedArgs &m, int which, int b) {
    int k = 0;
    for (int c = 1; c < m.classes; c++) k = b >= m.c[c].start[which] ? c : k;
    return k;
}

template <int COST_BLOCK>
__global__ __launch_bounds__(COST_BLOCK) void census_cost_mixed_kernel(MixedArgs m) {
    extern __shared__ __align__(16) uint8_t lds[];
    const MixedClass &c = m.c[mixed_class(m, MIX_COST, blockIdx.x)];
    const int local = blockIdx.x - c.start[MIX_COST];
    const int pair = local / c.H, y = local - pair * c.H;
    const PairDesc pd = m.table[c.first + pair];
    switch (c.D) {
        case 64: census_cost_row<COST_BLOCK, 64>(pd, c.W, c.H, y, m.keep_stages, lds); break;
        case 128: census_cost_row<COST_BLOCK, 128>(pd, c.W, c.H, y, m.keep_stages, lds); break;
        case 192: census_cost_row<COST_BLOCK, 192>(pd, c.W, c.H, y, m.keep_stages, lds); break;
        default: census_cost_row<COST_BLOCK, 256>(pd, c.W, c.H, y, m.keep_stages, lds); break;
    }
}

// Two waves of 16 lanes a line meeting by recomputation (path_rc_kernel), disparity counts up
// to 16 KMAX.
// Item `item` of the fused path launch: the block of that index in the launch's grid.
template <int KMAX, int S, bool STREAM>
__device__ __forceinline__ void mixed_path_item(const MixedArgs &m, int item) {
    const MixedClass &c = m.c[mixed_class(m, MIX_PATH, item)];
    const int local = item - c.start[MIX_PATH];
    const int pair = local / c.path_blocks, b = local - pair * c.path_blocks;
    const PathArgs a = c.path;  // in scalar registers: the class index is a run-time one
    const int index = c.first + pair;
    switch (c.D) {
        case 64: path_rc_block<4, S, 16, STREAM>(a, m.table, index, b); break;
        case 128: if constexpr (KMAX >= 8) path_rc_block<8, S, 16, STREAM>(a, m.table, index, b); break;
        case 192: if constexpr (KMAX >= 12) path_rc_block<12, S, 16, STREAM>(a, m.table, index, b); break;
        default: if constexpr (KMAX >= 16) path_rc_block<16, S, 16, STREAM>(a, m.table, index, b); break;
    }
}

template <int KMAX, int S, bool STREAM>
__global__ __launch_bounds__(TWO_BLOCK) void path_rc_mixed_kernel(MixedArgs m) {
    mixed_path_item<KMAX, S, STREAM>(m, blockIdx.x);
}

// The same items with persistent blocks: as many blocks as the device holds at once, each
// taking the next item (in the grid's order) from queue[0] until none is left. An item's
// two waves meet at most once, at one s_barrier both reach (or neither: a block whose lines
// are all empty), so the item loop's own barriers stay in step. The last block to leave
// (queue[1] counts them) sets both counters back to zero for the next launch.
template <int KMAX, int S, bool STREAM>
__global__ __launch_bounds__(TWO_BLOCK) void path_rc_queue_kernel(MixedArgs m, uint32_t *queue,
                                                                  int items) {
    __shared__ int next;
    for (;;) {
        if (threadIdx.x == 0) next = (int)atomicAdd(&queue[0], 1u);
        __syncthreads();
        const int item = next;
        __syncthreads();  // every thread has read `next` before it is written again
        if (item >= items) break;
        mixed_path_item<KMAX, S, STREAM>(m, item);
    }
    if (threadIdx.x == 0 && atomicAdd(&queue[1], 1u) == gridDim.x - 1) {
        atomicExch(&queue[0], 0u);
        atomicExch(&queue[1], 0u);
    }
}

// As select_kernel at each class's disparity count (the lanes a pixel launch_paths_and_select
// gives it).
template <int KMAX, bool STREAM>
__global__ __launch_bounds__(256) void select_mixed_kernel(MixedArgs m) {
    const MixedClass &c = m.c[mixed_class(m, MIX_SELECT, blockIdx.x)];
    const int local = blockIdx.x - c.start[MIX_SELECT];
    const int per_pair = c.select_x * c.H;
    const int pair = local / per_pair, r = local - pair * per_pair;
    const int y = r / c.select_x, bx = r - y * c.select_x;
    const PairDesc pd = m.table[c.first + pair];
    switch (c.D) {
        case 64: select_block<16, 4, STREAM>(pd, c.orientations, c.W, c.H, bx, y); break;
        case 128: if constexpr (KMAX >= 8) select_block<16, 8, STREAM>(pd, c.orientations, c.W, c.H, bx, y); break;
        case 192: if constexpr (KMAX >= 12) select_block<12, 16, STREAM>(pd, c.orientations, c.W, c.H, bx, y); break;
        default: if constexpr (KMAX >= 16) select_block<16, 16, STREAM>(pd, c.orientations, c.W, c.H, bx, y); break;
    }
}

__global__ __launch_bounds__(CHECK_BLOCK) void check_median_mixed_kernel(MixedArgs m) {
    extern __shared__ __align__(16) uint32_t check_lds[];
    const MixedClass &c = m.c[mixed_class(m, MIX_CHECK, blockIdx.x)];
    const int local = blockIdx.x - c.start[MIX_CHECK];
    const int pair = local / c.H, y = local - pair * c.H;
    check_median_row(m.table[c.first + pair], c.W, c.H, y, m.keep_stages, m.positive, check_lds);
}

template <int KMAX, bool STREAM>
static void launch_mixed_paths(const MixedArgs &m, int blocks, int segment, uint32_t *queue,
                               hipStream_t stream) {
    if (queue) {
        // Persistent blocks: what the device holds of this kernel at once.
        const void *kernel = segment == 4
                                 ? reinterpret_cast<const void *>(path_rc_queue_kernel<KMAX, 4, STREAM>)
                                 : reinterpret_cast<const void *>(path_rc_queue_kernel<KMAX, 8, STREAM>);
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, TWO_BLOCK, 0) != hipSuccess ||
            per_cu < 1)
            per_cu = 1;
        const int grid = std::min(blocks, per_cu * compute_unit_count());
        if (segment == 4)
            hipLaunchKernelGGL((path_rc_queue_kernel<KMAX, 4, STREAM>), dim3(grid), dim3(TWO_BLOCK), 0, stream, m, queue, blocks);
        else
            hipLaunchKernelGGL((path_rc_queue_kernel<KMAX, 8, STREAM>), dim3(grid), dim3(TWO_BLOCK), 0, stream, m, queue, blocks);
        return;
    }
    if (segment == 4)
        hipLaunchKernelGGL((path_rc_mixed_kernel<KMAX, 4, STREAM>), dim3(blocks), dim3(TWO_BLOCK), 0, stream, m);
    else
        hipLaunchKernelGGL((path_rc_mixed_kernel<KMAX, 8, STREAM>), dim3(blocks), dim3(TWO_BLOCK), 0, stream, m);
}

template <int KMAX>
static void launch_mixed_aggregation(const MixedArgs &m, int path_blocks, int select_blocks,
                                     int segment, uint32_t *queue, bool stream_volumes,
                                     hipStream_t stream) {
    int slot = profile_begin(PK_PATH, stream);
    if (stream_volumes)
        launch_mixed_paths<KMAX, true>(m, path_blocks, segment, queue, stream);
    else
        launch_mixed_paths<KMAX, false>(m, path_blocks, segment, queue, stream);
    profile_end(slot, stream);
    slot = profile_begin(PK_SELECT, stream);
    if (stream_volumes)
        hipLaunchKernelGGL((select_mixed_kernel<KMAX, true>), dim3(select_blocks), dim3(256), 0, stream, m);
    else
        hipLaunchKernelGGL((select_mixed_kernel<KMAX, false>), dim3(select_blocks), dim3(256), 0, stream, m);
    profile_end(slot, stream);
}

// One pair's blocks in the fused path launch, and its select blocks a row: as
// launch_paths_and_select lays out a pair with two waves of 16 lanes a line, four lines a
// block (eight down-only columns at 3 paths), and select with 16 disparities a lane (12 at
// 192).
static void mixed_pair_geometry(const MatchParams &p, int block_start[4], int *path_blocks,
                                int *select_x) {
    const int W = p.width, H = p.height;
    const int lines[4] = {H, W, W + H - 1, W + H - 1};
    int blocks = 0;
    for (int o = 0; o < 4; o++) {
        block_start[o] = blocks;
        const int per_block = o == 1 && p.paths != 8 ? 8 : 4;
        if (o < orientations_of(p.paths)) blocks += (lines[o] + per_block - 1) / per_block;
    }
    *path_blocks = blocks;
    const int K = p.disparities / 16;
    const int spl = K == 12 ? 16 : K;
    *select_x = (W + 256 / spl * SELECT_GROUPS - 1) / (256 / spl * SELECT_GROUPS);
}

bool mixed_launch_fits(const MixedMember *members, int n) {
    if (n < 1 || n > MAX_MIXED_CLASSES) return false;
    // Each launch's grid is one dimension: its threads stay under 2^31.
    const int64_t threads[4] = {1024, TWO_BLOCK, 256, CHECK_BLOCK};
    int64_t blocks[4] = {0, 0, 0, 0};
    for (int i = 0; i < n; i++) {
        const MatchParams &p = members[i].p;
        if (p.width < 1 || p.width > MAX_WIDTH || p.height < 1 || members[i].count < 1 ||
            (p.disparities != 64 && p.disparities != 128 && p.disparities != 192 &&
             p.disparities != 256))
            return false;
        int block_start[4], path_blocks, select_x;
        mixed_pair_geometry(p, block_start, &path_blocks, &select_x);
        const int64_t per_pair[4] = {p.height, path_blocks, (int64_t)select_x * p.height,
                                     p.height};
        for (int w = 0; w < 4; w++) blocks[w] += per_pair[w] * members[i].count;
    }
    for (int w = 0; w < 4; w++)
        if (blocks[w] * threads[w] > INT32_MAX) return false;
    return true;
}

hipError_t launch_match_mixed(const MixedMember *members, int n, const PairDesc *table,
                              int segment, uint32_t *queue, hipStream_t stream) {
    if (!table || (segment != 4 && segment != 8) || !mixed_launch_fits(members, n))
        return hipErrorInvalidValue;
    // The classes with the most work a pair first: their long lines start first.
    int order[MAX_MIXED_CLASSES];
    for (int i = 0; i < n; i++) order[i] = i;
    auto work = [&](int i) {
        const MatchParams &p = members[i].p;
        return (double)p.width * p.height * p.disparities;
    };
    std::stable_sort(order, order + n, [&](int a, int b) { return work(a) > work(b); });
    MixedArgs m{};
    m.table = table;
    m.classes = n;
    m.keep_stages = members[0].p.keep_stages ? 1 : 0;
    m.positive = members[0].p.positive ? 1 : 0;
    int starts[4] = {0, 0, 0, 0};
    int kmax = 0, rows = 0;
    size_t volumes = 0, cost_lds = 0, check_lds = 0;
    for (int k = 0; k < n; k++) {
        const MixedMember &mm = members[order[k]];
        const MatchParams &p = mm.p;
        MixedClass &c = m.c[k];
        const int W = p.width, H = p.height;
        c.W = W;
        c.H = H;
        c.D = p.disparities;
        c.orientations = orientations_of(p.paths);
        c.first = mm.first;
        PathArgs &a = c.path;
        a.table = nullptr;
        a.W = W;
        a.H = H;
        a.backward_columns = p.paths == 8;
        a.first = 0;
        a.last = c.orientations - 1;
        a.longest_first = p.paths == 8;
        mixed_pair_geometry(p, a.block_start, &c.path_blocks, &c.select_x);
        const int K = p.disparities / 16;
        const int per_pair[4] = {H, c.path_blocks, c.select_x * H, H};
        for (int w = 0; w < 4; w++) {
            c.start[w] = starts[w];
            starts[w] += per_pair[w] * mm.count;
        }
        kmax = std::max(kmax, K);
        rows += H * mm.count;
        volumes += (size_t)c.orientations * W * H * p.disparities * mm.count;
        cost_lds = std::max(cost_lds, census_cost_lds(W, p.disparities));
        check_lds = std::max(check_lds, check_median_lds(W));
    }
    static const hipError_t lds_ok = [] {
        const void *kernels[] = {reinterpret_cast<const void *>(census_cost_mixed_kernel<512>),
                                 reinterpret_cast<const void *>(census_cost_mixed_kernel<1024>)};
        hipError_t e = hipSuccess;
        for (const void *k : kernels)
            if (e == hipSuccess)
                e = hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize,
                                        (int)census_cost_lds(MAX_WIDTH, MAX_D));
        if (e == hipSuccess)
            e = hipFuncSetAttribute(reinterpret_cast<const void *>(check_median_mixed_kernel),
                                    hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)check_median_lds(MAX_WIDTH));
        return e;
    }();
    if (lds_ok != hipSuccess) return lds_ok;
    int slot = profile_begin(PK_COST, stream);
    if (rows <= 3 * compute_unit_count())
        hipLaunchKernelGGL(census_cost_mixed_kernel<512>, dim3(starts[MIX_COST]), dim3(512),
                           cost_lds, stream, m);
    else
        hipLaunchKernelGGL(census_cost_mixed_kernel<1024>, dim3(starts[MIX_COST]), dim3(1024),
                           cost_lds, stream, m);
    profile_end(slot, stream);
    const bool stream_volumes = volumes > ((size_t)256 << 20);
    switch (kmax) {
        case 4: launch_mixed_aggregation<4>(m, starts[MIX_PATH], starts[MIX_SELECT], segment, queue, stream_volumes, stream); break;
        case 8: launch_mixed_aggregation<8>(m, starts[MIX_PATH], starts[MIX_SELECT], segment, queue, stream_volumes, stream); break;
        case 12: launch_mixed_aggregation<12>(m, starts[MIX_PATH], starts[MIX_SELECT], segment, queue, stream_volumes, stream); break;
        default: launch_mixed_aggregation<16>(m, starts[MIX_PATH], starts[MIX_SELECT], segment, queue, stream_volumes, stream); break;
    }
    slot = profile_begin(PK_CHECK_MEDIAN, stream);
    hipLaunchKernelGGL(check_median_mixed_kernel, dim3(starts[MIX_CHECK]), dim3(CHECK_BLOCK),
                       check_lds, stream, m);
    profile_end(slot, stream);
    return hipGetLastError();
}

}  // namespace nxs
