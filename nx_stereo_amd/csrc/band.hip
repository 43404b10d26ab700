// Band search and coarse-to-fine matching on gfx950: NxStereo.sgm_band/4 and
// sgm_coarse_to_fine/3 (lib/nx_stereo.ex:164-471).
//
// Coarse-to-fine shrinks both images to half size (rounded 2x2 average), matches the
// halves over 64 disparities with the full-search kernels, and derives each pixel's band
// base from that match: twice the lowest match of its coarse 3x3 neighbourhood, less a
// margin. The fine pass then searches only `band` disparities per pixel, [base, base+band).
//
// Layout in HBM: base s32 {H,W}; band volumes pixel-major, volume[(y*W + x)*band + k], so
// one pixel's band is contiguous; one u8 path volume per aggregation direction.
//
// Aggregation: each line of a direction is owned by `band` consecutive lanes, lane k holding
// candidate base + k in a u32. A path moving to a pixel with another base moves its values
// by the difference (a DPP move for one, ds_bpermute within the line's lanes beyond); a
// candidate the previous band lacks is reached only by the p2 jump (or, one past its ends,
// from the end + p1). Path values stay <= p2 + 48 <= 127, so every direction stores bytes
// and selection adds them.
#include <type_traits>

#include "device.hpp"
#include "kernels.hpp"
#include "profile.hpp"

namespace nxs {

// Bases above this act as this: every candidate lies past the image's left edge then (W is
// at most 4096), and base + k stays far from overflow.
constexpr int BASE_CEILING = 1 << 16;

__device__ __forceinline__ int searched_of(int base) { return min(max(base, 0), BASE_CEILING); }

// ------------------------------------------------------------------------------------------
// nx_stereo.ex:236-247 (shrink): half[y][x] = (sum of the 2x2 block + 2) / 4. Grid y: the
// image (0 left, 1 right).

__global__ __launch_bounds__(256) void shrink_kernel(const uint8_t *left, const uint8_t *right,
                                                     int W, int H, uint8_t *half_left,
                                                     uint8_t *half_right) {
    const int cw = W / 2, ch = H / 2;
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= cw * ch) return;
    const int y = i / cw, x = i - y * cw;
    const uint8_t *top = (blockIdx.y ? right : left) + (size_t)(2 * y) * W + 2 * x;
    const uint32_t sum = top[0] + top[1] + top[W] + top[W + 1];
    (blockIdx.y ? half_right : half_left)[i] = (uint8_t)((sum + 2) >> 2);
}

// ------------------------------------------------------------------------------------------
// nx_stereo.ex:249-259 (band_base): the lowest match (>= 0) of the coarse 3x3
// neighbourhood, none outside the map; base = clip(floor(2 lowest) - margin, 0,
// 128 - band), -1 without a match, over the coarse pixel's 2x2 block.

__global__ __launch_bounds__(256) void band_base_kernel(const float *coarse, int cw, int ch,
                                                        int band, int margin, int32_t *base) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= cw * ch) return;
    const int y = i / cw, x = i - y * cw;
    const float none = 1.0e9f;
    float lowest = none;
#pragma unroll
    for (int dy = -1; dy <= 1; dy++)
#pragma unroll
        for (int dx = -1; dx <= 1; dx++) {
            const int sy = y + dy, sx = x + dx;
            if (sy < 0 || sy >= ch || sx < 0 || sx >= cw) continue;
            const float v = coarse[sy * cw + sx];
            lowest = fminf(lowest, v >= 0.0f ? v : none);
        }
    int b = -1;
    if (lowest < none) b = min(max((int)(floorf(lowest * 2.0f) - (float)margin), 0), 128 - band);
    const int W = 2 * cw;
    int2 *row0 = reinterpret_cast<int2 *>(base + (size_t)(2 * y) * W + 2 * x);
    int2 *row1 = reinterpret_cast<int2 *>(base + (size_t)(2 * y + 1) * W + 2 * x);
    *row0 = make_int2(b, b);
    *row1 = make_int2(b, b);
}

// ------------------------------------------------------------------------------------------
// Band census and costs (nx_stereo.ex:324-344, 490-525), one block per image row. The seven
// rows around row y of both images come into LDS (clamped), the block makes row y's census
// of both there, then C[y,x,k] = popcount(FL[y,x] ^ FR[y, x - base - k]), FR = 0 left of the
// image. A thread makes 16 bytes of one pixel (8 with a band of 8): consecutive threads
// store consecutive runs.

constexpr int BAND_COST_BLOCK = 512;

// The image rows in LDS carry CENSUS_REACH replicated pixels on each side, so the census
// reads its 9 columns without clamping each one, padded as the full search's census rows
// (census_row_pitch), so that sparse_9x7 takes its four-pixel packed form (census_quad).

// LDS bytes: the two census rows (a whole number of four-pixel groups each), then 7 padded
// rows of each image.
__host__ __device__ constexpr size_t band_cost_lds(int W) {
    return 16 * (size_t)((W + 3) & ~3) + 14 * (size_t)census_row_pitch(W);
}

// Census of the pixel at `img`, in the middle of 7 padded rows `pitch` apart. WINDOW 0:
// sparse_9x7 (dy in -3..3, dx in -4..4, both non-zero, 48 bits); 1: dense_9x5 (dy in
// -2..2, dx in -4..4 but the centre, 44 bits); 2: dense_9x5 against the blended anchor
// (1,2,1)x(1,2,1) over the 3x3 plus 16 centre plus 24, over 32. Bit order row-major over
// the window.
template <int WINDOW>
__device__ __forceinline__ uint64_t census_at(const uint8_t *img, int pitch) {
    const uint8_t *mid = img + 3 * pitch;
    int anchor = mid[0];
    if constexpr (WINDOW == 2) {
        int sum = 0;
#pragma unroll
        for (int dy = -1; dy <= 1; dy++)
#pragma unroll
            for (int dx = -1; dx <= 1; dx++)
                sum += (dy ? 1 : 2) * (dx ? 1 : 2) * mid[dy * pitch + dx];
        anchor = (sum + 16 * anchor + 24) >> 5;
    }
    constexpr int REACH = WINDOW == 0 ? 3 : 2;
    uint64_t f = 0;
    int bit = 0;
#pragma unroll
    for (int dy = -REACH; dy <= REACH; dy++) {
        if (WINDOW == 0 && dy == 0) continue;
        const uint8_t *row = img + (3 + dy) * pitch;
#pragma unroll
        for (int dx = -4; dx <= 4; dx++) {
            if (dx == 0 && (WINDOW == 0 || dy == 0)) continue;
            f |= (uint64_t)((int)row[dx] + 1 < anchor) << bit;
            bit++;
        }
    }
    return f;
}

// Grid x: the row.
template <int WINDOW, int B>
__global__ __launch_bounds__(BAND_COST_BLOCK) void band_cost_kernel(
    const uint8_t *left, const uint8_t *right, const int32_t *base, int W, int H, uint8_t *costs,
    uint64_t *census_left, uint64_t *census_right) {
    extern __shared__ __align__(16) uint8_t band_lds[];
    const int W4 = (W + 3) & ~3;
    uint64_t *cl = reinterpret_cast<uint64_t *>(band_lds);  // [W4]
    uint64_t *cr = cl + W4;                                   // [W4]
    const int pitch = census_row_pitch(W);
    uint8_t *rows = reinterpret_cast<uint8_t *>(cr + W4);     // [2][7][pitch]
    const int y = blockIdx.x;
    // A thread loads its byte of all 14 rows before it stores any, so that the loads are in
    // flight together (a thread past the padded row loads a byte it does not store); the
    // rows' offsets are scalars, the column the one vector offset.
    const Rsrc images[2] = {buffer_of(left, (uint32_t)W * H), buffer_of(right, (uint32_t)W * H)};
    for (int i0 = 0; i0 < pitch; i0 += BAND_COST_BLOCK) {
        const int i = i0 + (int)threadIdx.x;
        const int column = min(max(i - CENSUS_REACH, 0), W - 1);
        uint32_t v[14];
#pragma unroll
        for (int k = 0; k < 14; k++)
            v[k] = __builtin_amdgcn_raw_buffer_load_b8(images[k / 7], column,
                                                       min(max(y + k % 7 - 3, 0), H - 1) * W, 0);
        if (i < pitch) {
#pragma unroll
            for (int k = 0; k < 14; k++) rows[k * pitch + i] = (uint8_t)v[k];
        }
    }
    __syncthreads();
    if constexpr (WINDOW == 0) {
        // Four pixels a thread in packed compares; pixels past the row's end land in the
        // groups' spare slots, which nothing reads.
        const int quads = W4 >> 2;
        for (int i = threadIdx.x; i < 2 * quads; i += BAND_COST_BLOCK) {
            const int side = i >= quads ? 1 : 0, x0 = 4 * (i - side * quads);
            const uint32_t *img = reinterpret_cast<const uint32_t *>(rows + side * 7 * pitch + x0);
            const CensusQuad q = census_quad(img, pitch >> 2);
#pragma unroll
            for (int k = 0; k < 4; k++) {
                (side ? cr : cl)[x0 + k] = (uint64_t)q.lo[k] | (uint64_t)q.hi[k] << 32;
                if (census_left && x0 + k < W)
                    (side ? census_right : census_left)[(size_t)y * W + x0 + k] =
                        census_bits_in_order(q.lo[k], q.hi[k]);
            }
        }
    } else {
        for (int i = threadIdx.x; i < 2 * W; i += BAND_COST_BLOCK) {
            const int side = i >= W ? 1 : 0, x = i - side * W;
            const uint64_t f = census_at<WINDOW>(rows + side * 7 * pitch + CENSUS_REACH + x, pitch);
            (side ? cr : cl)[x] = f;
            if (census_left) (side ? census_right : census_left)[(size_t)y * W + x] = f;
        }
    }
    __syncthreads();
    constexpr int CHUNKS = B >= 16 ? B / 16 : 1;
    constexpr int WORDS = B >= 16 ? 4 : 2;
    for (int q = threadIdx.x; q < W * CHUNKS; q += BAND_COST_BLOCK) {
        const int x = q / CHUNKS, chunk = q - x * CHUNKS;
        const int first = x - searched_of(base[(size_t)y * W + x]) - chunk * 16;
        const uint64_t l = cl[x];
        uint32_t words[WORDS];
#pragma unroll
        for (int w = 0; w < WORDS; w++) {
            uint32_t word = 0;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int column = first - 4 * w - j;
                const uint64_t r = column >= 0 ? cr[column] : 0;
                word |= (uint32_t)__builtin_popcountll(l ^ r) << (8 * j);
            }
            words[w] = word;
        }
        uint8_t *out = costs + ((size_t)y * W + x) * B + chunk * 16;
        if constexpr (WORDS == 4)
            *reinterpret_cast<uint4 *>(out) = make_uint4(words[0], words[1], words[2], words[3]);
        else
            *reinterpret_cast<uint2 *>(out) = make_uint2(words[0], words[1]);
    }
}

// ------------------------------------------------------------------------------------------
// Band aggregation (nx_stereo.ex:346-419, scan_band), one direction per line. Lane k of a
// line holds L(p, base(p) + k). Written unnormalised like the full search: with v the
// previous pixel's values, m = min(v) and delta = base(p) - base(previous),
//   part[j] = min(v[j], min(v[j-1], v[j+1]) + p1)     (v outside the band: a sentinel)
//   L(p, k) = min(part[k + delta], m + p2) - m + C(p, k)   for 0 <= k + delta < band,
//             min(v[0] + p1, m + p2) - m + C               for k + delta = -1,
//             min(v[band-1] + p1, m + p2) - m + C          for k + delta = band,
//             p2 + C                                       further out,
// which is scan_band's min(window, p2, neighbours + p1) over the moved, normalised values.
// A line entering the image starts from v = 0 in its own band (delta 0). The walk keeps
// u = 1024 - v rather than v (BandLine below), and does the move only on steps where a line
// of the wave changes base.

// Neighbouring lanes of one line and the maximum over its B lanes (B = 8, 16, 32 or 64).
template <int B>
struct BandLanes {
    static __device__ __forceinline__ uint32_t from_left(uint32_t edge, uint32_t v, int lane) {
        const uint32_t t = B <= 16 ? dpp_or<DPP_ROW_SHR1>(edge, v) : dpp_or<DPP_WAVE_SHR1>(edge, v);
        return (B == 8 || B == 32) && lane == 0 ? edge : t;
    }
    static __device__ __forceinline__ uint32_t from_right(uint32_t edge, uint32_t v, int lane) {
        const uint32_t t = B <= 16 ? dpp_or<DPP_ROW_SHL1>(edge, v) : dpp_or<DPP_WAVE_SHL1>(edge, v);
        return (B == 8 || B == 32) && lane == B - 1 ? edge : t;
    }
    // The maximum over the line's lanes: quad_perm [1,0,3,2] and [2,3,0,1] (the quad),
    // row_half_mirror (8 lanes), row_mirror (16 lanes), then rows 2r and 2r+1, then the
    // wave's halves.
    static __device__ __forceinline__ uint32_t max(uint32_t v) {
        v = ::max(v, dpp<0xB1>(v));
        v = ::max(v, dpp<0x4E>(v));
        v = ::max(v, dpp<0x141>(v));
        if constexpr (B >= 16) v = ::max(v, dpp<0x140>(v));
        if constexpr (B >= 32) v = pair_rows_max(v);
        if constexpr (B == 64) v = halves_max(v);
        return v;
    }
};

struct BandPathArgs {
    const uint8_t *costs;
    const int32_t *base;
    uint8_t *paths[8];  // one volume per direction launched
    int W, H, p1, p2;
    int directions;     // 3 or 8
    int block_start[9]; // direction r's first block; block_start[directions] = the grid
};

// Direction r: orientation r >> 1 (rows, columns, x-y diagonals, x+y diagonals), backward
// when r is odd: (0,1) (0,-1) (1,0) (-1,0) (1,1) (-1,-1) (1,-1) (-1,1) in (dy, dx).
__host__ __device__ __forceinline__ int band_orientation(int r) { return r >> 1; }

constexpr int BAND_PATH_BLOCK = 64;

// One line's recurrence in B lanes, on u = K - L (K = 1024 above every path value): the
// minima become maxima, so a DPP move that reads past the band's end and takes 0 is
// neutral, and the neighbour terms fold into DPP-sourced max instructions.
template <int B>
struct BandLine {
    static constexpr uint32_t K = 0x400;
    uint32_t u = K;  // K - the previous pixel's values (0 entering the image)
    int delta = 0;   // the generic walk's base change at the coming step
    bool moving = false;

    // part[k] = min(v[k], min(v[k-1], v[k+1]) + p1) and the line's minimum m come from v
    // alone; the candidate k + delta of the previous band is part moved by delta (a DPP
    // move for one, where the band's end lane takes its own v + p1; ds_bpermute beyond;
    // further out, nothing: p2). Every candidate is >= m, so
    // min(sel, m + p2) - m = min(sel - m, p2). In u: sel - m = u(m) - u(sel).
    // `moving`: a line of the wave changes base here; `delta_of()` its change (called only
    // then).
    template <class DeltaOf>
    __device__ __forceinline__ uint32_t step(uint32_t c, int lane, int group, uint32_t p1,
                                             uint32_t p2, bool moving_here, DeltaOf delta_of) {
        const uint32_t hi = BandLanes<B>::from_right(0u, u, lane);
        const uint32_t near = ::max(BandLanes<B>::from_left(0u, u, lane), hi);
        const uint32_t part = ::max(u, near - p1);  // near >= K - 127 > p1
        const uint32_t m = BandLanes<B>::max(u);
        uint32_t sel = part;
        // Rare (coarse-to-fine's bases change on about 5% of a line's steps): laid out away
        // from the common step, which then falls through without a taken branch.
        if (__builtin_expect(moving_here, false)) {
            const int d = delta_of();
            const uint32_t vp = u - p1;
            const uint32_t up = BandLanes<B>::from_right(vp, part, lane);
            const uint32_t down = BandLanes<B>::from_left(vp, part, lane);
            sel = d == 1 ? up : (d == -1 ? down : part);
            if (__builtin_amdgcn_ballot_w64(d > 1 || d < -1)) {
                const int moved = lane + d;
                const int from = min(max(moved, 0), B - 1);
                const uint32_t f = (uint32_t)__builtin_amdgcn_ds_bpermute(
                    (group * B + from) * 4, (int)(part | (vp << 16)));
                const uint32_t moved_sel =
                    (uint32_t)moved < (uint32_t)B ? (f & 0xFFFFu)
                                                  : (moved == -1 || moved == B ? f >> 16 : 0u);
                sel = d > 1 || d < -1 ? moved_sel : sel;
            }
        }
        const uint32_t e = ::min(m - sel, p2);
        u = (K - c) - e;
        return e + c;  // this pixel's path values
    }

    // The generic walk's change from this step's base b to the next step's (none past the
    // line's end).
    __device__ __forceinline__ void next(int b, int following, bool live) {
        delta = live ? following - b : 0;
        moving = __builtin_amdgcn_ballot_w64(delta != 0) != 0;
    }
};

template <int B, int PREFETCH>
__global__ __launch_bounds__(BAND_PATH_BLOCK) void band_path_kernel(BandPathArgs a) {
    constexpr int G = BAND_PATH_BLOCK / B;  // lines per wave
    const int lane = threadIdx.x % B, group = threadIdx.x / B;
    // The direction, its first block and its volume, by constant indices (a dynamic index
    // into the argument struct would copy it to scratch).
    int r = 0, start = 0;
    uint8_t *volume_of_r = a.paths[0];
#pragma unroll
    for (int i = 1; i < 8; i++)
        if (i < a.directions && (int)blockIdx.x >= a.block_start[i]) {
            r = i;
            start = a.block_start[i];
            volume_of_r = a.paths[i];
        }
    const int o = band_orientation(r);
    const bool backward = r & 1;
    const int W = a.W, H = a.H;
    const int first_line = ((int)blockIdx.x - start) * G;
    const LineGeometry g = line_geometry(o, first_line + group, W, H);
    int longest = 0, shortest = 1 << 30;
#pragma unroll
    for (int i = 0; i < G; i++) {
        const int length = line_geometry(o, first_line + i, W, H).length;
        longest = max(longest, length);
        shortest = min(shortest, length);
    }
    if (longest == 0) return;  // wave-uniform

    const int dy = o == 0 ? 0 : 1;
    const int dx = o == 0 ? 1 : (o == 1 ? 0 : (o == 2 ? 1 : -1));
    const int L = g.length;
    int pixel = g.y0 * W + g.x0, step = dy * W + dx;
    if (backward) {
        pixel += max(L - 1, 0) * step;
        step = -step;
    }
    // Byte offsets (all under 2^31): this lane's cost and path byte, the pixel's base.
    const uint32_t volume = (uint32_t)W * H * B;
    const Rsrc costs = buffer_of(a.costs, volume);
    const Rsrc bases = buffer_of(a.base, (uint32_t)W * H * 4);
    const Rsrc out = buffer_of(volume_of_r, volume);
    const uint32_t p1 = (uint32_t)a.p1, p2 = (uint32_t)a.p2;

    uint32_t cost_ring[PREFETCH];
    {
        uint32_t ahead = (uint32_t)(pixel * B + lane);
#pragma unroll
        for (int p = 0; p < PREFETCH; p++) {
            cost_ring[p] = __builtin_amdgcn_raw_buffer_load_b8(costs, p < L ? ahead : MASKED, 0, 0);
            ahead += step * B;
        }
    }
    BandLine<B> line;

    if constexpr (B >= 16 && PREFETCH == 16) {
        // Blocks of 16 steps. A block loads its bases at once, lane j of a line the base of
        // step t0 + j (its last pixel's past the line's end), and turns their changes into
        // a lane mask: a step tests scalar bits. The bits of step j: lane j of every line.
        constexpr uint64_t LINE_BITS = B == 16 ? 0x0001000100010001ull
                                               : (B == 32 ? 0x0000000100000001ull : 1ull);
        const int j = lane < 16 ? lane : 15;
        int bases_now = 0, previous = 0;
        // Lane j: the change into step t0 + j (none into the first step).
        auto changes_of = [&](int t0) {
            const int before = t0 == 0 ? dpp_or<DPP_ROW_SHR1>(bases_now, bases_now)
                                       : dpp_or<DPP_ROW_SHR1>(dpp<dpp_row_ror(1)>(previous), bases_now);
            return bases_now - before;
        };
        auto delta_at = [&](int changes, int p) {
            return [=] { return __builtin_amdgcn_ds_bpermute((group * B + p) * 4, changes); };
        };
        if (o < 2 && shortest == longest) {
            // Rows and columns where every line of the wave has the same length n: offsets
            // are the lane's at the line's lowest address plus a wave-uniform position.
            // Past the line's end, loads read its last pixel (used by nothing) and stores
            // are skipped.
            const int n = longest;
            const uint32_t low = (uint32_t)(o == 0 ? g.y0 * W : g.x0);
            const uint32_t lane_cost = low * B + lane, lane_base = low * 4;
            const int stride = o == 0 ? 1 : W;
            const int move = backward ? -stride : stride;
            auto position = [&](int t) { return (backward ? n - 1 - t : t) * stride; };
            auto block_bases = [&](int t0) {
                return (int)__builtin_amdgcn_raw_buffer_load_b32(
                    bases, lane_base + (uint32_t)position(min(t0 + j, n - 1)) * 4u, 0, 0);
            };
            bases_now = searched_of(block_bases(0));
            // WHOLE: every load and store of the block lies inside the line.
            auto block = [&](int t0, auto whole) {
                constexpr bool WHOLE = decltype(whole)::value;
                const int bases_next = block_bases(t0 + 16);
                const int changes = changes_of(t0);
                const uint64_t moves = __builtin_amdgcn_ballot_w64(changes != 0 && lane < 16);
                // A block of steps where no line of the wave changes base (most of them) runs
                // without the per-step test.
                auto steps = [&](auto any) __attribute__((always_inline)) {
                    constexpr bool ANY = decltype(any)::value;
                    uint64_t m = moves;
                    int here = position(t0), there = position(min(t0 + PREFETCH, n - 1));
#pragma unroll
                    for (int p = 0; p < 16; p++) {
                        const int t = t0 + p;
                        __builtin_amdgcn_sched_barrier(0);
                        const uint32_t c = cost_ring[p];
                        cost_ring[p] = __builtin_amdgcn_raw_buffer_load_b8(costs, lane_cost, (uint32_t)there * B, 0);
                        const bool moving_here = ANY && (m & LINE_BITS) != 0;
                        m >>= 1;
                        const uint32_t v = line.step(c, lane, group, p1, p2, moving_here, delta_at(changes, p));
                        if (WHOLE || t < n)
                            __builtin_amdgcn_raw_buffer_store_b8((uint8_t)v, out, lane_cost, (uint32_t)here * B, 0);
                        here += move;
                        there = WHOLE ? there + move : position(min(t + 1 + PREFETCH, n - 1));
                    }
                };
                if (moves) steps(std::true_type{});
                else steps(std::false_type{});
                previous = bases_now;
                bases_now = searched_of(bases_next);
            };
            int t0 = 0;
            for (; t0 + 2 * PREFETCH <= n; t0 += 16) block(t0, std::true_type{});
            for (; t0 < n; t0 += 16) block(t0, std::false_type{});
            return;
        }
        // Lines of unequal length (diagonals, an orientation's last wave): per-lane offsets,
        // masked past each line's end.
        auto block_bases = [&](int t0) {
            const int t = min(t0 + j, L - 1);
            return (int)__builtin_amdgcn_raw_buffer_load_b32(
                bases, L > 0 ? (uint32_t)(pixel + t * step) * 4u : MASKED, 0, 0);
        };
        bases_now = searched_of(block_bases(0));
        uint32_t at = (uint32_t)(pixel * B + lane);
        uint32_t ahead = at + (uint32_t)(PREFETCH * step * B);
        for (int t0 = 0; t0 < longest; t0 += 16) {
            const int bases_next = block_bases(t0 + 16);
            const int changes = changes_of(t0);
            const uint64_t moves = __builtin_amdgcn_ballot_w64(changes != 0 && lane < 16);
            auto steps = [&](auto any) __attribute__((always_inline)) {
                constexpr bool ANY = decltype(any)::value;
                uint64_t m = moves;
#pragma unroll
                for (int p = 0; p < 16; p++) {
                    const int t = t0 + p;
                    __builtin_amdgcn_sched_barrier(0);
                    const uint32_t c = cost_ring[p];
                    cost_ring[p] = __builtin_amdgcn_raw_buffer_load_b8(costs, t + PREFETCH < L ? ahead : MASKED, 0, 0);
                    const bool moving_here = ANY && (m & LINE_BITS) != 0;
                    m >>= 1;
                    const uint32_t v = line.step(c, lane, group, p1, p2, moving_here, delta_at(changes, p));
                    __builtin_amdgcn_raw_buffer_store_b8((uint8_t)v, out, t < L ? at : MASKED, 0, 0);
                    at += step * B;
                    ahead += step * B;
                }
            };
            if (moves) steps(std::true_type{});
            else steps(std::false_type{});
            previous = bases_now;
            bases_now = searched_of(bases_next);
        }
        return;
    }

    // A band of 8: every step loads its line's base, per-lane offsets, masked past each
    // line's end.
    int base_ring[PREFETCH];
    uint32_t ahead_base = (uint32_t)pixel * 4u;
#pragma unroll
    for (int p = 0; p < PREFETCH; p++) {
        base_ring[p] = (int)__builtin_amdgcn_raw_buffer_load_b32(bases, p < L ? ahead_base : MASKED, 0, 0);
        ahead_base += step * 4;
    }
    uint32_t at = (uint32_t)(pixel * B + lane);
    uint32_t ahead = at + (uint32_t)(PREFETCH * step * B);
    for (int t0 = 0; t0 < longest; t0 += PREFETCH) {
#pragma unroll
        for (int p = 0; p < PREFETCH; p++) {
            const int t = t0 + p;
            __builtin_amdgcn_sched_barrier(0);
            const bool active = t < L;
            const int b = searched_of(base_ring[p]);
            const uint32_t c = cost_ring[p];
            const bool next = t + PREFETCH < L;
            cost_ring[p] = __builtin_amdgcn_raw_buffer_load_b8(costs, next ? ahead : MASKED, 0, 0);
            base_ring[p] = (int)__builtin_amdgcn_raw_buffer_load_b32(bases, next ? ahead_base : MASKED, 0, 0);
            const uint32_t v = line.step(c, lane, group, p1, p2, line.moving,
                                         [&] { return line.delta; });
            __builtin_amdgcn_raw_buffer_store_b8((uint8_t)v, out, active ? at : MASKED, 0, 0);
            // The next step's base sits in the next slot (slot 0 after the last one,
            // reloaded at this group's first step).
            line.next(b, searched_of(base_ring[(p + 1) % PREFETCH]), t + 1 < L);
            at += step * B;
            ahead += step * B;
            ahead_base += step * 4;
        }
    }
}

// ------------------------------------------------------------------------------------------
// nx_stereo.ex:431-471 (band_select), one pixel per thread: the sums of the directions'
// bytes, winner (ties to the lower k), second best outside winner +-1, subpixel step in
// f32 away from the band's ends, and the key (minimum << 16 | base + winner). A negative
// base or a winner past x gives no match.

struct BandSelectArgs {
    const uint8_t *paths[8];
    int directions;
    const int32_t *base;
    int W, H, D;
    float uniqueness;
    float *selected;
    uint32_t *keys;
    uint16_t *sums;  // optional: the sums, W*H*band u16
};

template <int B>
__global__ __launch_bounds__(256) void band_select_kernel(BandSelectArgs a) {
    const int W = a.W, D = a.D;
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= W * a.H) return;
    const int x = p % W;
    constexpr int WORDS = B / 4;
    uint32_t s[B / 2];  // u16 pairs (2k, 2k+1)
#pragma unroll
    for (int k = 0; k < B / 2; k++) s[k] = 0;
#pragma unroll
    for (int r = 0; r < 8; r++) {
        if (r >= a.directions) break;
        const uint32_t *src = reinterpret_cast<const uint32_t *>(a.paths[r] + (size_t)p * B);
        uint32_t w[WORDS];
        if constexpr (B == 8) {
            const uint2 q = *reinterpret_cast<const uint2 *>(src);
            w[0] = q.x;
            w[1] = q.y;
        } else {
#pragma unroll
            for (int c = 0; c < B / 16; c++) {
                const uint4 q = reinterpret_cast<const uint4 *>(src)[c];
                w[4 * c] = q.x;
                w[4 * c + 1] = q.y;
                w[4 * c + 2] = q.z;
                w[4 * c + 3] = q.w;
            }
        }
#pragma unroll
        for (int k = 0; k < WORDS; k++) {
            s[2 * k] = pk_add(s[2 * k], __builtin_amdgcn_perm(0u, w[k], 0x0C010C00u));
            s[2 * k + 1] = pk_add(s[2 * k + 1], __builtin_amdgcn_perm(0u, w[k], 0x0C030C02u));
        }
    }
    if (a.sums) {
        uint32_t *dst = reinterpret_cast<uint32_t *>(a.sums + (size_t)p * B);
#pragma unroll
        for (int k = 0; k < B / 2; k++) dst[k] = s[k];
    }
    const int raw = a.base[p];
    const int b = searched_of(raw);
    int values[B];
#pragma unroll
    for (int k = 0; k < B; k++) {
        const int sum = (int)((k & 1) ? s[k >> 1] >> 16 : s[k >> 1] & 0xFFFFu);
        values[k] = sum + (b + k > x ? 0x8000 : 0);
    }
    int minimum = values[0], winner = 0;
#pragma unroll
    for (int k = 1; k < B; k++)
        if (values[k] < minimum) {
            minimum = values[k];
            winner = k;
        }
    int second = 0x7FFFFFFF;
    float lower = 0.0f, upper = 0.0f;
#pragma unroll
    for (int k = 0; k < B; k++) {
        const int d = k - winner;
        if (d > 1 || d < -1) second = min(second, values[k]);
        if (k == max(winner - 1, 0)) lower = (float)values[k];
        if (k == min(winner + 1, B - 1)) upper = (float)values[k];
    }
    const float minimum_f = (float)minimum;
    const int winner_d = b + winner;
    const bool accepted =
        raw >= 0 && winner_d <= x && (float)second > minimum_f * a.uniqueness;
    const float denominator = lower + upper - 2.0f * minimum_f;
    float correction = (lower - upper) / (2.0f * fmaxf(denominator, 1.0f));
    correction = fminf(fmaxf(correction, -0.5f), 0.5f);
    const bool refine = denominator > 0.0f && winner > 0 && winner < B - 1 &&
                        winner_d < min(D - 1, x);
    const float value = (float)winner_d + (refine ? correction : 0.0f);
    a.selected[p] = accepted ? value : -1.0f;
    a.keys[p] = accepted ? ((uint32_t)minimum << 16) | (uint32_t)winner_d : INVALID_KEY;
}

// ------------------------------------------------------------------------------------------

static size_t up256(size_t b) { return (b + 255) & ~(size_t)255; }

static int band_directions(int paths) { return paths == 8 ? 8 : 3; }

size_t band_bytes(int W, int H, int band, int paths) {
    const size_t P = (size_t)W * H, N = P * band;
    return up256(8 * P) * 2 + up256(N) * (1 + band_directions(paths)) + up256(4 * P) * 3;
}

void band_carve(BandBuffers &b, void *base, int W, int H, int band, int paths) {
    const size_t P = (size_t)W * H, N = P * band;
    uint8_t *p = static_cast<uint8_t *>(base);
    b.census_left = reinterpret_cast<uint64_t *>(p);
    p += up256(8 * P);
    b.census_right = reinterpret_cast<uint64_t *>(p);
    p += up256(8 * P);
    b.costs = p;
    p += up256(N);
    for (int r = 0; r < 8; r++) {
        if (r < band_directions(paths)) {
            b.paths[r] = p;
            p += up256(N);
        } else {
            b.paths[r] = nullptr;
        }
    }
    b.selected = reinterpret_cast<float *>(p);
    p += up256(4 * P);
    b.keys = reinterpret_cast<uint32_t *>(p);
    p += up256(4 * P);
    b.checked = reinterpret_cast<float *>(p);
}

template <int WINDOW>
static void launch_band_cost_window(const BandParams &p, const uint8_t *left,
                                    const uint8_t *right, const int32_t *base, BandBuffers &b,
                                    hipStream_t stream) {
    uint64_t *cl = p.keep_stages ? b.census_left : nullptr;
    uint64_t *cr = p.keep_stages ? b.census_right : nullptr;
    const size_t lds = band_cost_lds(p.width);
    const dim3 grid(p.height), block(BAND_COST_BLOCK);
    switch (p.band) {
        case 8: hipLaunchKernelGGL((band_cost_kernel<WINDOW, 8>), grid, block, lds, stream, left, right, base, p.width, p.height, b.costs, cl, cr); break;
        case 16: hipLaunchKernelGGL((band_cost_kernel<WINDOW, 16>), grid, block, lds, stream, left, right, base, p.width, p.height, b.costs, cl, cr); break;
        case 32: hipLaunchKernelGGL((band_cost_kernel<WINDOW, 32>), grid, block, lds, stream, left, right, base, p.width, p.height, b.costs, cl, cr); break;
        default: hipLaunchKernelGGL((band_cost_kernel<WINDOW, 64>), grid, block, lds, stream, left, right, base, p.width, p.height, b.costs, cl, cr); break;
    }
}

// Above 64 KB of LDS a kernel must say so first (band_cost_lds(4096) = 120 KB).
template <int WINDOW, int B>
static hipError_t allow_band_cost_lds() {
    return hipFuncSetAttribute(reinterpret_cast<const void *>(band_cost_kernel<WINDOW, B>),
                               hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)band_cost_lds(MAX_IMAGE_WIDTH));
}

template <int WINDOW>
static hipError_t allow_band_cost_window() {
    hipError_t e = allow_band_cost_lds<WINDOW, 8>();
    if (e == hipSuccess) e = allow_band_cost_lds<WINDOW, 16>();
    if (e == hipSuccess) e = allow_band_cost_lds<WINDOW, 32>();
    if (e == hipSuccess) e = allow_band_cost_lds<WINDOW, 64>();
    return e;
}

template <int B>
static void launch_band_paths(const BandPathArgs &a, int blocks, hipStream_t stream) {
    hipLaunchKernelGGL((band_path_kernel<B, 16>), dim3(blocks), dim3(BAND_PATH_BLOCK), 0, stream, a);
}

template <int B>
static void launch_band_select(const BandParams &p, const int32_t *base, BandBuffers &b,
                               uint16_t *sums, hipStream_t stream) {
    const int pixels = p.width * p.height;
    BandSelectArgs a;
    for (int r = 0; r < 8; r++) a.paths[r] = b.paths[r];
    a.directions = band_directions(p.paths);
    a.base = base;
    a.W = p.width;
    a.H = p.height;
    a.D = p.disparities;
    a.uniqueness = p.uniqueness;
    a.selected = b.selected;
    a.keys = b.keys;
    a.sums = sums;
    hipLaunchKernelGGL(band_select_kernel<B>, dim3((pixels + 255) / 256), dim3(256), 0, stream, a);
}

hipError_t launch_band(const BandParams &p, const uint8_t *left, const uint8_t *right,
                       const int32_t *base, BandBuffers &b, float *disparity, uint16_t *sums,
                       hipStream_t stream) {
    const int W = p.width, H = p.height;
    if (W > MAX_IMAGE_WIDTH || W < 1 || H < 1) return hipErrorInvalidValue;
    if (p.band != 8 && p.band != 16 && p.band != 32 && p.band != 64) return hipErrorInvalidValue;
    static const hipError_t lds_ok = [] {
        hipError_t e = allow_band_cost_window<0>();
        if (e == hipSuccess) e = allow_band_cost_window<1>();
        if (e == hipSuccess) e = allow_band_cost_window<2>();
        return e;
    }();
    if (lds_ok != hipSuccess) return lds_ok;

    int slot = profile_begin(PK_BAND_COST, stream);
    switch (p.census) {
        case 1: launch_band_cost_window<1>(p, left, right, base, b, stream); break;
        case 2: launch_band_cost_window<2>(p, left, right, base, b, stream); break;
        default: launch_band_cost_window<0>(p, left, right, base, b, stream); break;
    }
    profile_end(slot, stream);

    BandPathArgs a;
    a.costs = b.costs;
    a.base = base;
    for (int r = 0; r < 8; r++) a.paths[r] = b.paths[r];
    a.W = W;
    a.H = H;
    a.p1 = p.p1;
    a.p2 = p.p2;
    a.directions = band_directions(p.paths);
    const int lines[4] = {H, W, W + H - 1, W + H - 1};
    const int per_block = BAND_PATH_BLOCK / p.band;
    int blocks = 0;
    for (int r = 0; r < a.directions; r++) {
        a.block_start[r] = blocks;
        blocks += (lines[band_orientation(r)] + per_block - 1) / per_block;
    }
    for (int r = a.directions; r <= 8; r++) a.block_start[r] = blocks;
    slot = profile_begin(PK_BAND_PATH, stream);
    switch (p.band) {
        case 8: launch_band_paths<8>(a, blocks, stream); break;
        case 16: launch_band_paths<16>(a, blocks, stream); break;
        case 32: launch_band_paths<32>(a, blocks, stream); break;
        default: launch_band_paths<64>(a, blocks, stream); break;
    }
    profile_end(slot, stream);

    slot = profile_begin(PK_BAND_SELECT, stream);
    switch (p.band) {
        case 8: launch_band_select<8>(p, base, b, sums, stream); break;
        case 16: launch_band_select<16>(p, base, b, sums, stream); break;
        case 32: launch_band_select<32>(p, base, b, sums, stream); break;
        default: launch_band_select<64>(p, base, b, sums, stream); break;
    }
    profile_end(slot, stream);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launch_check_median(b.selected, b.keys, b.checked, disparity, W, H, p.keep_stages,
                               p.positive, stream);
}

// ------------------------------------------------------------------------------------------
// Coarse-to-fine (nx_stereo.ex:164-259).

size_t coarse_to_fine_bytes(int W, int H, int band, int paths, int coarse_paths) {
    const int cw = W / 2, ch = H / 2;
    const size_t half = (size_t)cw * ch;
    return up256(half) * 2 + match_bytes(cw, ch, 64, coarse_paths) + up256(4 * half) +
           up256(4 * (size_t)W * H) + band_bytes(W, H, band, paths);
}

void coarse_to_fine_carve(CoarseBuffers &c, void *base, int W, int H, int band, int paths,
                          int coarse_paths) {
    const int cw = W / 2, ch = H / 2;
    const size_t half = (size_t)cw * ch;
    uint8_t *p = static_cast<uint8_t *>(base);
    c.half_left = p;
    p += up256(half);
    c.half_right = p;
    p += up256(half);
    match_carve(c.match, p, cw, ch, 64, coarse_paths);
    p += match_bytes(cw, ch, 64, coarse_paths);
    c.coarse = reinterpret_cast<float *>(p);
    p += up256(4 * half);
    c.base = reinterpret_cast<int32_t *>(p);
    p += up256(4 * (size_t)W * H);
    band_carve(c.band, p, W, H, band, paths);
}

hipError_t launch_coarse_to_fine(const BandParams &fine, int margin, int coarse_paths,
                                 const uint8_t *left, const uint8_t *right, CoarseBuffers &c,
                                 float *disparity, hipStream_t stream) {
    const int W = fine.width, H = fine.height, cw = W / 2, ch = H / 2;
    if ((W | H) & 1) return hipErrorInvalidValue;
    const int half = cw * ch;
    int slot = profile_begin(PK_SHRINK, stream);
    hipLaunchKernelGGL(shrink_kernel, dim3((half + 255) / 256, 2), dim3(256), 0, stream, left,
                       right, W, H, c.half_left, c.half_right);
    profile_end(slot, stream);
    MatchParams coarse;
    coarse.width = cw;
    coarse.height = ch;
    coarse.disparities = 64;
    coarse.paths = coarse_paths;
    coarse.p1 = fine.p1;
    coarse.p2 = fine.p2;
    coarse.uniqueness = fine.uniqueness;
    coarse.keep_stages = false;
    coarse.positive = false;  // sgm/3's own output: a match at 0 counts
    hipError_t e = launch_match(coarse, c.half_left, c.half_right, c.match, c.coarse, nullptr, stream);
    if (e != hipSuccess) return e;
    slot = profile_begin(PK_SHRINK, stream);
    hipLaunchKernelGGL(band_base_kernel, dim3((half + 255) / 256), dim3(256), 0, stream, c.coarse,
                       cw, ch, fine.band, margin, c.base);
    profile_end(slot, stream);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launch_band(fine, left, right, c.base, c.band, disparity, nullptr, stream);
}

}  // namespace nxs
