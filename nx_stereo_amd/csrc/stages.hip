// The stages around the matcher on gfx950: Rectification, SpeckleFilter, Depth,
// PointCloud, TemporalFilter and Quality. f32 arithmetic follows the
// reference's operation order with one rounding per operation (this file is
// built with -ffp-contract=off; the explicit __f*_rn intrinsics keep it so).
#include "device.hpp"
#include "kernels.hpp"

namespace nxs {

__global__ void fill_f32_kernel(float *dst, float value, size_t count) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) dst[i] = value;
}

// Whether this block is the last of `blocks` to finish. Each block's threads have written
// their share; thread 0 fences at device scope (the writes reach memory every XCD reads)
// and counts the block in. The block that brings the count to `blocks` fences again (so
// it sees every other block's writes) and sets the counter back to zero for the next
// launch, so no clear is needed between launches.
__device__ bool last_block(unsigned *counter, unsigned blocks) {
    __shared__ bool last;
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        last = atomicAdd(counter, 1u) == blocks - 1;
        if (last) atomicExch(counter, 0u);
    }
    __syncthreads();
    if (last) __threadfence();
    return last;
}

// ------------------------------------------------------------------------------------------
// Rectification (rectify.rs:34-128; rectify.ex:133-207): per output pixel the projection
// through inverse(P3 R), plumb_bob distortion and K, a centre-aligned scale to the frame,
// and a bilinear sample with 0 outside, rounded half away from zero and clamped.

struct RectifyArgs {
    Camera camera[2];
    const uint8_t *frame[2];
    uint8_t *out[2];
    int fw, fh, width, height;
    float scale_x, scale_y;
};

__global__ __launch_bounds__(256) void rectify_kernel(RectifyArgs a) {
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    const int v = blockIdx.y;
    const int side = blockIdx.z;
    if (u >= a.width) return;
    const Camera &c = a.camera[side];
    const float *ipr = c.inverse_pr;
    const float fu = (float)u, fv = (float)v;
    float x = __fadd_rn(__fadd_rn(__fmul_rn(ipr[0], fu), __fmul_rn(ipr[1], fv)), ipr[2]);
    float y = __fadd_rn(__fadd_rn(__fmul_rn(ipr[3], fu), __fmul_rn(ipr[4], fv)), ipr[5]);
    const float w = __fadd_rn(__fadd_rn(__fmul_rn(ipr[6], fu), __fmul_rn(ipr[7], fv)), ipr[8]);
    x = __fdiv_rn(x, w);
    y = __fdiv_rn(y, w);
    const float k1 = c.d[0], k2 = c.d[1], p1 = c.d[2], p2 = c.d[3], k3 = c.d[4];
    const float r2 = __fadd_rn(__fmul_rn(x, x), __fmul_rn(y, y));
    // 1 + (k3 r2 r2 r2 + (k2 r2 r2 + k1 r2)), left-associated products
    const float radial = __fadd_rn(
        1.0f, __fadd_rn(__fmul_rn(__fmul_rn(__fmul_rn(k3, r2), r2), r2),
                        __fadd_rn(__fmul_rn(__fmul_rn(k2, r2), r2), __fmul_rn(k1, r2))));
    // x * radial + 2 p1 x y + p2 (r2 + 2 x x)
    const float dxd = __fadd_rn(
        __fadd_rn(__fmul_rn(x, radial), __fmul_rn(__fmul_rn(__fmul_rn(2.0f, p1), x), y)),
        __fmul_rn(p2, __fadd_rn(r2, __fmul_rn(__fmul_rn(2.0f, x), x))));
    // y * radial + p1 (r2 + 2 y y) + 2 p2 x y
    const float dyd = __fadd_rn(
        __fadd_rn(__fmul_rn(y, radial), __fmul_rn(p1, __fadd_rn(r2, __fmul_rn(__fmul_rn(2.0f, y), y)))),
        __fmul_rn(__fmul_rn(__fmul_rn(2.0f, p2), x), y));
    const float mx = __fadd_rn(__fadd_rn(__fmul_rn(c.k[0], dxd), __fmul_rn(c.k[1], dyd)), c.k[2]);
    const float my = __fadd_rn(__fmul_rn(c.k[4], dyd), c.k[5]);
    const float sx = __fsub_rn(__fmul_rn(__fadd_rn(mx, 0.5f), a.scale_x), 0.5f);
    const float sy = __fsub_rn(__fmul_rn(__fadd_rn(my, 0.5f), a.scale_y), 0.5f);
    const float fx0 = floorf(sx), fy0 = floorf(sy);
    const int x0 = (int)fx0, y0 = (int)fy0;
    const float cw = __fsub_rn(sx, fx0), rw = __fsub_rn(sy, fy0);
    const uint8_t *frame = a.frame[side];
    float s[4];
#pragma unroll
    for (int t = 0; t < 4; t++) {
        const int xx = x0 + (t & 1), yy = y0 + (t >> 1);
        const bool inside = xx >= 0 && xx < a.fw && yy >= 0 && yy < a.fh;
        s[t] = inside ? (float)frame[(size_t)yy * a.fw + xx] : 0.0f;
    }
    const float top = __fadd_rn(__fmul_rn(__fsub_rn(1.0f, cw), s[0]), __fmul_rn(cw, s[1]));
    const float bottom = __fadd_rn(__fmul_rn(__fsub_rn(1.0f, cw), s[2]), __fmul_rn(cw, s[3]));
    const float value = __fadd_rn(__fmul_rn(__fsub_rn(1.0f, rw), top), __fmul_rn(rw, bottom));
    const float r = fminf(fmaxf(roundf(value), 0.0f), 255.0f);
    a.out[side][(size_t)v * a.width + u] = (uint8_t)r;
}

hipError_t launch_rectify(const uint8_t *lf, const uint8_t *rf, int fw, int fh, const Camera &left,
                          const Camera &right, int width, int height, uint8_t *lo, uint8_t *ro,
                          hipStream_t stream) {
    RectifyArgs a;
    a.camera[0] = left;
    a.camera[1] = right;
    a.frame[0] = lf;
    a.frame[1] = rf;
    a.out[0] = lo;
    a.out[1] = ro;
    a.fw = fw;
    a.fh = fh;
    a.width = width;
    a.height = height;
    a.scale_x = (float)((double)fw / (double)width);
    a.scale_y = (float)((double)fh / (double)height);
    hipLaunchKernelGGL(rectify_kernel, dim3((width + 255) / 256, height, 2), dim3(256), 0, stream, a);
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------------
// SpeckleFilter (filters.rs:222-268; filter.ex:48-70): keep c when c > 0 and at least
// min_support of the other pixels within `radius` are > 0 and within max_diff of c;
// outside the image reads -1.

__global__ __launch_bounds__(256) void despeckle_kernel(const float *__restrict__ in,
                                                        float *__restrict__ out, int W, int H,
                                                        float min_support, float max_diff,
                                                        int radius) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= W) return;
    const float c = in[(size_t)y * W + x];
    unsigned support = 0;
    const int y_lo = max(y - radius, 0), y_hi = min(y + radius, H - 1);
    const int x_lo = max(x - radius, 0), x_hi = min(x + radius, W - 1);
    for (int yy = y_lo; yy <= y_hi; yy++) {
        const float *row = in + (size_t)yy * W;
        for (int xx = x_lo; xx <= x_hi; xx++) {
            if (yy == y && xx == x) continue;
            const float n = row[xx];
            support += (n > 0.0f) && (fabsf(__fsub_rn(n, c)) <= max_diff);
        }
    }
    out[(size_t)y * W + x] = (c > 0.0f && (float)support >= min_support) ? c : -1.0f;
}

// The stage's own radius (3, stage/speckle_filter.ex): a 32x8 block of pixels reads its
// (32+2R)x(8+2R) window into LDS once (outside the image: -1, which never counts), then
// counts from LDS.
constexpr int SPECKLE_TX = 32, SPECKLE_TY = 8;

template <int R>
__global__ __launch_bounds__(SPECKLE_TX * SPECKLE_TY) void despeckle_tiled_kernel(
    const float *__restrict__ in, float *__restrict__ out, int W, int H, float min_support,
    float max_diff) {
    constexpr int TW = SPECKLE_TX + 2 * R, TH = SPECKLE_TY + 2 * R;
    __shared__ float tile[TH][TW];
    const int x0 = blockIdx.x * SPECKLE_TX, y0 = blockIdx.y * SPECKLE_TY;
    for (int i = threadIdx.x; i < TW * TH; i += SPECKLE_TX * SPECKLE_TY) {
        const int r = i / TW, c = i - r * TW;
        const int yy = y0 - R + r, xx = x0 - R + c;
        tile[r][c] = (yy >= 0 && yy < H && xx >= 0 && xx < W) ? in[(size_t)yy * W + xx] : -1.0f;
    }
    __syncthreads();
    const int tx = threadIdx.x % SPECKLE_TX, ty = threadIdx.x / SPECKLE_TX;
    const int x = x0 + tx, y = y0 + ty;
    if (x >= W || y >= H) return;
    const float c = tile[ty + R][tx + R];
    unsigned support = 0;
#pragma unroll
    for (int dy = -R; dy <= R; dy++) {
#pragma unroll
        for (int dx = -R; dx <= R; dx++) {
            if (dy == 0 && dx == 0) continue;
            const float n = tile[ty + R + dy][tx + R + dx];
            support += (n > 0.0f) && (fabsf(__fsub_rn(n, c)) <= max_diff);
        }
    }
    out[(size_t)y * W + x] = (c > 0.0f && (float)support >= min_support) ? c : -1.0f;
}

hipError_t launch_despeckle(const float *in, float *out, int W, int H, float min_support,
                            float max_diff, int radius, hipStream_t stream) {
    if (radius == 3) {
        hipLaunchKernelGGL(despeckle_tiled_kernel<3>,
                           dim3((W + SPECKLE_TX - 1) / SPECKLE_TX, (H + SPECKLE_TY - 1) / SPECKLE_TY),
                           dim3(SPECKLE_TX * SPECKLE_TY), 0, stream, in, out, W, H, min_support,
                           max_diff);
    } else {
        hipLaunchKernelGGL(despeckle_kernel, dim3((W + 255) / 256, H), dim3(256), 0, stream, in, out,
                           W, H, min_support, max_diff, radius);
    }
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------------
// Depth (depth.rs:12-44; depth.ex:26-32): m = f*B/d for d > 0 else 0; mm = round(m*1000),
// 0 beyond 65535, at least 0.

__global__ __launch_bounds__(256) void depth_kernel(const float *__restrict__ disparity, int pixels,
                                                    float focal_baseline,
                                                    uint16_t *__restrict__ mm,
                                                    float *__restrict__ m) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= pixels) return;
    const float d = disparity[i];
    const float metres = d > 0.0f ? __fdiv_rn(focal_baseline, d) : 0.0f;
    const float rounded = roundf(__fmul_rn(metres, 1000.0f));
    m[i] = metres;
    mm[i] = rounded > 65535.0f ? (uint16_t)0 : (uint16_t)(rounded > 0.0f ? rounded : 0.0f);
}

hipError_t launch_depth(const float *disparity, int pixels, float f, float b, uint16_t *mm, float *m,
                        hipStream_t stream) {
    const float fb = f * b;  // host f32 multiply, as depth.rs:20
    hipLaunchKernelGGL(depth_kernel, dim3((pixels + 255) / 256), dim3(256), 0, stream, disparity,
                       pixels, fb, mm, m);
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------------
// PointCloud (depth.rs:58-89; depth.ex:73-90): every decimation-th pixel unprojected where
// nearest < z < farthest, NaN otherwise, and the kept count.

// `partials` holds one count per block; the last block to finish adds them up into `kept`.
__global__ __launch_bounds__(256) void cloud_kernel(const float *__restrict__ depth_m, int W,
                                                    int gw, int count, int dec, float f, float cx,
                                                    float cy, float nearest, float farthest,
                                                    float *__restrict__ points,
                                                    int32_t *__restrict__ partials,
                                                    unsigned *__restrict__ counter,
                                                    int32_t *__restrict__ kept_total) {
    __shared__ int32_t wave_kept[4];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    int kept = 0;
    if (i < count) {
        const int gi = i / gw, gj = i - gi * gw;
        const int y = gi * dec, x = gj * dec;
        const float z = depth_m[(size_t)y * W + x];
        float *p = points + (size_t)i * 3;
        if (z > nearest && z < farthest) {
            const float zf = __fdiv_rn(z, f);
            p[0] = __fmul_rn(__fsub_rn((float)x, cx), zf);
            p[1] = __fmul_rn(__fsub_rn((float)y, cy), zf);
            p[2] = z;
            kept = 1;
        } else {
            const float nan = __int_as_float(0x7FC00000);
            p[0] = nan;
            p[1] = nan;
            p[2] = nan;
        }
    }
    // Block count: ballot per wave, then four waves.
    const unsigned long long ballot = __ballot(kept);
    if ((threadIdx.x & 63) == 0) wave_kept[threadIdx.x >> 6] = __popcll(ballot);
    __syncthreads();
    if (threadIdx.x == 0)
        partials[blockIdx.x] = wave_kept[0] + wave_kept[1] + wave_kept[2] + wave_kept[3];
    const unsigned blocks = gridDim.x;
    if (!last_block(counter, blocks)) return;
    __shared__ int32_t sums[256];
    int32_t total = 0;
    for (unsigned b = threadIdx.x; b < blocks; b += 256) total += partials[b];
    sums[threadIdx.x] = total;
    __syncthreads();
    for (int half = 128; half > 0; half >>= 1) {
        if ((int)threadIdx.x < half) sums[threadIdx.x] += sums[threadIdx.x + half];
        __syncthreads();
    }
    if (threadIdx.x == 0) *kept_total = sums[0];
}

int cloud_partials(int W, int H, int dec) {
    const int count = (W / dec) * (H / dec);
    return (count + 255) / 256 + 1;  // one count per block, then room for the counter
}

hipError_t launch_cloud(const float *depth_m, int W, int H, int dec, float f, float cx, float cy,
                        float nearest, float farthest, float *points, int32_t *partials,
                        int32_t *kept, hipStream_t stream) {
    const int gw = W / dec, count = gw * (H / dec);
    const int blocks = cloud_partials(W, H, dec) - 1;
    if (blocks > 0) {
        // The finishing counter sits just before `kept`, past every member's counts.
        hipLaunchKernelGGL(cloud_kernel, dim3(blocks), dim3(256), 0, stream, depth_m, W, gw, count,
                           dec, f, cx, cy, nearest, farthest, points, partials,
                           reinterpret_cast<unsigned *>(kept - 1), kept);
    } else {
        // No grid point: the count is zero.
        hipLaunchKernelGGL(fill_f32_kernel, dim3(1), dim3(64), 0, stream,
                           reinterpret_cast<float *>(kept), 0.0f, (size_t)1);
    }
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------------
// TemporalFilter (filters.rs:296-318; filter.ex:122-129), with the carried previous
// disparity and missed counts.

__global__ __launch_bounds__(256) void temporal_kernel(float *__restrict__ disparity,
                                                       float *__restrict__ previous,
                                                       float *__restrict__ missed, int pixels,
                                                       float alpha, float delta, float hold,
                                                       float max_missed) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= pixels) return;
    const float c = disparity[i], p = previous[i];
    const bool matched = c > 0.0f;
    const bool agree = matched && p > 0.0f && fabsf(__fsub_rn(c, p)) <= delta;
    const float smoothed =
        agree ? __fadd_rn(__fmul_rn(alpha, c), __fmul_rn(__fsub_rn(1.0f, alpha), p)) : c;
    const float pairs_missed = matched ? 0.0f : __fadd_rn(missed[i], 1.0f);
    const bool held = !matched && p > 0.0f && pairs_missed <= hold;
    const float filtered = held ? p : smoothed;
    disparity[i] = filtered;
    previous[i] = filtered;
    missed[i] = fminf(pairs_missed, max_missed);
}

hipError_t launch_temporal(float *disparity, float *previous, float *missed, int pixels,
                           float alpha, float delta, float hold, float max_missed,
                           hipStream_t stream) {
    hipLaunchKernelGGL(temporal_kernel, dim3((pixels + 255) / 256), dim3(256), 0, stream, disparity,
                       previous, missed, pixels, alpha, delta, hold, max_missed);
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------------
// Quality (measures.rs:125-154; diagnostics.ex:264-271): per column, f64 sums down the rows
// of [matched, v, v^2, both, |c-p| * both]; each column's sum cast to f32; then an f64 sum
// over columns, cast to f32. The carried previous disparity takes the current.
//
// Two launches, with no fence between blocks. Every term is an f32 value (matched and both
// are 0 or 1, and a matched v is at least 0.5) whose bits fall within an f64 column sum's 53
// for up to 4096 rows, so each column's f64 sum is exact in any order and its f32 cast is
// the reference's (past 4096 rows a sum may round, and its value then depends on the
// chunking, the same from run to run).
//  1. A grid of QUALITY_COLS columns by QUALITY_ROWS rows per block: its QUALITY_PHASES
//     thread rows sum every QUALITY_PHASES-th row of the chunk, and the block adds its
//     columns' chunk sums into per-column f64 totals with atomics (exact, so in any order).
//     It also moves the current disparity into the carried previous.
//  2. One block: each column's total cast to f32 (and the total set back to zero for the
//     next launch), then the columns of each group of QUALITY_COLS added by a fixed
//     pairwise tree (a wave's shuffles), and the groups in order.

constexpr int QUALITY_COLS = 64;
constexpr int QUALITY_PHASES = 8;
constexpr int QUALITY_ROWS = 16;
constexpr int QUALITY_FINISH = 1024;

__global__ __launch_bounds__(QUALITY_COLS * QUALITY_PHASES) void quality_chunks_kernel(
    const float *__restrict__ disparity, float *__restrict__ previous, int W, int H,
    double *__restrict__ totals) {
    __shared__ double column[QUALITY_PHASES][5][QUALITY_COLS];
    const int lx = threadIdx.x % QUALITY_COLS, phase = threadIdx.x / QUALITY_COLS;
    const int x = blockIdx.x * QUALITY_COLS + lx;
    const int y0 = blockIdx.y * QUALITY_ROWS;
    constexpr int PER_THREAD = QUALITY_ROWS / QUALITY_PHASES;
    double s0 = 0, s1 = 0, s2 = 0, s3 = 0, s4 = 0;
    if (x < W) {
        float c[PER_THREAD], p[PER_THREAD];
#pragma unroll
        for (int j = 0; j < PER_THREAD; j++) {
            const size_t i = (size_t)min(y0 + phase + j * QUALITY_PHASES, H - 1) * W + x;
            c[j] = disparity[i];
            p[j] = previous[i];
        }
#pragma unroll
        for (int j = 0; j < PER_THREAD; j++) {
            const int y = y0 + phase + j * QUALITY_PHASES;
            if (y >= H) break;
            const float matched = c[j] > 0.0f ? 1.0f : 0.0f;
            const float both = __fmul_rn(p[j] > 0.0f ? 1.0f : 0.0f, matched);
            const float value = __fmul_rn(c[j], matched);
            s0 += (double)matched;
            s1 += (double)value;
            s2 += (double)__fmul_rn(value, value);
            s3 += (double)both;
            s4 += (double)__fmul_rn(fabsf(__fsub_rn(c[j], p[j])), both);
            previous[(size_t)y * W + x] = c[j];
        }
    }
    column[phase][0][lx] = s0;
    column[phase][1][lx] = s1;
    column[phase][2][lx] = s2;
    column[phase][3][lx] = s3;
    column[phase][4][lx] = s4;
    __syncthreads();
    const int k = threadIdx.x / QUALITY_COLS;  // plane k of column lx, for the first 5*64 threads
    if (k < 5 && x < W) {
        double sum = 0;
        for (int q = 0; q < QUALITY_PHASES; q++) sum += column[q][k][lx];
        atomicAdd(&totals[(size_t)k * W + x], sum);
    }
}

constexpr int QUALITY_FINISH_GROUPS = 64;  // groups of QUALITY_COLS: up to 4096 columns

// One block of QUALITY_FINISH threads: the column totals into the five sums.
__device__ void finish_quality(double *__restrict__ totals, int W, float *__restrict__ sums5) {
    static_assert(QUALITY_COLS == 64, "a group of columns is one wave's lanes");
    __shared__ double group_sum[5][QUALITY_FINISH_GROUPS];
    const int groups = (W + QUALITY_COLS - 1) / QUALITY_COLS;
    const int wave = threadIdx.x / QUALITY_COLS, lane = threadIdx.x % QUALITY_COLS;
    // A wave per group: each column's total cast to f32 (a column past W is 0) and set back to
    // zero for the next launch, then the pairwise tree, lane i + half into lane i.
    for (int g = wave; g < groups; g += QUALITY_FINISH / QUALITY_COLS) {
        const int x = g * QUALITY_COLS + lane;
        double v[5];
#pragma unroll
        for (int k = 0; k < 5; k++) v[k] = x < W ? (double)(float)totals[(size_t)k * W + x] : 0.0;
#pragma unroll
        for (int k = 0; k < 5; k++)
            if (x < W) totals[(size_t)k * W + x] = 0.0;
        for (int half = QUALITY_COLS / 2; half > 0; half >>= 1)
#pragma unroll
            for (int k = 0; k < 5; k++) v[k] += __shfl_down(v[k], half, QUALITY_COLS);
        if (lane == 0)
#pragma unroll
            for (int k = 0; k < 5; k++) group_sum[k][g] = v[k];
    }
    __syncthreads();
    if (threadIdx.x < 5) {  // the groups in order
        double total = 0;
        for (int g = 0; g < groups; g++) total += group_sum[threadIdx.x][g];
        sums5[threadIdx.x] = (float)total;
    }
}

__global__ __launch_bounds__(QUALITY_FINISH) void quality_finish_kernel(
    double *__restrict__ totals, int W, float *__restrict__ sums5) {
    finish_quality(totals, W, sums5);
}

size_t quality_scratch_bytes(int W, int H) {
    (void)H;
    return 5 * (size_t)W * sizeof(double);  // the column totals
}

hipError_t launch_quality(const float *disparity, float *previous, int W, int H, double *partials,
                          float *sums5, hipStream_t stream) {
    if (W > QUALITY_FINISH_GROUPS * QUALITY_COLS) return hipErrorInvalidValue;
    const int groups = (W + QUALITY_COLS - 1) / QUALITY_COLS;
    const int n_chunks = (H + QUALITY_ROWS - 1) / QUALITY_ROWS;
    hipLaunchKernelGGL(quality_chunks_kernel, dim3(groups, n_chunks),
                       dim3(QUALITY_COLS * QUALITY_PHASES), 0, stream, disparity, previous, W, H,
                       partials);
    hipLaunchKernelGGL(quality_finish_kernel, dim3(1), dim3(QUALITY_FINISH), 0, stream, partials,
                       W, sums5);
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------------
// The per-pixel tail, fused: an ordered run of SpeckleFilter (radius 3), TemporalFilter,
// Depth, PointCloud and Quality in one launch over 64x16 tiles, each pixel through the run's
// stages in turn with the very operations of the kernels above, then one block that
// finishes what needs every tile (the quality sums, the kept count). A tile's 512 threads
// take two rows each, 8 apart, as quality_chunks_kernel does, so each column's quality
// sums gather the tile's 16 rows before one f64 atomic add per column and sum (exact in
// any order, as there). At 728x544 the launch takes 12.5 us against 22 for the four
// launches it replaces (64x8 tiles, twice the blocks and the atomics: 15 us). The kept count gathers by an integer atomic add into `kept_accum`,
// which the finishing block moves into `kept_total` and sets back to zero.

constexpr int TAIL_TX = 64, TAIL_PHASES = 8, TAIL_TY = 16, TAIL_R = 3;

__global__ __launch_bounds__(TAIL_TX * TAIL_PHASES) void tail_kernel(TailArgs a) {
    constexpr int TW = TAIL_TX + 2 * TAIL_R, TH = TAIL_TY + 2 * TAIL_R;
    __shared__ float tile[TH][TW];
    __shared__ double column[TAIL_PHASES][5][TAIL_TX];
    __shared__ int32_t block_kept;
    const int W = a.W, H = a.H;
    const int x0 = blockIdx.x * TAIL_TX, y0 = blockIdx.y * TAIL_TY;
    const int lx = threadIdx.x % TAIL_TX, phase = threadIdx.x / TAIL_TX;
    if (a.despeckle) {
        // The window with its radius-3 border; outside the image reads -1, which never counts.
        for (int i = threadIdx.x; i < TW * TH; i += TAIL_TX * TAIL_PHASES) {
            const int r = i / TW, c = i - r * TW;
            const int yy = y0 - TAIL_R + r, xx = x0 - TAIL_R + c;
            tile[r][c] = (yy >= 0 && yy < H && xx >= 0 && xx < W) ? a.in[(size_t)yy * W + xx] : -1.0f;
        }
        __syncthreads();
    }
    const int x = x0 + lx;
    double s0 = 0, s1 = 0, s2 = 0, s3 = 0, s4 = 0;
    int kept = 0;
#pragma unroll
    for (int j = 0; j < TAIL_TY / TAIL_PHASES; j++) {
        const int ty = phase + j * TAIL_PHASES, y = y0 + ty;
        if (x >= W || y >= H) continue;
        const size_t i = (size_t)y * W + x;
        float v;
        if (a.despeckle) {  // despeckle_tiled_kernel<3>
            const float c = tile[ty + TAIL_R][lx + TAIL_R];
            unsigned support = 0;
#pragma unroll
            for (int dy = -TAIL_R; dy <= TAIL_R; dy++) {
#pragma unroll
                for (int dx = -TAIL_R; dx <= TAIL_R; dx++) {
                    if (dy == 0 && dx == 0) continue;
                    const float n = tile[ty + TAIL_R + dy][lx + TAIL_R + dx];
                    support += (n > 0.0f) && (fabsf(__fsub_rn(n, c)) <= a.max_diff);
                }
            }
            v = (c > 0.0f && (float)support >= a.min_support) ? c : -1.0f;
        } else {
            v = a.in[i];
        }
        if (a.temporal) {  // temporal_kernel
            const float p = a.t_prev[i];
            const bool matched = v > 0.0f;
            const bool agree = matched && p > 0.0f && fabsf(__fsub_rn(v, p)) <= a.delta;
            const float smoothed =
                agree ? __fadd_rn(__fmul_rn(a.alpha, v), __fmul_rn(__fsub_rn(1.0f, a.alpha), p)) : v;
            const float pairs_missed = matched ? 0.0f : __fadd_rn(a.t_missed[i], 1.0f);
            const bool held = !matched && p > 0.0f && pairs_missed <= a.hold;
            v = held ? p : smoothed;
            a.t_prev[i] = v;
            a.t_missed[i] = fminf(pairs_missed, a.max_missed);
        }
        if (a.despeckle || a.temporal) a.out[i] = v;
        if (a.depth) {  // depth_kernel
            const float metres = v > 0.0f ? __fdiv_rn(a.focal_baseline, v) : 0.0f;
            const float rounded = roundf(__fmul_rn(metres, 1000.0f));
            a.m[i] = metres;
            a.mm[i] = rounded > 65535.0f ? (uint16_t)0 : (uint16_t)(rounded > 0.0f ? rounded : 0.0f);
            if (a.cloud && x % a.dec == 0 && y % a.dec == 0 && x / a.dec < a.gw && y / a.dec < a.gh) {
                // cloud_kernel, on this pixel's depth_m
                float *pt = a.points + ((size_t)(y / a.dec) * a.gw + x / a.dec) * 3;
                if (metres > a.nearest && metres < a.farthest) {
                    const float zf = __fdiv_rn(metres, a.f);
                    pt[0] = __fmul_rn(__fsub_rn((float)x, a.cx), zf);
                    pt[1] = __fmul_rn(__fsub_rn((float)y, a.cy), zf);
                    pt[2] = metres;
                    kept++;
                } else {
                    const float nan = __int_as_float(0x7FC00000);
                    pt[0] = nan;
                    pt[1] = nan;
                    pt[2] = nan;
                }
            }
        }
        if (a.quality) {  // quality_chunks_kernel
            const float p = a.q_prev[i];
            const float matched = v > 0.0f ? 1.0f : 0.0f;
            const float both = __fmul_rn(p > 0.0f ? 1.0f : 0.0f, matched);
            const float value = __fmul_rn(v, matched);
            s0 += (double)matched;
            s1 += (double)value;
            s2 += (double)__fmul_rn(value, value);
            s3 += (double)both;
            s4 += (double)__fmul_rn(fabsf(__fsub_rn(v, p)), both);
            a.q_prev[i] = v;
        }
    }
    if (a.cloud) {
        // The block's kept points (a thread's two rows may both hold one), then one add.
        if (threadIdx.x == 0) block_kept = 0;
        __syncthreads();
        if (kept) atomicAdd(&block_kept, kept);
        __syncthreads();
        if (threadIdx.x == 0 && block_kept) atomicAdd(a.kept_accum, block_kept);
    }
    if (a.quality) {
        column[phase][0][lx] = s0;
        column[phase][1][lx] = s1;
        column[phase][2][lx] = s2;
        column[phase][3][lx] = s3;
        column[phase][4][lx] = s4;
        __syncthreads();
        const int k = threadIdx.x / TAIL_TX;  // plane k of column lx, for the first 5*64 threads
        if (k < 5 && x < W) {
            double sum = 0;
            for (int q = 0; q < TAIL_PHASES; q++) sum += column[q][k][lx];
            atomicAdd(&a.totals[(size_t)k * W + x], sum);
        }
    }
}

// The tail's finish: the quality sums (as quality_finish_kernel) and the kept count.
__global__ __launch_bounds__(QUALITY_FINISH) void tail_finish_kernel(
    double *__restrict__ totals, int W, float *__restrict__ sums5,
    int32_t *__restrict__ kept_accum, int32_t *__restrict__ kept_total) {
    if (totals) finish_quality(totals, W, sums5);
    if (kept_accum && threadIdx.x == 0) {
        *kept_total = *kept_accum;
        *kept_accum = 0;
    }
}

hipError_t launch_tail(const TailArgs &a, hipStream_t stream) {
    if (a.quality && a.W > QUALITY_FINISH_GROUPS * QUALITY_COLS) return hipErrorInvalidValue;
    hipLaunchKernelGGL(tail_kernel, dim3((a.W + TAIL_TX - 1) / TAIL_TX, (a.H + TAIL_TY - 1) / TAIL_TY),
                       dim3(TAIL_TX * TAIL_PHASES), 0, stream, a);
    return hipGetLastError();
}

hipError_t launch_tail_finish(const TailArgs &a, hipStream_t stream) {
    if (!a.quality && !a.cloud) return hipSuccess;
    hipLaunchKernelGGL(tail_finish_kernel, dim3(1), dim3(QUALITY_FINISH), 0, stream,
                       a.quality ? a.totals : nullptr, a.W, a.sums5,
                       a.cloud ? a.kept_accum : nullptr, a.kept_total);
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------------
// MedianFilter (filters.rs:134-216; image.ex:479-505): the middle of each pixel's k x k
// window (3 or 5), borders replicated, in f32 comparisons. Every value but a NaN or a signed
// zero has one place in the order, so the middle is the reference's (the maps this stage
// filters hold -1 and positive disparities). 3x3: the classic 19-exchange median network.
// 5x5: Batcher's odd-even merge sorting network on 32 slots, the 7 past the window at
// +infinity, then the 13th value.

struct MergeNetwork32 {
    int count;
    unsigned char lo[192], hi[192];
    // Batcher's odd-even merge sort on n = 32 inputs (191 compare-exchanges).
    constexpr MergeNetwork32() : count(0), lo(), hi() {
        const int n = 32;
        for (int p = 1; p < n; p += p)
            for (int k = p; k > 0; k /= 2)
                for (int j = k % p; j + k < n; j += 2 * k)
                    for (int i = 0; i < k && i + j + k < n; i++)
                        if ((i + j) / (2 * p) == (i + j + k) / (2 * p)) {
                            lo[count] = (unsigned char)(i + j);
                            hi[count] = (unsigned char)(i + j + k);
                            count++;
                        }
    }
};
constexpr MergeNetwork32 MERGE32{};
static_assert(MERGE32.count == 191, "Batcher's network on 32 inputs has 191 exchanges");

template <int SIDE>
__global__ __launch_bounds__(256) void median_filter_kernel(const float *__restrict__ in,
                                                            float *__restrict__ out, int W,
                                                            int H) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= W) return;
    constexpr int R = SIDE / 2, SLOTS = SIDE == 3 ? 9 : 32;
    float v[SLOTS];
    int n = 0;
#pragma unroll
    for (int dy = -R; dy <= R; dy++) {
        const float *row = in + (size_t)min(max(y + dy, 0), H - 1) * W;
#pragma unroll
        for (int dx = -R; dx <= R; dx++) v[n++] = row[min(max(x + dx, 0), W - 1)];
    }
#define MX(a, b)                          \
    {                                     \
        const float lo_ = fminf(v[a], v[b]); \
        const float hi_ = fmaxf(v[a], v[b]); \
        v[a] = lo_;                       \
        v[b] = hi_;                       \
    }
    if constexpr (SIDE == 3) {
        MX(1, 2) MX(4, 5) MX(7, 8) MX(0, 1) MX(3, 4) MX(6, 7) MX(1, 2) MX(4, 5) MX(7, 8) MX(0, 3)
        MX(5, 8) MX(4, 7) MX(3, 6) MX(1, 4) MX(2, 5) MX(4, 7) MX(4, 2) MX(6, 4) MX(4, 2)
    } else {
#pragma unroll
        for (int i = SIDE * SIDE; i < SLOTS; i++) v[i] = __builtin_inff();
#pragma unroll
        for (int c = 0; c < MERGE32.count; c++) MX(MERGE32.lo[c], MERGE32.hi[c])
    }
#undef MX
    out[(size_t)y * W + x] = v[SIDE * SIDE / 2];
}

hipError_t launch_median_filter(const float *in, float *out, int W, int H, int kernel_size,
                                hipStream_t stream) {
    const dim3 grid((W + 255) / 256, H);
    if (kernel_size == 3)
        hipLaunchKernelGGL(median_filter_kernel<3>, grid, dim3(256), 0, stream, in, out, W, H);
    else if (kernel_size == 5)
        hipLaunchKernelGGL(median_filter_kernel<5>, grid, dim3(256), 0, stream, in, out, W, H);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------------
// Sharpness (measures.rs:11-51): per column x < W-1 and image, over the rows, the pairs
// (x, x+1) with both above black: their squared difference, the left grey and the count.
// Then, per plane, each column's sum cast to f32, those summed in f64 and cast to f32. The
// column sums are integers and their f32 values integers under 2^33, so the f64 sum of at
// most 4095 of them is exact in any order. Planes: [left energy, right energy, left grey,
// right grey, left pairs, right pairs]. Two launches: row chunks add into u64 column totals
// (atomics), then one block finishes and sets the totals back to zero.

constexpr int SHARP_ROWS = 64;
constexpr int SHARP_FINISH = 1024;

__global__ __launch_bounds__(256) void sharpness_chunks_kernel(
    const uint8_t *__restrict__ left, const uint8_t *__restrict__ right, int W, int H,
    unsigned long long *__restrict__ totals) {
    const int n = W - 1, x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= n) return;
    const int y0 = blockIdx.y * SHARP_ROWS, y1 = min(H, y0 + SHARP_ROWS);
    uint32_t energy[2] = {0, 0}, grey[2] = {0, 0}, pairs[2] = {0, 0};
    for (int y = y0; y < y1; y++) {
#pragma unroll
        for (int side = 0; side < 2; side++) {
            const uint8_t *row = (side ? right : left) + (size_t)y * W;
            const int a = row[x], b = row[x + 1];
            if (a > 0 && b > 0) {
                energy[side] += (uint32_t)((b - a) * (b - a));
                grey[side] += (uint32_t)a;
                pairs[side] += 1;
            }
        }
    }
#pragma unroll
    for (int side = 0; side < 2; side++) {
        if (energy[side]) atomicAdd(&totals[(size_t)(0 + side) * n + x], (unsigned long long)energy[side]);
        if (grey[side]) atomicAdd(&totals[(size_t)(2 + side) * n + x], (unsigned long long)grey[side]);
        if (pairs[side]) atomicAdd(&totals[(size_t)(4 + side) * n + x], (unsigned long long)pairs[side]);
    }
}

__global__ __launch_bounds__(SHARP_FINISH) void sharpness_finish_kernel(
    unsigned long long *__restrict__ totals, int n, float *__restrict__ sums6) {
    __shared__ double part[SHARP_FINISH];
    for (int plane = 0; plane < 6; plane++) {
        double sum = 0.0;
        for (int x = threadIdx.x; x < n; x += SHARP_FINISH) {
            sum += (double)(float)totals[(size_t)plane * n + x];
            totals[(size_t)plane * n + x] = 0;
        }
        part[threadIdx.x] = sum;
        __syncthreads();
        for (int half = SHARP_FINISH / 2; half > 0; half >>= 1) {
            if ((int)threadIdx.x < half) part[threadIdx.x] += part[threadIdx.x + half];
            __syncthreads();
        }
        if (threadIdx.x == 0) sums6[plane] = (float)part[0];
        __syncthreads();
    }
}

size_t sharpness_scratch_bytes(int W) { return 6 * (size_t)max(W - 1, 1) * sizeof(uint64_t); }

hipError_t launch_sharpness(const uint8_t *left, const uint8_t *right, int W, int H,
                            uint64_t *totals, float *sums6, hipStream_t stream) {
    auto *t = reinterpret_cast<unsigned long long *>(totals);
    const int n = W - 1;
    if (n > 0)
        hipLaunchKernelGGL(sharpness_chunks_kernel, dim3((n + 255) / 256, (H + SHARP_ROWS - 1) / SHARP_ROWS),
                           dim3(256), 0, stream, left, right, W, H, t);
    hipLaunchKernelGGL(sharpness_finish_kernel, dim3(1), dim3(SHARP_FINISH), 0, stream, t, max(n, 0),
                       sums6);
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------------
// RowAlignment (measures.rs:53-108): per band of W / bands columns, over rows 0, row_step,
// ...: each matched left pixel's absolute difference with the right pixel at its rounded
// disparity (half away from zero, clamped to the image) on each of the 2 reach + 1 rows
// from reach above to reach below (clamped), then the matched count; each sum cast to f32.
// A block takes one band and ALIGN_ROWS sampled rows; each thread sums up to ALIGN_PLANES
// planes at a time in registers (planes past that in further passes), the waves reduce
// them, and the block adds them into u64 totals. One more launch casts the totals and sets
// them back to zero.

constexpr int ALIGN_ROWS = 8;
constexpr int ALIGN_PLANES = 16;

__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

__global__ __launch_bounds__(256) void alignment_chunks_kernel(
    const uint8_t *__restrict__ left, const uint8_t *__restrict__ right,
    const float *__restrict__ disparity, int W, int H, int bands, int reach, int row_step,
    unsigned long long *__restrict__ totals) {
    const int planes = 2 * reach + 2, band = blockIdx.x, bw = W / bands;
    const int x0 = band * bw, rows = (H + row_step - 1) / row_step;
    const int r0 = blockIdx.y * ALIGN_ROWS, r1 = min(rows, r0 + ALIGN_ROWS);
    for (int p0 = 0; p0 < planes; p0 += ALIGN_PLANES) {
        uint32_t acc[ALIGN_PLANES];
#pragma unroll
        for (int p = 0; p < ALIGN_PLANES; p++) acc[p] = 0;
        for (int r = r0; r < r1; r++) {
            const int y = r * row_step;
            for (int x = x0 + (int)threadIdx.x; x < x0 + bw; x += blockDim.x) {
                const float value = disparity[(size_t)y * W + x];
                if (!(value > 0.0f)) continue;
                const float sx = fminf(fmaxf(__fsub_rn((float)x, roundf(value)), 0.0f), (float)(W - 1));
                const int source_x = (int)sx;
                const int grey = left[(size_t)y * W + x];
#pragma unroll
                for (int p = 0; p < ALIGN_PLANES; p++) {
                    const int plane = p0 + p;
                    if (plane < 2 * reach + 1) {
                        const int sy = min(max(y + plane - reach, 0), H - 1);
                        acc[p] += (uint32_t)abs((int)right[(size_t)sy * W + source_x] - grey);
                    } else if (plane == 2 * reach + 1) {
                        acc[p] += 1;
                    }
                }
            }
        }
#pragma unroll
        for (int p = 0; p < ALIGN_PLANES; p++) {
            const uint32_t v = wave_sum(acc[p]);
            if ((threadIdx.x & 63) == 0 && v != 0 && p0 + p < planes)
                atomicAdd(&totals[(size_t)band * planes + p0 + p], (unsigned long long)v);
        }
    }
}

__global__ __launch_bounds__(256) void alignment_finish_kernel(unsigned long long *__restrict__ totals,
                                                               int count, float *__restrict__ sums) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    sums[i] = __ull2float_rn(totals[i]);
    totals[i] = 0;
}

size_t alignment_scratch_bytes(int bands, int reach) {
    return (size_t)bands * (2 * reach + 2) * sizeof(uint64_t);
}

hipError_t launch_alignment(const uint8_t *left, const uint8_t *right, const float *disparity,
                            int W, int H, int bands, int reach, int row_step, uint64_t *totals,
                            float *sums, hipStream_t stream) {
    auto *t = reinterpret_cast<unsigned long long *>(totals);
    const int rows = (H + row_step - 1) / row_step, count = bands * (2 * reach + 2);
    if (W / bands > 0)
        hipLaunchKernelGGL(alignment_chunks_kernel, dim3(bands, (rows + ALIGN_ROWS - 1) / ALIGN_ROWS),
                           dim3(256), 0, stream, left, right, disparity, W, H, bands, reach,
                           row_step, t);
    hipLaunchKernelGGL(alignment_finish_kernel, dim3((count + 255) / 256), dim3(256), 0, stream, t,
                       count, sums);
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------------
// Decode (stage/decode.ex:24-57; image.ex:402-476, resize_area): the f32 area resize of both
// frames to the working size. The reference resizes rows first, into an f32 image of the
// frame's width, then columns; a thread makes one output pixel, recomputing the row pass at
// its column taps with the same operations in the same order (taps in order, the first
// product starting each sum), so the f32 values are the reference's. Then rounded half away
// from zero and clipped to 0..255. Grid y: the image (0 left, 1 right).

__global__ __launch_bounds__(256) void resize_area_kernel(const uint8_t *left,
                                                          const uint8_t *right, int fw, int W,
                                                          int H, AreaTaps rows, AreaTaps cols,
                                                          uint8_t *left_out, uint8_t *right_out) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= W * H) return;
    const int y = i / W, x = i - y * W;
    const uint8_t *frame = blockIdx.y ? right : left;
    const int rn = rows.count[y], cn = cols.count[x];
    const int *ri = rows.index + y * rows.most, *ci = cols.index + x * cols.most;
    const float *rw = rows.weight + y * rows.most, *cw = cols.weight + x * cols.most;
    float total = 0.0f;
    for (int k = 0; k < cn; k++) {
        const uint8_t *column = frame + ci[k];
        float t = 0.0f;
        for (int j = 0; j < rn; j++) {
            const float product = rw[j] * (float)column[(size_t)ri[j] * fw];
            t = j == 0 ? product : t + product;
        }
        const float product = cw[k] * t;
        total = k == 0 ? product : total + product;
    }
    const float r = fminf(fmaxf(roundf(total), 0.0f), 255.0f);
    (blockIdx.y ? right_out : left_out)[i] = (uint8_t)r;
}

hipError_t launch_resize_area(const uint8_t *left, const uint8_t *right, int fw, int W, int H,
                              const AreaTaps &rows, const AreaTaps &cols, uint8_t *left_out,
                              uint8_t *right_out, hipStream_t stream) {
    hipLaunchKernelGGL(resize_area_kernel, dim3((W * H + 255) / 256, 2), dim3(256), 0, stream,
                       left, right, fw, W, H, rows, cols, left_out, right_out);
    return hipGetLastError();
}

hipError_t launch_fill_f32(float *dst, float value, size_t count, hipStream_t stream) {
    if (count == 0) return hipSuccess;
    hipLaunchKernelGGL(fill_f32_kernel, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, stream,
                       dst, value, count);
    return hipGetLastError();
}

}  // namespace nxs
