/*
 * nxs_hip: the C ABI of the MI355X (gfx950) NxStereo segment backend.
 *
 * It replaces the native stereo_match process that NxStereo.NativeSegment
 * drives over an Erlang Port (lib/nx_stereo/native_segment.ex,
 * lib/nx_stereo/matcher.ex:248-251), in-process: the records are those of
 * native/stereo_match/src/protocol.rs:1-23, byte for byte, so the Elixir
 * side's encoder and decoder (native_segment.ex:256-386) serve unchanged.
 *
 *   request  1 configure  -> reply tag 7 (configured) or 9 (refused, UTF-8 reason)
 *   request  2 run        -> reply tag 8 (member times in us, output tensors) or 9
 *   request  3 release    -> no reply record (the ticket completes with an empty reply)
 *
 * Tensors: name, type (0 u8, 1 u16, 2 f32, 3 s32), rank, u32 dimensions,
 * row-major little-endian elements (tensor.rs:25-40). A frame or image entry may also come
 * as type 4, u32 {height, width / 4}: four pixels a word, pixel 0 in the low byte, as
 * NxSegment's pipelines without byte storage pack them (packing.ex; segment.ex:207-241).
 * It is read as the u8 image it holds. Besides segment.rs's kinds, "decode"
 * (stage/decode.ex: both frames area-resized to the working size) is built.
 *
 * Threading: every function may be called from any thread. nxs_submit copies
 * the record and returns at once; a context-owned thread computes requests
 * in submission order per segment and posts completions. The process never
 * aborts on a request: configuration errors and run errors become tag 9.
 */
#ifndef NXS_HIP_H
#define NXS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct nxs_ctx nxs_ctx;

/* Status codes of the calls below. */
#define NXS_OK 0
#define NXS_TIMEOUT 1
#define NXS_EINVAL (-1)
#define NXS_EDEVICE (-2)
#define NXS_ENOMEM (-3)
#define NXS_EUNKNOWN (-4)

/* A completed request's reply record, valid only during the call. */
typedef void (*nxs_done_fn)(void *user, int64_t ticket, const uint8_t *reply, size_t len);

/* Opens a context on HIP device `device`: the stream owner and device-memory
 * owner of every segment configured through it. Replaces opening the Port
 * (matcher.ex:248-255). On failure writes a reason into err. */
int nxs_open(int device, nxs_ctx **out, char *err, size_t errlen);

/* Fails every pending ticket, releases every segment, frees the device
 * memory and joins the context's thread (matcher.ex:148-160: the owner's
 * exit releases its segments). */
void nxs_close(nxs_ctx *ctx);

/* Queues one request record (protocol.rs:107-168). Returns a ticket >= 1, or
 * a negative status. With `done`, the reply is posted to done(user, ...)
 * from the context's thread; without, it is kept for nxs_wait/nxs_reply.
 * Replaces Matcher.send_run/6 + the Port write (matcher.ex:66-115). */
int64_t nxs_submit(nxs_ctx *ctx, const uint8_t *record, size_t len, nxs_done_fn done,
                   void *user);

/* Waits for `ticket` (submitted without a callback): NXS_OK once done,
 * NXS_TIMEOUT after timeout_ms (< 0 waits forever), NXS_EINVAL if unknown. */
int nxs_wait(nxs_ctx *ctx, int64_t ticket, int timeout_ms);

/* The kept reply of a completed ticket; valid until nxs_release_reply. */
int nxs_reply(nxs_ctx *ctx, int64_t ticket, const uint8_t **reply, size_t *len);
void nxs_release_reply(nxs_ctx *ctx, int64_t ticket);

/* How many kept replies are complete and not yet released (non-blocking). */
int nxs_poll(nxs_ctx *ctx, int max);

/* Serving: runs of different segments overlap on the device. Each run holds
 * one of its segment's run slots (3 pairs in flight per segment) and, with a
 * Disparity member, a matcher workspace (costs and path volumes, about
 * 5*W*H*D bytes) from a pool whose HBM budget this sets: a run waits for its
 * workspace when the budget is spent, and is admitted when a run in flight
 * completes. The default is half the device memory free at nxs_open, or
 * NXS_HBM_BUDGET_MB. No counterpart in the reference (the Port computes one
 * request at a time). */
int nxs_set_hbm_budget(nxs_ctx *ctx, size_t bytes);
/* The workspace bytes allocated and the runs in flight, for monitoring. */
int nxs_stats(nxs_ctx *ctx, size_t *workspace_bytes, int *runs_in_flight);

/* Coalescing: runs of other segments queued behind the run being started,
 * whose first Disparity member searches the full range at the same size,
 * disparities and paths, share its matcher launch (one descriptor table),
 * up to `max_pairs` a launch (at most 16; 0 or 1 turns it off; negative
 * leaves the setting). A run joins only when its run slot and workspace are
 * free at once, never ahead of an earlier request naming its segment; its
 * stages before and after the matcher stay on its segment's stream, and its
 * results are those of a run alone. The default is 8, or NXS_COALESCE.
 * `launches` and `pairs` (either may be NULL) receive how many coalesced
 * launches ran and the runs in them. No counterpart in the reference (the
 * Port computes one request at a time). */
int nxs_coalescing(nxs_ctx *ctx, int max_pairs, int64_t *launches, int64_t *pairs);

/* ---- The matcher on device buffers (the hot path, without records) ----
 *
 * Stage.Disparity's computation (stage/disparity.ex:135-163, full search;
 * Matcher::disparity, native/stereo_match/src/lib.rs:107-144): u8 left/right
 * {height,width} in device memory -> f32 disparity {height,width}, -1.0 where
 * unmatched, enqueued on `stream` (a hipStream_t; NULL for the null stream).
 * disparities in {64, 128, 192, 256} (the reference takes 64 and 128),
 * paths in {3, 8}, 0 <= p1 < p2 <= 79, uniqueness >= 1, 8 < width <= 4096,
 * height > 6. `workspace` is device memory of nxs_match_workspace_size bytes,
 * owned by the caller; calls that overlap must not share one. */
size_t nxs_match_workspace_size(int width, int height, int disparities, int paths);
int nxs_match_device(const uint8_t *left, const uint8_t *right, int width, int height,
                     int disparities, int paths, int p1, int p2, float uniqueness,
                     float *disparity, void *workspace, size_t workspace_bytes, void *stream);

/* The matcher of nxs_match_device with no workspace argument and no allocation: for
 * callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc). No counterpart in the reference. Its
 * workspace is a scratch array of NXS_NOALLOC_SCRATCH_BYTES in the library's device
 * memory, one per device, held from the library's load until the process ends; sizes
 * whose nxs_match_workspace_size exceeds it are refused (NXS_EINVAL, with the reason).
 * It calls no hipMalloc or hipFree. Every call on one device shares that device's array:
 * calls that overlap on a device (on different streams, or from different threads
 * without ordering) race and give wrong results; calls on one stream are ordered. */
#define NXS_NOALLOC_SCRATCH_BYTES (64u << 20)
int nxs_match_device_noalloc(const uint8_t *left, const uint8_t *right, int width, int height,
                             int disparities, int paths, int p1, int p2, float uniqueness,
                             float *disparity, void *stream);

/* The same on host buffers, returning the intermediate products for
 * inspection (any output pointer may be NULL): census features u64 {H,W}
 * of each image, matching costs u8 {H,W,D}, aggregated path costs u16
 * {H,W,D}, the selected disparity f32 and key u32 {H,W} before the
 * left-right check, the checked disparity f32 {H,W}, and the stage's output.
 * Synchronous; allocates and frees its own device memory. */
int nxs_match_host(const uint8_t *left, const uint8_t *right, int width, int height,
                   int disparities, int paths, int p1, int p2, float uniqueness,
                   uint64_t *census_left, uint64_t *census_right, uint8_t *costs,
                   uint16_t *sums, float *selected, uint32_t *keys, float *checked,
                   float *disparity);

/* ---- Band search and coarse-to-fine ----
 *
 * NxStereo.sgm_band/4 (lib/nx_stereo.ex:262-322): each pixel searches only the `band`
 * disparities [base, base + band) of its s32 base {height,width} (device memory; a negative
 * base: no estimate, no match). u8 images as nxs_match_device; disparities 64 or 128; band
 * 8, 16, 32 or 64; census 0 (sparse_9x7), 1 (dense_9x5) or 2 (dense_9x5_blend). The output
 * is sgm_band/4's: the 3x3 median of the checked map (a match at 0 stays 0). Bases above
 * 65536 act as 65536. `workspace`: nxs_band_workspace_size bytes of device memory. */
size_t nxs_band_workspace_size(int width, int height, int band, int paths);
int nxs_match_band_device(const uint8_t *left, const uint8_t *right, const int32_t *base,
                          int width, int height, int disparities, int band, int paths,
                          int census, int p1, int p2, float uniqueness, float *disparity,
                          void *workspace, size_t workspace_bytes, void *stream);

/* The same on host buffers with the intermediate products (any output may be NULL): census
 * u64 {H,W} of each image, band costs u8 {H,W,band}, summed path costs u16 {H,W,band}, the
 * selected map f32 and keys u32 before the check, the checked map, the output. */
int nxs_match_band_host(const uint8_t *left, const uint8_t *right, const int32_t *base,
                        int width, int height, int disparities, int band, int paths, int census,
                        int p1, int p2, float uniqueness, uint64_t *census_left,
                        uint64_t *census_right, uint8_t *costs, uint16_t *sums, float *selected,
                        uint32_t *keys, float *checked, float *disparity);

/* NxStereo.sgm_coarse_to_fine/3 (lib/nx_stereo.ex:164-259): both images shrunk to half size
 * (rounded 2x2 average) and matched over 64 disparities with coarse_paths; each pixel's base
 * is twice the lowest match of its coarse 3x3 neighbourhood less `margin` (0..127), clipped
 * to [0, 128 - band]; then the band search at 128 disparities. Even width and height, and
 * halves that nxs_match_device takes. The output is sgm_coarse_to_fine/3's (the Disparity
 * stage then maps values <= 0 to -1). */
size_t nxs_coarse_to_fine_workspace_size(int width, int height, int band, int paths,
                                         int coarse_paths);
int nxs_coarse_to_fine_device(const uint8_t *left, const uint8_t *right, int width, int height,
                              int band, int margin, int paths, int coarse_paths, int census,
                              int p1, int p2, float uniqueness, float *disparity, void *workspace,
                              size_t workspace_bytes, void *stream);
/* On host buffers, with the half images u8 {H/2,W/2}, the coarse map f32 {H/2,W/2} and the
 * bases s32 {H,W} (any output may be NULL). */
int nxs_coarse_to_fine_host(const uint8_t *left, const uint8_t *right, int width, int height,
                            int band, int margin, int paths, int coarse_paths, int census, int p1,
                            int p2, float uniqueness, uint8_t *half_left, uint8_t *half_right,
                            float *coarse, int32_t *base, float *disparity);

/* ---- The batched form ----
 *
 * Many pairs of uneven sizes and disparity ranges through the matcher
 * (as nxs_match_device) in one graph: pairs are grouped into size classes
 * (width, height, disparities, paths) and read their settings from a
 * descriptor table in device memory; the whole batch is captured once into
 * a hipGraph that nxs_batch_run replays. `mode`:
 *   NXS_BATCH_DEFAULT: each kernel one launch per class, with the class
 *     size's own aggregation schedule, the classes on two parallel branches
 *     of the graph;
 *   NXS_BATCH_TUNED: each class's aggregation schedule (prefetch depth
 *     2/4/8, lanes per line 16/32/64, one or two waves per line, the two
 *     waves meeting through memory or by recomputation, orientations in one
 *     launch or one each) is timed on the batch itself and the fastest kept
 *     for the batch (single-pair launches of that size keep their own
 *     schedule); then the plans (one, two or four branches, each class in
 *     one launch or in launches of 8 or 2 pairs, and the fused plan below
 *     with checkpoints every 4 or 8 steps, on one or two branches, its
 *     aggregation in one block per line group or in persistent blocks on a
 *     device work queue) are timed and the fastest kept;
 *   NXS_BATCH_FUSED: every class in one launch per kernel (census and
 *     costs, aggregation, selection, check and median: four launches for the
 *     batch), the aggregation as two waves of 16 lanes a line meeting by
 *     recomputation from checkpoints every 8 steps; up to 8 size classes.
 * The results are the same in every mode. The batch owns one matcher
 * workspace per pair; the pairs' images and outputs stay the caller's, read
 * and written at every replay. No counterpart in the reference, whose
 * matcher takes one pair at a time. */
#define NXS_BATCH_DEFAULT 0
#define NXS_BATCH_TUNED 1
#define NXS_BATCH_FUSED 2
typedef struct nxs_pair {
    const uint8_t *left, *right; /* device, u8 {height, width} */
    float *disparity;            /* device, f32 {height, width} */
    int width, height, disparities, paths, p1, p2;
    float uniqueness;
} nxs_pair;

typedef struct nxs_batch nxs_batch;

/* The aggregation schedule matcher launches of one size use: the autotuner's
 * pick, or what nxs_schedule_set says. Results are the same under every
 * schedule; only the time differs. */
typedef struct nxs_schedule {
    int prefetch; /* steps of costs loaded ahead: 2, 4 or 8 */
    int split;    /* 1: one launch per line orientation */
    int lanes;    /* lanes sharing one line: 16, 32 or 64, where disparities / lanes is
                     2, 4, 8, 12 or 16 */
    int waves;    /* waves per line (16 lanes, or 32 with 2 waves at 64, 128 or 256
                     disparities): 2, a wave per direction of a line (at
                     3 paths, where columns go down only, four columns a wave); 4 or 8
                     (8 paths only), each direction's disparities over 2 or 4 waves as
                     well, where disparities / (8 x waves) is 2, 4 or 8. Without a pick:
                     2 waves of 16 lanes, prefetch 8 (4 at 192 disparities); at 128
                     disparities and 8 paths, 2 waves of 32 lanes, prefetch 4 up to
                     500,000 pixels and 8 above; at 8 paths where 4 x width x height x
                     disparities passes 256 MB, 2 waves of 16 lanes meeting by
                     recomputation, checkpoints every 8 steps */
    int meet;     /* with 2 waves: 0, each direction stores its bytes of the line's first
                     half and the other reads them back; 1, each direction stores its path
                     every `prefetch` steps of its first half and the other recomputes the
                     bytes from there (a third of the traffic, one more recurrence step per
                     pixel of the second half) */
} nxs_schedule;

int nxs_schedule_set(int width, int height, int disparities, int paths, const nxs_schedule *s);
int nxs_schedule_get(int width, int height, int disparities, int paths, nxs_schedule *s);

int nxs_batch_create(const nxs_pair *pairs, int count, int mode, nxs_batch **out);

/* A batch's plan, as nxs_batch_describe's first line names it. */
typedef struct nxs_batch_plan {
    int branches; /* 1, 2 or 4 parallel branches of the graph */
    int chunk;    /* pairs a launch, 0: each class in one launch (unused when fused) */
    int fused;    /* 0; or 4 or 8: each branch's classes (whole classes, the largest first to
                     the least loaded branch; at most 8 classes) in one launch per kernel,
                     checkpoints every `fused` steps */
    int queue;    /* with fused: 1 = the aggregation launch in persistent blocks (as many as
                     the device holds at once) taking its line groups from a device work
                     queue, 0 = one block per line group */
} nxs_batch_plan;

/* nxs_batch_create with the plan pinned, untuned, and each class's schedule its
 * size's (nxs_schedule_get): for a caller that has tuned once and keeps the plan it
 * picked, so that a later batch of the same mix skips tuning. No counterpart in the
 * reference. */
int nxs_batch_create_planned(const nxs_pair *pairs, int count, const nxs_batch_plan *plan,
                             nxs_batch **out);
int nxs_batch_run(nxs_batch *batch, void *stream);
/* The classes and their schedules, one line each; valid while the batch lives. */
const char *nxs_batch_describe(nxs_batch *batch);
void nxs_batch_destroy(nxs_batch *batch);

/* ---- A configured segment on device-resident entries ----
 *
 * Segment::run (native/stereo_match/src/segment.rs:206-239) with the entry
 * tensors already in device memory: `record` is a run record (kind 2) of a
 * segment configured through nxs_submit, holding each member's live inputs
 * (its entry count may be 0); `entries` are device tensors read in place
 * (a disparity entry is copied first, since filters change it in place).
 * Every member is enqueued on `stream` (NULL: the segment's own stream) and
 * the call returns without waiting. `outputs` receives up to max_outputs
 * filled tensors in first-filled order, as device memory of the segment,
 * valid until its next run; *n_outputs is their count. Run errors return
 * NXS_EINVAL with the reason in nxs_last_error. No counterpart in the
 * reference, whose runs always cross a pipe: it serves callers whose frames
 * are already on the GPU (the benchmark, the batched and served forms). */
typedef struct nxs_tensor {
    const char *name;
    uint8_t type; /* 0 u8, 1 u16, 2 f32, 3 s32, 4 u32 (a packed frame or image) */
    uint8_t rank;
    uint32_t shape[4];
    void *data;
} nxs_tensor;

int nxs_run_device(nxs_ctx *ctx, const uint8_t *record, size_t len, const nxs_tensor *entries,
                   int n_entries, nxs_tensor *outputs, int max_outputs, int *n_outputs,
                   void *stream);

/* ---- Kernel timing ----
 * Between start and stop, every matcher launch is bracketed by HIP events on
 * its own stream (up to `capacity` launches). nxs_profile_stop waits for
 * them and adds up milliseconds and launch counts per kernel, in the order
 * census_cost, path (the aggregation), select, check_median (the left-right check and the
 * median), band_cost, band_path, band_select, shrink (coarse-to-fine's shrink and band
 * bases). No
 * counterpart in the reference (test/bench/pipeline.exs --kernels times its
 * Vulkan kernels one by one). */
int nxs_profile_start(int capacity);
int nxs_profile_stop(double *total_ms, int *launches, int kernels);

/* Why the last failing call on this thread failed (a static buffer per thread). */
const char *nxs_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
