// The segment runtime behind the C ABI (include/nxs_hip.h): the request
// records of native/stereo_match/src/protocol.rs, the configured stage lists
// of segment.rs, device arenas and streams per segment, and a context thread
// that computes requests in order and posts replies.
#include "../../include/nxs_hip.h"

#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "kernels.hpp"
#include "profile.hpp"

namespace nxs {

static thread_local std::string g_last_error;

static int fail(int code, const std::string &why) {
    g_last_error = why;
    return code;
}

// ------------------------------------------------------------------------------------------
// Kernel timing (profile.hpp): a pool of event pairs, filled in launch order.

struct Profiler {
    std::mutex mu;
    bool on = false;
    int used = 0;
    std::vector<hipEvent_t> begin, end;
    std::vector<int> kernel;
};
static Profiler g_profiler;
// Launches captured into a graph are not bracketed by profiling events.
static thread_local bool t_capturing = false;

int profile_begin(ProfiledKernel k, hipStream_t stream) {
    if (t_capturing) return -1;
    std::lock_guard<std::mutex> lock(g_profiler.mu);
    if (!g_profiler.on || g_profiler.used >= (int)g_profiler.begin.size()) return -1;
    const int slot = g_profiler.used++;
    g_profiler.kernel[slot] = k;
    if (hipEventRecord(g_profiler.begin[slot], stream) != hipSuccess) return -1;
    return slot;
}

void profile_end(int slot, hipStream_t stream) {
    if (slot >= 0) (void)hipEventRecord(g_profiler.end[slot], stream);
}

// ------------------------------------------------------------------------------------------
// Aggregation schedules per size class (kernels.hpp): the autotuner's picks.

static std::mutex g_schedules_mu;
static std::map<std::tuple<int, int, int, int>, Schedule> g_schedules;

// Without a pick: two waves per line (one per direction), which halves a line's sequential
// walk, loading 8 steps ahead; at 192 disparities (12 bytes a lane) 4, where eight steps'
// registers cost occupancy. Measured on MI355X at 8 paths: 728x544x128 170 us (prefetch 4:
// 179), 1280x720x192 637 us (prefetch 8: 666), 1920x1080x256 fastest with 8, 960x540x64
// even.
static Schedule default_schedule(int width, int height, int disparities, int paths) {
    Schedule s;
    // At 3 paths the rows' walk bounds the kernel (the down-only columns take four lines a
    // wave), and a prefetch of 8 covers it best: 38 us at 364x272x64 (coarse-to-fine's half
    // images) and 92 us at 728x544x128, against 71 and 156 us with one wave a line, and 40
    // and 110 us with 32 lanes a line.
    s.prefetch = disparities == 192 ? 4 : 8;
    s.lanes = 16;
    s.waves = 2;
    // At 8 paths and 128 disparities, 32 lanes a line (half the instructions a step in each
    // wave, twice the waves) walks the lines as fast or faster: 364x272 51-54 against 68 us,
    // 728x544 159 against 169-172, 1024x768 332 against 348, 640x480, 1280x720 and 1920x1080
    // the same; a prefetch of 4 is best up to 728x544 and 8 above (1024x768: 368 against
    // 332). At 64 disparities 16 lanes stay (640x480: 76 against 103-105 us; equal or a
    // little slower below 400x300, though 480x270 alone walks 53 against 91 us with 32). At
    // 256 the kernel is bound by its traffic (1920x1080: 1916 against 1981 us with 32).
    if (paths == 8 && disparities == 128) {
        s.lanes = 32;
        s.prefetch = (size_t)width * height <= 500000 ? 4 : 8;
    }
    // At 8 paths, where the four orientations' byte volumes pass the 256 MB Infinity Cache,
    // the two waves meet by recomputation from a checkpoint every 8 steps, blocks longest
    // lines first: the path kernel then moves about 13N bytes instead of 20N, for about half
    // again the recurrence work. Measured on MI355X: 1920x1080 at 64, 128 and 256
    // disparities 520, 712 and 1207 us against 544, 924 and 1824; 1280x720x192 457 against
    // 605; 1024x768x128 302 against 346. Where the volumes stay in the cache the recurrence
    // work costs more than the traffic saves (728x544x128 198 against 160 us, 960x540x64 139
    // against 123).
    if (paths == 8 && (size_t)4 * width * height * disparities > ((size_t)256 << 20)) {
        s.lanes = 16;
        s.prefetch = 8;
        s.meet = 1;
    }
    return s;
}

Schedule schedule_for(int width, int height, int disparities, int paths) {
    std::lock_guard<std::mutex> lock(g_schedules_mu);
    auto it = g_schedules.find(std::make_tuple(width, height, disparities, paths));
    return it == g_schedules.end() ? default_schedule(width, height, disparities, paths)
                                   : it->second;
}

void remember_schedule(int width, int height, int disparities, int paths, const Schedule &s) {
    std::lock_guard<std::mutex> lock(g_schedules_mu);
    g_schedules[std::make_tuple(width, height, disparities, paths)] = s;
}

// ------------------------------------------------------------------------------------------
// Records (protocol.rs:1-23, 55-168; tensor.rs:25-75)

enum : uint8_t { CONFIGURE = 1, RUN = 2, RELEASE = 3 };
enum : uint8_t { TAG_CONFIGURED = 7, TAG_OUTPUTS = 8, TAG_REFUSED = 9 };
// tensor.rs:25-40's element types, and u32 (4): frames and images packed as words of four
// pixels, pixel 0 in the low byte (packing.ex), as NxSegment's pipelines without byte
// storage hold them.
enum : uint8_t { TYPE_U8 = 0, TYPE_U16 = 1, TYPE_F32 = 2, TYPE_S32 = 3, TYPE_U32 = 4 };

static int element_bytes(uint8_t type) {
    switch (type) {
        case TYPE_U8: return 1;
        case TYPE_U16: return 2;
        case TYPE_F32:
        case TYPE_S32:
        case TYPE_U32: return 4;
        default: return -1;
    }
}

struct Error {
    std::string reason;
};

struct Reader {
    const uint8_t *p;
    size_t n;

    const uint8_t *take(size_t count) {
        if (count > n) {
            std::ostringstream s;
            s << "the request ends " << (count - n) << " bytes early";
            throw Error{s.str()};
        }
        const uint8_t *t = p;
        p += count;
        n -= count;
        return t;
    }
    uint8_t u8() { return *take(1); }
    uint16_t u16() {
        const uint8_t *b = take(2);
        return (uint16_t)(b[0] | (b[1] << 8));
    }
    uint32_t u32() {
        const uint8_t *b = take(4);
        return (uint32_t)b[0] | ((uint32_t)b[1] << 8) | ((uint32_t)b[2] << 16) | ((uint32_t)b[3] << 24);
    }
    double f64() {
        double v;
        std::memcpy(&v, take(8), 8);
        return v;
    }
    std::string name() {
        size_t length = u8();
        const uint8_t *b = take(length);
        return std::string(reinterpret_cast<const char *>(b), length);
    }
};

struct HostTensor {
    uint8_t type = 0;
    std::vector<uint32_t> shape;
    const uint8_t *bytes = nullptr;  // into the request record
    size_t size = 0;
};

using Options = std::map<std::string, double>;
using Inputs = std::map<std::string, std::vector<float>>;

struct Geometry {
    int width, height, frame_width, frame_height;
};

struct Request {
    uint8_t kind;
    uint32_t segment;
    Geometry geometry{};
    std::vector<std::pair<std::string, Options>> members;
    std::vector<Inputs> inputs;
    std::vector<std::pair<std::string, HostTensor>> entries;
};

static Request parse(const uint8_t *record, size_t len) {
    Reader r{record, len};
    Request q;
    q.kind = r.u8();
    q.segment = r.u32();
    if (q.kind == CONFIGURE) {
        q.geometry.width = r.u16();
        q.geometry.height = r.u16();
        q.geometry.frame_width = r.u16();
        q.geometry.frame_height = r.u16();
        int count = r.u8();
        for (int m = 0; m < count; m++) {
            std::string kind = r.name();
            Options options;
            int n = r.u8();
            for (int i = 0; i < n; i++) {
                std::string key = r.name();
                options[key] = r.f64();
            }
            q.members.emplace_back(kind, options);
        }
    } else if (q.kind == RUN) {
        int count = r.u8();
        for (int m = 0; m < count; m++) {
            Inputs inputs;
            int n = r.u8();
            for (int i = 0; i < n; i++) {
                std::string key = r.name();
                int values = r.u16();
                const uint8_t *b = r.take(4 * (size_t)values);
                std::vector<float> v(values);
                std::memcpy(v.data(), b, 4 * (size_t)values);
                inputs[key] = std::move(v);
            }
            q.inputs.push_back(std::move(inputs));
        }
        int entries = r.u8();
        for (int e = 0; e < entries; e++) {
            std::string key = r.name();
            HostTensor t;
            t.type = r.u8();
            int rank = r.u8();
            size_t elements = 1;
            for (int d = 0; d < rank; d++) {
                t.shape.push_back(r.u32());
                elements *= t.shape.back();
            }
            int eb = element_bytes(t.type);
            if (eb < 0) throw Error{"unknown element type " + std::to_string(t.type)};
            t.size = elements * eb;
            t.bytes = r.take(t.size);
            bool replaced = false;
            for (auto &existing : q.entries)
                if (existing.first == key) {
                    existing.second = t;
                    replaced = true;
                }
            if (!replaced) q.entries.emplace_back(key, t);
        }
    } else if (q.kind != RELEASE) {
        throw Error{"unknown request kind " + std::to_string(q.kind)};
    }
    if (r.n != 0) throw Error{std::to_string(r.n) + " bytes follow the request"};
    return q;
}

static std::vector<uint8_t> refused_record(const std::string &reason) {
    std::vector<uint8_t> v;
    v.push_back(TAG_REFUSED);
    v.insert(v.end(), reason.begin(), reason.end());
    return v;
}

static std::string shape_text(const std::vector<uint32_t> &shape) {
    std::ostringstream s;
    s << "[";
    for (size_t i = 0; i < shape.size(); i++) s << (i ? ", " : "") << shape[i];
    s << "]";
    return s.str();
}

static std::string number_text(double v) {
    std::ostringstream s;
    if (v == std::floor(v) && std::fabs(v) < 1e15) s << (long long)v;
    else s << v;
    return s.str();
}

// ------------------------------------------------------------------------------------------
// Stages (segment.rs:244-333): kinds, build settings and the constants they share with Nx.

enum class Kind {
    Decode,
    Rectification,
    Sharpness,
    Disparity,
    MedianFilter,
    SpeckleFilter,
    RowAlignment,
    TemporalFilter,
    Depth,
    PointCloud,
    Quality
};

struct Member {
    Kind kind;
    std::string name;
    int disparities = 0, paths = 0;
    // Disparity with search coarse_to_fine (sgm_coarse_to_fine/3): the half-size match's
    // paths, the fine band, its margin and census window (0 sparse_9x7, 1 dense_9x5,
    // 2 dense_9x5_blend).
    bool coarse_to_fine = false;
    int coarse_paths = 0, band = 0, margin = 0, fine_census = 0;
    int radius = 0;
    int kernel_size = 0;
    int bands = 0, reach = 0, row_step = 0;
    int decimation = 0;
    float nearest = 0, farthest = 0;
    float max_missed = 0;
};

static double option(const Options &o, const std::string &kind, const char *name) {
    auto it = o.find(name);
    if (it == o.end()) throw Error{kind + " needs its " + name + " setting"};
    return it->second;
}

static void takes_only(const Options &o, const std::string &kind,
                       std::initializer_list<const char *> names) {
    std::vector<std::string> unknown;
    for (auto &kv : o) {
        bool known = false;
        for (const char *n : names) known = known || kv.first == n;
        if (!known) unknown.push_back(kv.first);
    }
    if (unknown.empty()) return;
    std::string list;
    for (size_t i = 0; i < unknown.size(); i++) list += (i ? ", " : "") + unknown[i];
    throw Error{kind + " has no setting " + list};
}

static int count_option(const Options &o, const std::string &kind, const char *name, int lowest,
                        int highest) {
    double v = option(o, kind, name);
    if (v == std::floor(v) && v >= lowest && v <= highest) return (int)v;
    throw Error{kind + " " + name + " must be an integer from " + std::to_string(lowest) + " to " +
                std::to_string(highest) + ", got " + number_text(v)};
}

static int integer_option(const Options &o, const std::string &kind, const char *name,
                          std::initializer_list<int> allowed) {
    double v = option(o, kind, name);
    std::string list;
    for (int a : allowed) {
        if ((double)a == v) return a;
        list += (list.empty() ? "" : ", ") + std::to_string(a);
    }
    throw Error{kind + " " + name + " must be one of [" + list + "], got " + number_text(v)};
}

// Parameters::check (lib.rs:41-73), with 192 and 256 disparities added.
static void check_matching(int W, int H, int D, int paths, int p1, int p2, float uniqueness) {
    if (!(W > 8 && H > 6 && W <= 4096))
        throw Error{"expected an image wider than 8, higher than 6 and at most 4096 wide, got " +
                    std::to_string(W) + "x" + std::to_string(H)};
    if (D != 64 && D != 128 && D != 192 && D != 256)
        throw Error{"disparities must be 64, 128, 192 or 256, got " + std::to_string(D)};
    if (paths != 3 && paths != 8) throw Error{"paths must be 3 or 8, got " + std::to_string(paths)};
    // The aggregation addresses each volume with 32-bit buffer offsets, and keeps 2^31 and
    // above as the offset of a masked access.
    if ((uint64_t)W * H * D >= (1ull << 31))
        throw Error{"the cost volume (width x height x disparities) must stay under 2^31 bytes, got " +
                    std::to_string(W) + "x" + std::to_string(H) + "x" + std::to_string(D)};
    if (!(0 <= p1 && p1 < p2 && p2 + 48 <= 127))
        throw Error{"expected 0 <= p1 < p2 <= 79, got " + std::to_string(p1) + " and " +
                    std::to_string(p2)};
    if (std::isnan(uniqueness) || uniqueness < 1.0f) {
        std::ostringstream s;
        s << "uniqueness must be at least 1.0, got " << uniqueness;
        throw Error{s.str()};
    }
}

// sgm_band/4's checks (nx_stereo.ex:88-131, 286-308): the matching's, with 64 or 128
// disparities, a band of 8, 16, 32 or 64 and a census window index.
static void check_band(int W, int H, int D, int band, int paths, int census, int p1, int p2,
                       float uniqueness) {
    check_matching(W, H, D, paths, p1, p2, uniqueness);
    if (D != 64 && D != 128) throw Error{"disparities must be 64 or 128, got " + std::to_string(D)};
    if (band != 8 && band != 16 && band != 32 && band != 64)
        throw Error{"band must be 8, 16, 32 or 64, got " + std::to_string(band)};
    if (census < 0 || census > 2)
        throw Error{"census must be 0 (sparse_9x7), 1 (dense_9x5) or 2 (dense_9x5_blend), got " +
                    std::to_string(census)};
}

// sgm_coarse_to_fine/3's (nx_stereo.ex:164-185): even sizes whose halves match at 64.
static void check_coarse_to_fine(int W, int H, int band, int margin, int paths, int coarse_paths,
                                 int census, int p1, int p2, float uniqueness) {
    check_band(W, H, 128, band, paths, census, p1, p2, uniqueness);
    if ((W | H) & 1)
        throw Error{"coarse_to_fine expects an even height and width, got " + std::to_string(W) +
                    "x" + std::to_string(H)};
    check_matching(W / 2, H / 2, 64, coarse_paths, p1, p2, uniqueness);
    if (margin < 0 || margin > 127)
        throw Error{"margin must be an integer from 0 to 127, got " + std::to_string(margin)};
}

static BandParams band_params(int W, int H, int D, int band, int paths, int census, int p1,
                              int p2, float uniqueness) {
    BandParams bp;
    bp.width = W;
    bp.height = H;
    bp.disparities = D;
    bp.band = band;
    bp.paths = paths;
    bp.census = census;
    bp.p1 = p1;
    bp.p2 = p2;
    bp.uniqueness = uniqueness;
    return bp;
}


static Member make_member(const Geometry &g, const std::string &kind, const Options &o) {
    Member m;
    m.name = kind;
    const int W = g.width, H = g.height;
    if (kind == "decode") {
        // stage/decode.ex:18-21 takes no settings; image.ex:410-412 only shrinks.
        takes_only(o, kind, {});
        if (g.frame_width < W || g.frame_height < H)
            throw Error{"resize_area only shrinks, got " + std::to_string(g.frame_height) + "x" +
                        std::to_string(g.frame_width) + " to " + std::to_string(H) + "x" +
                        std::to_string(W)};
        m.kind = Kind::Decode;
    } else if (kind == "rectification") {
        takes_only(o, kind, {});
        if (g.frame_width < W || g.frame_height < H)
            throw Error{"rectification only shrinks: " + std::to_string(g.frame_width) + "x" +
                        std::to_string(g.frame_height) + " frames are smaller than the " +
                        std::to_string(W) + "x" + std::to_string(H) + " working size"};
        m.kind = Kind::Rectification;
    } else if (kind == "disparity") {
        takes_only(o, kind, {"num_disparities", "paths", "search", "coarse_paths", "band", "margin",
                             "fine_census"});
        m.disparities = integer_option(o, kind, "num_disparities", {64, 128, 192, 256});
        m.paths = integer_option(o, kind, "paths", {3, 8});
        // search 0: full; 1: coarse_to_fine, with the settings stage/disparity.ex:71-106
        // declares for it and the checks of its init (stage/disparity.ex:121-195).
        m.coarse_to_fine = o.count("search") && integer_option(o, kind, "search", {0, 1}) == 1;
        if (m.coarse_to_fine) {
            if (m.disparities != 128)
                throw Error{"coarse_to_fine searches 128 disparities, got " +
                            std::to_string(m.disparities)};
            if ((W | H) & 1)
                throw Error{"coarse_to_fine needs an even width and height, got " +
                            std::to_string(W) + "x" + std::to_string(H)};
            m.coarse_paths =
                o.count("coarse_paths") ? integer_option(o, kind, "coarse_paths", {3, 8}) : m.paths;
            m.band = o.count("band") ? integer_option(o, kind, "band", {8, 16, 32, 64}) : 16;
            m.margin = o.count("margin") ? count_option(o, kind, "margin", 0, 63) : 4;
            if (m.margin >= m.band)
                throw Error{"margin must be below band (" + std::to_string(m.band) + "), got " +
                            std::to_string(m.margin)};
            m.fine_census = o.count("fine_census") ? integer_option(o, kind, "fine_census", {0, 1, 2}) : 0;
            check_coarse_to_fine(W, H, m.band, m.margin, m.paths, m.coarse_paths, m.fine_census, 8,
                                 64, 1.0f);
        } else {
            if (o.count("fine_census") && option(o, kind, "fine_census") != 0.0)
                throw Error{"fine_census requires coarse_to_fine search"};
            for (const char *only : {"coarse_paths", "band", "margin"})
                if (o.count(only))
                    throw Error{std::string(only) + " requires coarse_to_fine search"};
            check_matching(W, H, m.disparities, m.paths, 8, 64, 1.0f);
        }
        m.kind = Kind::Disparity;
    } else if (kind == "sharpness") {
        takes_only(o, kind, {});
        m.kind = Kind::Sharpness;
    } else if (kind == "median_filter") {
        takes_only(o, kind, {"kernel_size"});
        m.kernel_size = integer_option(o, kind, "kernel_size", {3, 5});
        m.kind = Kind::MedianFilter;
    } else if (kind == "row_alignment") {
        takes_only(o, kind, {"bands", "reach", "row_step"});
        m.bands = count_option(o, kind, "bands", 1, W);
        m.reach = count_option(o, kind, "reach", 0, H);
        m.row_step = count_option(o, kind, "row_step", 1, H);
        m.kind = Kind::RowAlignment;
    } else if (kind == "speckle_filter") {
        takes_only(o, kind, {"radius"});
        m.radius = count_option(o, kind, "radius", 1, std::min(W, H));
        m.kind = Kind::SpeckleFilter;
    } else if (kind == "temporal_filter") {
        takes_only(o, kind, {"max_missed"});
        m.max_missed = (float)option(o, kind, "max_missed");
        m.kind = Kind::TemporalFilter;
    } else if (kind == "depth") {
        takes_only(o, kind, {});
        m.kind = Kind::Depth;
    } else if (kind == "point_cloud") {
        takes_only(o, kind, {"decimation", "min_depth_m", "max_depth_m"});
        m.decimation = count_option(o, kind, "decimation", 1, std::min(W, H));
        m.nearest = (float)option(o, kind, "min_depth_m");
        m.farthest = (float)option(o, kind, "max_depth_m");
        m.kind = Kind::PointCloud;
    } else if (kind == "quality") {
        takes_only(o, kind, {});
        m.kind = Kind::Quality;
    } else {
        throw Error{"no HIP implementation of " + kind};
    }
    return m;
}

static const std::vector<const char *> &member_outputs(Kind k) {
    static const std::vector<const char *> rect = {"left", "right"}, disp = {"disparity"},
                                           sharp = {"sharpness_sums"},
                                           align = {"alignment_sums"}, depth = {"depth_mm"},
                                           cloud = {"cloud_grid", "cloud_points"},
                                           quality = {"quality_sums"};
    switch (k) {
        case Kind::Decode:
        case Kind::Rectification: return rect;
        case Kind::Sharpness: return sharp;
        case Kind::RowAlignment: return align;
        case Kind::Depth: return depth;
        case Kind::PointCloud: return cloud;
        case Kind::Quality: return quality;
        default: return disp;
    }
}

// ------------------------------------------------------------------------------------------
// Segments: configured stage lists with their device memory and carries.

#define HIP_TRY(expr)                                                                       \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess) throw Error{std::string("HIP error: ") + hipGetErrorString(e_)}; \
    } while (0)

struct Slot {
    uint8_t type;
    std::vector<uint32_t> shape;
    void *dev;  // nullptr when the tensor is of a shape or type no stage reads
};

// Per-run resources of a segment, one set per pair in flight: pinned staging both ways
// and the events that time its members.
struct RunSlot {
    uint8_t *upload = nullptr, *download = nullptr;
    size_t upload_bytes = 0, download_bytes = 0;
    // Device staging of the same sizes: entries land there on the segment's upload stream
    // and outputs leave from there on its download stream, so that a pair's transfers
    // overlap the next pair's kernels on the segment's stream.
    uint8_t *device_upload = nullptr, *device_download = nullptr;
    hipEvent_t uploaded = nullptr, computed = nullptr, done = nullptr;
    std::vector<hipEvent_t> events;  // before the first member, then after each
    bool busy = false;
};

// A matcher workspace (costs and path volumes) from the context's pool.
struct Workspace {
    void *ptr = nullptr;
    size_t bytes = 0;
    bool busy = false;
};

constexpr int RUN_SLOTS = 3;  // pairs of one segment in flight at once

struct Segment {
    uint32_t id = 0;
    Geometry g{};
    std::vector<Member> members;
    hipStream_t stream = nullptr;
    hipStream_t upload_stream = nullptr, download_stream = nullptr;  // submitted runs' transfers
    void *arena = nullptr;
    std::mutex mu;  // held while work is enqueued on the segment
    RunSlot slots[RUN_SLOTS];
    int next_slot = 0;
    size_t workspace_bytes = 0;  // of its largest matcher member, 0 without one
    // The first Disparity member when it searches the full range, else -1: served runs of
    // segments whose matchers are of one size class may share that member's launches.
    int matcher = -1;
    // nxs_run_device's matcher workspace: the segment's own, outside the serving pool and
    // its budget, held until the segment is released (a run on a caller's stream may still
    // use it when nxs_run_device returns).
    void *device_workspace = nullptr;
    // Runs of one segment keep submission order across streams: the stream the last run was
    // enqueued on, and `order`, recorded after each run on a stream not the segment's own
    // (and on the segment's own stream before a run on another follows it).
    hipStream_t last_stream = nullptr;
    hipEvent_t order = nullptr;
    // Runs of consecutive per-pixel stages in one launch (launch_tail); NXS_FUSE_TAIL=0 at
    // configure keeps every stage a launch of its own.
    bool fuse_tail = true;
    bool foreign = false;  // a run has been enqueued on a caller's stream

    // device buffers carved from the arena
    uint8_t *frames[2]{}, *images[2]{};
    float *disparity[2]{};
    uint16_t *depth_mm = nullptr;
    float *depth_m = nullptr, *cloud = nullptr, *quality_sums = nullptr;
    int32_t *tail_kept = nullptr;  // the fused tail's kept-point accumulator (left at zero)
    uint64_t *sharpness_totals = nullptr, *alignment_totals = nullptr;
    float *sharpness_sums = nullptr, *alignment_sums = nullptr;
    double *quality_partials = nullptr;
    int32_t *cloud_partials = nullptr, *cloud_kept = nullptr;
    AreaTaps decode_rows{}, decode_cols{};  // Decode's area-resize taps (on the device)
    std::vector<float *> carry_a, carry_b;  // per member: temporal previous/missed, quality previous

    ~Segment() {
        if (upload_stream) (void)hipStreamSynchronize(upload_stream);
        if (stream) (void)hipStreamSynchronize(stream);
        if (download_stream) (void)hipStreamSynchronize(download_stream);
        if (foreign && order) (void)hipEventSynchronize(order);  // the last caller-stream run
        if (order) (void)hipEventDestroy(order);
        if (device_workspace) (void)hipFree(device_workspace);
        for (auto &slot : slots) {
            for (auto e : slot.events) (void)hipEventDestroy(e);
            for (auto e : {slot.uploaded, slot.computed, slot.done})
                if (e) (void)hipEventDestroy(e);
            if (slot.upload) (void)hipHostFree(slot.upload);
            if (slot.download) (void)hipHostFree(slot.download);
            if (slot.device_upload) (void)hipFree(slot.device_upload);
            if (slot.device_download) (void)hipFree(slot.device_download);
        }
        if (arena) (void)hipFree(arena);
        if (stream) (void)hipStreamDestroy(stream);
        if (upload_stream) (void)hipStreamDestroy(upload_stream);
        if (download_stream) (void)hipStreamDestroy(download_stream);
    }
};

static size_t up256(size_t b) { return (b + 255) & ~(size_t)255; }
static uint8_t *up256_ptr(uint8_t *p) {
    return reinterpret_cast<uint8_t *>((reinterpret_cast<uintptr_t>(p) + 255) & ~(uintptr_t)255);
}

// image.ex:458-476 (area_weights): output o of `size` covers the input span
// [o scale, (o + 1) scale), scale = in_size / size; each input it overlaps is a tap with
// weight share / scale, computed in f64 and rounded to f32 (the padding taps of zero weight
// the reference adds change no sum and are left out).
struct HostTaps {
    int size = 0, most = 0;
    std::vector<int> count, index;
    std::vector<float> weight;
    size_t bytes() const { return (size_t)size * (1 + 2 * (size_t)most) * 4; }
    // Copies the taps to `dst` (bytes() bytes of device memory) and points at them.
    AreaTaps upload(uint8_t *dst) const {
        std::vector<uint8_t> host(bytes());
        uint8_t *p = host.data();
        std::memcpy(p, count.data(), 4 * (size_t)size);
        std::memcpy(p + 4 * (size_t)size, index.data(), 4 * index.size());
        std::memcpy(p + 4 * ((size_t)size + index.size()), weight.data(), 4 * weight.size());
        if (hipMemcpy(dst, host.data(), host.size(), hipMemcpyHostToDevice) != hipSuccess)
            throw Error{"copying the decode taps failed"};
        AreaTaps t;
        t.count = reinterpret_cast<const int *>(dst);
        t.index = t.count + size;
        t.weight = reinterpret_cast<const float *>(t.index + index.size());
        t.most = most;
        return t;
    }
};

static HostTaps area_taps(int size, int in_size) {
    HostTaps t;
    t.size = size;
    const double scale = (double)in_size / (double)size;
    t.most = (int)std::ceil(scale) + 1;
    t.count.assign(size, 0);
    t.index.assign((size_t)size * t.most, 0);
    t.weight.assign((size_t)size * t.most, 0.0f);
    for (int o = 0; o < size; o++) {
        const double first = o * scale, last = (o + 1) * scale;
        int n = 0;
        for (int input = (int)std::floor(first); input <= (int)std::ceil(last) - 1; input++) {
            const double share = std::min((double)input + 1, last) - std::max((double)input, first);
            if (share > 0 && n < t.most) {
                // (o + 1) scale can round past in_size in f64: a last tap one past the input
                // with a weight near 1e-15, which changes no rounded output; it reads the
                // last input instead.
                t.index[(size_t)o * t.most + n] = std::min(input, in_size - 1);
                t.weight[(size_t)o * t.most + n] = (float)(share / scale);
                n++;
            }
        }
        t.count[o] = n;
    }
    return t;
}

static std::shared_ptr<Segment> configure(uint32_t id, const Geometry &g,
                                          const std::vector<std::pair<std::string, Options>> &ms) {
    if (g.width < 2 || g.height < 2)
        throw Error{"a segment works at 2x2 or more, got " + std::to_string(g.width) + "x" +
                    std::to_string(g.height)};
    auto s = std::make_shared<Segment>();
    s->id = id;
    s->g = g;
    for (auto &m : ms) s->members.push_back(make_member(g, m.first, m.second));

    const size_t P = (size_t)g.width * g.height, F = (size_t)g.frame_width * g.frame_height;
    int max_dec = 0;
    size_t match = 0, cloud = 0, partials = 0;
    for (auto &m : s->members) {
        if (m.kind == Kind::Disparity)
            match = std::max(match, m.coarse_to_fine
                                        ? coarse_to_fine_bytes(g.width, g.height, m.band, m.paths,
                                                               m.coarse_paths)
                                        : match_bytes(g.width, g.height, m.disparities, m.paths));
        if (m.kind == Kind::PointCloud) {
            max_dec = m.decimation;
            cloud = std::max(cloud, (size_t)(g.width / m.decimation) * (g.height / m.decimation) * 3);
            partials = std::max(partials, (size_t)cloud_partials(g.width, g.height, m.decimation));
        }
    }
    (void)max_dec;
    size_t sharp = 0, align = 0, align_sums = 0;  // scratch bytes; the alignment sums (floats)
    for (auto &m : s->members) {
        if (m.kind == Kind::Sharpness) sharp = sharpness_scratch_bytes(g.width);
        if (m.kind == Kind::RowAlignment) {
            align = std::max(align, alignment_scratch_bytes(m.bands, m.reach));
            align_sums = std::max(align_sums, (size_t)m.bands * (2 * m.reach + 2));
        }
    }
    int carries = 0;
    for (auto &m : s->members) carries += m.kind == Kind::TemporalFilter ? 2 : (m.kind == Kind::Quality);
    // Decode's taps: rows (H outputs of the frame's height), then columns (W outputs of its
    // width), each as counts, indices and weights.
    bool decode = false;
    for (auto &m : s->members) decode = decode || m.kind == Kind::Decode;
    HostTaps rows_taps, cols_taps;
    if (decode) {
        rows_taps = area_taps(g.height, g.frame_height);
        cols_taps = area_taps(g.width, g.frame_width);
    }
    const size_t taps_bytes = rows_taps.bytes() + cols_taps.bytes();
    size_t bytes = up256(F) * 2 + up256(P) * 2 + up256(4 * P) * 2 + up256(2 * P) + up256(4 * P) +
                   up256(4 * cloud) + up256(4 * (partials + 1)) +
                   up256(quality_scratch_bytes(g.width, g.height)) + up256(32) + up256(32) +
                   up256(sharp) + up256(32) + up256(align) + up256(4 * align_sums) +
                   up256(4 * P) * carries + up256(taps_bytes) + 256;
    HIP_TRY(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    // NXS_TRANSFER_STREAMS=0 keeps a submitted run's transfers on the segment's stream.
    const char *transfer = std::getenv("NXS_TRANSFER_STREAMS");
    if (!transfer || std::atoi(transfer) != 0) {
        HIP_TRY(hipStreamCreateWithFlags(&s->upload_stream, hipStreamNonBlocking));
        HIP_TRY(hipStreamCreateWithFlags(&s->download_stream, hipStreamNonBlocking));
    }
    HIP_TRY(hipEventCreateWithFlags(&s->order, hipEventDisableTiming | hipEventDisableSystemFence));
    HIP_TRY(hipMalloc(&s->arena, bytes));
    uint8_t *p = static_cast<uint8_t *>(s->arena);
    auto carve = [&](size_t n) {
        uint8_t *q = p;
        p += up256(n);
        return q;
    };
    s->frames[0] = carve(F);
    s->frames[1] = carve(F);
    s->images[0] = carve(P);
    s->images[1] = carve(P);
    s->disparity[0] = reinterpret_cast<float *>(carve(4 * P));
    s->disparity[1] = reinterpret_cast<float *>(carve(4 * P));
    s->depth_mm = reinterpret_cast<uint16_t *>(carve(2 * P));
    s->depth_m = reinterpret_cast<float *>(carve(4 * P));
    s->cloud = reinterpret_cast<float *>(carve(4 * cloud));
    s->cloud_partials = reinterpret_cast<int32_t *>(carve(4 * (partials + 1)));
    s->cloud_kept = s->cloud_partials + partials;
    s->quality_partials = reinterpret_cast<double *>(carve(quality_scratch_bytes(g.width, g.height)));
    // The cloud and quality kernels' finishing counters start at zero (each launch leaves
    // them at zero again).
    HIP_TRY(launch_fill_f32(reinterpret_cast<float *>(s->cloud_partials), 0.0f, partials + 1,
                            s->stream));
    HIP_TRY(launch_fill_f32(reinterpret_cast<float *>(s->quality_partials), 0.0f,
                            quality_scratch_bytes(g.width, g.height) / sizeof(float), s->stream));
    s->quality_sums = reinterpret_cast<float *>(carve(4 * 8));
    s->tail_kept = reinterpret_cast<int32_t *>(carve(4 * 8));
    HIP_TRY(launch_fill_f32(reinterpret_cast<float *>(s->tail_kept), 0.0f, 8, s->stream));
    // The sharpness and alignment totals start at zero (each launch leaves them at zero).
    s->sharpness_totals = reinterpret_cast<uint64_t *>(carve(sharp));
    s->sharpness_sums = reinterpret_cast<float *>(carve(4 * 8));
    s->alignment_totals = reinterpret_cast<uint64_t *>(carve(align));
    s->alignment_sums = reinterpret_cast<float *>(carve(4 * align_sums));
    if (sharp)
        HIP_TRY(launch_fill_f32(reinterpret_cast<float *>(s->sharpness_totals), 0.0f,
                                sharp / sizeof(float), s->stream));
    if (align)
        HIP_TRY(launch_fill_f32(reinterpret_cast<float *>(s->alignment_totals), 0.0f,
                                align / sizeof(float), s->stream));
    if (decode) {
        uint8_t *taps = carve(taps_bytes);
        s->decode_rows = rows_taps.upload(taps);
        s->decode_cols = cols_taps.upload(taps + rows_taps.bytes());
    }
    s->workspace_bytes = match;
    for (size_t i = 0; i < s->members.size(); i++)
        if (s->members[i].kind == Kind::Disparity) {
            s->matcher = s->members[i].coarse_to_fine ? -1 : (int)i;
            break;
        }
    if (const char *fuse = std::getenv("NXS_FUSE_TAIL")) s->fuse_tail = std::atoi(fuse) != 0;
    for (auto &m : s->members) {
        float *a = nullptr, *b = nullptr;
        if (m.kind == Kind::TemporalFilter) {
            a = reinterpret_cast<float *>(carve(4 * P));
            b = reinterpret_cast<float *>(carve(4 * P));
            HIP_TRY(launch_fill_f32(a, -1.0f, P, s->stream));  // previous: unmatched
            HIP_TRY(launch_fill_f32(b, 0.0f, P, s->stream));   // missed: none
        } else if (m.kind == Kind::Quality) {
            a = reinterpret_cast<float *>(carve(4 * P));
            HIP_TRY(launch_fill_f32(a, -1.0f, P, s->stream));
        }
        s->carry_a.push_back(a);
        s->carry_b.push_back(b);
    }
    // Pinned staging per pair in flight. Up: the frames, images and maps an entry may bring.
    // Down: the reply record: left and right (u8), disparity (f32), depth_mm (u16), the
    // cloud and two small tensors, with the record's and each tensor's header.
    for (auto &slot : s->slots) {
        slot.upload_bytes = 2 * F + 2 * P + 8 * P + 256;
        slot.download_bytes = 2 * P + 4 * P + 2 * P + 4 * cloud + 4 * align_sums + 4096;
        HIP_TRY(hipHostMalloc((void **)&slot.upload, slot.upload_bytes));
        HIP_TRY(hipHostMalloc((void **)&slot.download, slot.download_bytes));
        HIP_TRY(hipMalloc((void **)&slot.device_upload, slot.upload_bytes));
        HIP_TRY(hipMalloc((void **)&slot.device_download, slot.download_bytes));
        // The events between members only time them: no system-scope fence (each one would
        // write back and invalidate the caches between the members' kernels). `done`, after
        // the downloads, tells the host the run is done: it keeps the fence.
        slot.events.resize(s->members.size() + 1);
        for (size_t k = 0; k < slot.events.size(); k++)
            HIP_TRY(hipEventCreateWithFlags(&slot.events[k], hipEventDisableSystemFence));
        HIP_TRY(hipEventCreateWithFlags(&slot.uploaded, hipEventDisableTiming | hipEventDisableSystemFence));
        HIP_TRY(hipEventCreateWithFlags(&slot.computed, hipEventDisableTiming | hipEventDisableSystemFence));
        HIP_TRY(hipEventCreateWithFlags(&slot.done, hipEventDisableTiming));
    }
    HIP_TRY(hipStreamSynchronize(s->stream));
    return s;
}

static const std::vector<float> &input(const Inputs &in, const std::string &name, size_t length) {
    auto it = in.find(name);
    if (it == in.end()) throw Error{"no " + name + " among the stage's inputs"};
    if (it->second.size() != length)
        throw Error{name + " has " + std::to_string(it->second.size()) + " values, expected " +
                    std::to_string(length)};
    return it->second;
}

static float scalar(const Inputs &in, const std::string &name) { return input(in, name, 1)[0]; }

// segment.rs:146-156: a penalty is an integer byte.
static int penalty(const Inputs &in, const std::string &name) {
    float v = scalar(in, name);
    if (v == std::floor(v) && v >= 0.0f && v <= 255.0f) return (int)v;
    std::ostringstream s;
    s << name << " must be an integer from 0 to 255, got " << v;
    throw Error{s.str()};
}

static Camera camera(const Inputs &in, const std::string &side) {
    Camera c;
    std::memcpy(c.k, input(in, side + ".k", 9).data(), 36);
    std::memcpy(c.d, input(in, side + ".d", 5).data(), 20);
    std::memcpy(c.inverse_pr, input(in, side + ".inverse_pr", 9).data(), 36);
    return c;
}

// tensor.rs:123-183: the tensor `name` of `type` and `shape`, or why not.
static void *shaped(const std::map<std::string, Slot> &tensors, const std::string &name,
                    uint8_t type, std::vector<uint32_t> shape, const char *what) {
    auto it = tensors.find(name);
    if (it == tensors.end()) throw Error{"no " + name + " among the segment's tensors"};
    const Slot &s = it->second;
    if (s.shape != shape || s.type != type || !s.dev)
        throw Error{name + " is a " + shape_text(s.shape != shape ? s.shape : shape) +
                    " tensor of type " + std::to_string(s.type) + ", expected a " +
                    shape_text(shape) + " " + what};
    return s.dev;
}

// An entry already in device memory (nxs_run_device).
struct DeviceEntry {
    std::string name;
    uint8_t type;
    std::vector<uint32_t> shape;
    void *data;
};

struct Enqueued {
    std::map<std::string, Slot> tensors;
    std::vector<const char *> filled;
    int free_disparity = 0;  // which of the segment's two maps the next filter writes
};

// The fused per-pixel tail (launch_tail): SpeckleFilter of radius 3, TemporalFilter, Depth,
// PointCloud and Quality may run as one launch (and one finishing block) where consecutive
// members take them in this order, PointCloud only after Depth.
static int tail_rank(const Member &m) {
    switch (m.kind) {
        case Kind::SpeckleFilter: return m.radius == 3 ? 0 : -1;
        case Kind::TemporalFilter: return 1;
        case Kind::Depth: return 2;
        case Kind::PointCloud: return 3;
        case Kind::Quality: return 4;
        default: return -1;
    }
}

// How many members from i on form one fused tail.
static size_t tail_length(const std::vector<Member> &members, size_t i) {
    size_t n = 0;
    int last = -1;
    bool depth = false;
    while (i + n < members.size()) {
        const Member &m = members[i + n];
        const int rank = tail_rank(m);
        if (rank <= last || (m.kind == Kind::PointCloud && !depth)) break;
        depth = depth || m.kind == Kind::Depth;
        last = rank;
        n++;
    }
    return n;
}

// Runs of one segment share its arena, carries and workspace, so each waits for the one
// before it, whichever stream that went on: before a run on `stream`, the stream waits for
// the segment's last run when that ran elsewhere (the segment's own stream marks its work
// then, a caller's stream was marked after its run). Called with the segment's lock held.
static void order_before(Segment &s, hipStream_t stream) {
    if (!s.last_stream || s.last_stream == stream) return;
    if (s.last_stream == s.stream) HIP_TRY(hipEventRecord(s.order, s.stream));
    HIP_TRY(hipStreamWaitEvent(stream, s.order, 0));
}

static void order_after(Segment &s, hipStream_t stream) {
    if (stream != s.stream) {
        HIP_TRY(hipEventRecord(s.order, stream));
        s.foreign = true;
    }
    s.last_stream = stream;
}

// Enqueues every member on `stream`: entries go to the segment's buffers (device entries
// a stage only reads are read in place), then each stage in order (segment.rs:206-239).
// Host entries go through the slot's pinned upload buffer (or, without a slot, a
// synchronous copy); `workspace` holds workspace_bytes for the matcher, if any.
static Enqueued enqueue_entries(Segment &s, const std::vector<Inputs> &inputs,
                                const std::vector<std::pair<std::string, HostTensor>> &entries,
                                const std::vector<DeviceEntry> &device_entries,
                                hipStream_t stream, RunSlot *run_slot,
                                const std::vector<hipEvent_t> &events) {
    if (inputs.size() != s.members.size())
        throw Error{std::to_string(s.members.size()) + " stages, got the inputs of " +
                    std::to_string(inputs.size())};
    const Geometry &g = s.g;
    const uint32_t W = g.width, H = g.height;
    const size_t P = (size_t)W * H;
    const std::vector<uint32_t> image = {H, W},
                                frame = {(uint32_t)g.frame_height, (uint32_t)g.frame_width};
    Enqueued q;
    std::map<std::string, Slot> &tensors = q.tensors;
    int &free_disparity = q.free_disparity;

    // A frame or image packed as u32 words (segment.ex:207-241) holds the u8 image's bytes:
    // it is read as that image.
    auto packed = [](uint8_t type, const std::vector<uint32_t> &shape, const std::vector<uint32_t> &as) {
        return type == TYPE_U32 && shape.size() == 2 && as[1] % 4 == 0 && shape[0] == as[0] &&
               shape[1] * 4 == as[1];
    };
    auto unpacked = [&](const std::string &name, uint8_t &type, std::vector<uint32_t> &shape) {
        const bool is_frame = name == "left_frame" || name == "right_frame";
        const bool is_image = name == "left" || name == "right";
        if ((is_frame && packed(type, shape, frame)) || (is_image && packed(type, shape, image))) {
            type = TYPE_U8;
            shape = is_frame ? frame : image;
        }
    };
    // Where an entry a stage can read goes, or nullptr.
    auto target_of = [&](const std::string &name, uint8_t type, const std::vector<uint32_t> &shape) -> void * {
        if ((name == "left_frame" || name == "right_frame") && type == TYPE_U8 && shape == frame)
            return s.frames[name == "right_frame"];
        if ((name == "left" || name == "right") && type == TYPE_U8 && shape == image)
            return s.images[name == "right"];
        if (name == "disparity" && type == TYPE_F32 && shape == image) return s.disparity[0];
        if (name == "depth_m" && type == TYPE_F32 && shape == image) return s.depth_m;
        return nullptr;
    };
    // Host entries through the run slot's staging: pinned, then its device staging on the
    // upload stream, then (once that is done) into place on the run's stream.
    hipStream_t up = s.upload_stream ? s.upload_stream : stream;
    size_t staged = 0;
    struct Landing {
        void *target;
        size_t from, size;
    };
    std::vector<Landing> landings;
    for (auto &e : entries) {
        const HostTensor &t = e.second;
        Slot slot{t.type, t.shape, nullptr};
        unpacked(e.first, slot.type, slot.shape);
        if (void *target = target_of(e.first, slot.type, slot.shape)) {
            if (run_slot && staged + t.size <= run_slot->upload_bytes) {
                std::memcpy(run_slot->upload + staged, t.bytes, t.size);
                HIP_TRY(hipMemcpyAsync(run_slot->device_upload + staged, run_slot->upload + staged,
                                       t.size, hipMemcpyHostToDevice, up));
                landings.push_back(Landing{target, staged, t.size});
                staged += (t.size + 255) & ~(size_t)255;
            } else {
                HIP_TRY(hipStreamSynchronize(stream));
                HIP_TRY(hipMemcpy(target, t.bytes, t.size, hipMemcpyHostToDevice));
            }
            slot.dev = target;
            if (e.first == "disparity") free_disparity = 1;
        }
        tensors[e.first] = slot;
    }
    if (!landings.empty() && up != stream) {
        HIP_TRY(hipEventRecord(run_slot->uploaded, up));
        HIP_TRY(hipStreamWaitEvent(stream, run_slot->uploaded, 0));
    }
    if (!landings.empty()) {
        for (const Landing &l : landings)
            HIP_TRY(hipMemcpyAsync(l.target, run_slot->device_upload + l.from, l.size,
                                   hipMemcpyDeviceToDevice, stream));
    }
    for (auto &e : device_entries) {
        Slot slot{e.type, e.shape, nullptr};
        unpacked(e.name, slot.type, slot.shape);
        if (void *target = target_of(e.name, slot.type, slot.shape)) {
            if (e.name == "disparity") {
                // Filters may change the disparity in place: work on the segment's copy.
                HIP_TRY(hipMemcpyAsync(target, e.data, 4 * P, hipMemcpyDeviceToDevice, stream));
                slot.dev = target;
                free_disparity = 1;
            } else {
                slot.dev = e.data;
            }
        }
        tensors[e.name] = slot;
    }

    if (!events.empty()) HIP_TRY(hipEventRecord(events[0], stream));
    return q;
}

// A full-search Disparity member's matching, as its step reads it (inputs checked, images
// taken, the workspace carved, the output map chosen), and the tensors it leaves: the
// pair's descriptor for the matcher, launched alone (enqueue_members) or with other
// segments' pairs of its size class in one launch (the served form's coalescing).
static PairDesc matcher_pair(Segment &s, const Member &m, const Inputs &in, Enqueued &q,
                             void *workspace, MatchParams &mp) {
    const uint32_t W = s.g.width, H = s.g.height;
    const std::vector<uint32_t> image = {H, W};
    const int p1 = penalty(in, "p1"), p2 = penalty(in, "p2");
    const float u = scalar(in, "uniqueness");
    check_matching(W, H, m.disparities, m.paths, p1, p2, u);
    PairDesc pd{};
    pd.left = (const uint8_t *)shaped(q.tensors, "left", TYPE_U8, image, "u8 image");
    pd.right = (const uint8_t *)shaped(q.tensors, "right", TYPE_U8, image, "u8 image");
    pd.disparity = s.disparity[q.free_disparity];
    if (!workspace) throw Error{"no matcher workspace"};
    match_carve(pd.b, workspace, W, H, m.disparities, m.paths);
    pd.sums = nullptr;
    pd.p1 = p1;
    pd.p2 = p2;
    pd.uniqueness = u;
    mp = MatchParams{(int)W, (int)H, m.disparities, m.paths, p1, p2, u};
    q.tensors["disparity"] = Slot{TYPE_F32, image, pd.disparity};
    q.free_disparity ^= 1;
    return pd;
}

static void note_outputs(Enqueued &q, const Member &m) {
    for (const char *name : member_outputs(m.kind)) {
        bool seen = false;
        for (const char *f : q.filled) seen = seen || std::strcmp(f, name) == 0;
        if (!seen) q.filled.push_back(name);
    }
}

// Enqueues members [from, to) on `stream` (segment.rs:206-239), after enqueue_entries;
// `workspace` holds workspace_bytes for the matcher, if any.
static void enqueue_members(Segment &s, const std::vector<Inputs> &inputs, Enqueued &q,
                            size_t from, size_t to, hipStream_t stream,
                            const std::vector<hipEvent_t> &events, void *workspace) {
    const Geometry &g = s.g;
    const uint32_t W = g.width, H = g.height;
    const size_t P = (size_t)W * H;
    const std::vector<uint32_t> image = {H, W},
                                frame = {(uint32_t)g.frame_height, (uint32_t)g.frame_width};
    std::map<std::string, Slot> &tensors = q.tensors;
    int &free_disparity = q.free_disparity;
    auto note_outputs = [&](const Member &m) { nxs::note_outputs(q, m); };
    for (size_t i = from; i < to; i++) {
        const Member &m = s.members[i];
        const Inputs &in = inputs[i];
        if (const size_t run = s.fuse_tail ? std::min(tail_length(s.members, i), to - i) : 0;
            run >= 2) {
            // Members i .. i+run-1 in one launch. Each member's inputs are read in member
            // order, as its own step reads them; the first member's time is the launch's,
            // the last one's the finish's.
            TailArgs a;
            a.W = (int)W;
            a.H = (int)H;
            for (size_t k = i; k < i + run; k++) {
                const Member &mk = s.members[k];
                const Inputs &ink = inputs[k];
                switch (mk.kind) {
                    case Kind::SpeckleFilter: {
                        auto *d = (const float *)shaped(tensors, "disparity", TYPE_F32, image, "f32 map");
                        a.despeckle = 1;
                        a.min_support = scalar(ink, "min_support");
                        a.max_diff = scalar(ink, "max_diff");
                        a.in = d;
                        a.out = s.disparity[0] == d ? s.disparity[1] : s.disparity[0];
                        tensors["disparity"] = Slot{TYPE_F32, image, a.out};
                        free_disparity = s.disparity[0] == a.out ? 1 : 0;
                        break;
                    }
                    case Kind::TemporalFilter: {
                        a.alpha = scalar(ink, "alpha");
                        a.delta = scalar(ink, "delta");
                        a.hold = scalar(ink, "hold");
                        auto *d = (float *)shaped(tensors, "disparity", TYPE_F32, image, "f32 map");
                        if (!a.in) a.in = a.out = d;
                        a.temporal = 1;
                        a.t_prev = s.carry_a[k];
                        a.t_missed = s.carry_b[k];
                        a.max_missed = mk.max_missed;
                        break;
                    }
                    case Kind::Depth: {
                        auto *d = (const float *)shaped(tensors, "disparity", TYPE_F32, image, "f32 map");
                        const float f = scalar(ink, "focal_length_px"), b = scalar(ink, "baseline_m");
                        if (!a.in) a.in = d;
                        a.depth = 1;
                        a.focal_baseline = f * b;  // host f32 multiply, as depth.rs:20
                        a.mm = s.depth_mm;
                        a.m = s.depth_m;
                        tensors["depth_mm"] = Slot{TYPE_U16, image, s.depth_mm};
                        tensors["depth_m"] = Slot{TYPE_F32, image, s.depth_m};
                        break;
                    }
                    case Kind::PointCloud: {
                        a.f = scalar(ink, "focal_length_px");
                        a.cx = scalar(ink, "cx");
                        a.cy = scalar(ink, "cy");
                        (void)shaped(tensors, "depth_m", TYPE_F32, image, "f32 map");
                        a.cloud = 1;
                        a.dec = mk.decimation;
                        a.gw = (int)W / mk.decimation;
                        a.gh = (int)H / mk.decimation;
                        a.nearest = mk.nearest;
                        a.farthest = mk.farthest;
                        a.points = s.cloud;
                        a.kept_accum = s.tail_kept;
                        a.kept_total = s.cloud_kept;
                        tensors["cloud_grid"] =
                            Slot{TYPE_F32, {H / mk.decimation, W / mk.decimation, 3}, s.cloud};
                        tensors["cloud_points"] = Slot{TYPE_S32, {}, s.cloud_kept};
                        break;
                    }
                    case Kind::Quality: {
                        auto *d = (const float *)shaped(tensors, "disparity", TYPE_F32, image, "f32 map");
                        if (!a.in) a.in = d;
                        a.quality = 1;
                        a.q_prev = s.carry_a[k];
                        a.totals = s.quality_partials;
                        a.sums5 = s.quality_sums;
                        tensors["quality_sums"] = Slot{TYPE_F32, {5}, s.quality_sums};
                        break;
                    }
                    default: throw Error{"a member that cannot join a fused tail"};
                }
            }
            HIP_TRY(launch_tail(a, stream));
            for (size_t k = i; k + 1 < i + run; k++)
                if (!events.empty()) HIP_TRY(hipEventRecord(events[k + 1], stream));
            HIP_TRY(launch_tail_finish(a, stream));
            if (!events.empty()) HIP_TRY(hipEventRecord(events[i + run], stream));
            for (size_t k = i; k < i + run; k++) note_outputs(s.members[k]);
            i += run - 1;
            continue;
        }
        switch (m.kind) {
            case Kind::Decode: {
                auto *lf = (const uint8_t *)shaped(tensors, "left_frame", TYPE_U8, frame, "u8 image");
                auto *rf = (const uint8_t *)shaped(tensors, "right_frame", TYPE_U8, frame, "u8 image");
                HIP_TRY(launch_resize_area(lf, rf, g.frame_width, W, H, s.decode_rows,
                                           s.decode_cols, s.images[0], s.images[1], stream));
                tensors["left"] = Slot{TYPE_U8, image, s.images[0]};
                tensors["right"] = Slot{TYPE_U8, image, s.images[1]};
                break;
            }
            case Kind::Rectification: {
                Camera left = camera(in, "left"), right = camera(in, "right");
                auto *lf = (const uint8_t *)shaped(tensors, "left_frame", TYPE_U8, frame, "u8 image");
                auto *rf = (const uint8_t *)shaped(tensors, "right_frame", TYPE_U8, frame, "u8 image");
                HIP_TRY(launch_rectify(lf, rf, g.frame_width, g.frame_height, left, right, W, H,
                                       s.images[0], s.images[1], stream));
                tensors["left"] = Slot{TYPE_U8, image, s.images[0]};
                tensors["right"] = Slot{TYPE_U8, image, s.images[1]};
                break;
            }
            case Kind::Disparity: {
                if (!m.coarse_to_fine) {
                    MatchParams mp;
                    PairDesc pd = matcher_pair(s, m, in, q, workspace, mp);
                    HIP_TRY(launch_match(mp, pd.left, pd.right, pd.b, pd.disparity, nullptr, stream));
                    break;
                }
                const int p1 = penalty(in, "p1"), p2 = penalty(in, "p2");
                const float u = scalar(in, "uniqueness");
                check_coarse_to_fine(W, H, m.band, m.margin, m.paths, m.coarse_paths,
                                     m.fine_census, p1, p2, u);
                auto *l = (const uint8_t *)shaped(tensors, "left", TYPE_U8, image, "u8 image");
                auto *r = (const uint8_t *)shaped(tensors, "right", TYPE_U8, image, "u8 image");
                float *out = s.disparity[free_disparity];
                if (!workspace) throw Error{"no matcher workspace"};
                CoarseBuffers c;
                coarse_to_fine_carve(c, workspace, W, H, m.band, m.paths, m.coarse_paths);
                BandParams bp = band_params(W, H, 128, m.band, m.paths, m.fine_census, p1, p2, u);
                bp.positive = true;  // the stage's "> 0 else -1" (stage/disparity.ex:162)
                HIP_TRY(launch_coarse_to_fine(bp, m.margin, m.coarse_paths, l, r, c, out, stream));
                tensors["disparity"] = Slot{TYPE_F32, image, out};
                free_disparity ^= 1;
                break;
            }
            case Kind::Sharpness: {
                auto *l = (const uint8_t *)shaped(tensors, "left", TYPE_U8, image, "u8 image");
                auto *r = (const uint8_t *)shaped(tensors, "right", TYPE_U8, image, "u8 image");
                HIP_TRY(launch_sharpness(l, r, W, H, s.sharpness_totals, s.sharpness_sums, stream));
                tensors["sharpness_sums"] = Slot{TYPE_F32, {6}, s.sharpness_sums};
                break;
            }
            case Kind::MedianFilter: {
                auto *d = (const float *)shaped(tensors, "disparity", TYPE_F32, image, "f32 map");
                float *out = s.disparity[0] == d ? s.disparity[1] : s.disparity[0];
                HIP_TRY(launch_median_filter(d, out, W, H, m.kernel_size, stream));
                tensors["disparity"] = Slot{TYPE_F32, image, out};
                free_disparity = s.disparity[0] == out ? 1 : 0;
                break;
            }
            case Kind::RowAlignment: {
                auto *l = (const uint8_t *)shaped(tensors, "left", TYPE_U8, image, "u8 image");
                auto *r = (const uint8_t *)shaped(tensors, "right", TYPE_U8, image, "u8 image");
                auto *d = (const float *)shaped(tensors, "disparity", TYPE_F32, image, "f32 map");
                HIP_TRY(launch_alignment(l, r, d, W, H, m.bands, m.reach, m.row_step,
                                         s.alignment_totals, s.alignment_sums, stream));
                tensors["alignment_sums"] =
                    Slot{TYPE_F32, {(uint32_t)m.bands, (uint32_t)(2 * m.reach + 2)}, s.alignment_sums};
                break;
            }
            case Kind::SpeckleFilter: {
                // segment.rs:437-447 reads the map before the scalars.
                auto *d = (const float *)shaped(tensors, "disparity", TYPE_F32, image, "f32 map");
                const float support = scalar(in, "min_support"), diff = scalar(in, "max_diff");
                float *out = s.disparity[0] == d ? s.disparity[1] : s.disparity[0];
                HIP_TRY(launch_despeckle(d, out, W, H, support, diff, m.radius, stream));
                tensors["disparity"] = Slot{TYPE_F32, image, out};
                free_disparity = s.disparity[0] == out ? 1 : 0;
                break;
            }
            case Kind::TemporalFilter: {
                const float alpha = scalar(in, "alpha"), delta = scalar(in, "delta"),
                            hold = scalar(in, "hold");
                auto *d = (float *)shaped(tensors, "disparity", TYPE_F32, image, "f32 map");
                HIP_TRY(launch_temporal(d, s.carry_a[i], s.carry_b[i], (int)P, alpha, delta, hold,
                                        m.max_missed, stream));
                break;
            }
            case Kind::Depth: {
                // segment.rs:475-482 reads the map before the scalars.
                auto *d = (const float *)shaped(tensors, "disparity", TYPE_F32, image, "f32 map");
                const float f = scalar(in, "focal_length_px"), b = scalar(in, "baseline_m");
                HIP_TRY(launch_depth(d, (int)P, f, b, s.depth_mm, s.depth_m, stream));
                tensors["depth_mm"] = Slot{TYPE_U16, image, s.depth_mm};
                tensors["depth_m"] = Slot{TYPE_F32, image, s.depth_m};
                break;
            }
            case Kind::PointCloud: {
                const float f = scalar(in, "focal_length_px"), cx = scalar(in, "cx"),
                            cy = scalar(in, "cy");
                auto *d = (const float *)shaped(tensors, "depth_m", TYPE_F32, image, "f32 map");
                HIP_TRY(launch_cloud(d, W, H, m.decimation, f, cx, cy, m.nearest, m.farthest,
                                     s.cloud, s.cloud_partials, s.cloud_kept, stream));
                tensors["cloud_grid"] = Slot{TYPE_F32, {H / m.decimation, W / m.decimation, 3}, s.cloud};
                tensors["cloud_points"] = Slot{TYPE_S32, {}, s.cloud_kept};
                break;
            }
            case Kind::Quality: {
                auto *d = (const float *)shaped(tensors, "disparity", TYPE_F32, image, "f32 map");
                HIP_TRY(launch_quality(d, s.carry_a[i], W, H, s.quality_partials, s.quality_sums,
                                       stream));
                tensors["quality_sums"] = Slot{TYPE_F32, {5}, s.quality_sums};
                break;
            }
        }
        if (!events.empty()) HIP_TRY(hipEventRecord(events[i + 1], stream));
        note_outputs(m);
    }
}

// Enqueues every member on `stream`: entries go to the segment's buffers (device entries
// a stage only reads are read in place), then each stage in order (segment.rs:206-239).
// Host entries go through the slot's pinned upload buffer (or, without a slot, a
// synchronous copy); `workspace` holds workspace_bytes for the matcher, if any.
static Enqueued enqueue(Segment &s, const std::vector<Inputs> &inputs,
                        const std::vector<std::pair<std::string, HostTensor>> &entries,
                        const std::vector<DeviceEntry> &device_entries, hipStream_t stream,
                        RunSlot *run_slot, const std::vector<hipEvent_t> &events, void *workspace) {
    Enqueued q = enqueue_entries(s, inputs, entries, device_entries, stream, run_slot, events);
    enqueue_members(s, inputs, q, 0, s.members.size(), stream, events, workspace);
    return q;
}

// Lays out a run's reply record in its slot's download staging (protocol.rs:182-202): tag 8,
// the member count, each member's GPU time in us (written when the run completes), the
// output count, then each filled tensor in first-filled order (segment.rs:220-234), its
// name, type and shape written now and its bytes downloaded straight after them. The reply
// is then the staging itself, with nothing copied on the host. Returns the record's length.
static size_t enqueue_downloads(Segment &s, Enqueued &q, RunSlot &slot, size_t members,
                                hipStream_t stream) {
    size_t at = 3 + 4 * members;
    if (at > slot.download_bytes || q.filled.size() > 255) throw Error{"output staging too small"};
    slot.download[0] = TAG_OUTPUTS;
    slot.download[1] = (uint8_t)members;
    slot.download[at - 1] = (uint8_t)q.filled.size();
    std::vector<std::pair<size_t, size_t>> spans;  // each tensor's bytes: offset, length
    for (const char *name : q.filled) {
        const Slot &t = q.tensors[name];
        size_t elements = 1;
        for (uint32_t d : t.shape) elements *= d;
        const size_t bytes = elements * element_bytes(t.type);
        const size_t length = std::strlen(name);
        const size_t header = 3 + length + 4 * t.shape.size();
        if (at + header + bytes > slot.download_bytes) throw Error{"output staging too small"};
        uint8_t *h = slot.download + at;
        *h++ = (uint8_t)length;
        std::memcpy(h, name, length);
        h += length;
        *h++ = t.type;
        *h++ = (uint8_t)t.shape.size();
        for (uint32_t d : t.shape)
            for (int b = 0; b < 4; b++) *h++ = (uint8_t)(d >> (8 * b));
        at += header;
        spans.push_back({at, bytes});
        HIP_TRY(hipMemcpyAsync(slot.device_download + at, t.dev, bytes, hipMemcpyDeviceToDevice,
                               stream));
        at += bytes;
    }
    // The next run's kernels may now overwrite the segment's buffers: the bytes leave from
    // the slot's device staging on the download stream.
    hipStream_t down = s.download_stream ? s.download_stream : stream;
    if (down != stream) {
        HIP_TRY(hipEventRecord(slot.computed, stream));
        HIP_TRY(hipStreamWaitEvent(down, slot.computed, 0));
    }
    for (const auto &span : spans)
        HIP_TRY(hipMemcpyAsync(slot.download + span.first, slot.device_download + span.first,
                               span.second, hipMemcpyDeviceToHost, down));
    HIP_TRY(hipEventRecord(slot.done, down));
    return at;
}

// The member times of a completed run, into its reply record.
static void finish_record(RunSlot &slot, size_t members) {
    for (size_t i = 0; i < members; i++) {
        float ms = 0;
        HIP_TRY(hipEventElapsedTime(&ms, slot.events[i], slot.events[i + 1]));
        const uint32_t us = (uint32_t)std::min(4294967295.0, std::floor((double)ms * 1000.0));
        for (int k = 0; k < 4; k++) slot.download[2 + 4 * i + k] = (uint8_t)(us >> (8 * k));
    }
}

// ------------------------------------------------------------------------------------------
// The context. A submitting thread queues records; the context's worker thread parses
// them and enqueues each run's work on its segment's stream without waiting; the
// completion thread watches the runs in flight and posts each reply as its pair
// finishes. Runs of one segment stay in order (one stream, FIFO); runs of different
// segments overlap on the device, admitted as their segment's run slots and the
// matcher workspace pool's HBM budget allow (continuous batching).

struct Job {
    int64_t ticket;
    std::vector<uint8_t> record;
    nxs_done_fn done;
    void *user;
    // The record parsed (its host tensors point into `record`), once the worker has read it.
    std::unique_ptr<Request> parsed;
};

struct Kept {
    bool complete = false;
    std::vector<uint8_t> reply;
};

// A run in flight: what the completion thread needs to reply once its last event fires.
struct Pending {
    int64_t ticket;
    std::shared_ptr<Segment> segment;
    int slot;
    Workspace *workspace;
    size_t length;  // of the reply record in the slot's download staging
    nxs_done_fn done;
    void *user;
};

}  // namespace nxs

struct nxs_ctx {
    int device = 0;
    std::mutex mu;  // everything below but the segments' device work
    std::condition_variable cv_jobs, cv_done, cv_free, cv_pending;
    std::deque<nxs::Job> jobs;
    std::unordered_map<int64_t, nxs::Kept> kept;
    std::map<uint32_t, std::shared_ptr<nxs::Segment>> segments;
    std::deque<nxs::Pending> pending;
    std::vector<std::unique_ptr<nxs::Workspace>> workspaces;
    size_t budget = 0, allocated = 0;  // the matcher workspaces' HBM budget and use
    int64_t next_ticket = 1;
    bool stopping = false, worker_done = false;
    std::thread worker, completer;

    // Coalescing: at most `coalesce_max` runs share one matcher launch (1: none), and the
    // launches and runs so coalesced.
    static constexpr int MAX_COALESCE = 16, GATHER_SCAN = 64, TABLES = 8;
    int coalesce_max = 8;
    int64_t coalesced_launches = 0, coalesced_pairs = 0;
    // The coalesced launches' descriptor tables: pinned host staging and a device copy
    // each, taken in turn, a table reused once the launch that read it has finished.
    struct TableSlot {
        nxs::PairDesc *host = nullptr, *dev = nullptr;
        hipEvent_t done = nullptr;
        bool used = false;
    };
    TableSlot tables[TABLES];
    int next_table_index = 0;

    TableSlot *next_table() {
        using namespace nxs;
        TableSlot &t = tables[next_table_index];
        if (!t.host) {
            PairDesc *host = nullptr, *dev = nullptr;
            hipEvent_t done = nullptr;
            hipError_t e = hipHostMalloc((void **)&host, MAX_COALESCE * sizeof(PairDesc));
            if (e == hipSuccess) e = hipMalloc((void **)&dev, MAX_COALESCE * sizeof(PairDesc));
            if (e == hipSuccess) e = hipEventCreateWithFlags(&done, hipEventDisableTiming);
            if (e != hipSuccess) {
                if (host) (void)hipHostFree(host);
                if (dev) (void)hipFree(dev);
                throw Error{std::string("HIP error: ") + hipGetErrorString(e)};
            }
            t.host = host;
            t.dev = dev;
            t.done = done;
        } else if (t.used) {
            HIP_TRY(hipEventSynchronize(t.done));
        }
        next_table_index = (next_table_index + 1) % TABLES;
        return &t;
    }

    // A reply that lives in a run slot's staging: handed to the callback as it is, or
    // copied for nxs_reply.
    void post_bytes(int64_t ticket, nxs_done_fn done, void *user, const uint8_t *reply,
                    size_t length) {
        if (done) {
            done(user, ticket, reply, length);
            return;
        }
        std::vector<uint8_t> copy;
        try {
            copy.assign(reply, reply + length);
        } catch (const std::bad_alloc &) {
            copy = nxs::refused_record("out of host memory");
        }
        post(ticket, done, user, std::move(copy));
    }

    void post(int64_t ticket, nxs_done_fn done, void *user, std::vector<uint8_t> reply) {
        if (done) {
            done(user, ticket, reply.data(), reply.size());
        } else {
            std::lock_guard<std::mutex> lock(mu);
            auto &k = kept[ticket];
            k.reply = std::move(reply);
            k.complete = true;
            cv_done.notify_all();
        }
    }

    // A free workspace of `bytes`, allocated within the budget, or waited for (with `wait`
    // false, nullptr where it would wait).
    nxs::Workspace *acquire_workspace(size_t bytes, std::unique_lock<std::mutex> &lock,
                                      bool wait = true) {
        using namespace nxs;
        for (;;) {
            for (auto &w : workspaces)
                if (!w->busy && w->bytes == bytes) {
                    w->busy = true;
                    return w.get();
                }
            bool any_busy = false;
            for (auto &w : workspaces) any_busy = any_busy || w->busy;
            // Within the budget, or alone on the device: allocate.
            if (allocated + bytes <= budget || !any_busy) {
                void *ptr = nullptr;
                if (hipMalloc(&ptr, bytes) == hipSuccess) {
                    workspaces.push_back(std::make_unique<Workspace>(Workspace{ptr, bytes, true}));
                    allocated += bytes;
                    return workspaces.back().get();
                }
                (void)hipGetLastError();
                if (!any_busy) {
                    // Free idle workspaces of other sizes and try once more.
                    bool freed = false;
                    for (auto it = workspaces.begin(); it != workspaces.end();) {
                        if (!(*it)->busy) {
                            (void)hipFree((*it)->ptr);
                            allocated -= (*it)->bytes;
                            it = workspaces.erase(it);
                            freed = true;
                        } else {
                            ++it;
                        }
                    }
                    if (!freed) throw Error{"out of device memory for the matcher"};
                    continue;
                }
            } else {
                // Over the budget: give back an idle workspace of another size first.
                for (auto it = workspaces.begin(); it != workspaces.end(); ++it)
                    if (!(*it)->busy) {
                        (void)hipFree((*it)->ptr);
                        allocated -= (*it)->bytes;
                        workspaces.erase(it);
                        break;
                    }
                if (allocated + bytes <= budget) continue;
            }
            // Admission: a run in flight completes and frees memory. Every busy workspace
            // belongs to such a run (device runs keep workspaces of their own), so the wait
            // ends, unless the context closes first.
            if (!wait) return nullptr;
            if (stopping) throw Error{"the context is closed"};
            cv_free.wait(lock);
        }
    }

    // A run taken from the queue with what it holds: a run slot of its segment and, with a
    // matcher, a workspace from the pool.
    struct Admitted {
        nxs::Job job;
        std::shared_ptr<nxs::Segment> seg;
        int slot = -1;
        nxs::Workspace *ws = nullptr;
    };
    enum Admission { ADMITTED, REFUSED, BUSY };

    // Takes a run slot of segment `id` and its matcher workspace for `a`, waiting for them;
    // with `wait` false it takes them only if both are free now, else nothing (BUSY).
    // Called with `lock` (on mu) held; a refusal's reason goes to `why`.
    Admission admit(uint32_t id, Admitted &a, std::unique_lock<std::mutex> &lock, bool wait,
                    std::string &why) {
        using namespace nxs;
        auto it = segments.find(id);
        if (it == segments.end()) {
            why = "segment " + std::to_string(id) + " is not configured";
            return REFUSED;
        }
        std::shared_ptr<Segment> seg = it->second;
        int slot = -1;
        for (;;) {
            for (int k = 0; k < RUN_SLOTS && slot < 0; k++) {
                const int s = (seg->next_slot + k) % RUN_SLOTS;
                if (!seg->slots[s].busy) slot = s;
            }
            if (slot >= 0) break;
            if (!wait) return BUSY;
            if (stopping) {
                why = "the context is closed";
                return REFUSED;
            }
            cv_free.wait(lock);
        }
        seg->slots[slot].busy = true;
        Workspace *ws = nullptr;
        try {
            if (seg->workspace_bytes) {
                ws = acquire_workspace(seg->workspace_bytes, lock, wait);
                if (!ws) {
                    seg->slots[slot].busy = false;
                    return BUSY;
                }
            }
        } catch (const Error &e) {
            why = e.reason;
        } catch (const std::bad_alloc &) {
            why = "out of host memory";
        }
        if (!why.empty()) {
            seg->slots[slot].busy = false;
            cv_free.notify_all();
            return REFUSED;
        }
        seg->next_slot = (slot + 1) % RUN_SLOTS;
        a.seg = std::move(seg);
        a.slot = slot;
        a.ws = ws;
        return ADMITTED;
    }

    // Gives back what an admitted run holds.
    void release(Admitted &a) {
        std::lock_guard<std::mutex> lock(mu);
        a.seg->slots[a.slot].busy = false;
        if (a.ws) a.ws->busy = false;
        cv_free.notify_all();
    }

    // Refuses an admitted run whose enqueueing failed: the work it already enqueued (on its
    // segment's stream, and on `also` if not null) may still use its slot and workspace, so
    // that finishes first.
    void refuse_admitted(Admitted &a, const std::string &why, hipStream_t also = nullptr) {
        if (a.seg->upload_stream) (void)hipStreamSynchronize(a.seg->upload_stream);
        (void)hipStreamSynchronize(a.seg->stream);
        if (a.seg->download_stream) (void)hipStreamSynchronize(a.seg->download_stream);
        if (also) (void)hipStreamSynchronize(also);
        release(a);
        post(a.job.ticket, a.job.done, a.job.user, nxs::refused_record(why));
    }

    static std::string reason_of(std::exception_ptr e) {
        try {
            std::rethrow_exception(e);
        } catch (const nxs::Error &x) {
            return x.reason;
        } catch (const std::bad_alloc &) {
            return "out of host memory";
        } catch (const std::exception &x) {
            return std::string("internal error: ") + x.what();
        } catch (...) {
            return "internal error";
        }
    }

    // Enqueues a run; the reply comes from the completion thread.
    void start_run(nxs::Job &job) {
        using namespace nxs;
        std::vector<Admitted> group(1);
        group[0].job = std::move(job);
        std::string why;
        Admission admitted;
        {
            std::unique_lock<std::mutex> lock(mu);
            admitted = admit(group[0].job.parsed->segment, group[0], lock, true, why);
        }
        if (admitted != ADMITTED) {
            Job &j = group[0].job;
            post(j.ticket, j.done, j.user, refused_record(why));
            return;
        }
        if (coalesce_max > 1 && group[0].seg->matcher >= 0) gather(group);
        if (group.size() == 1)
            run_one(group[0]);
        else
            run_group(group);
    }

    void run_one(Admitted &a) {
        using namespace nxs;
        Segment &seg = *a.seg;
        RunSlot &slot = seg.slots[a.slot];
        const Request &q = *a.job.parsed;
        std::exception_ptr failed;
        {
            std::lock_guard<std::mutex> guard(seg.mu);
            try {
                order_before(seg, seg.stream);
                Enqueued e = enqueue(seg, q.inputs, q.entries, {}, seg.stream, &slot,
                                     slot.events, a.ws ? a.ws->ptr : nullptr);
                const size_t length =
                    enqueue_downloads(seg, e, slot, seg.members.size(), seg.stream);
                order_after(seg, seg.stream);
                std::lock_guard<std::mutex> lock(mu);
                pending.push_back(Pending{a.job.ticket, a.seg, a.slot, a.ws, length, a.job.done,
                                          a.job.user});
                cv_pending.notify_all();
                return;
            } catch (...) {
                seg.last_stream = seg.stream;  // whatever was enqueued went there
                failed = std::current_exception();
            }
        }
        refuse_admitted(a, reason_of(failed));
    }

    // Whether segment `seg` can join a group led by `lead`: its first Disparity member
    // searches the full range at the lead's size, disparities and paths.
    static bool same_matcher(const nxs::Segment &seg, const nxs::Segment &lead) {
        if (seg.matcher < 0 || lead.matcher < 0) return false;
        const nxs::Member &m = seg.members[seg.matcher], &l = lead.members[lead.matcher];
        return seg.g.width == lead.g.width && seg.g.height == lead.g.height &&
               m.disparities == l.disparities && m.paths == l.paths;
    }

    // Coalescing: runs of other segments queued behind the group's first, whose matchers
    // are of its size class, join it while their run slot and workspace are free now (they
    // never wait, so they hold nothing another run waits for). Jobs are taken in queue
    // order and never ahead of an earlier job naming their own segment.
    void gather(std::vector<Admitted> &group) {
        using namespace nxs;
        std::unique_lock<std::mutex> lock(mu);
        std::vector<uint32_t> passed = {group[0].job.parsed->segment};
        auto seen = [&](uint32_t id) {
            return std::find(passed.begin(), passed.end(), id) != passed.end();
        };
        size_t scanned = 0;
        for (auto it = jobs.begin();
             it != jobs.end() && group.size() < (size_t)coalesce_max && scanned < GATHER_SCAN;
             scanned++) {
            Job &job = *it;
            if (!job.parsed) {
                try {
                    job.parsed = std::make_unique<Request>(parse(job.record.data(), job.record.size()));
                } catch (...) {
                    ++it;  // refused in its turn; it names no segment
                    continue;
                }
            }
            const Request &q = *job.parsed;
            bool take = false;
            if (q.kind == RUN && !seen(q.segment)) {
                auto s = segments.find(q.segment);
                take = s != segments.end() && same_matcher(*s->second, *group[0].seg);
            }
            if (!seen(q.segment)) passed.push_back(q.segment);
            Admitted a;
            std::string why;
            if (!take || admit(q.segment, a, lock, false, why) != ADMITTED) {
                ++it;
                continue;
            }
            a.job = std::move(job);
            it = jobs.erase(it);
            group.push_back(std::move(a));
        }
    }

    // One group: each run's stages before its matcher on its own segment's stream, then the
    // matchers as one launch over a descriptor table on the first run's stream once every
    // run is ready for it, then each run's remaining stages and downloads on its own
    // stream after that launch. A run whose enqueueing fails is refused alone.
    void run_group(std::vector<Admitted> &group) {
        using namespace nxs;
        struct Part {
            Enqueued q;
            PairDesc pd{};
            MatchParams mp{};
            bool live = false;
        };
        std::vector<Part> parts(group.size());
        std::vector<std::pair<size_t, std::string>> refused;
        TableSlot *table = nullptr;
        hipStream_t lead_stream = nullptr;
        {
            // Only this thread holds more than one segment lock at a time.
            std::vector<std::unique_lock<std::mutex>> locks;
            for (auto &a : group) locks.emplace_back(a.seg->mu);
            for (size_t k = 0; k < group.size(); k++) {
                Admitted &a = group[k];
                Segment &seg = *a.seg;
                RunSlot &slot = seg.slots[a.slot];
                const Request &q = *a.job.parsed;
                const size_t m = (size_t)seg.matcher;
                try {
                    order_before(seg, seg.stream);
                    parts[k].q = enqueue_entries(seg, q.inputs, q.entries, {}, seg.stream, &slot,
                                                 slot.events);
                    enqueue_members(seg, q.inputs, parts[k].q, 0, m, seg.stream, slot.events,
                                    a.ws ? a.ws->ptr : nullptr);
                    parts[k].pd = matcher_pair(seg, seg.members[m], q.inputs[m], parts[k].q,
                                               a.ws ? a.ws->ptr : nullptr, parts[k].mp);
                    note_outputs(parts[k].q, seg.members[m]);
                    parts[k].live = true;
                } catch (...) {
                    seg.last_stream = seg.stream;
                    refused.emplace_back(k, reason_of(std::current_exception()));
                }
            }
            // The matchers. A run's stages before its matcher end with its event m (the
            // event after its uploads when the matcher is its first member).
            std::vector<size_t> live;
            for (size_t k = 0; k < group.size(); k++)
                if (parts[k].live) live.push_back(k);
            try {
                if (live.size() == 1) {
                    Part &p = parts[live[0]];
                    HIP_TRY(launch_match(p.mp, p.pd.left, p.pd.right, p.pd.b, p.pd.disparity,
                                         nullptr, group[live[0]].seg->stream));
                } else if (live.size() > 1) {
                    Admitted &lead = group[live[0]];
                    lead_stream = lead.seg->stream;
                    for (size_t j = 1; j < live.size(); j++) {
                        const Admitted &a = group[live[j]];
                        HIP_TRY(hipStreamWaitEvent(
                            lead_stream, a.seg->slots[a.slot].events[a.seg->matcher], 0));
                    }
                    table = next_table();
                    for (size_t j = 0; j < live.size(); j++) table->host[j] = parts[live[j]].pd;
                    HIP_TRY(hipMemcpyAsync(table->dev, table->host, live.size() * sizeof(PairDesc),
                                           hipMemcpyHostToDevice, lead_stream));
                    const MatchParams &mp = parts[live[0]].mp;
                    HIP_TRY(launch_match_batch(mp, table->dev, (int)live.size(),
                                               schedule_for(mp.width, mp.height, mp.disparities,
                                                            mp.paths),
                                               lead_stream));
                    HIP_TRY(hipEventRecord(table->done, lead_stream));
                    table->used = true;
                    std::lock_guard<std::mutex> lock(mu);
                    coalesced_launches++;
                    coalesced_pairs += (int64_t)live.size();
                }
            } catch (...) {
                const std::string why = reason_of(std::current_exception());
                for (size_t k : live) {
                    parts[k].live = false;
                    group[k].seg->last_stream = group[k].seg->stream;
                    refused.emplace_back(k, why);
                }
            }
            // Each run's remaining stages and downloads, after the matchers' launch.
            for (size_t k : live) {
                if (!parts[k].live) continue;
                Admitted &a = group[k];
                Segment &seg = *a.seg;
                RunSlot &slot = seg.slots[a.slot];
                const Request &q = *a.job.parsed;
                const size_t m = (size_t)seg.matcher;
                try {
                    if (table && seg.stream != lead_stream)
                        HIP_TRY(hipStreamWaitEvent(seg.stream, table->done, 0));
                    HIP_TRY(hipEventRecord(slot.events[m + 1], seg.stream));
                    enqueue_members(seg, q.inputs, parts[k].q, m + 1, seg.members.size(), seg.stream,
                                    slot.events, a.ws ? a.ws->ptr : nullptr);
                    const size_t length =
                        enqueue_downloads(seg, parts[k].q, slot, seg.members.size(), seg.stream);
                    order_after(seg, seg.stream);
                    std::lock_guard<std::mutex> lock(mu);
                    pending.push_back(Pending{a.job.ticket, a.seg, a.slot, a.ws, length, a.job.done,
                                              a.job.user});
                    cv_pending.notify_all();
                } catch (...) {
                    parts[k].live = false;
                    seg.last_stream = seg.stream;
                    refused.emplace_back(k, reason_of(std::current_exception()));
                }
            }
        }
        // Refusals after the segment locks are given back (a reply callback may run a
        // segment's device run).
        for (auto &r : refused) refuse_admitted(group[r.first], r.second, lead_stream);
    }

    // Forgets a segment. Runs in flight keep it alive until they are answered; its own
    // device memory (arena, device-run workspace) goes when the last of them is, after its
    // streams' work (the segment's destructor).
    void drop_segment(uint32_t id) {
        std::shared_ptr<nxs::Segment> old;
        {
            std::lock_guard<std::mutex> lock(mu);
            auto it = segments.find(id);
            if (it == segments.end()) return;
            old = it->second;
            segments.erase(it);
        }
    }

    void handle(nxs::Job &job) {
        using namespace nxs;
        try {
            if (!job.parsed)
                job.parsed = std::make_unique<Request>(parse(job.record.data(), job.record.size()));
        } catch (const Error &e) {
            post(job.ticket, job.done, job.user, refused_record(e.reason));
            return;
        } catch (const std::bad_alloc &) {
            post(job.ticket, job.done, job.user, refused_record("out of host memory"));
            return;
        }
        const Request &q = *job.parsed;
        if (q.kind == RUN) {
            start_run(job);
            return;
        }
        std::vector<uint8_t> reply;
        try {
            // main.rs:146-160: a refused (re)configuration leaves the segment unconfigured.
            drop_segment(q.segment);
            if (q.kind == CONFIGURE) {
                std::shared_ptr<Segment> s = configure(q.segment, q.geometry, q.members);
                std::lock_guard<std::mutex> lock(mu);
                segments[q.segment] = std::move(s);
                reply = {TAG_CONFIGURED};
            }
        } catch (const Error &e) {
            reply = refused_record(e.reason);
        } catch (const std::bad_alloc &) {
            reply = refused_record("out of host memory");
        }
        post(job.ticket, job.done, job.user, std::move(reply));
    }

    void work() {
        (void)hipSetDevice(device);
        for (;;) {
            nxs::Job job;
            {
                std::unique_lock<std::mutex> lock(mu);
                cv_jobs.wait(lock, [&] { return stopping || !jobs.empty(); });
                if (jobs.empty()) {
                    worker_done = true;
                    cv_pending.notify_all();
                    return;
                }
                job = std::move(jobs.front());
                jobs.pop_front();
            }
            handle(job);
        }
    }

    // Replies to runs as their last events complete, in any order across segments.
    void complete() {
        using namespace nxs;
        (void)hipSetDevice(device);
        for (;;) {
            std::vector<int64_t> tickets;
            {
                std::unique_lock<std::mutex> lock(mu);
                cv_pending.wait(lock, [&] { return !pending.empty() || worker_done; });
                if (pending.empty()) return;
            }
            bool progressed = false;
            for (;;) {
                Pending p;
                bool found = false;
                hipError_t status = hipSuccess;
                {
                    std::lock_guard<std::mutex> lock(mu);
                    for (auto it = pending.begin(); it != pending.end(); ++it) {
                        status = hipEventQuery(it->segment->slots[it->slot].done);
                        if (status != hipErrorNotReady) {
                            p = std::move(*it);
                            pending.erase(it);
                            found = true;
                            break;
                        }
                    }
                }
                if (!found) break;
                progressed = true;
                std::vector<uint8_t> refusal;
                RunSlot &slot = p.segment->slots[p.slot];
                try {
                    if (status != hipSuccess)
                        throw Error{std::string("HIP error: ") + hipGetErrorString(status)};
                    finish_record(slot, p.segment->members.size());
                } catch (const Error &e) {
                    refusal = refused_record(e.reason);
                } catch (const std::bad_alloc &) {
                    refusal = refused_record("out of host memory");
                } catch (const std::exception &e) {
                    refusal = refused_record(std::string("internal error: ") + e.what());
                }
                // The reply is the slot's staging: the slot is given back after it is posted.
                if (refusal.empty())
                    post_bytes(p.ticket, p.done, p.user, slot.download, p.length);
                else
                    post(p.ticket, p.done, p.user, std::move(refusal));
                {
                    std::lock_guard<std::mutex> lock(mu);
                    slot.busy = false;
                    if (p.workspace) p.workspace->busy = false;
                    cv_free.notify_all();
                }
            }
            if (!progressed) std::this_thread::sleep_for(std::chrono::microseconds(20));
        }
    }
};

extern "C" {

int nxs_open(int device, nxs_ctx **out, char *err, size_t errlen) {
    auto report = [&](const std::string &why) {
        if (err && errlen) {
            std::strncpy(err, why.c_str(), errlen - 1);
            err[errlen - 1] = 0;
        }
        nxs::g_last_error = why;
    };
    if (!out) {
        report("no output pointer");
        return NXS_EINVAL;
    }
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || device < 0 || device >= count) {
        report(std::string("no HIP device ") + std::to_string(device) +
               (e != hipSuccess ? std::string(": ") + hipGetErrorString(e) : std::string()));
        return NXS_EDEVICE;
    }
    e = hipSetDevice(device);
    if (e != hipSuccess) {
        report(std::string("hipSetDevice: ") + hipGetErrorString(e));
        return NXS_EDEVICE;
    }
    nxs_ctx *ctx = nullptr;
    try {
        ctx = new nxs_ctx();
        ctx->device = device;
        // The matcher workspaces may take half the device memory free now, unless
        // NXS_HBM_BUDGET_MB says otherwise (nxs_set_hbm_budget changes it later).
        size_t free_bytes = 0, total_bytes = 0;
        if (hipMemGetInfo(&free_bytes, &total_bytes) != hipSuccess) free_bytes = (size_t)8 << 30;
        ctx->budget = free_bytes / 2;
        if (const char *mb = std::getenv("NXS_HBM_BUDGET_MB"))
            ctx->budget = (size_t)std::atoll(mb) << 20;
        if (const char *c = std::getenv("NXS_COALESCE"))
            ctx->coalesce_max = std::max(1, std::min(nxs_ctx::MAX_COALESCE, std::atoi(c)));
        ctx->worker = std::thread([ctx] { ctx->work(); });
        ctx->completer = std::thread([ctx] { ctx->complete(); });
    } catch (const std::exception &x) {
        // Out of memory or threads: stop what was started and report.
        if (ctx) {
            {
                std::lock_guard<std::mutex> lock(ctx->mu);
                ctx->stopping = true;
            }
            ctx->cv_jobs.notify_all();
            if (ctx->worker.joinable()) ctx->worker.join();
            delete ctx;
        }
        report(std::string("nxs_open: ") + x.what());
        return NXS_ENOMEM;
    }
    *out = ctx;
    return NXS_OK;
}

int nxs_set_hbm_budget(nxs_ctx *ctx, size_t bytes) {
    if (!ctx) return nxs::fail(NXS_EINVAL, "no context");
    std::lock_guard<std::mutex> lock(ctx->mu);
    ctx->budget = bytes;
    ctx->cv_free.notify_all();
    return NXS_OK;
}

int nxs_stats(nxs_ctx *ctx, size_t *workspace_bytes, int *runs_in_flight) {
    if (!ctx) return nxs::fail(NXS_EINVAL, "no context");
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (workspace_bytes) *workspace_bytes = ctx->allocated;
    if (runs_in_flight) *runs_in_flight = (int)ctx->pending.size();
    return NXS_OK;
}

int nxs_coalescing(nxs_ctx *ctx, int max_pairs, int64_t *launches, int64_t *pairs) {
    if (!ctx) return nxs::fail(NXS_EINVAL, "no context");
    if (max_pairs > nxs_ctx::MAX_COALESCE)
        return nxs::fail(NXS_EINVAL, "at most " + std::to_string(nxs_ctx::MAX_COALESCE) +
                                         " pairs share a launch");
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (max_pairs >= 0) ctx->coalesce_max = std::max(1, max_pairs);
    if (launches) *launches = ctx->coalesced_launches;
    if (pairs) *pairs = ctx->coalesced_pairs;
    return NXS_OK;
}

void nxs_close(nxs_ctx *ctx) {
    if (!ctx) return;
    std::deque<nxs::Job> pending;
    {
        std::lock_guard<std::mutex> lock(ctx->mu);
        ctx->stopping = true;
        pending.swap(ctx->jobs);
    }
    ctx->cv_jobs.notify_all();
    ctx->cv_free.notify_all();  // a run waiting for admission is refused
    // Requests not started fail; runs already on the device finish and are answered.
    for (auto &job : pending)
        ctx->post(job.ticket, job.done, job.user, nxs::refused_record("the context is closed"));
    if (ctx->worker.joinable()) ctx->worker.join();
    if (ctx->completer.joinable()) ctx->completer.join();
    (void)hipSetDevice(ctx->device);
    ctx->segments.clear();
    for (auto &w : ctx->workspaces) (void)hipFree(w->ptr);
    ctx->workspaces.clear();
    for (auto &t : ctx->tables) {
        if (t.done) {
            (void)hipEventSynchronize(t.done);
            (void)hipEventDestroy(t.done);
        }
        if (t.dev) (void)hipFree(t.dev);
        if (t.host) (void)hipHostFree(t.host);
    }
    delete ctx;
}

int64_t nxs_submit(nxs_ctx *ctx, const uint8_t *record, size_t len, nxs_done_fn done, void *user) {
    if (!ctx || (!record && len)) return nxs::fail(NXS_EINVAL, "no context or record");
    try {
        // The copy (megabytes with frames) is made before the context's lock is taken.
        nxs::Job job{0, std::vector<uint8_t>(record, record + len), done, user, nullptr};
        std::lock_guard<std::mutex> lock(ctx->mu);
        if (ctx->stopping) return nxs::fail(NXS_EINVAL, "the context is closing");
        const int64_t ticket = ctx->next_ticket++;
        job.ticket = ticket;
        ctx->jobs.push_back(std::move(job));
        if (!done) ctx->kept[ticket];
        ctx->cv_jobs.notify_one();
        return ticket;
    } catch (const std::bad_alloc &) {
        return nxs::fail(NXS_ENOMEM, "out of host memory for the record");
    }
}

int nxs_wait(nxs_ctx *ctx, int64_t ticket, int timeout_ms) {
    if (!ctx) return NXS_EINVAL;
    std::unique_lock<std::mutex> lock(ctx->mu);
    auto it = ctx->kept.find(ticket);
    if (it == ctx->kept.end()) return nxs::fail(NXS_EINVAL, "unknown ticket");
    auto ready = [&] { return ctx->kept[ticket].complete; };
    if (timeout_ms < 0) {
        ctx->cv_done.wait(lock, ready);
        return NXS_OK;
    }
    return ctx->cv_done.wait_for(lock, std::chrono::milliseconds(timeout_ms), ready) ? NXS_OK
                                                                                       : NXS_TIMEOUT;
}

int nxs_reply(nxs_ctx *ctx, int64_t ticket, const uint8_t **reply, size_t *len) {
    if (!ctx || !reply || !len) return NXS_EINVAL;
    std::lock_guard<std::mutex> lock(ctx->mu);
    auto it = ctx->kept.find(ticket);
    if (it == ctx->kept.end() || !it->second.complete) return nxs::fail(NXS_EINVAL, "no reply");
    *reply = it->second.reply.data();
    *len = it->second.reply.size();
    return NXS_OK;
}

void nxs_release_reply(nxs_ctx *ctx, int64_t ticket) {
    if (!ctx) return;
    std::lock_guard<std::mutex> lock(ctx->mu);
    auto it = ctx->kept.find(ticket);
    if (it != ctx->kept.end() && it->second.complete) ctx->kept.erase(it);
}

int nxs_poll(nxs_ctx *ctx, int max) {
    if (!ctx) return NXS_EINVAL;
    std::lock_guard<std::mutex> lock(ctx->mu);
    int n = 0;
    for (auto &kv : ctx->kept)
        if (kv.second.complete && (max <= 0 || n < max)) n++;
    return n;
}

size_t nxs_match_workspace_size(int width, int height, int disparities, int paths) {
    return nxs::match_bytes(width, height, disparities, paths);
}

int nxs_match_device(const uint8_t *left, const uint8_t *right, int width, int height,
                     int disparities, int paths, int p1, int p2, float uniqueness,
                     float *disparity, void *workspace, size_t workspace_bytes, void *stream) {
    try {
        nxs::check_matching(width, height, disparities, paths, p1, p2, uniqueness);
    } catch (const nxs::Error &e) {
        return nxs::fail(NXS_EINVAL, e.reason);
    }
    if (!left || !right || !disparity || !workspace)
        return nxs::fail(NXS_EINVAL, "a device pointer is null");
    if (workspace_bytes < nxs::match_bytes(width, height, disparities, paths))
        return nxs::fail(NXS_EINVAL, "the workspace is smaller than nxs_match_workspace_size");
    nxs::MatchBuffers b;
    nxs::match_carve(b, workspace, width, height, disparities, paths);
    nxs::MatchParams mp{width, height, disparities, paths, p1, p2, uniqueness};
    hipError_t e = nxs::launch_match(mp, left, right, b, disparity, nullptr, (hipStream_t)stream);
    if (e != hipSuccess) return nxs::fail(NXS_EDEVICE, hipGetErrorString(e));
    return NXS_OK;
}

}  // extern "C"

// ---- The matcher without allocation ----
// The workspace of nxs_match_device_noalloc: one array per device, in the code object's
// device memory, which each device holds from the library's load until the process ends.
// Host code never uses its name as an address: that is a host-side stand-in, and a kernel
// given it would fault. Its device address, per device, comes from hipGetSymbolAddress with
// that device current, on the entry point's first call there.
static __device__ __align__(256) uint8_t scratch[NXS_NOALLOC_SCRATCH_BYTES];

namespace nxs {

static std::mutex g_scratch_mu;
static std::map<int, void *> g_scratch;  // device -> the array's device address there

// The array's address on the current device (taken once per device), or nullptr with the
// reason in *why.
static void *scratch_address(std::string *why) {
    int device = 0;
    hipError_t e = hipGetDevice(&device);
    if (e != hipSuccess) {
        *why = std::string("hipGetDevice: ") + hipGetErrorString(e);
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    auto it = g_scratch.find(device);
    if (it != g_scratch.end()) return it->second;
    void *ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
        *why = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
        return nullptr;
    }
    g_scratch[device] = ptr;
    return ptr;
}

}  // namespace nxs

extern "C" {

int nxs_match_device_noalloc(const uint8_t *left, const uint8_t *right, int width, int height,
                             int disparities, int paths, int p1, int p2, float uniqueness,
                             float *disparity, void *stream) {
    try {
        nxs::check_matching(width, height, disparities, paths, p1, p2, uniqueness);
    } catch (const nxs::Error &e) {
        return nxs::fail(NXS_EINVAL, e.reason);
    }
    if (!left || !right || !disparity) return nxs::fail(NXS_EINVAL, "a device pointer is null");
    const size_t need = nxs::match_bytes(width, height, disparities, paths);
    if (need > (size_t)NXS_NOALLOC_SCRATCH_BYTES)
        return nxs::fail(NXS_EINVAL,
                         "a " + std::to_string(width) + "x" + std::to_string(height) + "x" +
                             std::to_string(disparities) + " match needs " + std::to_string(need) +
                             " bytes of workspace, more than the " +
                             std::to_string((size_t)NXS_NOALLOC_SCRATCH_BYTES) +
                             " of nxs_match_device_noalloc's scratch: use nxs_match_device");
    std::string why;
    void *workspace = nxs::scratch_address(&why);
    if (!workspace) return nxs::fail(NXS_EDEVICE, why);
    // Every buffer carved from it is written before it is read, so what an earlier call left
    // there needs no clearing.
    nxs::MatchBuffers b;
    nxs::match_carve(b, workspace, width, height, disparities, paths);
    nxs::MatchParams mp{width, height, disparities, paths, p1, p2, uniqueness};
    hipError_t e = nxs::launch_match(mp, left, right, b, disparity, nullptr, (hipStream_t)stream);
    if (e != hipSuccess) return nxs::fail(NXS_EDEVICE, hipGetErrorString(e));
    return NXS_OK;
}

int nxs_match_host(const uint8_t *left, const uint8_t *right, int width, int height,
                   int disparities, int paths, int p1, int p2, float uniqueness,
                   uint64_t *census_left, uint64_t *census_right, uint8_t *costs, uint16_t *sums,
                   float *selected, uint32_t *keys, float *checked, float *disparity) {
    try {
        nxs::check_matching(width, height, disparities, paths, p1, p2, uniqueness);
    } catch (const nxs::Error &e) {
        return nxs::fail(NXS_EINVAL, e.reason);
    }
    if (!left || !right) return nxs::fail(NXS_EINVAL, "no input image");
    const size_t P = (size_t)width * height, N = P * disparities;
    const size_t ws = nxs::match_bytes(width, height, disparities, paths);
    uint8_t *dev = nullptr;
    const size_t total = ws + 2 * P + 4 * P + 2 * N + 1024;
    hipError_t e = hipMalloc(&dev, total);
    if (e != hipSuccess) return nxs::fail(NXS_ENOMEM, hipGetErrorString(e));
    uint8_t *d_left = dev + ws, *d_right = d_left + P;
    float *d_out = reinterpret_cast<float *>(nxs::up256_ptr(d_right + P));
    uint16_t *d_sums = reinterpret_cast<uint16_t *>(nxs::up256_ptr((uint8_t *)(d_out + P)));
    nxs::MatchBuffers b;
    nxs::match_carve(b, dev, width, height, disparities, paths);
    nxs::MatchParams mp{width, height, disparities, paths, p1, p2, uniqueness};
    mp.keep_stages = true;
    int rc = NXS_OK;
    auto check = [&](hipError_t err) {
        if (err != hipSuccess && rc == NXS_OK) rc = nxs::fail(NXS_EDEVICE, hipGetErrorString(err));
    };
    check(hipMemcpy(d_left, left, P, hipMemcpyHostToDevice));
    check(hipMemcpy(d_right, right, P, hipMemcpyHostToDevice));
    if (rc == NXS_OK) check(nxs::launch_match(mp, d_left, d_right, b, d_out, sums ? d_sums : nullptr, nullptr));
    check(hipDeviceSynchronize());
    auto down = [&](void *host, const void *devp, size_t n) {
        if (host && rc == NXS_OK) check(hipMemcpy(host, devp, n, hipMemcpyDeviceToHost));
    };
    down(census_left, b.census_left, 8 * P);
    down(census_right, b.census_right, 8 * P);
    down(costs, b.costs, N);
    down(sums, d_sums, 2 * N);
    down(selected, b.selected, 4 * P);
    down(keys, b.keys, 4 * P);
    down(checked, b.checked, 4 * P);
    down(disparity, d_out, 4 * P);
    (void)hipFree(dev);
    return rc;
}

size_t nxs_band_workspace_size(int width, int height, int band, int paths) {
    return nxs::band_bytes(width, height, band, paths);
}

int nxs_match_band_device(const uint8_t *left, const uint8_t *right, const int32_t *base,
                          int width, int height, int disparities, int band, int paths,
                          int census, int p1, int p2, float uniqueness, float *disparity,
                          void *workspace, size_t workspace_bytes, void *stream) {
    try {
        nxs::check_band(width, height, disparities, band, paths, census, p1, p2, uniqueness);
    } catch (const nxs::Error &e) {
        return nxs::fail(NXS_EINVAL, e.reason);
    }
    if (!left || !right || !base || !disparity || !workspace)
        return nxs::fail(NXS_EINVAL, "a device pointer is null");
    if (workspace_bytes < nxs::band_bytes(width, height, band, paths))
        return nxs::fail(NXS_EINVAL, "the workspace is smaller than nxs_band_workspace_size");
    nxs::BandBuffers b;
    nxs::band_carve(b, workspace, width, height, band, paths);
    const nxs::BandParams bp =
        nxs::band_params(width, height, disparities, band, paths, census, p1, p2, uniqueness);
    hipError_t e = nxs::launch_band(bp, left, right, base, b, disparity, nullptr, (hipStream_t)stream);
    if (e != hipSuccess) return nxs::fail(NXS_EDEVICE, hipGetErrorString(e));
    return NXS_OK;
}

int nxs_match_band_host(const uint8_t *left, const uint8_t *right, const int32_t *base,
                        int width, int height, int disparities, int band, int paths, int census,
                        int p1, int p2, float uniqueness, uint64_t *census_left,
                        uint64_t *census_right, uint8_t *costs, uint16_t *sums, float *selected,
                        uint32_t *keys, float *checked, float *disparity) {
    try {
        nxs::check_band(width, height, disparities, band, paths, census, p1, p2, uniqueness);
    } catch (const nxs::Error &e) {
        return nxs::fail(NXS_EINVAL, e.reason);
    }
    if (!left || !right || !base) return nxs::fail(NXS_EINVAL, "no input image or base");
    const size_t P = (size_t)width * height, N = P * band;
    const size_t ws = nxs::band_bytes(width, height, band, paths);
    uint8_t *dev = nullptr;
    const size_t total = ws + 2 * P + 4 * P + 4 * P + 2 * N + 1024;
    hipError_t e = hipMalloc(&dev, total);
    if (e != hipSuccess) return nxs::fail(NXS_ENOMEM, hipGetErrorString(e));
    uint8_t *d_left = dev + ws, *d_right = d_left + P;
    int32_t *d_base = reinterpret_cast<int32_t *>(nxs::up256_ptr(d_right + P));
    float *d_out = reinterpret_cast<float *>(nxs::up256_ptr((uint8_t *)(d_base + P)));
    uint16_t *d_sums = reinterpret_cast<uint16_t *>(nxs::up256_ptr((uint8_t *)(d_out + P)));
    nxs::BandBuffers b;
    nxs::band_carve(b, dev, width, height, band, paths);
    nxs::BandParams bp =
        nxs::band_params(width, height, disparities, band, paths, census, p1, p2, uniqueness);
    bp.keep_stages = true;
    int rc = NXS_OK;
    auto check = [&](hipError_t err) {
        if (err != hipSuccess && rc == NXS_OK) rc = nxs::fail(NXS_EDEVICE, hipGetErrorString(err));
    };
    check(hipMemcpy(d_left, left, P, hipMemcpyHostToDevice));
    check(hipMemcpy(d_right, right, P, hipMemcpyHostToDevice));
    check(hipMemcpy(d_base, base, 4 * P, hipMemcpyHostToDevice));
    if (rc == NXS_OK)
        check(nxs::launch_band(bp, d_left, d_right, d_base, b, d_out, sums ? d_sums : nullptr, nullptr));
    check(hipDeviceSynchronize());
    auto down = [&](void *host, const void *devp, size_t n) {
        if (host && rc == NXS_OK) check(hipMemcpy(host, devp, n, hipMemcpyDeviceToHost));
    };
    down(census_left, b.census_left, 8 * P);
    down(census_right, b.census_right, 8 * P);
    down(costs, b.costs, N);
    down(sums, d_sums, 2 * N);
    down(selected, b.selected, 4 * P);
    down(keys, b.keys, 4 * P);
    down(checked, b.checked, 4 * P);
    down(disparity, d_out, 4 * P);
    (void)hipFree(dev);
    return rc;
}

size_t nxs_coarse_to_fine_workspace_size(int width, int height, int band, int paths,
                                         int coarse_paths) {
    return nxs::coarse_to_fine_bytes(width, height, band, paths, coarse_paths);
}

int nxs_coarse_to_fine_device(const uint8_t *left, const uint8_t *right, int width, int height,
                              int band, int margin, int paths, int coarse_paths, int census,
                              int p1, int p2, float uniqueness, float *disparity, void *workspace,
                              size_t workspace_bytes, void *stream) {
    try {
        nxs::check_coarse_to_fine(width, height, band, margin, paths, coarse_paths, census, p1,
                                  p2, uniqueness);
    } catch (const nxs::Error &e) {
        return nxs::fail(NXS_EINVAL, e.reason);
    }
    if (!left || !right || !disparity || !workspace)
        return nxs::fail(NXS_EINVAL, "a device pointer is null");
    if (workspace_bytes < nxs::coarse_to_fine_bytes(width, height, band, paths, coarse_paths))
        return nxs::fail(NXS_EINVAL, "the workspace is smaller than nxs_coarse_to_fine_workspace_size");
    nxs::CoarseBuffers c;
    nxs::coarse_to_fine_carve(c, workspace, width, height, band, paths, coarse_paths);
    const nxs::BandParams bp =
        nxs::band_params(width, height, 128, band, paths, census, p1, p2, uniqueness);
    hipError_t e = nxs::launch_coarse_to_fine(bp, margin, coarse_paths, left, right, c, disparity,
                                              (hipStream_t)stream);
    if (e != hipSuccess) return nxs::fail(NXS_EDEVICE, hipGetErrorString(e));
    return NXS_OK;
}

int nxs_coarse_to_fine_host(const uint8_t *left, const uint8_t *right, int width, int height,
                            int band, int margin, int paths, int coarse_paths, int census, int p1,
                            int p2, float uniqueness, uint8_t *half_left, uint8_t *half_right,
                            float *coarse, int32_t *base, float *disparity) {
    try {
        nxs::check_coarse_to_fine(width, height, band, margin, paths, coarse_paths, census, p1,
                                  p2, uniqueness);
    } catch (const nxs::Error &e) {
        return nxs::fail(NXS_EINVAL, e.reason);
    }
    if (!left || !right) return nxs::fail(NXS_EINVAL, "no input image");
    const size_t P = (size_t)width * height, half = P / 4;
    const size_t ws = nxs::coarse_to_fine_bytes(width, height, band, paths, coarse_paths);
    uint8_t *dev = nullptr;
    hipError_t e = hipMalloc(&dev, ws + 2 * P + 4 * P + 1024);
    if (e != hipSuccess) return nxs::fail(NXS_ENOMEM, hipGetErrorString(e));
    uint8_t *d_left = dev + ws, *d_right = d_left + P;
    float *d_out = reinterpret_cast<float *>(nxs::up256_ptr(d_right + P));
    nxs::CoarseBuffers c;
    nxs::coarse_to_fine_carve(c, dev, width, height, band, paths, coarse_paths);
    const nxs::BandParams bp =
        nxs::band_params(width, height, 128, band, paths, census, p1, p2, uniqueness);
    int rc = NXS_OK;
    auto check = [&](hipError_t err) {
        if (err != hipSuccess && rc == NXS_OK) rc = nxs::fail(NXS_EDEVICE, hipGetErrorString(err));
    };
    check(hipMemcpy(d_left, left, P, hipMemcpyHostToDevice));
    check(hipMemcpy(d_right, right, P, hipMemcpyHostToDevice));
    if (rc == NXS_OK)
        check(nxs::launch_coarse_to_fine(bp, margin, coarse_paths, d_left, d_right, c, d_out, nullptr));
    check(hipDeviceSynchronize());
    auto down = [&](void *host, const void *devp, size_t n) {
        if (host && rc == NXS_OK) check(hipMemcpy(host, devp, n, hipMemcpyDeviceToHost));
    };
    down(half_left, c.half_left, half);
    down(half_right, c.half_right, half);
    down(coarse, c.coarse, 4 * half);
    down(base, c.base, 4 * P);
    down(disparity, d_out, 4 * P);
    (void)hipFree(dev);
    return rc;
}

int nxs_run_device(nxs_ctx *ctx, const uint8_t *record, size_t len, const nxs_tensor *entries,
                   int n_entries, nxs_tensor *outputs, int max_outputs, int *n_outputs,
                   void *stream) {
    using namespace nxs;
    if (!ctx || !record || (n_entries > 0 && !entries) || !n_outputs)
        return fail(NXS_EINVAL, "no context, record or output count");
    try {
        Request q = parse(record, len);
        if (q.kind != RUN) return fail(NXS_EINVAL, "nxs_run_device takes a run record");
        std::shared_ptr<Segment> held;
        {
            std::lock_guard<std::mutex> lock(ctx->mu);
            auto it = ctx->segments.find(q.segment);
            if (it == ctx->segments.end())
                return fail(NXS_EINVAL, "segment " + std::to_string(q.segment) + " is not configured");
            held = it->second;  // alive (with its workspace) until this call returns
        }
        Segment &seg = *held;
        std::lock_guard<std::mutex> guard(seg.mu);
        // Device runs of a segment keep a workspace of their own, outside the serving pool.
        if (seg.workspace_bytes && !seg.device_workspace) {
            hipError_t e = hipMalloc(&seg.device_workspace, seg.workspace_bytes);
            if (e != hipSuccess) {
                (void)hipGetLastError();
                seg.device_workspace = nullptr;
                return fail(NXS_ENOMEM, std::string("out of device memory for the matcher: ") +
                                            hipGetErrorString(e));
            }
        }
        void *workspace = seg.device_workspace;
        std::vector<DeviceEntry> device;
        for (int e = 0; e < n_entries; e++) {
            const nxs_tensor &t = entries[e];
            if (!t.name || t.rank > 4) return fail(NXS_EINVAL, "an entry has no name or rank > 4");
            device.push_back(DeviceEntry{t.name, t.type, std::vector<uint32_t>(t.shape, t.shape + t.rank), t.data});
        }
        hipStream_t s = stream ? (hipStream_t)stream : seg.stream;
        order_before(seg, s);
        Enqueued out;
        try {
            out = enqueue(seg, q.inputs, q.entries, device, s, nullptr, {}, workspace);
        } catch (...) {
            order_after(seg, s);  // what was enqueued before the error still orders
            throw;
        }
        order_after(seg, s);
        *n_outputs = (int)out.filled.size();
        for (int k = 0; k < (int)out.filled.size() && k < max_outputs; k++) {
            const Slot &slot = out.tensors[out.filled[k]];
            nxs_tensor &t = outputs[k];
            t.name = out.filled[k];
            t.type = slot.type;
            t.rank = (uint8_t)slot.shape.size();
            for (int d = 0; d < 4; d++) t.shape[d] = d < t.rank ? slot.shape[d] : 0;
            t.data = slot.dev;
        }
        return NXS_OK;
    } catch (const Error &e) {
        return fail(NXS_EINVAL, e.reason);
    } catch (const std::bad_alloc &) {
        return fail(NXS_ENOMEM, "out of host memory");
    } catch (const std::exception &e) {
        return fail(NXS_EUNKNOWN, std::string("internal error: ") + e.what());
    }
}

int nxs_profile_start(int capacity) {
    using namespace nxs;
    std::lock_guard<std::mutex> lock(g_profiler.mu);
    try {
        g_profiler.begin.reserve(std::max(capacity, 0));
        g_profiler.end.reserve(std::max(capacity, 0));
        g_profiler.kernel.reserve(std::max(capacity, 0));
    } catch (const std::bad_alloc &) {
        return fail(NXS_ENOMEM, "out of host memory for the profiler");
    }
    while ((int)g_profiler.begin.size() < capacity) {
        hipEvent_t a, b;
        if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess)
            return fail(NXS_EDEVICE, "hipEventCreate failed");
        g_profiler.begin.push_back(a);
        g_profiler.end.push_back(b);
        g_profiler.kernel.push_back(0);
    }
    g_profiler.used = 0;
    g_profiler.on = true;
    return NXS_OK;
}

int nxs_profile_stop(double *total_ms, int *launches, int kernels) {
    using namespace nxs;
    std::lock_guard<std::mutex> lock(g_profiler.mu);
    g_profiler.on = false;
    for (int k = 0; k < kernels; k++) {
        if (total_ms) total_ms[k] = 0;
        if (launches) launches[k] = 0;
    }
    for (int i = 0; i < g_profiler.used; i++) {
        if (hipEventSynchronize(g_profiler.end[i]) != hipSuccess)
            return fail(NXS_EDEVICE, "hipEventSynchronize failed");
        float ms = 0;
        if (hipEventElapsedTime(&ms, g_profiler.begin[i], g_profiler.end[i]) != hipSuccess)
            return fail(NXS_EDEVICE, "hipEventElapsedTime failed");
        const int k = g_profiler.kernel[i];
        if (k < kernels) {
            if (total_ms) total_ms[k] += ms;
            if (launches) launches[k]++;
        }
    }
    g_profiler.used = 0;
    return NXS_OK;
}

}  // extern "C"

// ------------------------------------------------------------------------------------------
// The batched form: many pairs, grouped into size classes, each kernel one launch per
// class over a descriptor table in device memory, the whole batch captured once into a
// hipGraph and replayed. The autotuner times every aggregation schedule per class on the
// batch itself and keeps the fastest.

namespace nxs {

struct BatchClass {
    MatchParams p;
    int first = 0, count = 0;  // its descriptors: tables[first .. first+count)
    Schedule schedule;
    float tuned_ms = 0;
};

// How the batch spreads over the device: each class's pairs go in launches of `chunk` pairs
// (0: the whole class in one), and the launches are dealt to `branches` concurrent streams
// (parallel branches of the captured graph), the largest first to the least loaded. Fused
// (4 or 8): every class in one launch per kernel (launch_match_mixed), the aggregation
// meeting by recomputation every `fused` steps, whatever the classes' own schedules.
struct BatchPlan {
    int branches = 1;
    int chunk = 0;
    int fused = 0;
    int queue = 0;  // fused: 1 = the aggregation in persistent blocks taking line groups from a
                    // device work queue (one per branch)
};

// One launch chain of the batch: pairs tables[first .. first+count) of class `cls`.
struct BatchLaunch {
    int cls, first, count, branch;
};

constexpr int MAX_BRANCHES = 4;

inline std::string plan_text(const BatchPlan &p) {
    const std::string queue = p.queue ? ", the aggregation's blocks persistent on a work queue" : "";
    if (p.fused && p.branches <= 1)
        return "every class in one launch per kernel, checkpoints every " +
               std::to_string(p.fused) + " steps" + queue;
    if (p.fused)
        return "the classes on " + std::to_string(p.branches) +
               " branches, each branch's in one launch per kernel, checkpoints every " +
               std::to_string(p.fused) + " steps" + queue;
    return std::to_string(p.branches) + " branches, " +
           (p.chunk ? std::to_string(p.chunk) + " pairs a launch" : std::string("a class a launch"));
}

}  // namespace nxs

struct nxs_batch {
    std::vector<nxs::BatchClass> classes;
    void *workspace = nullptr;
    nxs::PairDesc *tables = nullptr;
    uint32_t *queues = nullptr;  // two counters a branch, for fused plans with a work queue
    hipStream_t stream = nullptr;
    // The graph's parallel branches, with the events that fork them from `stream` and join
    // them back (all at the default priority).
    hipStream_t branch[nxs::MAX_BRANCHES] = {};
    hipEvent_t fork = nullptr, join[nxs::MAX_BRANCHES] = {};
    nxs::BatchPlan plan;
    std::vector<nxs::BatchLaunch> launches;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::string description;

    ~nxs_batch() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (stream) (void)hipStreamSynchronize(stream);
        for (auto b : branch)
            if (b) (void)hipStreamSynchronize(b);
        if (tables) (void)hipFree(tables);
        if (queues) (void)hipFree(queues);
        if (workspace) (void)hipFree(workspace);
        if (stream) (void)hipStreamDestroy(stream);
        for (auto b : branch)
            if (b) (void)hipStreamDestroy(b);
        if (fork) (void)hipEventDestroy(fork);
        for (auto e : join)
            if (e) (void)hipEventDestroy(e);
    }

    // The launches of a plan: each class cut into chunks, dealt largest first (by cost
    // evaluations) to the least loaded branch.
    std::vector<nxs::BatchLaunch> deal(const nxs::BatchPlan &plan) const {
        using namespace nxs;
        std::vector<BatchLaunch> out;
        for (int c = 0; c < (int)classes.size(); c++) {
            const int step = plan.chunk > 0 ? plan.chunk : classes[c].count;
            for (int k = 0; k < classes[c].count; k += step)
                out.push_back(BatchLaunch{c, classes[c].first + k, std::min(step, classes[c].count - k), 0});
        }
        auto work = [&](const BatchLaunch &l) {
            const MatchParams &p = classes[l.cls].p;
            return (double)p.width * p.height * p.disparities * l.count;
        };
        std::stable_sort(out.begin(), out.end(),
                         [&](const BatchLaunch &a, const BatchLaunch &b) { return work(a) > work(b); });
        double load[MAX_BRANCHES] = {};
        for (auto &l : out) {
            int best = 0;
            for (int b = 1; b < plan.branches; b++)
                if (load[b] < load[best]) best = b;
            l.branch = best;
            load[best] += work(l);
        }
        return out;
    }

    // Every launch of the plan, forked from `s` onto the branches and joined back into it.
    void enqueue_all(hipStream_t s, const nxs::BatchPlan &p,
                     const std::vector<nxs::BatchLaunch> &ls) {
        using namespace nxs;
        if (p.fused) {
            // Whole classes dealt largest first (by cost evaluations) to the least loaded
            // branch; each branch's classes in one launch per kernel.
            const int nb = std::max(1, std::min(p.branches, (int)classes.size()));
            std::vector<int> order(classes.size());
            for (size_t c = 0; c < order.size(); c++) order[c] = (int)c;
            auto work = [&](int c) {
                const MatchParams &q = classes[c].p;
                return (double)q.width * q.height * q.disparities * classes[c].count;
            };
            std::stable_sort(order.begin(), order.end(),
                             [&](int a, int b) { return work(a) > work(b); });
            std::vector<MixedMember> groups[MAX_BRANCHES];
            double load[MAX_BRANCHES] = {};
            for (int c : order) {
                int best = 0;
                for (int b = 1; b < nb; b++)
                    if (load[b] < load[best]) best = b;
                groups[best].push_back(MixedMember{classes[c].p, classes[c].first, classes[c].count});
                load[best] += work(c);
            }
            auto queue = [&](int b) { return p.queue ? queues + 2 * b : nullptr; };
            if (nb == 1) {
                HIP_TRY(launch_match_mixed(groups[0].data(), (int)groups[0].size(), tables, p.fused,
                                           queue(0), s));
                return;
            }
            HIP_TRY(hipEventRecord(fork, s));
            for (int b = 0; b < nb; b++) {
                HIP_TRY(hipStreamWaitEvent(branch[b], fork, 0));
                HIP_TRY(launch_match_mixed(groups[b].data(), (int)groups[b].size(), tables,
                                           p.fused, queue(b), branch[b]));
                HIP_TRY(hipEventRecord(join[b], branch[b]));
                HIP_TRY(hipStreamWaitEvent(s, join[b], 0));
            }
            return;
        }
        if (p.branches <= 1) {
            for (const auto &l : ls)
                HIP_TRY(launch_match_batch(classes[l.cls].p, tables + l.first, l.count,
                                           classes[l.cls].schedule, s));
            return;
        }
        HIP_TRY(hipEventRecord(fork, s));
        for (int b = 0; b < p.branches; b++) HIP_TRY(hipStreamWaitEvent(branch[b], fork, 0));
        for (const auto &l : ls)
            HIP_TRY(launch_match_batch(classes[l.cls].p, tables + l.first, l.count,
                                       classes[l.cls].schedule, branch[l.branch]));
        for (int b = 0; b < p.branches; b++) {
            HIP_TRY(hipEventRecord(join[b], branch[b]));
            HIP_TRY(hipStreamWaitEvent(s, join[b], 0));
        }
    }

    // Times each plan on the whole batch (one warm run, then three) and keeps the fastest.
    void tune_plan() {
        using namespace nxs;
        hipEvent_t start, stop;
        HIP_TRY(hipEventCreate(&start));
        HIP_TRY(hipEventCreate(&stop));
        float best = 1e30f;
        std::vector<BatchPlan> plans;
        for (int branches : {1, 2, 4})
            for (int chunk : {0, 8, 2}) plans.push_back(BatchPlan{branches, chunk, 0});
        std::vector<MixedMember> all;
        for (const auto &c : classes) all.push_back(MixedMember{c.p, c.first, c.count});
        if (mixed_launch_fits(all.data(), (int)all.size()))  // at most 8 classes, and room
            for (int branches : {1, 2}) {
                for (int segment : {4, 8}) plans.push_back(BatchPlan{branches, 0, segment, 0});
                plans.push_back(BatchPlan{branches, 0, 8, 1});
            }
        for (const BatchPlan &p : plans) {
            const std::vector<BatchLaunch> ls = deal(p);
            enqueue_all(stream, p, ls);
            HIP_TRY(hipEventRecord(start, stream));
            for (int r = 0; r < 3; r++) enqueue_all(stream, p, ls);
            HIP_TRY(hipEventRecord(stop, stream));
            HIP_TRY(hipEventSynchronize(stop));
            float ms = 0;
            HIP_TRY(hipEventElapsedTime(&ms, start, stop));
            trials += "  batch: " + plan_text(p) + ": " + std::to_string(ms / 3) + " ms\n";
            if (ms < best) {
                best = ms;
                plan = p;
            }
        }
        (void)hipEventDestroy(start);
        (void)hipEventDestroy(stop);
    }

    std::string trials;  // every candidate's time, for the description

    // Times each candidate schedule on this class's pairs: one warm run, then three.
    void tune(nxs::BatchClass &c) {
        using namespace nxs;
        hipEvent_t start, stop;
        HIP_TRY(hipEventCreate(&start));
        HIP_TRY(hipEventCreate(&stop));
        float best = 1e30f;
        // Every prefetch depth and lane count in one launch, and one launch per orientation.
        std::vector<Schedule> candidates;
        for (int lanes : {16, 32, 64})
            if (lanes_fit(c.p.disparities, lanes))
                for (int prefetch : {2, 4, 8}) candidates.push_back(Schedule{prefetch, false, lanes, 1});
        for (int waves : {2, 4, 8})
            if (waves == 2 || (c.p.paths == 8 && slices_fit(c.p.disparities, waves / 2) > 1))
                for (int prefetch : {2, 4, 8}) candidates.push_back(Schedule{prefetch, false, 16, waves});
        if (two_lanes_fit(c.p.disparities))
            for (int prefetch : {2, 4, 8}) candidates.push_back(Schedule{prefetch, false, 32, 2});
        // The two waves meeting by recomputation, checkpoints every 4 or 8 steps.
        for (int prefetch : {4, 8}) candidates.push_back(Schedule{prefetch, false, 16, 2, 1});
        if (two_lanes_fit(c.p.disparities)) candidates.push_back(Schedule{8, false, 32, 2, 1});
        candidates.push_back(Schedule{4, true, 16, 1});
        for (const Schedule &s : candidates) {
            HIP_TRY(launch_match_batch(c.p, tables + c.first, c.count, s, stream));
            HIP_TRY(hipEventRecord(start, stream));
            for (int r = 0; r < 3; r++)
                HIP_TRY(launch_match_batch(c.p, tables + c.first, c.count, s, stream));
            HIP_TRY(hipEventRecord(stop, stream));
            HIP_TRY(hipEventSynchronize(stop));
            float ms = 0;
            HIP_TRY(hipEventElapsedTime(&ms, start, stop));
            trials += "  prefetch " + std::to_string(s.prefetch) + ", " +
                      std::to_string(s.lanes) + " lanes, " + std::to_string(s.waves) +
                      (s.waves > 1 ? " waves" : " wave") + (s.meet ? ", recomputed" : "") +
                      (s.split ? ", split" : "") + ": " + std::to_string(ms / 3) + " ms\n";
            if (ms < best) {
                best = ms;
                c.schedule = s;
            }
        }
        c.tuned_ms = best / 3;
        (void)hipEventDestroy(start);
        (void)hipEventDestroy(stop);
        // The pick stays the batch's own: a schedule timed on a whole class of pairs at once
        // is not the one a single pair of that size runs fastest with (at 728x544x128 the
        // batch picks the recomputing form, which a single pair runs 20% slower).
    }
};

extern "C" {

}  // extern "C"

namespace nxs {

// nxs_batch_create in `mode`, or with the plan `pinned` (untuned) where given.
static int create_batch(const nxs_pair *pairs, int count, int mode, const BatchPlan *pinned,
                        nxs_batch **out) {
    const bool autotune = mode == NXS_BATCH_TUNED && !pinned;
    const bool fused = pinned ? pinned->fused != 0 : mode == NXS_BATCH_FUSED;
    auto batch = std::make_unique<nxs_batch>();
    try {
        // Size classes in order of first appearance; each pair's workspace in one allocation.
        std::vector<std::vector<int>> members;
        for (int i = 0; i < count; i++) {
            const nxs_pair &q = pairs[i];
            if (!q.left || !q.right || !q.disparity) throw Error{"a pair has a null pointer"};
            check_matching(q.width, q.height, q.disparities, q.paths, q.p1, q.p2, q.uniqueness);
            int c = 0;
            for (; c < (int)batch->classes.size(); c++) {
                const MatchParams &p = batch->classes[c].p;
                if (p.width == q.width && p.height == q.height && p.disparities == q.disparities &&
                    p.paths == q.paths)
                    break;
            }
            if (c == (int)batch->classes.size()) {
                BatchClass cls;
                cls.p = MatchParams{q.width, q.height, q.disparities, q.paths, q.p1, q.p2, q.uniqueness};
                cls.schedule = schedule_for(q.width, q.height, q.disparities, q.paths);
                batch->classes.push_back(cls);
                members.emplace_back();
            }
            members[c].push_back(i);
        }
        if (fused && batch->classes.size() > (size_t)MAX_MIXED_CLASSES)
            throw Error{"a fused batch takes at most " + std::to_string(MAX_MIXED_CLASSES) +
                        " size classes, got " + std::to_string(batch->classes.size())};
        if (fused) {
            std::vector<MixedMember> all;
            for (size_t c = 0; c < batch->classes.size(); c++)
                all.push_back(MixedMember{batch->classes[c].p, 0, (int)members[c].size()});
            if (!mixed_launch_fits(all.data(), (int)all.size()))
                throw Error{"a fused batch's launches would pass 2^31 threads: too many pairs"};
        }
        size_t total = 0;
        for (size_t c = 0; c < batch->classes.size(); c++) {
            const MatchParams &p = batch->classes[c].p;
            total += up256(match_bytes(p.width, p.height, p.disparities, p.paths)) * members[c].size();
        }
        HIP_TRY(hipMalloc(&batch->workspace, total));
        HIP_TRY(hipMalloc((void **)&batch->tables, sizeof(PairDesc) * count));
        HIP_TRY(hipMalloc((void **)&batch->queues, sizeof(uint32_t) * 2 * MAX_BRANCHES));
        HIP_TRY(hipMemset(batch->queues, 0, sizeof(uint32_t) * 2 * MAX_BRANCHES));
        HIP_TRY(hipStreamCreateWithFlags(&batch->stream, hipStreamNonBlocking));
        for (int b = 0; b < MAX_BRANCHES; b++) {
            HIP_TRY(hipStreamCreateWithFlags(&batch->branch[b], hipStreamNonBlocking));
            HIP_TRY(hipEventCreateWithFlags(&batch->join[b], hipEventDisableTiming));
        }
        HIP_TRY(hipEventCreateWithFlags(&batch->fork, hipEventDisableTiming));
        std::vector<PairDesc> host(count);
        uint8_t *ws = static_cast<uint8_t *>(batch->workspace);
        int next = 0;
        for (size_t c = 0; c < batch->classes.size(); c++) {
            BatchClass &cls = batch->classes[c];
            cls.first = next;
            cls.count = (int)members[c].size();
            const size_t bytes = up256(match_bytes(cls.p.width, cls.p.height, cls.p.disparities, cls.p.paths));
            for (int i : members[c]) {
                const nxs_pair &q = pairs[i];
                PairDesc &d = host[next++];
                d.left = q.left;
                d.right = q.right;
                d.disparity = q.disparity;
                d.sums = nullptr;
                match_carve(d.b, ws, q.width, q.height, q.disparities, q.paths);
                ws += bytes;
                d.p1 = q.p1;
                d.p2 = q.p2;
                d.uniqueness = q.uniqueness;
            }
        }
        HIP_TRY(hipMemcpy(batch->tables, host.data(), sizeof(PairDesc) * count, hipMemcpyHostToDevice));
        if (autotune)
            for (auto &c : batch->classes) {
                batch->trials += std::to_string(c.count) + " x " + std::to_string(c.p.width) + "x" +
                                 std::to_string(c.p.height) + "x" + std::to_string(c.p.disparities) +
                                 ":\n";
                batch->tune(c);
            }
        // Without tuning: the classes in two branches, a class a launch; or fused, checkpoints
        // every 8 steps.
        batch->plan = pinned ? *pinned
                             : fused ? BatchPlan{1, 0, 8, 0}
                                     : BatchPlan{std::min(2, (int)batch->classes.size()), 0, 0, 0};
        if (autotune) batch->tune_plan();
        batch->launches = batch->deal(batch->plan);
        // One graph for the whole batch.
        t_capturing = true;
        hipError_t e = hipStreamBeginCapture(batch->stream, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
            try {
                batch->enqueue_all(batch->stream, batch->plan, batch->launches);
            } catch (const Error &) {
                hipGraph_t broken = nullptr;
                (void)hipStreamEndCapture(batch->stream, &broken);
                if (broken) (void)hipGraphDestroy(broken);
                t_capturing = false;
                throw;
            }
            e = hipStreamEndCapture(batch->stream, &batch->graph);
        }
        t_capturing = false;
        HIP_TRY(e);
        HIP_TRY(hipGraphInstantiate(&batch->exec, batch->graph, nullptr, nullptr, 0));
        std::ostringstream text;
        text << "plan: " << plan_text(batch->plan) << ", "
             << (batch->plan.fused
                     ? std::to_string(4 * std::min(batch->plan.branches,
                                                   (int)batch->classes.size())) + " launches"
                                   : std::to_string(batch->launches.size()) + " launch chains")
             << "\n";
        for (auto &c : batch->classes) {
            text << c.count << " x " << c.p.width << "x" << c.p.height << "x" << c.p.disparities
                 << " paths " << c.p.paths << ": ";
            if (batch->plan.fused) {
                // The fused launch's aggregation, whatever the class's own pick.
                text << "16 lanes per line, 2 waves per line, meeting by recomputation"
                     << (autotune ? " (its own pick took " + std::to_string(c.tuned_ms) + " ms)"
                                  : "")
                     << "\n";
                continue;
            }
            text << "prefetch " << c.schedule.prefetch << ", " << c.schedule.lanes
                 << " lanes per line, " << c.schedule.waves
                 << (c.schedule.waves == 2 ? " waves" : " wave") << " per line, "
                 << (c.schedule.meet ? "meeting by recomputation, " : "")
                 << (c.schedule.split ? "one launch per orientation" : "orientations in one launch")
                 << (autotune ? " (tuned, " + std::to_string(c.tuned_ms) + " ms)" : "") << "\n";
        }
        batch->description = text.str() + batch->trials;
    } catch (const Error &e) {
        return fail(NXS_EDEVICE, e.reason);
    }
    *out = batch.release();
    return NXS_OK;
}

}  // namespace nxs

extern "C" {

int nxs_batch_create(const nxs_pair *pairs, int count, int mode, nxs_batch **out) {
    using namespace nxs;
    if (!pairs || count < 1 || !out) return fail(NXS_EINVAL, "no pairs or output pointer");
    if (mode < NXS_BATCH_DEFAULT || mode > NXS_BATCH_FUSED)
        return fail(NXS_EINVAL, "mode must be NXS_BATCH_DEFAULT, NXS_BATCH_TUNED or NXS_BATCH_FUSED");
    return create_batch(pairs, count, mode, nullptr, out);
}

int nxs_batch_create_planned(const nxs_pair *pairs, int count, const nxs_batch_plan *plan,
                             nxs_batch **out) {
    using namespace nxs;
    if (!pairs || count < 1 || !out || !plan)
        return fail(NXS_EINVAL, "no pairs, plan or output pointer");
    if (plan->branches != 1 && plan->branches != 2 && plan->branches != 4)
        return fail(NXS_EINVAL, "branches must be 1, 2 or 4");
    if (plan->chunk < 0) return fail(NXS_EINVAL, "chunk must be 0 or more pairs");
    if (plan->fused != 0 && plan->fused != 4 && plan->fused != 8)
        return fail(NXS_EINVAL, "fused must be 0, 4 or 8");
    if (plan->queue != 0 && plan->queue != 1) return fail(NXS_EINVAL, "queue must be 0 or 1");
    if (plan->queue && !plan->fused) return fail(NXS_EINVAL, "a work queue takes a fused plan");
    const BatchPlan pinned{plan->branches, plan->fused ? 0 : plan->chunk, plan->fused,
                           plan->queue};
    return create_batch(pairs, count, NXS_BATCH_DEFAULT, &pinned, out);
}

int nxs_schedule_set(int width, int height, int disparities, int paths, const nxs_schedule *s) {
    if (!s) return nxs::fail(NXS_EINVAL, "no schedule");
    if (s->prefetch != 2 && s->prefetch != 4 && s->prefetch != 8)
        return nxs::fail(NXS_EINVAL, "prefetch must be 2, 4 or 8");
    if (!nxs::lanes_fit(disparities, s->lanes))
        return nxs::fail(NXS_EINVAL, std::to_string(s->lanes) + " lanes per line do not fit " +
                                         std::to_string(disparities) + " disparities");
    if (s->waves != 1 && s->waves != 2 && s->waves != 4 && s->waves != 8)
        return nxs::fail(NXS_EINVAL, "waves must be 1, 2, 4 or 8");
    if (s->waves > 2 && s->lanes != 16)
        return nxs::fail(NXS_EINVAL, "four or more waves per line take 16 lanes per line");
    if (s->waves == 2 && s->lanes != 16 && !(s->lanes == 32 && nxs::two_lanes_fit(disparities)))
        return nxs::fail(NXS_EINVAL, "two waves per line take 16 lanes per line, or 32 at 64, "
                                     "128 or 256 disparities");
    if (s->waves > 2 && paths != 8)
        return nxs::fail(NXS_EINVAL, std::to_string(s->waves) + " waves per line take 8 paths");
    if (s->waves > 2 && nxs::slices_fit(disparities, s->waves / 2) == 1)
        return nxs::fail(NXS_EINVAL, std::to_string(s->waves) + " waves per line do not fit " +
                                         std::to_string(disparities) + " disparities");
    if (s->meet != 0 && s->meet != 1) return nxs::fail(NXS_EINVAL, "meet must be 0 or 1");
    if (s->meet != 0 && s->waves != 2)
        return nxs::fail(NXS_EINVAL, "meeting by recomputation takes two waves per line");
    nxs::remember_schedule(width, height, disparities, paths,
                           nxs::Schedule{s->prefetch, s->split != 0, s->lanes, s->waves, s->meet});
    return NXS_OK;
}

int nxs_schedule_get(int width, int height, int disparities, int paths, nxs_schedule *s) {
    if (!s) return nxs::fail(NXS_EINVAL, "no schedule");
    const nxs::Schedule k = nxs::schedule_for(width, height, disparities, paths);
    *s = nxs_schedule{k.prefetch, k.split ? 1 : 0, k.lanes, k.waves, k.meet};
    return NXS_OK;
}

int nxs_batch_run(nxs_batch *batch, void *stream) {
    if (!batch) return nxs::fail(NXS_EINVAL, "no batch");
    hipError_t e = hipGraphLaunch(batch->exec, (hipStream_t)stream);
    if (e != hipSuccess) return nxs::fail(NXS_EDEVICE, std::string("hipGraphLaunch: ") + hipGetErrorString(e));
    return NXS_OK;
}

const char *nxs_batch_describe(nxs_batch *batch) {
    return batch ? batch->description.c_str() : "";
}

void nxs_batch_destroy(nxs_batch *batch) { delete batch; }

const char *nxs_last_error(void) { return nxs::g_last_error.c_str(); }

}  // extern "C"
