// stereo_match_hip: the Port executable of NxStereo.Matcher on the GPU.
//
// It speaks what native/stereo_match/src/main.rs speaks, so NxStereo.Matcher started with
// `path:` pointing here (matcher.ex:22-27, 248-255) and an unchanged NxStereo.NativeSegment
// drive libnxs_hip:
//
// * records in both directions carry a 4-byte big-endian length (Erlang's {:packet, 4};
//   framing.rs:29-50, 105-131);
// * requests are protocol.rs's configure (1), run (2) and release (3) records, handed to
//   nxs_submit as they are; replies are tag 7, 8 and 9 records, written in the order of the
//   requests (main.rs:7-9), and a release has none (main.rs:181-184);
// * log records (tag 2, a level byte, UTF-8 text; framing.rs:134-140) report the device at
//   start and a failure to open it;
// * fd 1 is pointed at stderr and records go to a duplicate of the real stdout, so nothing
//   a library prints can cut a record (framing.rs:52-71);
// * a closed pipe ends the program quietly (main.rs:21-26);
// * `--threads N`, which NxStereo.Matcher passes from its :threads option
//   (matcher.ex:263-266), is checked as main.rs:28-38 checks it and otherwise unused: the
//   work runs on the GPU.
//
// Usage: stereo_match_hip [--threads N] [--device N]
#include <errno.h>
#include <signal.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>

#include <condition_variable>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "nxs_hip.h"

namespace {

constexpr uint8_t TAG_LOG = 2;
constexpr uint8_t LOG_INFO = 0, LOG_ERROR = 2;

// Reads exactly n bytes; false on end of input (or a record cut short by it).
bool read_exactly(int fd, uint8_t *p, size_t n) {
    while (n > 0) {
        const ssize_t got = read(fd, p, n);
        if (got == 0) return false;
        if (got < 0) {
            if (errno == EINTR) continue;
            return false;
        }
        p += got;
        n -= (size_t)got;
    }
    return true;
}

bool write_all(int fd, const uint8_t *p, size_t n) {
    while (n > 0) {
        const ssize_t put = write(fd, p, n);
        if (put < 0) {
            if (errno == EINTR) continue;
            return false;
        }
        p += put;
        n -= (size_t)put;
    }
    return true;
}

std::mutex g_output;  // records are written whole, by the writer or the reader

// One framed record: the big-endian length, then the body.
bool write_record(int fd, const uint8_t *body, size_t len) {
    const uint8_t length[4] = {(uint8_t)(len >> 24), (uint8_t)(len >> 16), (uint8_t)(len >> 8),
                               (uint8_t)len};
    std::lock_guard<std::mutex> lock(g_output);
    return write_all(fd, length, 4) && write_all(fd, body, len);
}

bool write_log(int fd, uint8_t level, const std::string &text) {
    std::vector<uint8_t> record = {TAG_LOG, level};
    record.insert(record.end(), text.begin(), text.end());
    return write_record(fd, record.data(), record.size());
}

// Tickets in request order; 0 marks the end of the input.
struct Order {
    std::mutex mu;
    std::condition_variable cv;
    std::deque<int64_t> tickets;

    void push(int64_t ticket) {
        std::lock_guard<std::mutex> lock(mu);
        tickets.push_back(ticket);
        cv.notify_one();
    }
    int64_t pop() {
        std::unique_lock<std::mutex> lock(mu);
        cv.wait(lock, [&] { return !tickets.empty(); });
        const int64_t t = tickets.front();
        tickets.pop_front();
        return t;
    }
};

}  // namespace

// An integer of at least `lowest`, all of `text`.
bool integer(const char *text, int lowest, int *value) {
    char *end = nullptr;
    const long v = strtol(text, &end, 10);
    if (end == text || *end != '\0' || v < lowest || v > 1 << 30) return false;
    *value = (int)v;
    return true;
}

int main(int argc, char **argv) {
    signal(SIGPIPE, SIG_IGN);  // a closed Port shows as a failed write
    const int out = dup(1);
    if (out < 0 || dup2(2, 1) < 0) return 1;
    int device = 0, threads = 0;
    for (int i = 1; i < argc; i++) {
        const bool valued = i + 1 < argc;
        if (strcmp(argv[i], "--threads") == 0 && valued) {
            if (!integer(argv[++i], 1, &threads)) {
                write_log(out, LOG_ERROR,
                          std::string("--threads takes a positive integer, got ") + argv[i]);
                return 2;
            }
        } else if (strcmp(argv[i], "--device") == 0 && valued) {
            if (!integer(argv[++i], 0, &device)) {
                write_log(out, LOG_ERROR, std::string("--device takes a device index, got ") + argv[i]);
                return 2;
            }
        } else {
            std::string got;
            for (int k = 1; k < argc; k++) got += std::string(k > 1 ? " " : "") + argv[k];
            write_log(out, LOG_ERROR, "usage: stereo_match_hip [--threads N] [--device N], got " + got);
            return 2;
        }
    }

    nxs_ctx *ctx = nullptr;
    char err[512] = {0};
    if (nxs_open(device, &ctx, err, sizeof err) != NXS_OK) {
        write_log(out, LOG_ERROR, std::string("cannot open HIP device ") + std::to_string(device) +
                                      ": " + err);
        return 1;
    }
    if (!write_log(out, LOG_INFO, "computing on HIP device " + std::to_string(device) +
                                      (threads ? " (--threads " + std::to_string(threads) +
                                                     ": host threads are not used)"
                                               : ""))) {
        nxs_close(ctx);
        return 0;
    }

    Order order;
    // The writer: each reply in request order, as soon as it is complete. After a failed
    // write (the Port is gone) it keeps releasing replies until the input ends.
    std::thread writer([&] {
        bool broken = false;
        for (;;) {
            const int64_t ticket = order.pop();
            if (ticket == 0) return;
            if (nxs_wait(ctx, ticket, -1) != NXS_OK) continue;
            const uint8_t *reply = nullptr;
            size_t len = 0;
            if (nxs_reply(ctx, ticket, &reply, &len) == NXS_OK && len > 0 && !broken)
                broken = !write_record(out, reply, len);
            nxs_release_reply(ctx, ticket);
        }
    });

    // The reader (this thread): records are queued as they arrive, computed by the context.
    std::vector<uint8_t> record;
    for (;;) {
        uint8_t length[4];
        if (!read_exactly(0, length, 4)) break;
        const size_t len = (size_t)length[0] << 24 | (size_t)length[1] << 16 |
                           (size_t)length[2] << 8 | (size_t)length[3];
        record.resize(len);
        if (!read_exactly(0, record.data(), len)) break;
        const int64_t ticket = nxs_submit(ctx, record.data(), len, nullptr, nullptr);
        if (ticket <= 0) {
            write_log(out, LOG_ERROR, std::string("request not queued: ") + nxs_last_error());
            continue;
        }
        order.push(ticket);  // a release's ticket completes with an empty reply: nothing written
    }
    order.push(0);
    writer.join();
    nxs_close(ctx);
    close(out);
    return 0;
}
