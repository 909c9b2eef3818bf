// staging.cpp -- host<->device copies staged through pinned memory.
//
// A copy from pageable host memory runs at about half the PCIe rate a
// pinned buffer reaches (MI355X host link: pageable 27-30 GB/s up and
// 17 GB/s down, pinned 57 GB/s both ways).  The large transfers of a run (the
// phantom, the coil maps, the signal) therefore go through two pinned chunk
// buffers per plan: host threads copy chunk k+1 between the caller's memory
// and one buffer while the DMA engine moves chunk k through the other, which
// brings both directions to 48-50 GB/s end to end.
#include "hom_internal.h"

#include <algorithm>
#include <atomic>
#include <cstring>
#include <system_error>
#include <thread>

namespace hom {
namespace {

constexpr size_t kChunk = size_t(8) << 20;      // bytes per pinned buffer
constexpr size_t kStageMin = size_t(4) << 20;   // smaller copies go direct
constexpr int kThreads = 4;                     // host threads per copy

// A rows x width byte array with row stride pitch, addressed by its packed
// byte offsets [pos, end): f(region offset, packed offset, bytes) per
// contiguous piece.
template <typename F>
void for_pieces(size_t pos, size_t end, size_t width, size_t pitch, F &&f) {
    if (pitch == width) {
        if (end > pos) f(pos, pos, end - pos);
        return;
    }
    while (pos < end) {
        const size_t r = pos / width, c = pos % width;
        const size_t n = std::min(width - c, end - pos);
        f(r * pitch + c, pos, n);
        pos += n;
    }
}

bool ensure_buffers(hom_plan *p) {
    Staging &s = p->staging;
    if (s.buf[0]) return true;
    for (int b = 0; b < 2; ++b) {
        if (hipHostMalloc(&s.buf[b], kChunk, hipHostMallocDefault) != hipSuccess ||
            hipEventCreateWithFlags(&s.ev[b], hipEventDisableTiming) != hipSuccess) {
            (void)hipGetLastError();
            staging_release(p);
            return false;
        }
    }
    return true;
}

// Host threads sharing one staged copy.  Chunk k is published by the main
// thread (go = k); every thread copies its share of it (thread i of n takes
// the i-th n-th of the chunk's bytes) and counts itself done.  A worker
// leaves when the main thread sets abort.
class Crew {
public:
    template <typename F>
    Crew(size_t chunks, F work) {
        const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
        const int want = std::min<int>(kThreads, (int)hw) - 1;
        for (int i = 0; i < want; ++i) {
            try {
                pool_.emplace_back([this, chunks, work, i] {
                    for (size_t k = 0; k < chunks; ++k) {
                        long g;
                        while ((g = go_.load(std::memory_order_acquire)) < (long)k && !abort_.load())
                            std::this_thread::yield();
                        if (g < (long)k) return;  // aborted
                        work(k, i + 1, n_);
                        done_.fetch_add(1, std::memory_order_acq_rel);
                    }
                });
            } catch (const std::system_error &) {
                break;  // fewer threads: the shares are fixed below
            }
        }
        n_ = (int)pool_.size() + 1;
    }
    // Runs the main thread's share of chunk k and waits for the others'.
    template <typename F>
    void run(size_t k, F &&work) {
        go_.store((long)k, std::memory_order_release);
        work(k, 0, n_);
        const long want = (long)(n_ - 1) * (long)(k + 1);
        while (done_.load(std::memory_order_acquire) < want) std::this_thread::yield();
    }
    ~Crew() {
        abort_.store(true);
        for (auto &t : pool_) t.join();
    }

private:
    std::vector<std::thread> pool_;
    std::atomic<long> go_{-1}, done_{0};
    std::atomic<bool> abort_{false};
    int n_ = 1;
};

}  // namespace

void staging_release(hom_plan *p) {
    Staging &s = p->staging;
    for (int b = 0; b < 2; ++b) {
        if (s.ev[b]) (void)hipEventDestroy(s.ev[b]);
        if (s.buf[b]) (void)hipHostFree(s.buf[b]);
        s.ev[b] = nullptr;
        s.buf[b] = nullptr;
    }
}

int staged_upload(hom_plan *p, void *dev, size_t dev_pitch, const void *host, size_t host_pitch, size_t width,
                  size_t rows) {
    const size_t total = width * rows;
    if (total == 0) return HOM_OK;
    if (total < kStageMin || !ensure_buffers(p)) {
        int rc = check_hip(p, hipMemcpy2DAsync(dev, dev_pitch, host, host_pitch, width, rows, hipMemcpyHostToDevice,
                                               p->stream), "hipMemcpy2DAsync");
        return rc ? rc : check_hip(p, hipStreamSynchronize(p->stream), "hipStreamSynchronize");
    }
    Staging &s = p->staging;
    const size_t nk = (total + kChunk - 1) / kChunk;
    const char *src = static_cast<const char *>(host);
    char *dst = static_cast<char *>(dev);
    // thread i of n packs its share of chunk k into buffer k % 2
    auto pack = [&s, src, total, width, host_pitch](size_t k, int i, int n) {
        const size_t a = k * kChunk, len = std::min(kChunk, total - a);
        const size_t lo = a + len * (size_t)i / (size_t)n, hi = a + len * (size_t)(i + 1) / (size_t)n;
        char *buf = static_cast<char *>(s.buf[k & 1]);
        for_pieces(lo, hi, width, host_pitch,
                   [&](size_t off, size_t pos, size_t bytes) { std::memcpy(buf + (pos - a), src + off, bytes); });
    };
    int rc = HOM_OK;
    {
        Crew crew(nk, pack);
        for (size_t k = 0; k < nk && !rc; ++k) {
            // buffer k % 2 is free once chunk k - 2 has left it
            if (k >= 2) rc = check_hip(p, hipEventSynchronize(s.ev[k & 1]), "hipEventSynchronize");
            if (rc) break;
            crew.run(k, pack);
            const size_t a = k * kChunk, len = std::min(kChunk, total - a);
            const char *buf = static_cast<const char *>(s.buf[k & 1]);
            for_pieces(a, a + len, width, dev_pitch, [&](size_t off, size_t pos, size_t bytes) {
                if (!rc)
                    rc = check_hip(p, hipMemcpyAsync(dst + off, buf + (pos - a), bytes, hipMemcpyHostToDevice,
                                                     p->stream), "hipMemcpyAsync");
            });
            if (!rc) rc = check_hip(p, hipEventRecord(s.ev[k & 1], p->stream), "hipEventRecord");
        }
    }
    const int rs = check_hip(p, hipStreamSynchronize(p->stream), "hipStreamSynchronize");
    return rc ? rc : rs;
}

int staged_download(hom_plan *p, void *host, size_t host_pitch, const void *dev, size_t dev_pitch, size_t width,
                    size_t rows) {
    const size_t total = width * rows;
    if (total == 0) return HOM_OK;
    if (total < kStageMin || !ensure_buffers(p)) {
        int rc = check_hip(p, hipMemcpy2DAsync(host, host_pitch, dev, dev_pitch, width, rows, hipMemcpyDeviceToHost,
                                               p->stream), "hipMemcpy2DAsync");
        return rc ? rc : check_hip(p, hipStreamSynchronize(p->stream), "hipStreamSynchronize");
    }
    Staging &s = p->staging;
    const size_t nk = (total + kChunk - 1) / kChunk;
    char *dst = static_cast<char *>(host);
    const char *src = static_cast<const char *>(dev);
    // thread i of n copies its share of chunk k out of buffer k % 2
    auto unpack = [&s, dst, total, width, host_pitch](size_t k, int i, int n) {
        const size_t a = k * kChunk, len = std::min(kChunk, total - a);
        const size_t lo = a + len * (size_t)i / (size_t)n, hi = a + len * (size_t)(i + 1) / (size_t)n;
        const char *buf = static_cast<const char *>(s.buf[k & 1]);
        for_pieces(lo, hi, width, host_pitch,
                   [&](size_t off, size_t pos, size_t bytes) { std::memcpy(dst + off, buf + (pos - a), bytes); });
    };
    int rc = HOM_OK;
    // device -> buffer k % 2 for chunk k, then an event the host waits on
    auto issue = [&](size_t k) {
        const size_t a = k * kChunk, len = std::min(kChunk, total - a);
        char *buf = static_cast<char *>(s.buf[k & 1]);
        for_pieces(a, a + len, width, dev_pitch, [&](size_t off, size_t pos, size_t bytes) {
            if (!rc)
                rc = check_hip(p, hipMemcpyAsync(buf + (pos - a), src + off, bytes, hipMemcpyDeviceToHost, p->stream),
                               "hipMemcpyAsync");
        });
        if (!rc) rc = check_hip(p, hipEventRecord(s.ev[k & 1], p->stream), "hipEventRecord");
    };
    {
        Crew crew(nk, unpack);
        issue(0);
        if (nk > 1) issue(1);
        for (size_t k = 0; k < nk && !rc; ++k) {
            rc = check_hip(p, hipEventSynchronize(s.ev[k & 1]), "hipEventSynchronize");
            if (rc) break;
            crew.run(k, unpack);
            if (k + 2 < nk) issue(k + 2);
        }
    }
    const int rs = check_hip(p, hipStreamSynchronize(p->stream), "hipStreamSynchronize");
    return rc ? rc : rs;
}

}  // namespace hom
