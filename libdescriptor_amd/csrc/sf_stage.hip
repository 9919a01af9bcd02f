// Host staging: the copy-thread pool and the pinned-chunk copies of
// sf_stage.h.
#include "sf_stage.h"

#include <algorithm>
#include <atomic>
#include <climits>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <thread>

namespace ldsf {

namespace {

constexpr size_t kChunk = size_t(4) << 20;      // bytes per DMA and per pinned slot
constexpr size_t kPerWorker = size_t(8) << 20;  // a copy is split into parts of at least this

// f(0..n-1) on `workers` detached threads and the caller.  The pool lives
// until the process ends; a process forked from this one has no workers and
// runs every call on the caller.
class Pool {
public:
    explicit Pool(int workers) : workers_(workers)
    {
        for (int i = 0; i < workers; ++i) std::thread([this] { loop(); }).detach();
    }
    int threads() const { return workers_ + 1; }
    void run(int n, const std::function<void(int)> &f)
    {
        std::lock_guard<std::mutex> serial(run_mu_);
        std::unique_lock<std::mutex> lk(mu_);
        job_ = &f;
        next_ = 0;
        n_ = n;
        ++gen_;
        cv_work_.notify_all();
        work(lk);
        cv_done_.wait(lk, [&] { return next_ >= n_ && active_ == 0; });
        job_ = nullptr;
    }

private:
    void work(std::unique_lock<std::mutex> &lk)
    {
        while (next_ < n_) {
            const int t = next_++;
            const std::function<void(int)> *f = job_;
            ++active_;
            lk.unlock();
            (*f)(t);
            lk.lock();
            if (--active_ == 0 && next_ >= n_) cv_done_.notify_all();
        }
    }
    void loop()
    {
        std::unique_lock<std::mutex> lk(mu_);
        long seen = 0;
        for (;;) {
            cv_work_.wait(lk, [&] { return gen_ != seen; });
            seen = gen_;
            work(lk);
        }
    }
    const int workers_;
    std::mutex run_mu_, mu_;
    std::condition_variable cv_work_, cv_done_;
    const std::function<void(int)> *job_ = nullptr;
    int next_ = 0, n_ = 0, active_ = 0;
    long gen_ = 0;
};

// LDSF_STAGING_THREADS threads in all, else up to 8
Pool &pool()
{
    static Pool *p = [] {
        int t = 0;
        if (const char *v = std::getenv("LDSF_STAGING_THREADS")) t = std::atoi(v);
        if (t <= 0) t = (int)std::min(8u, std::max(1u, std::thread::hardware_concurrency()));
        return new Pool(std::min(t, 64) - 1);
    }();
    return *p;
}

void set_error(std::atomic<int> &err, hipError_t e)
{
    int ok = hipSuccess;
    if (e != hipSuccess) err.compare_exchange_strong(ok, (int)e);
}

}  // namespace

int staging_threads() { return pool().threads(); }

void parallel_for(int n, const std::function<void(int)> &f)
{
    if (n <= 1) {
        if (n == 1) f(0);
        return;
    }
    pool().run(n, f);
}

Stager::~Stager()
{
    for (Slot &sl : slots_) {
        if (sl.pending) (void)hipEventSynchronize(sl.ev);
        if (sl.ev) (void)hipEventDestroy(sl.ev);
        if (sl.p) (void)hipHostFree(sl.p);
    }
}

int Stager::workers_for(size_t bytes) const
{
    const size_t w = std::min<size_t>(bytes / kPerWorker, (size_t)staging_threads());
    return (int)std::max<size_t>(w, 1);
}

hipError_t Stager::ensure(int workers)
{
    while ((int)slots_.size() < 2 * workers) {
        Slot sl;
        hipError_t e = hipHostMalloc((void **)&sl.p, kChunk, hipHostMallocDefault);
        if (e != hipSuccess) return e;
        e = hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming);
        if (e != hipSuccess) {
            (void)hipHostFree(sl.p);
            return e;
        }
        slots_.push_back(sl);
    }
    return hipSuccess;
}

// Worker w copies its part [w * part, (w + 1) * part) chunk by chunk: wait
// for the slot's previous DMA, fill it (copy), enqueue its DMA.  Parts are
// multiples of 64 bytes, so an int array splits on element boundaries.
template <class Copy>
hipError_t Stager::h2d_impl(char *dst, const char *src, size_t bytes, Copy &&copy)
{
    if (bytes == 0) return hipSuccess;
    const int P = workers_for(bytes);
    if (hipError_t e = ensure(P)) return e;
    const size_t part = ((bytes + P - 1) / P + 63) & ~size_t(63);
    std::atomic<int> err{hipSuccess};
    parallel_for(P, [&](int w) {
        const size_t lo = std::min(bytes, (size_t)w * part), hi = std::min(bytes, lo + part);
        int k = 0;
        for (size_t off = lo; off < hi && err.load() == hipSuccess; off += kChunk, ++k) {
            const size_t len = std::min(kChunk, hi - off);
            Slot &sl = slots_[2 * w + (k & 1)];
            hipError_t e = hipSuccess;
            if (sl.pending) {
                e = hipEventSynchronize(sl.ev);
                sl.pending = false;
            }
            if (e == hipSuccess) {
                copy(sl.p, src + off, len, w);
                e = hipMemcpyAsync(dst + off, sl.p, len, hipMemcpyHostToDevice, s_);
                // the slot is in use from here on, whatever the record returns
                sl.pending = e == hipSuccess;
                if (e == hipSuccess) e = hipEventRecord(sl.ev, s_);
            }
            set_error(err, e);
        }
    });
    if (err.load() != hipSuccess) {
        // leave no DMA in flight from a slot whose event was not recorded
        (void)hipStreamSynchronize(s_);
        for (Slot &sl : slots_) sl.pending = false;
    }
    return (hipError_t)err.load();
}

hipError_t Stager::h2d(void *dst, const void *src, size_t bytes)
{
    return h2d_impl((char *)dst, (const char *)src, bytes,
                    [](char *d, const char *s, size_t len, int) { std::memcpy(d, s, len); });
}

hipError_t Stager::h2d_minmax(int *dst, const int *src, size_t n, int *mn, int *mx)
{
    const int P = workers_for(n * sizeof(int));
    std::vector<int> lo(P, INT_MAX), hi(P, INT_MIN);
    hipError_t e = h2d_impl((char *)dst, (const char *)src, n * sizeof(int),
                            [&](char *d, const char *s, size_t len, int w) {
                                const int *a = (const int *)s;
                                int *b = (int *)d;
                                int l = lo[w], h = hi[w];
                                for (size_t i = 0, m = len / sizeof(int); i < m; ++i) {
                                    const int v = a[i];
                                    b[i] = v;
                                    l = v < l ? v : l;
                                    h = v > h ? v : h;
                                }
                                lo[w] = l;
                                hi[w] = h;
                            });
    *mn = *std::min_element(lo.begin(), lo.end());
    *mx = *std::max_element(hi.begin(), hi.end());
    return e;
}

// Worker w: enqueue the DMA of chunk k into one slot, then copy chunk k - 1
// out of the other once its DMA has landed.
hipError_t Stager::d2h(void *dst_, const void *src_, size_t bytes)
{
    if (bytes == 0) return hipSuccess;
    char *dst = (char *)dst_;
    const char *src = (const char *)src_;
    const int P = workers_for(bytes);
    if (hipError_t e = ensure(P)) return e;
    const size_t part = ((bytes + P - 1) / P + 63) & ~size_t(63);
    std::atomic<int> err{hipSuccess};
    parallel_for(P, [&](int w) {
        const size_t lo = std::min(bytes, (size_t)w * part), hi = std::min(bytes, lo + part);
        const size_t nk = (hi - lo + kChunk - 1) / kChunk;
        for (size_t k = 0; k <= nk; ++k) {
            if (k < nk && err.load() == hipSuccess) {
                const size_t off = lo + k * kChunk, len = std::min(kChunk, hi - off);
                Slot &sl = slots_[2 * w + (k & 1)];
                hipError_t e = hipSuccess;
                if (sl.pending) {
                    e = hipEventSynchronize(sl.ev);
                    sl.pending = false;
                }
                if (e == hipSuccess) e = hipMemcpyAsync(sl.p, src + off, len, hipMemcpyDeviceToHost, s_);
                sl.pending = e == hipSuccess;
                if (e == hipSuccess) e = hipEventRecord(sl.ev, s_);
                set_error(err, e);
            }
            if (k > 0) {
                const size_t off = lo + (k - 1) * kChunk, len = std::min(kChunk, hi - off);
                Slot &sl = slots_[2 * w + ((k - 1) & 1)];
                hipError_t e = hipSuccess;
                if (sl.pending) {
                    e = hipEventSynchronize(sl.ev);
                    sl.pending = false;
                }
                set_error(err, e);
                if (err.load() == hipSuccess) std::memcpy(dst + off, sl.p, len);
            }
        }
    });
    if (err.load() != hipSuccess) {
        (void)hipStreamSynchronize(s_);
        for (Slot &sl : slots_) sl.pending = false;
    }
    return (hipError_t)err.load();
}

}  // namespace ldsf
