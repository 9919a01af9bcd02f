// Host staging for the host-pointer entry points: a small process-wide pool of
// copy threads, device buffers that keep the largest size asked of them, and
// copies between pageable caller memory and the device through pinned chunks
// (the CPU copy of one chunk overlaps the DMA of the one before it, and a large
// copy is split over the pool's threads).
#ifndef LDSF_SF_STAGE_H_
#define LDSF_SF_STAGE_H_

#include <hip/hip_runtime.h>

#include <cstddef>
#include <functional>
#include <vector>

namespace ldsf {

// threads that parallel_for runs on, the calling thread included
int staging_threads();

// f(0), ..., f(n - 1) on the pool's threads and the calling thread; returns
// when every call has returned.  Calls from several threads run one at a time.
void parallel_for(int n, const std::function<void(int)> &f);

// a device array that grows to the largest size reserved and keeps it
template <typename T>
struct Grow {
    T *p = nullptr;
    size_t cap = 0;
    Grow() = default;
    Grow(const Grow &) = delete;
    Grow &operator=(const Grow &) = delete;
    ~Grow() { release(); }
    hipError_t reserve(size_t n)
    {
        if (n <= cap && p) return hipSuccess;
        release();
        const size_t m = n > 0 ? n : 1;
        hipError_t e = hipMalloc((void **)&p, sizeof(T) * m);
        if (e != hipSuccess) {
            p = nullptr;
            return e;
        }
        cap = m;
        return hipSuccess;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
};

// Copies between pageable host memory and the device on one stream, through
// two pinned chunks per thread.  Not thread-safe: one caller at a time.
class Stager {
public:
    explicit Stager(hipStream_t s) : s_(s) {}
    ~Stager();
    Stager(const Stager &) = delete;
    Stager &operator=(const Stager &) = delete;

    // enqueues the copy of src[0, bytes) to dst on the stream; returns once
    // src may be reused (its last chunks may still be in flight from pinned
    // memory, which the next call waits for before refilling)
    hipError_t h2d(void *dst, const void *src, size_t bytes);
    // the same for ints, with the smallest and largest value copied (n > 0)
    hipError_t h2d_minmax(int *dst, const int *src, size_t n, int *mn, int *mx);
    // copies src[0, bytes) on the device, after everything enqueued on the
    // stream so far, into dst; returns once dst holds it
    hipError_t d2h(void *dst, const void *src, size_t bytes);

private:
    struct Slot {
        char *p = nullptr;
        hipEvent_t ev = nullptr;
        bool pending = false;  // a DMA from or to p was enqueued after ev's last wait
    };
    hipError_t ensure(int workers);
    int workers_for(size_t bytes) const;
    template <class Copy>
    hipError_t h2d_impl(char *dst, const char *src, size_t bytes, Copy &&copy);
    hipStream_t s_;
    std::vector<Slot> slots_;  // 2 per worker
};

}  // namespace ldsf

#endif  // LDSF_SF_STAGE_H_
