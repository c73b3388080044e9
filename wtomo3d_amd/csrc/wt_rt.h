// Runtime calls that must not overlap a stream capture on another thread.
//
// While any stream of the process is being captured into a graph, HIP
// refuses the calls that order against the legacy stream or the whole
// device (synchronous hipMemcpy / hipMemset, hipMalloc / hipFree, device
// synchronisation, launches on the null stream) with
// hipErrorStreamCaptureImplicit, and the capture fails with them.  A host
// that drives several handles from several threads (a serving loop: one
// thread captures a batch's solve while another uploads a new model) hits
// this whenever the two overlap.
//
// The library therefore holds one process-wide reader/writer lock: each
// capture (begin .. instantiate) holds it exclusively, each of the calls
// above holds it shared.  Shared holders run concurrently with one another;
// a capture waits for them and they wait for a capture, which takes
// milliseconds.  The shared side is re-entrant on one thread (a depth count),
// so an entry point that launches on the null stream can hold it around its
// whole body while the wrappers below lock again inside.  A capture must not
// be begun by a thread that holds the shared side (it would wait for
// itself); the capture sites are only reached from entry points that do not.
#pragma once
#include <hip/hip_runtime.h>

#include <shared_mutex>

std::shared_mutex& wt_rt_mutex();

class wt_rt_shared {
public:
    wt_rt_shared()
    {
        if (depth()++ == 0) wt_rt_mutex().lock_shared();
    }
    ~wt_rt_shared()
    {
        if (--depth() == 0) wt_rt_mutex().unlock_shared();
    }
    wt_rt_shared(const wt_rt_shared&) = delete;
    wt_rt_shared& operator=(const wt_rt_shared&) = delete;
    static int& depth()
    {
        static thread_local int d = 0;
        return d;
    }
};

// Held while a stream is captured; exclusive against every wt_rt_shared.
class wt_rt_capture {
public:
    wt_rt_capture() { wt_rt_mutex().lock(); }
    ~wt_rt_capture() { wt_rt_mutex().unlock(); }
    wt_rt_capture(const wt_rt_capture&) = delete;
    wt_rt_capture& operator=(const wt_rt_capture&) = delete;
};

// Shared lock only for work that goes to the null (legacy) stream.
class wt_rt_stream_guard {
public:
    explicit wt_rt_stream_guard(const void* stream) : on_(stream == nullptr)
    {
        if (on_ && wt_rt_shared::depth()++ == 0) wt_rt_mutex().lock_shared();
    }
    ~wt_rt_stream_guard()
    {
        if (on_ && --wt_rt_shared::depth() == 0) wt_rt_mutex().unlock_shared();
    }
    wt_rt_stream_guard(const wt_rt_stream_guard&) = delete;
    wt_rt_stream_guard& operator=(const wt_rt_stream_guard&) = delete;

private:
    bool on_;
};

template <typename T>
inline hipError_t rt_malloc(T** p, size_t bytes)
{
    wt_rt_shared g;
    return hipMalloc((void**)p, bytes);
}
inline hipError_t rt_free(void* p)
{
    wt_rt_shared g;
    return hipFree(p);
}
template <typename T>
inline hipError_t rt_malloc_async(T** p, size_t bytes, hipStream_t s)
{
    wt_rt_stream_guard g(s);
    return hipMallocAsync((void**)p, bytes, s);
}
inline hipError_t rt_free_async(void* p, hipStream_t s)
{
    wt_rt_stream_guard g(s);
    return hipFreeAsync(p, s);
}
inline hipError_t rt_memcpy(void* dst, const void* src, size_t bytes, hipMemcpyKind kind)
{
    wt_rt_shared g;
    return hipMemcpy(dst, src, bytes, kind);
}
inline hipError_t rt_memset(void* dst, int v, size_t bytes)
{
    wt_rt_shared g;
    return hipMemset(dst, v, bytes);
}
inline hipError_t rt_device_sync()
{
    wt_rt_shared g;
    return hipDeviceSynchronize();
}
