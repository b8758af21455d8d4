// Host helpers shared by the global solve (solver.cpp) and the batched
// blocks (batch.cpp): the HIP status check, owned device buffers, a
// monotonic clock and the quaternion re-orthonormalisation of a camera.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

#include <sched.h>

#include "../../include/rpba.h"

#define HIPCHK(x)                                                                          \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess) {                                                            \
            std::fprintf(stderr, "rpba: HIP error %s at %s:%d\n", hipGetErrorString(e_), __FILE__, __LINE__); \
            return RPBA_E_HIP;                                                             \
        }                                                                                  \
    } while (0)

namespace rpba {

template <class T>
struct DBuf {
    T *p = nullptr;
    size_t n = 0;
    int alloc(size_t cnt)
    {
        if (cnt <= n && p) return RPBA_OK;
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        if (hipMalloc((void **)&p, sizeof(T) * std::max<size_t>(cnt, 1)) != hipSuccess) return RPBA_E_NOMEM;
        n = cnt;
        return RPBA_OK;
    }
    int put(const T *h, size_t cnt, hipStream_t st)
    {
        int rc = alloc(cnt);
        if (rc) return rc;
        if (cnt) HIPCHK(hipMemcpyAsync(p, h, sizeof(T) * cnt, hipMemcpyHostToDevice, st));
        return RPBA_OK;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

struct Timed {
    std::string name;
    hipEvent_t a, b;
    int pcg_j;   // launch index inside a PCG solve, or -1
    bool skip;   // a PCG launch after the stopping rule fired (passes the state on only)
};

// host threads for the bookkeeping loops: the processors of the calling
// process (its CPU affinity), at most 16
inline int host_threads()
{
    cpu_set_t cs;
    CPU_ZERO(&cs);
    int n = 0;
    if (sched_getaffinity(0, sizeof cs, &cs) == 0) n = CPU_COUNT(&cs);
    if (n <= 0) n = (int)std::thread::hardware_concurrency();
    return std::max(1, std::min(n, 16));
}

// f(begin, end, r) over nr contiguous ranges r of [0, n) (range r is the r-th
// of nr equal cuts), on nr threads; the ranges depend on n and nr only, so a
// loop whose ranges write disjoint outputs gives the same result as one thread
template <class F>
inline void parallel_ranges(long long n, int nr, F f)
{
    if (nr <= 1 || n < 4096) {
        for (int r = 0; r < nr; ++r) f(n * r / nr, n * (r + 1) / nr, r);
        return;
    }
    std::vector<std::thread> th;
    th.reserve((size_t)nr - 1);
    for (int r = 1; r < nr; ++r) th.emplace_back([=, &f] { f(n * r / nr, n * (r + 1) / nr, r); });
    f(0, n / nr, 0);
    for (std::thread &t : th) t.join();
}

// dst = src for large vectors of trivially copyable elements, the bytes
// copied in nr ranges
template <class D, class S>
inline void parallel_assign(D &dst, const S &src, int nr)
{
    using T = typename S::value_type;
    static_assert(std::is_same<T, typename D::value_type>::value, "same element type");
    dst.resize(src.size());
    const long long n = (long long)src.size();
    parallel_ranges(n, nr, [&](long long i0, long long i1, int) {
        if (i1 > i0) std::memcpy(dst.data() + i0, src.data() + i0, sizeof(T) * (size_t)(i1 - i0));
    });
}

inline double cpu_seconds()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// quaternion re-orthonormalisation (RectifyRotationMatrix4, rpba.h:923-960) of
// the 3x3 part of a row-major 3x4 matrix
inline void rectify34(double *P)
{
#define E(r, c) P[(r)*4 + (c)]
    double a0 = E(2, 1) - E(1, 2), a1 = E(0, 2) - E(2, 0), a2 = E(1, 0) - E(0, 1);
    if (std::fabs(a0) < 1.e-20) a0 = 1.e-20;
    if (std::fabs(a1) < 1.e-20) a1 = 1.e-20;
    if (std::fabs(a2) < 1.e-20) a2 = 1.e-20;
    const double al = std::sqrt(a0 * a0 + a1 * a1 + a2 * a2);
    a0 /= al;
    a1 /= al;
    a2 /= al;
    const double phi = std::atan2(al, E(0, 0) + E(1, 1) + E(2, 2) - 1.);
    const double q0 = std::cos(phi * 0.5), sh = std::sin(phi * 0.5);
    const double q1 = a0 * sh, q2 = a1 * sh, q3 = a2 * sh;
    E(0, 0) = q0 * q0 + q1 * q1 - q2 * q2 - q3 * q3;
    E(0, 1) = 2 * (q1 * q2 - q0 * q3);
    E(0, 2) = 2 * (q1 * q3 + q0 * q2);
    E(1, 0) = 2 * (q2 * q1 + q0 * q3);
    E(1, 1) = q0 * q0 - q1 * q1 + q2 * q2 - q3 * q3;
    E(1, 2) = 2 * (q2 * q3 - q0 * q1);
    E(2, 0) = 2 * (q3 * q1 - q0 * q2);
    E(2, 1) = 2 * (q3 * q2 + q0 * q1);
    E(2, 2) = q0 * q0 - q1 * q1 - q2 * q2 + q3 * q3;
#undef E
}

}  // namespace rpba
