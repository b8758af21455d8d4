// Exteval from device-resident tracks with no device allocation
// (rpba_exteval_device).  The partial sums of the Exteval kernels live in a
// file-scope device array, `scratch`, of fixed size; the library finds its
// address on each device once, with hipGetSymbolAddress, and passes that
// address to the kernels.  Nothing here calls hipMalloc or hipFree.
//
// Every call on one device shares that device's array: calls that overlap on
// one device (two streams, two host threads) race on it.  Each launch writes
// every partial slot it later reads, so nothing has to be cleared between
// calls.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/rpba.h"
#include "kernels.h"

namespace {

// Partial slots: one per 256-point workgroup of the Exteval kernel, so up to
// kMaxGroups * 256 = 8,388,608 points.  Layout: kMaxGroups doubles (point
// term sums), then 2 * kMaxGroups ints (track-length sums, point counts).
constexpr int kMaxGroups = 32768;
constexpr size_t kScratchBytes = kMaxGroups * (sizeof(double) + 2 * sizeof(int));

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void *g_ptr[kMaxDevices];  // device address of `scratch` per device, set once

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Device address of `scratch` on the current device (looked up on the first
// call on that device); null on failure.
static void *scratch_ptr(int *status)
{
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
        *status = RPBA_E_HIP;
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[dev]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
            *status = RPBA_E_HIP;
            return nullptr;
        }
        g_ptr[dev] = p;
    }
    *status = RPBA_OK;
    return g_ptr[dev];
}

extern "C" int rpba_exteval_device_max_points(void) { return kMaxGroups * 256; }

extern "C" int rpba_exteval_device(int32_t n_cams, int32_t n_pts, const int32_t *toff, const int32_t *tcam,
                                   const float *txy, const float *tw, const float *cam_P, const float *cam_f,
                                   const float *cam_d, const float *pts, const int32_t *img_w,
                                   const int32_t *img_h, int32_t est_k2, int32_t est_k4, int32_t est_f,
                                   double *out, void *stream)
{
    if (n_cams < 1 || n_pts < 0 || !toff || !tcam || !txy || !cam_P || !cam_f || !cam_d || !pts || !out)
        return RPBA_E_INVALID;
    if (rpba::exteval_groups(n_pts) > kMaxGroups) return RPBA_E_SCRATCH;
    int rc = RPBA_OK;
    unsigned char *base = static_cast<unsigned char *>(scratch_ptr(&rc));
    if (!base) return rc;
    rpba::ExtArgs a;
    a.npts = n_pts;
    a.ncams = n_cams;
    a.nraddp = (est_k2 != 0) + (est_k4 != 0) + (est_f != 0);
    a.toff = toff;
    a.cam = tcam;
    a.xy = txy;
    a.w = tw;
    a.P = cam_P;
    a.cf = cam_f;
    a.cd = cam_d;
    a.pts = pts;
    a.img_w = img_w;
    a.img_h = img_h;
    a.partial = reinterpret_cast<double *>(base);
    a.tpartial = reinterpret_cast<int *>(base + kMaxGroups * sizeof(double));
    a.cpartial = a.tpartial + kMaxGroups;
    rpba::launch_exteval(a, out, static_cast<hipStream_t>(stream));
    return hipGetLastError() == hipSuccess ? RPBA_OK : RPBA_E_HIP;
}
