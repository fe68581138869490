// cctamd_bs_field_noalloc: the field evaluation of cctamd_bs_field without a
// caller-supplied partial-sum workspace, for callers that cannot allocate
// device memory (code under HIP graph capture, tight serving loops, harnesses
// that forbid hipMalloc).  The split-K partial sums live in `scratch`, a
// file-scope device array of kScratchDoubles doubles that every GPU the library
// runs on holds until the process ends.  Nothing here allocates or frees.
//
// Host code never passes `scratch` itself to a kernel (its host-side name is a
// stand-in, not a device address): on the first call on a device, with that
// device current, hipGetSymbolAddress gives that device's copy, which is kept
// per device and used from then on.
//
// All calls on one GPU share its array.  The Biot-Savart pass writes every
// partial the reduction reads in the same call, so nothing needs clearing;
// two calls that overlap on one GPU (different streams or threads) race.

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>

#include "common.hpp"

namespace {

// 16 MiB: up to 10,922 work items of 64-point blocks (every shape of the
// 1,000-point single-CCT field or a 512-proton AG-CCT section field).
constexpr int64_t kScratchDoubles = CCTAMD_NOALLOC_PARTIAL_DOUBLES;
static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
double* g_scratch_ptr[kMaxDevices] = {};

// Device address of the current device's `scratch` (looked up once per device).
int scratch_for_current_device(double** out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return (int)e;
    if (dev < 0 || dev >= kMaxDevices) return CCTAMD_ERANGE;
    std::lock_guard<std::mutex> lock(g_mutex);
    if (g_scratch_ptr[dev] == nullptr) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return (int)e;
        if (p == nullptr) return (int)hipErrorInvalidSymbol;
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    *out = g_scratch_ptr[dev];
    return CCTAMD_OK;
}

}  // namespace

using namespace cctamd_detail;

extern "C" int cctamd_bs_field_noalloc(int precision, int ppl, const void* seg, const double* tile_org,
                                       const double* seg_near, const int32_t* work, int64_t nwork,
                                       const int64_t* design_wbase, const int32_t* design_nchunk,
                                       const int64_t* design_pt_off, int64_t ndesign, const double* qs,
                                       const int64_t* design_qs_off, const double* bconst, const double* points,
                                       int64_t npoints, double* B, void* stream) {
    if (!valid_ppl(ppl) || nwork < 0) return CCTAMD_EINVAL;
    if (nwork > kScratchDoubles / (3 * (int64_t)pb_of(ppl))) return CCTAMD_ERANGE;  // needs nwork * 3 * PB doubles
    double* partial = nullptr;
    const int rc = scratch_for_current_device(&partial);
    if (rc != CCTAMD_OK) return rc;
    const DeviceSet ds{precision, ppl, seg, tile_org, seg_near, (const Work*)work, nwork, design_wbase,
                       design_nchunk, design_pt_off, ndesign, qs, design_qs_off, bconst};
    return field_checked(ds, points, npoints, partial, B, (hipStream_t)stream);
}
