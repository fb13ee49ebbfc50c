// ksw2_noalloc.hip -- ksw_extd2_hip_noalloc: one alignment whose device
// buffers live in a fixed scratch array of the library instead of memory
// allocated per call.
//
// `scratch` is a file-scope device array.  Host code never uses its name as
// an address (in host code the name is a host-side stand-in): the set-up
// asks the runtime for the array's device address once per device, with that
// device current, and every later use (the plan's arenas, the fill kernel)
// goes through that address.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <mutex>

#include "../../include/ksw2_hip.h"
#include "ksw2_common.h"
#include "ksw2_kernels.h"
#include "ksw2_plan.h"

namespace kswhip {

namespace {

constexpr uint64_t kScratchBytes = KSW2_HIP_NOALLOC_SCRATCH_BYTES;
static_assert(kScratchBytes % 256 == 0, "scratch size");

static __device__ __align__(256) uint8_t scratch[kScratchBytes];

__global__ void __launch_bounds__(256) ksw_fill_zero_kernel(uint4 *p, uint64_t n16)
{
    const uint4 z = {0u, 0u, 0u, 0u};
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = z;
}

// Per device: the scratch array's device address and the stream and events
// its calls use, created once at set-up.
constexpr int kMaxDevices = 64;
struct NoAllocDevice {
    std::mutex mu;              // calls share the scratch: one at a time
    bool ready = false;
    uint8_t *ptr = nullptr;     // device address of `scratch` on this device
    hipStream_t stream = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
};
NoAllocDevice g_na[kMaxDevices];
int32_t g_host_stop = 0;        // never raised: these calls poll no progress callback

// The whole scratch as one region: one plan at a time (calls hold the
// device's lock).
struct ScratchPool : DevicePool {
    bool busy = false;
    bool alloc(uint64_t n, uint64_t *off) override
    {
        if (busy || n > bytes) return false;
        busy = true, *off = 0;
        return true;
    }
    void free(uint64_t) override { busy = false; }
};

// Set-up of one device; the caller holds D.mu and the device is current.
int prepare_locked(NoAllocDevice &D)
{
    if (D.ready) return KSW2_HIP_OK;
    void *ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return KSW2_HIP_EDEVICE;
    if (!D.stream && hipStreamCreateWithFlags(&D.stream, hipStreamNonBlocking) != hipSuccess) {
        D.stream = nullptr;
        return KSW2_HIP_EDEVICE;
    }
    for (hipEvent_t &e : D.ev)
        if (!e && hipEventCreate(&e) != hipSuccess) {
            e = nullptr;
            return KSW2_HIP_EDEVICE;
        }
    // waiting launches (band kernel, wide teams) record process-wide events
    if (!wait_slots_prepare()) return KSW2_HIP_EDEVICE;
    D.ptr = (uint8_t *)ptr;
    D.ready = true;
    return KSW2_HIP_OK;
}

}  // namespace

hipError_t launch_fill_zero(void *p, uint64_t bytes, hipStream_t st)
{
    if (bytes == 0) return hipSuccess;
    if (((uintptr_t)p & 15) || (bytes & 15)) return hipErrorInvalidValue;
    const uint64_t n16 = bytes / 16;
    const uint64_t blocks = std::min<uint64_t>((n16 + 255) / 256, 2048);
    hipLaunchKernelGGL(ksw_fill_zero_kernel, dim3((unsigned)blocks), dim3(256), 0, st, (uint4 *)p, n16);
    return hipGetLastError();
}

bool noalloc_fits(const ksw2_pair_t &pair, int8_t m, const ScoreClass &cls)
{
    // the plan's block: the arenas plus the job tables (a few KB)
    return pair_footprint(pair, takes_general(pair, m, cls)) + (64u << 10) <= kScratchBytes;
}

int noalloc_prepare()
{
    const int dev = library_device();
    if (dev < 0 || dev >= kMaxDevices) return KSW2_HIP_ENODEV;
    if (hipSetDevice(dev) != hipSuccess) return KSW2_HIP_ENODEV;
    NoAllocDevice &D = g_na[dev];
    std::lock_guard<std::mutex> lk(D.mu);
    return prepare_locked(D);
}

// One pair (already checked by the C entry point: lengths positive, scoring
// normalised, CIGAR buffer large enough) through a plan whose memory is the
// scratch array, stream and events the device's own, sequences copied
// straight from the caller's buffers.  Result into *ez; the CIGAR into
// ez->cigar (capacity ez->m_cigar words).
int noalloc_run(const ksw2_pair_t &pair, int8_t m, const ScoreClass &cls, const GenScore &gsc, ksw_extz_t *ez)
{
    const int dev = library_device();
    if (dev < 0 || dev >= kMaxDevices) return KSW2_HIP_ENODEV;
    if (hipSetDevice(dev) != hipSuccess) return KSW2_HIP_ENODEV;
    NoAllocDevice &D = g_na[dev];
    std::lock_guard<std::mutex> lk(D.mu);
    if (!noalloc_fits(pair, m, cls)) return KSW2_HIP_ETOOBIG;
    int st = prepare_locked(D);
    if (st != KSW2_HIP_OK) return st;
    ScratchPool pool;
    pool.base = D.ptr, pool.bytes = kScratchBytes;
    PlanEnv env;
    env.pool = &pool;
    env.stream = D.stream;
    env.events = D.ev;
    env.direct_upload = true;
    env.h_stop = &g_host_stop, env.d_stop = nullptr;  // no stop point polled
    Plan *p = nullptr;
    const int idx = 0;
    st = plan_create(&p, &pair, &idx, 1, m, cls, gsc, &env);
    if (st == KSW2_HIP_ENOMEM) st = KSW2_HIP_ETOOBIG;  // the block did not fit the scratch
    if (st == KSW2_HIP_OK) st = plan_launch(p, nullptr);
    if (st == KSW2_HIP_OK) st = plan_wait(p, nullptr);
    KswOut o;
    memset(&o, 0, sizeof(o));
    if (st == KSW2_HIP_OK) {
        if (hipMemcpyAsync(&o, p->d_out, sizeof(KswOut), hipMemcpyDeviceToHost, D.stream) != hipSuccess ||
            hipStreamSynchronize(D.stream) != hipSuccess)
            st = KSW2_HIP_EDEVICE;
    }
    if (st == KSW2_HIP_OK && o.n_cigar > 0) {
        if (o.n_cigar > ez->m_cigar) {
            st = KSW2_HIP_EINVAL;  // cannot happen: n_cigar <= qlen + tlen <= m_cigar
        } else if (hipMemcpyAsync(ez->cigar, p->d_cig + p->cig_off[0], (size_t)o.n_cigar * 4,
                                  hipMemcpyDeviceToHost, D.stream) != hipSuccess ||
                   hipStreamSynchronize(D.stream) != hipSuccess) {
            st = KSW2_HIP_EDEVICE;
        }
    }
    if (p) plan_destroy(p);  // waits for the stream; the region goes back to the pool
    if (st != KSW2_HIP_OK) return st;
    ez->max = (uint32_t)o.max, ez->zdropped = o.zdropped;
    ez->max_q = o.max_q, ez->max_t = o.max_t;
    ez->mqe = o.mqe, ez->mqe_t = o.mqe_t, ez->mte = o.mte, ez->mte_q = o.mte_q;
    ez->score = o.score, ez->reach_end = o.reach_end, ez->stopped = o.stopped;
    if (o.n_cigar < 0) {  // failed Singletrack walk: no path (the buffer stays the caller's)
        ez->n_cigar = 0, ez->zdropped = 1;
    } else {
        ez->n_cigar = o.n_cigar;
    }
    return KSW2_HIP_OK;
}

}  // namespace kswhip
