// hbs_noalloc.hip -- hbs_dot_f64_noalloc: the dot product of hbs_dot_f64
// for callers that cannot allocate device memory (tight serving loops,
// harnesses that forbid hipMalloc).  Its kernels are hbs_dot_f64's own
// (k_dot, then k_finalize over the per-workgroup partials, in the same fixed
// order, so the two give the same result bit for bit); what hbs_dot_f64
// keeps in memory it allocates per context -- the vector-workgroup plan, the
// partials and the result state -- lives here in one file-scope device
// array, `scratch`, of kScratchBytes (513 KiB) on every GPU the library runs
// on, until the process ends.
//
// Every call on a GPU shares that GPU's array and leaves its contents
// behind: calls that overlap on one GPU (two host threads, or two streams
// with a call on each in flight) race.  Nothing in the array needs clearing
// between calls: each call rewrites every partial it reads, and the plan is
// rewritten whenever n or the vector-workgroup size changes.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>

#include "hbs_objects.h"

namespace hbs {
namespace {

constexpr int kNoallocMaxN = 1 << 26;          // elements
constexpr int kNoallocMaxBlocks = 16400;        // vector workgroups (n / 4096 at 2^26, with room)
constexpr size_t kStateOff = 0;                 // SolverState: the result
constexpr size_t kSysOff = 256;                 // SysRange {0, 0, 0, nvb}
constexpr size_t kBlocksOff = 512;              // VecBlock[kNoallocMaxBlocks]
constexpr size_t kPartOff = kBlocksOff + sizeof(VecBlock) * (size_t)kNoallocMaxBlocks;   // partials
constexpr size_t kScratchBytes = kPartOff + sizeof(double) * kMaxDots * (size_t)kNoallocMaxBlocks;
static_assert(sizeof(SolverState) <= kSysOff - kStateOff, "state slot");
static_assert(sizeof(SysRange) <= kBlocksOff - kSysOff, "system slot");
static_assert(kPartOff % 256 == 0 && kScratchBytes % 256 == 0, "slot alignment");

// In host code this name is a host-side stand-in, not a device address:
// every host use goes through the per-device address from
// hipGetSymbolAddress (device_scratch below).
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;

std::mutex g_mu;
unsigned char *g_ptr[kMaxDevices] = {};   // this device's copy of `scratch`
int g_plan_n[kMaxDevices];                // plan held in that copy (n, elements per block)
int g_plan_epb[kMaxDevices];
VecBlock g_host_blocks[kNoallocMaxBlocks];

// The device address of `scratch` on `device` (current), looked up on the
// first call there.  Caller holds g_mu.
int device_scratch(int device, unsigned char **out)
{
    if (device < 0 || device >= kMaxDevices) return set_error(HBS_ERR_ARG, "device %d out of range", device);
    if (!g_ptr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_error(HBS_ERR_HIP, "scratch address on device %d: %s", device, hipGetErrorString(e));
        g_ptr[device] = static_cast<unsigned char *>(p);
        g_plan_n[device] = -1;
        g_plan_epb[device] = -1;
    }
    *out = g_ptr[device];
    return HBS_OK;
}

}  // namespace
}  // namespace hbs

using namespace hbs;

extern "C" int hbs_dot_f64_noalloc(hbs_ctx ctx, int n, const double *x, const double *y, double *out_host)
{
    if (!ctx || !out_host || n < 0 || (n > 0 && (!x || !y))) return set_error(HBS_ERR_ARG, "bad argument");
    if (n > kNoallocMaxN)
        return set_error(HBS_ERR_ARG, "n = %d exceeds the allocation-free dot's maximum of %d elements", n,
                         kNoallocMaxN);
    if (n == 0) {
        *out_host = 0.0;
        return HBS_OK;
    }
    const int epb = vec_elems_per_block(n);
    const long long nvb_ll = ((long long)n + epb - 1) / epb + 1;   // plan_vec's count, at most one more
    if (nvb_ll > kNoallocMaxBlocks)
        return set_error(HBS_ERR_ARG, "n = %d needs %lld vector workgroups, above the allocation-free dot's %d", n,
                         nvb_ll, kNoallocMaxBlocks);
    DevGuard g(ctx->device);
    hipStream_t q = ctx->stream;
    std::lock_guard<std::mutex> lock(g_mu);
    unsigned char *base = nullptr;
    CHK(device_scratch(ctx->device, &base));
    int nvb = 0;
    if (g_plan_n[ctx->device] != n || g_plan_epb[ctx->device] != epb) {
        nvb = plan_vec_into(0, n, epb, 0, g_host_blocks, kNoallocMaxBlocks);
        if (nvb < 0) return set_error(HBS_ERR_ARG, "vector plan of n = %d does not fit", n);
        const SysRange sr{0, 0, 0, nvb};
        HIPCHK(hipMemcpyAsync(base + kBlocksOff, g_host_blocks, sizeof(VecBlock) * (size_t)nvb,
                              hipMemcpyHostToDevice, q));
        HIPCHK(hipMemcpyAsync(base + kSysOff, &sr, sizeof sr, hipMemcpyHostToDevice, q));
        // the host copies are reused by the next call: finish the uploads
        HIPCHK(hipStreamSynchronize(q));
        g_plan_n[ctx->device] = n;
        g_plan_epb[ctx->device] = epb;
    } else {
        nvb = plan_vec_into(0, n, epb, 0, nullptr, kNoallocMaxBlocks);
    }
    VecArgs<double> a{};
    a.blocks = reinterpret_cast<const VecBlock *>(base + kBlocksOff);
    a.sys = reinterpret_cast<const SysRange *>(base + kSysOff);
    a.nvb = nvb;
    a.part_out = reinterpret_cast<double *>(base + kPartOff);
    SolverState *st = reinterpret_cast<SolverState *>(base + kStateOff);
    HIPCHK(launch_dot(a, x, y, q));
    HIPCHK(launch_finalize(st, a.sys, 1, a.part_out, nvb, 0, FIN_STORE, q));
    HIPCHK(hipMemcpyAsync(out_host, &st->out, sizeof(double), hipMemcpyDeviceToHost, q));
    HIPCHK(hipStreamSynchronize(q));
    return HBS_OK;
}
