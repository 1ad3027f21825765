// lcs_comb_device_noalloc: lcs_comb_device for callers that cannot allocate
// device memory (code under stream capture, tight serving loops, harnesses
// that forbid hipMalloc).  The launch's device workspace -- control words,
// the pair descriptor, the band task list and the strand ring -- lives in a
// file-scope device array, one copy per GPU, held until the process ends.
// The descriptor and tasks are written by a kernel from its arguments, so
// the call makes no copy from host memory either.  The kernels are those of
// lcs_comb_device (prep, band chain, extract); only the workspace differs.
//
// Every call on one GPU shares that GPU's array: calls whose kernels overlap
// on one GPU (two streams, two threads) race and must be ordered by the
// caller.  The array keeps the last call's contents; the prep kernel clears
// the control words and seeds the ring on every call.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "comb_kernels.h"
#include "lcs_hip.h"

namespace {

// Workspace layout (bytes): [0, 256) control words (ticket, abort flag),
// [256, 512) the PairDesc, then the task list
// (kNoAllocMaxTasks int2), then the ring (2n + 64 words of 8 bytes).
constexpr int32_t kNoAllocMaxN = 1 << 21;
constexpr int32_t kNoAllocMaxTasks = 1 << 16;
constexpr size_t kCtrlOff = 0, kDescOff = 256, kTaskOff = 512;
static_assert(sizeof(PairDesc) <= kTaskOff - kDescOff, "the pair descriptor must fit its scratch slot");
constexpr size_t kRingOff = kTaskOff + sizeof(int2) * (size_t)kNoAllocMaxTasks;
constexpr size_t kScratchBytes = kRingOff + 8 * (2 * (size_t)kNoAllocMaxN + 64);
constexpr unsigned long long kTimeoutTicks = 500000000ull;   // 5 s of the 100 MHz device clock

__device__ __align__(256) unsigned char scratch[kScratchBytes];

// Per-device address of `scratch` (hipGetSymbolAddress once per device, with
// that device current) and compute-unit count.
constexpr int kMaxDevices = 64;
std::mutex g_mu;
unsigned char *g_ptr[kMaxDevices];
int g_cus[kMaxDevices];

int device_scratch(int dev, unsigned char **ptr, int *cus)
{
    if (dev < 0 || dev >= kMaxDevices) return LCS_ERANGE;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[dev]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return LCS_EIO;
        int n = 0;
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0)
            return LCS_EIO;
        g_ptr[dev] = (unsigned char *)p;
        g_cus[dev] = n;
    }
    *ptr = g_ptr[dev];
    *cus = g_cus[dev];
    return LCS_OK;
}

// Writes the pair descriptor and the band tasks (band g = task g; groups of
// four bands for the LDS variant) into the workspace.
__global__ void noalloc_setup_kernel(PairDesc P, PairDesc *d_pair, int2 *tasks, int32_t ntasks)
{
    const int32_t t = (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);
    if (t == 0) *d_pair = P;
    if (t < ntasks) tasks[t] = make_int2(0, t);
}

// The launch's outcome for the caller: LCS_ETIMEDOUT if a wait gave up.
__global__ void noalloc_status_kernel(const unsigned *ctrl, int32_t *d_status)
{
    *d_status = (ctrl && ctrl[1]) ? LCS_ETIMEDOUT : LCS_OK;
}

}  // namespace

extern "C" int lcs_comb_device_noalloc(const int32_t *d_a, int32_t m, const int32_t *d_b, int32_t n,
                                       const lcs_schedule *sched, int32_t *d_h_out, int32_t *d_v_out,
                                       int32_t *d_status, void *stream)
{
    if (m < 0 || n < 0) return LCS_EINVAL;
    if ((m > 0 && (!d_a || !d_h_out)) || (n > 0 && (!d_b || !d_v_out))) return LCS_EINVAL;
    if ((long long)m + n > 0x7fffffffLL) return LCS_ERANGE;
    hipStream_t s = (hipStream_t)stream;
    if (m == 0 || n == 0) {
        if (lcsk_launch_identity(d_h_out, m, d_v_out, n, s)) return LCS_EIO;
        if (d_status) hipLaunchKernelGGL(noalloc_status_kernel, dim3(1), dim3(1), 0, s, (const unsigned *)nullptr,
                                         d_status);
        return hipGetLastError() == hipSuccess ? LCS_OK : LCS_EIO;
    }
    if (n > kNoAllocMaxN) return LCS_ERANGE;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return LCS_EIO;
    unsigned char *base = nullptr;
    int cus = 0;
    int rc = device_scratch(dev, &base, &cus);
    if (rc) return rc;
    // schedule: the band chain in one strip (no merges: they need more memory)
    if (sched && sched->strips > 1) return LCS_EINVAL;
    int tm = sched && sched->tile_m ? sched->tile_m : 0;
    if (tm && tm != 1 && tm != 2 && tm != 4 && tm != 8 && tm != 16 && tm != 32) return LCS_EINVAL;
    const int wps = sched && sched->waves_per_simd ? sched->waves_per_simd : 1;
    if (wps < 1 || wps > 8) return LCS_EINVAL;
    const int variant = sched ? sched->variant : LCS_VARIANT_AUTO;
    if (variant != LCS_VARIANT_AUTO && variant != LCS_VARIANT_CHAIN && variant != LCS_VARIANT_CHAIN_LDS)
        return LCS_EINVAL;
    const int group_w = variant == LCS_VARIANT_CHAIN_LDS ? 4 : 1;
    if (!tm) {   // the largest power of two <= 16 that gives every SIMD a band
        tm = 16;
        while (tm > 1 && (long long)m < 4LL * cus * 64LL * tm) tm >>= 1;
    }
    const long long nbands = ((long long)m + 64LL * tm - 1) / (64LL * tm);
    const long long ntasks = (nbands + group_w - 1) / group_w;
    if (ntasks > kNoAllocMaxTasks) return LCS_ERANGE;

    unsigned *ctrl = (unsigned *)(base + kCtrlOff);
    PairDesc *d_pair = (PairDesc *)(base + kDescOff);
    int2 *tasks = (int2 *)(base + kTaskOff);
    PairDesc P{};
    P.a = d_a;
    P.b = d_b;
    P.h_out = d_h_out;
    P.v_out = d_v_out;
    P.E = (unsigned long long *)(base + kRingOff);
    P.m = m;
    P.n = n;
    P.nbands = (int32_t)nbands;
    P.tile_m = tm;
    hipLaunchKernelGGL(noalloc_setup_kernel, dim3((unsigned)((ntasks + 255) / 256)), dim3(256), 0, s, P, d_pair,
                       tasks, (int32_t)ntasks);
    ChainArgs A{};
    A.pairs = d_pair;
    A.tasks = tasks;
    A.ntasks = (int32_t)ntasks;
    A.group_w = group_w;
    A.ticket = ctrl;
    A.abort_flag = ctrl + 1;
    A.sticky = nullptr;
    A.timeout_ticks = kTimeoutTicks;
    int blocks;
    if (group_w == 1) {
        const long long waves = std::min<long long>(ntasks, 4LL * cus * wps);
        blocks = (int)std::max<long long>(1, (waves + 3) / 4);
    } else {
        blocks = (int)std::max<long long>(1, std::min<long long>(ntasks, (long long)cus * wps));
    }
    rc = lcsk_launch_chain(A, d_pair, 1, std::max(m, n), tm, blocks, 256, s);
    if (rc) return rc == -22 ? LCS_EINVAL : LCS_EIO;
    if (d_status) hipLaunchKernelGGL(noalloc_status_kernel, dim3(1), dim3(1), 0, s, ctrl, d_status);
    return hipGetLastError() == hipSuccess ? LCS_OK : LCS_EIO;
}
