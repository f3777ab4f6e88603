// lbm_noalloc.cpp -- lbm_conservation_noalloc: the FP64 conservation totals
// of lbm_conservation without any device allocation, for callers that cannot
// allocate (code between HIP graph captures, tight serving loops, harnesses
// that forbid hipMalloc).
//
// The block partial sums live in `scratch`, a file-scope device array: one
// (mass, jx, jy, jz) per block of the conservation reduction, then the four
// totals.  Each GPU the library runs on holds its own copy until the process
// ends.  Host code never passes the array's name to a kernel or a copy (on
// the host that name is a stand-in, not device memory): it uses the device
// address hipGetSymbolAddress returns, looked up once per device on the
// first call made with that device current.  Calls that overlapped on one GPU
// would race on the array; the calls a process makes are serialised by a lock
// held until their totals are on the host.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/mesobasic_lbm.h"
#include "lbm_handle.h"
#include "lbm_internal.h"
#include "lbm_lowp.h"

namespace {

constexpr int64_t kScratchBlocks = int64_t(1) << 16;
constexpr int64_t kScratchMaxNodes =
    kScratchBlocks * lbm::kConsBlock * lbm::kConsPerThread;   // 2^28 nodes
static __device__ __align__(256) double scratch[(kScratchBlocks + 1) * 4];

constexpr int kMaxDevices = 64;
std::mutex g_mu;                        // guards g_scratch and serialises the calls
double* g_scratch[kMaxDevices] = {};    // device address of `scratch`, per device

// Totals of the nb block partials, added in block order from 0.0 -- the
// additions lbm_conservation makes on the host, so the same bits (one lane
// per component; the file is compiled without contraction).
__global__ void k_sum_partials(double* __restrict__ part, int nb)
{
    const int k = threadIdx.x;
    if (k >= 4) return;
    double s = 0.0;
    for (int b = 0; b < nb; b++) s += part[static_cast<int64_t>(b) * 4 + k];
    part[kScratchBlocks * 4 + k] = s;
}

#define NA_HIP(expr)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return lbm::set_error(LBM_EHIP, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_),   \
                                  __FILE__, __LINE__);                                        \
    } while (0)

// Device address of `scratch` on device dev (caller holds g_mu).
int scratch_on(int dev, double** out)
{
    if (dev < 0 || dev >= kMaxDevices)
        return lbm::set_error(LBM_ENODEV, "lbm_conservation_noalloc: device %d out of range", dev);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        NA_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch[dev] = static_cast<double*>(p);
    }
    *out = g_scratch[dev];
    return LBM_OK;
}

}  // namespace

extern "C" int lbm_conservation_noalloc(lbm_handle* h, double out[4])
{
    if (!h || !out) return lbm::set_error(LBM_EINVAL, "lbm_conservation_noalloc: NULL argument");
    if (h->g.n > kScratchMaxNodes)
        return lbm::set_error(LBM_EINVAL,
                              "lbm_conservation_noalloc: %lld nodes exceed its limit of %lld",
                              (long long)h->g.n, (long long)kScratchMaxNodes);
    if (!h->initialised)
        return lbm::set_error(LBM_ESTATE, "lbm_conservation_noalloc before initialisation");
    int dev = -1;
    NA_HIP(hipGetDevice(&dev));
    if (dev != h->device)
        return lbm::set_error(LBM_ESTATE,
                              "lbm_conservation_noalloc: device %d is current, the handle is on %d",
                              dev, h->device);
    std::lock_guard<std::mutex> lock(g_mu);
    double* part = nullptr;
    int rc = scratch_on(dev, &part);
    if (rc) return rc;
    int nb = 0;
    if (!is_lowp(h))
        NA_HIP(lbm::fp64_conservation(h->a, h->g, static_cast<const double*>(h->buf[h->cur]),
                                      h->kind[h->cur], part, &nb, h->stream));
    else
        NA_HIP(lbm::lowp_conservation(h->d.precision, h->a, h->g, h->d.rho_shift, h->buf[h->cur],
                                      h->kind[h->cur], h->rref[h->cur], part, &nb, h->stream));
    if (nb > kScratchBlocks)   // the node limit above keeps nb within the array
        return lbm::set_error(LBM_EINVAL, "lbm_conservation_noalloc: %d blocks", nb);
    k_sum_partials<<<1, 64, 0, h->stream>>>(part, nb);
    NA_HIP(hipGetLastError());
    NA_HIP(hipMemcpyAsync(out, part + kScratchBlocks * 4, 4 * sizeof(double), hipMemcpyDeviceToHost,
                          h->stream));
    NA_HIP(hipStreamSynchronize(h->stream));
    return LBM_OK;
}
