// qjx_noalloc.hip -- the ensemble density matrix without device allocation.
//
// qjx_density allocates its scratch (per-trajectory weights, split-K partial Gram
// matrices, rho) on every call.  Callers that cannot call hipMalloc (code under
// HIP graph capture, tight serving loops, harnesses that forbid it) use
// qjx_density_noalloc, which runs the same kernels on a fixed device array instead.
//
// The array below lives once per GPU for the life of the process.  Its host-side
// name is only a stand-in: the kernels get its device address, looked up once per
// device with hipGetSymbolAddress.  Every call on one GPU uses the same array, so
// calls that overlap on one GPU (two host threads, or two streams in flight) race
// and must be serialised by the caller.  The kernels write every scratch element
// they later read (weights for every trajectory, every entry of every split's
// Gram matrix, every rho entry), so nothing needs clearing between calls.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/qjx.h"
#include "qjx_ctx.h"
#include "qjx_kernels.h"

namespace {

constexpr size_t kScratchDoubles = size_t(8) << 20;   // 64 MiB
constexpr int kMaxDim = 1024;                           // largest Hilbert dimension
constexpr int kMaxTrajectories = 1 << 21;
constexpr int kMaxDevices = 64;

std::mutex g_mu;
double *g_scratch[kMaxDevices] = {};

int noalloc_fail(int code, const char *msg)
{
    qjx::set_last_error(msg);
    return code;
}

}  // namespace

static __device__ __align__(256) double scratch[kScratchDoubles];

// Device address of this GPU's scratch array (looked up on the first call per device).
static int scratch_on_current_device(double **out)
{
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
        return noalloc_fail(QJX_ERR_HIP, "no usable current HIP device");
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            qjx::set_last_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
            return QJX_ERR_HIP;
        }
        g_scratch[dev] = static_cast<double *>(p);
    }
    *out = g_scratch[dev];
    return QJX_OK;
}

static size_t round_up(size_t n, size_t a) { return (n + a - 1) / a * a; }

extern "C" int qjx_density_noalloc(qjx_ctx *c, double *rho, double *purity)
{
    if (!c || !c->have_ens) return noalloc_fail(QJX_ERR_ARG, "ensemble not initialised");
    if (c->d > kMaxDim) return noalloc_fail(QJX_ERR_ARG, "qjx_density_noalloc: Hilbert dimension above 1024");
    if (c->M > kMaxTrajectories)
        return noalloc_fail(QJX_ERR_ARG, "qjx_density_noalloc: more than 2^21 trajectories");
    double *base = nullptr;
    int rc = scratch_on_current_device(&base);
    if (rc) return rc;
    // the same split count as qjx_density when it fits (identical results), fewer
    // splits otherwise (the same sums, grouped differently)
    const size_t L2 = (size_t)c->L * c->L;
    const size_t wn = round_up((size_t)c->M + 1, 32);
    const size_t rn = 2 * (size_t)c->d * c->d;
    int splits = qjx::density_splits(c->M, c->L);
    while (splits > 1 && wn + (size_t)splits * L2 + rn > kScratchDoubles) splits--;
    if (wn + (size_t)splits * L2 + rn > kScratchDoubles)
        return noalloc_fail(QJX_ERR_ARG, "qjx_density_noalloc: ensemble does not fit the 64 MiB scratch");
    double *w = base, *part = base + wn, *drho = part + (size_t)splits * L2;
    qjx::launch_density(c->prec, c->slot[0], c->slot[1], c->st, c->M, c->d, c->dp, c->L, w, part, splits, drho,
                        w + c->M, c->stream);
    if (hipGetLastError() != hipSuccess ||
        (rho && hipMemcpyAsync(rho, drho, rn * sizeof(double), hipMemcpyDeviceToHost, c->stream) != hipSuccess) ||
        (purity && hipMemcpyAsync(purity, w + c->M, sizeof(double), hipMemcpyDeviceToHost, c->stream) != hipSuccess) ||
        hipStreamSynchronize(c->stream) != hipSuccess)
        return noalloc_fail(QJX_ERR_HIP, "qjx_density_noalloc: launch failed");
    return QJX_OK;
}
