// m2d_noalloc.hip -- the entry point that allocates nothing.
//
// m2d_accumulate_2d sums the CIC weights in a 64-bit fixed-point grid that
// the context allocates on first use (and grows with the grid).  Callers that
// must not allocate device memory -- code under HIP graph capture, tight
// serving loops, harnesses that forbid hipMalloc -- call
// m2d_accumulate_2d_noalloc instead: the same kernels, with the fixed-point
// grid in a file-scope device array of at most kStaticNodes nodes.  Each GPU
// holds its own copy of the array until the process ends; its device address
// is resolved once per device when a context is created on that device.
// Calls on one GPU share the array, so calls that overlap on one GPU race.
#include "m2d_internal.hpp"
#include <mutex>

namespace m2d {

constexpr size_t kStaticNodes = (size_t)1 << 20;  // 1024 x 1024 nodes: 8 MiB
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned long long scratch[kStaticNodes];

// per device: the address of `scratch` in that device's memory
static unsigned long long *g_scratch[kMaxDevices];
static std::mutex g_scratch_mu;

// called by m2d_ctx_create with `device` current
int noalloc_init_device(int device)
{
    M2D_CHECK(device >= 0 && device < kMaxDevices, M2D_ERR_BAD_ARG, "device %d beyond %d", device, kMaxDevices);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch[device]) return M2D_OK;
    void *ptr = nullptr;
    M2D_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    M2D_CHECK(ptr, M2D_ERR_HIP, "hipGetSymbolAddress(scratch) gave NULL on device %d", device);
    g_scratch[device] = (unsigned long long *)ptr;
    return M2D_OK;
}

}  // namespace m2d

using namespace m2d;

extern "C" {

int m2d_accumulate_2d_noalloc(m2d_ctx *ctx, const m2d_grid2d *g, const m2d_particles2d *p, double charge,
                              double *rho)
{
    M2D_CHECK(ctx && g && rho, M2D_ERR_BAD_ARG, "m2d_accumulate_2d_noalloc: NULL argument");
    M2D_CHECK(g->nx >= 2 && g->nz >= 2 && g->dx > 0 && g->dz > 0, M2D_ERR_BAD_ARG,
              "m2d_accumulate_2d_noalloc: bad grid");
    M2D_TRY(check_parts(p));
    const size_t n = (size_t)g->nx * (size_t)g->nz;
    M2D_CHECK(n <= kStaticNodes, M2D_ERR_BAD_ARG,
              "m2d_accumulate_2d_noalloc: a %d x %d grid exceeds the static scratch of %zu nodes", g->nx, g->nz,
              kStaticNodes);
    unsigned long long *acc = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        acc = g_scratch[ctx->device];
    }
    M2D_CHECK(acc, M2D_ERR_HIP, "m2d_accumulate_2d_noalloc: scratch of device %d not resolved", ctx->device);
    // the array keeps what the last call on this GPU left: start from zero
    // (a fill kernel, which graph capture records as a kernel node)
    M2D_TRY(zero_words(ctx, acc, n));
    if (p->n > 0) {
        k_accumulate2d<<<grid_for(p->n, 1), kBlock, 0, ctx->stream>>>(p->x, p->z, p->xl, p->zl, p->n, g->nx, g->nz,
                                                                     1.0 / g->dx, 1.0 / g->dz, g->xmin, g->zmin,
                                                                     acc, ctx->counters);
        M2D_HIP(hipGetLastError());
    }
    k_flush<<<(unsigned)((n + 255) / 256), 256, 0, ctx->stream>>>(acc, rho, charge, n);
    M2D_HIP(hipGetLastError());
    return M2D_OK;
}

}  // extern "C"
