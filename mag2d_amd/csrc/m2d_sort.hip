// m2d_sort.hip -- stable reordering of a species' particles by cell.
//
// The reference keeps particles in insertion order with holes
// (particles.hpp:223-249).  Here the store is periodically rebuilt in
// tile-major cell order, dropping empty slots: the E gather of neighbouring
// threads then reads neighbouring table entries, and a block's deposits fall
// in one LDS window.  Keys are (tile, cell-in-tile); the order among equal
// keys is the old slot order, so the result is deterministic.  The key/slot
// pairs are sorted with rocPRIM's LSD radix sort.
#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>
#include "m2d_objects.hpp"

namespace m2d {

struct SortGeom {
    int nx, nz, ntz;
    double idx, idz, xmin, zmin;
    uint32_t dead_key;
};

__device__ __forceinline__ uint32_t cell_key(const SortGeom &g, float x, float z)
{
    double X = ((double)x - g.xmin) * g.idx, Z = ((double)z - g.zmin) * g.idz;
    int i = X >= 0.0 ? (int)fmin(X, (double)(g.nx - 2)) : 0;
    int j = Z >= 0.0 ? (int)fmin(Z, (double)(g.nz - 2)) : 0;
    uint32_t tile = (uint32_t)((i / kTile) * g.ntz + (j / kTile));
    return (tile << 8) | ((uint32_t)(i % kTile) << 4) | (uint32_t)(j % kTile);
}

__global__ void k_sort_keys(const float *xs, const float *zs, int64_t n, SortGeom g, uint32_t *keys,
                            uint32_t *vals)
{
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    float x = xs[p];
    keys[p] = is_dead(x) ? g.dead_key : cell_key(g, x, zs[p]);
    vals[p] = (uint32_t)p;
}

__global__ void k_sort_gather(m2d_particles2d src, m2d_particles2d dst, const uint32_t *keys,
                              const uint32_t *vals, uint32_t dead_key, unsigned long long *n_live)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= src.n) return;
    uint32_t k = keys[i];
    if (k == dead_key) {
        dst.x[i] = dead_value();
        return;
    }
    if (i == src.n - 1 || keys[i + 1] == dead_key) *n_live = (unsigned long long)(i + 1);
    int64_t s = vals[i];
    dst.x[i] = src.x[s];
    dst.z[i] = src.z[s];
    dst.vx[i] = src.vx[s];
    dst.vy[i] = src.vy[s];
    dst.vz[i] = src.vz[s];
}

}  // namespace m2d

using namespace m2d;

extern "C" {

int m2d_sort_2d(m2d_ctx *ctx, m2d_fields2d *f, const m2d_particles2d *src, const m2d_particles2d *dst,
                int64_t *n_live)
{
    M2D_CHECK(ctx && f && src && dst, M2D_ERR_BAD_ARG, "m2d_sort_2d: NULL argument");
    M2D_CHECK(src->n >= 0 && src->n < (1ll << 32), M2D_ERR_BAD_ARG, "m2d_sort_2d: %lld slots",
              (long long)src->n);
    M2D_CHECK(dst->n >= src->n, M2D_ERR_BAD_ARG, "m2d_sort_2d: dst holds %lld < %lld slots",
              (long long)dst->n, (long long)src->n);
    M2D_CHECK(src->n == 0 || (src->x && src->z && src->vx && src->vy && src->vz && dst->x && dst->z &&
                              dst->vx && dst->vy && dst->vz),
              M2D_ERR_BAD_ARG, "m2d_sort_2d: NULL particle array");
    M2D_CHECK(src->x != dst->x, M2D_ERR_BAD_ARG, "m2d_sort_2d: src and dst must differ");
    const size_t n = (size_t)src->n;
    SortGeom g;
    g.nx = f->nx; g.nz = f->nz;
    int ntx = (f->nx - 2) / kTile + 1;
    g.ntz = (f->nz - 2) / kTile + 1;
    g.idx = f->idx; g.idz = f->idz;
    g.xmin = f->p.grid.xmin; g.zmin = f->p.grid.zmin;
    uint64_t keyspace = (uint64_t)ntx * g.ntz * 256;
    int bits = 0;
    while ((1ull << bits) <= keyspace) bits++;
    M2D_CHECK(bits <= 32, M2D_ERR_BAD_ARG, "m2d_sort_2d: grid too large for 32-bit keys");
    g.dead_key = bits == 32 ? 0xFFFFFFFFu : (uint32_t)((1ull << bits) - 1);
    if (n == 0) {
        if (n_live) *n_live = 0;
        return M2D_OK;
    }
    size_t temp = 0;
    hipError_t e = rocprim::radix_sort_pairs(nullptr, temp, (const uint32_t *)nullptr, (uint32_t *)nullptr,
                                             (const uint32_t *)nullptr, (uint32_t *)nullptr, n, 0, bits,
                                             ctx->stream);
    M2D_CHECK(e == hipSuccess, M2D_ERR_HIP, "radix_sort_pairs size query: %s", hipGetErrorString(e));
    const size_t arr = ((n * 4 + 255) / 256) * 256;
    M2D_TRY(ensure(ctx, ctx->sortbuf, 4 * arr + temp + 256));
    uint8_t *b = (uint8_t *)ctx->sortbuf.ptr;
    uint32_t *k0 = (uint32_t *)b, *k1 = (uint32_t *)(b + arr), *v0 = (uint32_t *)(b + 2 * arr),
             *v1 = (uint32_t *)(b + 3 * arr);
    void *tmp = b + 4 * arr;
    unsigned long long *live = ctx->counters + CNT_LIVE;
    M2D_TRY(zero_words(ctx, live, 1));
    k_sort_keys<<<(unsigned)((n + 255) / 256), 256, 0, ctx->stream>>>(src->x, src->z, src->n, g, k0, v0);
    M2D_HIP(hipGetLastError());
    e = rocprim::radix_sort_pairs(tmp, temp, k0, k1, v0, v1, n, 0, bits, ctx->stream);
    M2D_CHECK(e == hipSuccess, M2D_ERR_HIP, "radix_sort_pairs: %s", hipGetErrorString(e));
    k_sort_gather<<<(unsigned)((n + 255) / 256), 256, 0, ctx->stream>>>(*src, *dst, k1, v1, g.dead_key, live);
    M2D_HIP(hipGetLastError());
    if (n_live) {
        M2D_TRY(read_counters(ctx));
        *n_live = (int64_t)ctx->h_counters[CNT_LIVE];
    }
    return M2D_OK;
}

}  // extern "C"
