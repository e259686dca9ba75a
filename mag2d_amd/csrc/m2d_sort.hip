// m2d_sort.hip -- stable reordering of a species' particles by tile, and the
// tile schedule of the sorted store.
//
// The reference keeps particles in insertion order with holes
// (particles.hpp:223-249).  Here the store is periodically rebuilt in tile
// order (8 x 8-cell tiles, row-major), dropping empty slots: each block of
// the particle pass then gathers from and deposits into one small window of
// the grid, held in LDS.  Inside a tile the particles keep their old slot
// order, so the result is deterministic.  They are deliberately not ordered
// by cell: the 64 lanes of a wave would then share one or two cells, and
// their LDS atomics on the same four nodes would serialise (a cell-ordered
// store runs the pass about 1.5x slower until the particles have mixed).
// The key holds the cell too (tile << 6 | cell) and the radix sort ranks the
// tile bits only (rocPRIM's LSD radix sort of the key/slot pairs).
#include <cstring>
#include <new>
#include <rocprim/device/device_radix_sort.hpp>
#include "m2d_objects.hpp"

namespace m2d {

struct SortGeom {
    int nx, nz, ntz;
    double idx, idz, xmin, zmin;
    uint32_t dead_key;
};

__device__ __forceinline__ uint32_t cell_key(const SortGeom &g, double x, double z)
{
    double X = (x - g.xmin) * g.idx, Z = (z - g.zmin) * g.idz;
    int i = X >= 0.0 ? (int)fmin(X, (double)(g.nx - 2)) : 0;
    int j = Z >= 0.0 ? (int)fmin(Z, (double)(g.nz - 2)) : 0;
    uint32_t tile = (uint32_t)((i / kTile) * g.ntz + (j / kTile));
    return (tile << kTileShift) | ((uint32_t)(i % kTile) << 3) | (uint32_t)(j % kTile);
}

// the key of a long-time position (xls, zls not null) is that of the FP64 sum
__global__ void k_sort_keys(const float *xs, const float *zs, const float *xls, const float *zls, int64_t n,
                            SortGeom g, uint32_t *keys, uint32_t *vals)
{
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    float x = xs[p];
    keys[p] = is_dead(x) ? g.dead_key
                         : xls ? cell_key(g, pos64(x, xls[p]), pos64(zs[p], zls[p])) : cell_key(g, x, zs[p]);
    vals[p] = (uint32_t)p;
}

__global__ void k_sort_gather(m2d_particles2d src, m2d_particles2d dst, const uint32_t *keys,
                              const uint32_t *vals, uint32_t dead_key, unsigned long long *n_live)
{
    // destination ranges in XCD-contiguous order: a nearly sorted store's
    // sources of neighbouring ranges share cache lines, and then one L2
    const unsigned nb = (unsigned)((src.n + blockDim.x - 1) / blockDim.x);
    const unsigned b = xcd_block(blockIdx.x, nb);
    if (b >= nb) return;
    int64_t i = (int64_t)b * blockDim.x + threadIdx.x;
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
    if (src.xl) {
        dst.xl[i] = src.xl[s];
        dst.zl[i] = src.zl[s];
    }
}

// [start, end) of every non-empty tile in the sorted keys
__global__ void k_tile_bounds(const uint32_t *keys, int64_t n, uint32_t dead_key, int *tstart, int *tend)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t k = keys[i];
    if (k == dead_key) return;
    uint32_t t = k >> kTileShift;
    if (i == 0 || (keys[i - 1] >> kTileShift) != t) tstart[t] = (int)i;
    if (i == n - 1 || keys[i + 1] == dead_key || (keys[i + 1] >> kTileShift) != t) tend[t] = (int)(i + 1);
}

__global__ void k_tile_chunks(const int *tstart, const int *tend, int ntiles, int cap, int *count)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntiles) return;
    int c = tend[t] - tstart[t];
    count[t] = c > 0 ? (c + cap - 1) / cap : 0;
}

// The schedule's order class of a chunk [a, e): the rounds of `round`
// slots the pass takes over it (its rounds start at a rounded down to 64),
// less one, at most kSchedClasses - 1; round = 0: class 0 for all
constexpr int kSchedClasses = 8;
__device__ __forceinline__ int chunk_class(int a, int e, int round)
{
    if (round <= 0) return 0;
    const int r = (e - (a & ~63) + round - 1) / round - 1;
    return min(max(r, 0), kSchedClasses - 1);
}

// exclusive scans of the per-tile chunk counts (one block) and the
// schedule: chunks in descending class (the longest first), each class in
// tile order.  Blocks start in index order as earlier ones retire, so the
// launch ends on short blocks.
__global__ void __launch_bounds__(1024) k_tile_schedule(const int *tstart, const int *tend,
                                                        const int *count, int ntiles, int cap, int round,
                                                        int4 *sched, unsigned int *n_chunks)
{
    __shared__ int s[kSchedClasses][1024];  // per class: this thread's chunks, then their first index
    const int tid = threadIdx.x;
    const int per = (ntiles + 1023) / 1024;
    const int lo = tid * per, hi = min(ntiles, lo + per);
#pragma unroll
    for (int c = 0; c < kSchedClasses; c++) s[c][tid] = 0;
    for (int t = lo; t < hi; t++) {
        for (int k = 0; k < count[t]; k++) {
            const int a = tstart[t] + k * cap;
            s[chunk_class(a, min(tend[t], a + cap), round)][tid]++;
        }
    }
    int own[kSchedClasses];
#pragma unroll
    for (int c = 0; c < kSchedClasses; c++) own[c] = s[c][tid];
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        int v[kSchedClasses];
#pragma unroll
        for (int c = 0; c < kSchedClasses; c++) v[c] = tid >= o ? s[c][tid - o] : 0;
        __syncthreads();
#pragma unroll
        for (int c = 0; c < kSchedClasses; c++) s[c][tid] += v[c];
        __syncthreads();
    }
    // class c starts after every chunk of the classes above it
    int base = 0;
    int first[kSchedClasses];
#pragma unroll
    for (int c = kSchedClasses - 1; c >= 0; c--) {
        first[c] = base + s[c][tid] - own[c];
        base += s[c][1023];
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < kSchedClasses; c++) s[c][tid] = first[c];
    for (int t = lo; t < hi; t++) {
        for (int k = 0; k < count[t]; k++) {
            const int a = tstart[t] + k * cap, e = min(tend[t], a + cap);
            sched[s[chunk_class(a, e, round)][tid]++] = make_int4(t, a, e, 0);
        }
    }
    if (tid == 1023) *n_chunks = (unsigned int)base;
}

m2d_tiling *find_tiling(m2d_ctx *ctx, const float *x, int64_t n)
{
    for (m2d_tiling *t : ctx->tilings)
        if (t->x == x && t->n == n) return t;
    return nullptr;
}

void forget_tiling(m2d_ctx *ctx, const float *x)
{
    for (m2d_tiling *t : ctx->tilings)
        if (t->x == x) t->x = nullptr;
}

void free_tilings(m2d_ctx *ctx)
{
    for (m2d_tiling *t : ctx->tilings) {
        if (t->buf.ptr) (void)hipFree(t->buf.ptr);
        delete t;
    }
    ctx->tilings.clear();
}

// build the schedule of the store at x (n slots) from its sorted keys; tiles
// number ntx * nty * ntz (nty = 1 in 2D), 64 cells each; a tile's slots are
// cut into chunks of at most cap, ordered longest first by the rounds of
// `round` slots the pass takes over them (round = 0: tile order only)
int build_tiling(m2d_ctx *ctx, const float *x, int64_t n, const uint32_t *keys, uint32_t dead_key, int ntx,
                 int nty, int ntz, int cap, int round)
{
    forget_tiling(ctx, x);
    m2d_tiling *T = nullptr;
    for (m2d_tiling *t : ctx->tilings)
        if (t->x == nullptr) { T = t; break; }
    if (!T) {
        if (ctx->tilings.size() >= 8) {  // reuse the oldest record
            T = ctx->tilings.front();
            ctx->tilings.erase(ctx->tilings.begin());
        } else {
            T = new (std::nothrow) m2d_tiling();
            M2D_CHECK(T, M2D_ERR_NO_MEMORY, "tiling: host allocation");
        }
        ctx->tilings.push_back(T);
    }
    T->x = nullptr;
    const int ntiles = ntx * nty * ntz;
    const int max_chunks = ntiles + (int)(n / cap) + 1;
    const size_t words = 3 * (size_t)ntiles + 4 * (size_t)max_chunks + 4;
    M2D_TRY(ensure(ctx, T->buf, words * sizeof(int)));
    int *b = (int *)T->buf.ptr;
    T->sched = (int4 *)b;  // first: 16-byte aligned
    T->tile_start = b + 4 * (size_t)max_chunks;
    T->tile_end = T->tile_start + ntiles;
    T->chunk_count = T->tile_end + ntiles;
    T->n_chunks = (unsigned int *)(T->chunk_count + ntiles);
    M2D_TRY(fill_bytes(ctx, T->tile_start, 0, sizeof(int) * 2 * (size_t)ntiles));
    k_tile_bounds<<<(unsigned)((n + 255) / 256), 256, 0, ctx->stream>>>(keys, n, dead_key, T->tile_start,
                                                                        T->tile_end);
    k_tile_chunks<<<(unsigned)((ntiles + 255) / 256), 256, 0, ctx->stream>>>(T->tile_start, T->tile_end,
                                                                            ntiles, cap, T->chunk_count);
    k_tile_schedule<<<1, 1024, 0, ctx->stream>>>(T->tile_start, T->tile_end, T->chunk_count, ntiles, cap, round,
                                                 T->sched, T->n_chunks);
    M2D_HIP(hipGetLastError());
    T->ntx = ntx;
    T->nty = nty;
    T->ntz = ntz;
    T->max_chunks = max_chunks;
    T->n = n;
    T->x = x;
    return M2D_OK;
}

// Sort n 32-bit keys in ascending order from `in` into `out` (stream-ordered,
// temporaries in the context's sort buffer).
int sort_keys_u32(m2d_ctx *ctx, const uint32_t *in, uint32_t *out, size_t n)
{
    if (n == 0) return M2D_OK;
    size_t temp = 0;
    hipError_t e = rocprim::radix_sort_keys(nullptr, temp, in, out, n, 0, 32, ctx->stream);
    M2D_CHECK(e == hipSuccess, M2D_ERR_HIP, "radix_sort_keys size query: %s", hipGetErrorString(e));
    M2D_TRY(ensure(ctx, ctx->sortbuf, temp + 256));
    e = rocprim::radix_sort_keys(ctx->sortbuf.ptr, temp, in, out, n, 0, 32, ctx->stream);
    M2D_CHECK(e == hipSuccess, M2D_ERR_HIP, "radix_sort_keys: %s", hipGetErrorString(e));
    return M2D_OK;
}

}  // namespace m2d

using namespace m2d;

extern "C" {

int m2d_sort_2d(m2d_ctx *ctx, m2d_fields2d *f, const m2d_particles2d *src, const m2d_particles2d *dst,
                int64_t *n_live)
{
    M2D_CHECK(ctx && f && src && dst, M2D_ERR_BAD_ARG, "m2d_sort_2d: NULL argument");
    M2D_CHECK(src->n >= 0 && src->n < (1ll << 31), M2D_ERR_BAD_ARG, "m2d_sort_2d: %lld slots",
              (long long)src->n);
    M2D_CHECK(dst->n == src->n, M2D_ERR_BAD_ARG, "m2d_sort_2d: dst holds %lld slots, src %lld",
              (long long)dst->n, (long long)src->n);
    M2D_CHECK(src->n == 0 || (src->x && src->z && src->vx && src->vy && src->vz && dst->x && dst->z &&
                              dst->vx && dst->vy && dst->vz),
              M2D_ERR_BAD_ARG, "m2d_sort_2d: NULL particle array");
    M2D_CHECK(src->x != dst->x, M2D_ERR_BAD_ARG, "m2d_sort_2d: src and dst must differ");
    M2D_CHECK((src->xl == nullptr) == (src->zl == nullptr) && (dst->xl == nullptr) == (dst->zl == nullptr) &&
                  (src->xl == nullptr) == (dst->xl == nullptr),
              M2D_ERR_BAD_ARG, "m2d_sort_2d: src and dst must both hold position remainders or neither");
    const size_t n = (size_t)src->n;
    SortGeom g;
    g.nx = f->nx; g.nz = f->nz;
    const int ntx = (f->nx - 2) / kTile + 1;
    g.ntz = (f->nz - 2) / kTile + 1;
    g.idx = f->idx; g.idz = f->idz;
    g.xmin = f->p.grid.xmin; g.zmin = f->p.grid.zmin;
    uint64_t keyspace = ((uint64_t)ntx * g.ntz) << kTileShift;
    int bits = 0;
    while ((1ull << bits) <= keyspace) bits++;
    M2D_CHECK(bits <= 32, M2D_ERR_BAD_ARG, "m2d_sort_2d: grid too large for 32-bit keys");
    g.dead_key = bits == 32 ? 0xFFFFFFFFu : (uint32_t)((1ull << bits) - 1);
    forget_tiling(ctx, src->x);
    if (n == 0) {
        forget_tiling(ctx, dst->x);
        if (n_live) *n_live = 0;
        return M2D_OK;
    }
    size_t temp = 0;
    hipError_t e = rocprim::radix_sort_pairs(nullptr, temp, (const uint32_t *)nullptr, (uint32_t *)nullptr,
                                             (const uint32_t *)nullptr, (uint32_t *)nullptr, n, kTileShift, bits,
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
    k_sort_keys<<<(unsigned)((n + 255) / 256), 256, 0, ctx->stream>>>(src->x, src->z, src->xl, src->zl, src->n, g,
                                                                      k0, v0);
    M2D_HIP(hipGetLastError());
    e = rocprim::radix_sort_pairs(tmp, temp, k0, k1, v0, v1, n, kTileShift, bits, ctx->stream);
    M2D_CHECK(e == hipSuccess, M2D_ERR_HIP, "radix_sort_pairs: %s", hipGetErrorString(e));
    k_sort_gather<<<xcd_grid((unsigned)((n + 255) / 256)), 256, 0, ctx->stream>>>(*src, *dst, k1, v1, g.dead_key,
                                                                              live);
    M2D_HIP(hipGetLastError());
    M2D_TRY(build_tiling(ctx, dst->x, dst->n, k1, g.dead_key, ntx, 1, g.ntz, kChunkCap, kRound2));
    if (n_live) {
        M2D_TRY(read_counters(ctx));
        *n_live = (int64_t)ctx->h_counters[CNT_LIVE];
    }
    return M2D_OK;
}

int m2d_sort_forget_2d(m2d_ctx *ctx, const m2d_particles2d *p)
{
    M2D_CHECK(ctx && p, M2D_ERR_BAD_ARG, "m2d_sort_forget_2d: NULL argument");
    forget_tiling(ctx, p->x);
    return M2D_OK;
}

}  // extern "C"
