// m2d_core.hip -- context, error reporting, device memory, host geometry
// builders (t_grid) of libmag2d_amd.
#include "m2d_internal.hpp"
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>

namespace m2d {

static thread_local std::string g_last_error;

int set_error(int code, const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

int ensure(m2d_ctx *ctx, Scratch &s, size_t bytes)
{
    if (s.bytes >= bytes) return M2D_OK;
    if (s.ptr) {
        M2D_HIP(hipStreamSynchronize(ctx->stream));
        M2D_HIP(hipFree(s.ptr));
        s.ptr = nullptr;
        s.bytes = 0;
    }
    size_t want = bytes < 4096 ? 4096 : bytes;
    hipError_t e = hipMalloc(&s.ptr, want);
    if (e != hipSuccess) {
        s.ptr = nullptr;
        return set_error(M2D_ERR_NO_MEMORY, "hipMalloc(%zu): %s", want, hipGetErrorString(e));
    }
    s.bytes = want;
    return M2D_OK;
}

__global__ void k_fill_bytes(uint8_t *dst, uint8_t v, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = v;
}

__global__ void k_fill_u64(unsigned long long *dst, unsigned long long v, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = v;
}

int fill_bytes(m2d_ctx *ctx, void *dst, int value, size_t bytes)
{
    if (bytes == 0) return M2D_OK;
    if ((bytes % 8) == 0 && ((uintptr_t)dst % 8) == 0) {
        unsigned long long v = (uint8_t)value;
        v |= v << 8; v |= v << 16; v |= v << 32;
        size_t n = bytes / 8;
        int g = (int)std::min<size_t>((n + 255) / 256, 4096);
        k_fill_u64<<<g, 256, 0, ctx->stream>>>((unsigned long long *)dst, v, n);
    } else {
        int g = (int)std::min<size_t>((bytes + 255) / 256, 4096);
        k_fill_bytes<<<g, 256, 0, ctx->stream>>>((uint8_t *)dst, (uint8_t)value, bytes);
    }
    M2D_HIP(hipGetLastError());
    return M2D_OK;
}

int zero_words(m2d_ctx *ctx, unsigned long long *dst, size_t n)
{
    if (n == 0) return M2D_OK;
    int g = (int)std::min<size_t>((n + 255) / 256, 4096);
    k_fill_u64<<<g, 256, 0, ctx->stream>>>(dst, 0ull, n);
    M2D_HIP(hipGetLastError());
    return M2D_OK;
}

int read_counters(m2d_ctx *ctx)
{
    M2D_HIP(hipMemcpyAsync(ctx->h_counters, ctx->counters, CNT_N * sizeof(unsigned long long),
                           hipMemcpyDeviceToHost, ctx->stream));
    M2D_HIP(hipStreamSynchronize(ctx->stream));
    return M2D_OK;
}

static hipEvent_t pool_get(m2d_ctx *ctx)
{
    if (!ctx->ev_pool.empty()) {
        hipEvent_t e = ctx->ev_pool.back();
        ctx->ev_pool.pop_back();
        return e;
    }
    // timing only: read by hipEventElapsedTime after a stream synchronisation,
    // so the events need no system-scope fence -- with one, every timed
    // launch ends in an L2 write-back and the next waits for it (about 4 us
    // a step at 4M ions)
    hipEvent_t e = nullptr;
    if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) return nullptr;
    return e;
}

bool prof_pair(m2d_ctx *ctx, hipEvent_t &start, hipEvent_t &stop)
{
    if (!ctx->profile) return false;
    start = pool_get(ctx);
    stop = pool_get(ctx);
    if (!start || !stop) {
        if (start) ctx->ev_pool.push_back(start);
        if (stop) ctx->ev_pool.push_back(stop);
        return false;
    }
    ctx->ev_pending.push_back({start, stop});
    return true;
}

}  // namespace m2d

using namespace m2d;

namespace m2d {
// dst += scale * src (Array2D::add and Array2D::multiply, Array.hpp:51-72)
__global__ void k_axpy(double *dst, const double *src, double scale, size_t n)
{
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (size_t)gridDim.x * blockDim.x)
        dst[k] += scale * src[k];
}
}  // namespace m2d

extern "C" {

int m2d_ctx_profile(m2d_ctx *c, int enable)
{
    M2D_CHECK(c, M2D_ERR_BAD_ARG, "m2d_ctx_profile: ctx is NULL");
    c->profile = enable != 0;
    return M2D_OK;
}

int m2d_ctx_profile_read(m2d_ctx *c, double *ms, int64_t *launches)
{
    M2D_CHECK(c && ms && launches, M2D_ERR_BAD_ARG, "m2d_ctx_profile_read: NULL argument");
    M2D_HIP(hipStreamSynchronize(c->stream));
    double total = 0;
    for (auto &pr : c->ev_pending) {
        float t = 0;
        M2D_HIP(hipEventElapsedTime(&t, pr.first, pr.second));
        total += t;
        c->ev_pool.push_back(pr.first);
        c->ev_pool.push_back(pr.second);
    }
    *ms = total;
    *launches = (int64_t)c->ev_pending.size();
    c->ev_pending.clear();
    return M2D_OK;
}

int m2d_ctx_last_pass(m2d_ctx *c, uint32_t *variant)
{
    M2D_CHECK(c && variant, M2D_ERR_BAD_ARG, "m2d_ctx_last_pass: NULL argument");
    *variant = c->last_pass;
    return M2D_OK;
}

const char *m2d_last_error(void) { return g_last_error.c_str(); }
int m2d_abi_version(void) { return M2D_ABI_VERSION; }

int m2d_ctx_create(int device, uint64_t seed, m2d_ctx **out)
{
    M2D_CHECK(out, M2D_ERR_BAD_ARG, "m2d_ctx_create: out is NULL");
    *out = nullptr;
    int ndev = 0;
    M2D_HIP(hipGetDeviceCount(&ndev));
    M2D_CHECK(device >= 0 && device < ndev, M2D_ERR_BAD_ARG,
              "m2d_ctx_create: device %d of %d", device, ndev);
    M2D_HIP(hipSetDevice(device));
    m2d_ctx *c = new (std::nothrow) m2d_ctx();
    M2D_CHECK(c, M2D_ERR_NO_MEMORY, "m2d_ctx_create: host allocation");
    c->device = device;
    c->seed = seed;
    hipError_t e = hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking);
    if (e != hipSuccess) { delete c; return set_error(M2D_ERR_HIP, "hipStreamCreate: %s", hipGetErrorString(e)); }
    c->stream = c->own;
    e = hipMalloc(&c->counters, CNT_N * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipHostMalloc(&c->h_counters, CNT_N * sizeof(unsigned long long));
    if (e != hipSuccess) {
        if (c->counters) (void)hipFree(c->counters);
        (void)hipStreamDestroy(c->own);
        delete c;
        return set_error(M2D_ERR_HIP, "m2d_ctx_create: %s", hipGetErrorString(e));
    }
    int r = zero_words(c, c->counters, CNT_N);
    if (r == M2D_OK) e = hipStreamSynchronize(c->stream);
    if (r == M2D_OK && e == hipSuccess) r = noalloc_init_device(device);
    if (r != M2D_OK || e != hipSuccess) {
        (void)hipFree(c->counters);
        (void)hipHostFree(c->h_counters);
        (void)hipStreamDestroy(c->own);
        delete c;
        return r != M2D_OK ? r : set_error(M2D_ERR_HIP, "m2d_ctx_create: %s", hipGetErrorString(e));
    }
    *out = c;
    return M2D_OK;
}

int m2d_ctx_destroy(m2d_ctx *c)
{
    if (!c) return M2D_OK;
    ctx_release(c);
    return M2D_OK;
}

}  // extern "C"

namespace m2d {

void ctx_retain(m2d_ctx *c) { c->refs++; }

void ctx_release(m2d_ctx *c)
{
    if (--c->refs > 0) return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (Scratch *s : {&c->acc, &c->red, &c->tmp, &c->sortbuf, &c->cq})
        if (s->ptr) (void)hipFree(s->ptr);
    free_tilings(c);
    for (auto &pr : c->ev_pending) c->ev_pool.insert(c->ev_pool.end(), {pr.first, pr.second});
    for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
    if (c->counters) (void)hipFree(c->counters);
    if (c->h_counters) (void)hipHostFree(c->h_counters);
    if (c->own) (void)hipStreamDestroy(c->own);
    delete c;
}

}  // namespace m2d

extern "C" {

int m2d_ctx_set_stream(m2d_ctx *c, void *stream)
{
    M2D_CHECK(c, M2D_ERR_BAD_ARG, "m2d_ctx_set_stream: ctx is NULL");
    M2D_HIP(hipStreamSynchronize(c->stream));
    c->stream = stream ? (hipStream_t)stream : c->own;
    return M2D_OK;
}

void *m2d_ctx_stream(m2d_ctx *c) { return c ? (void *)c->stream : nullptr; }
uint64_t m2d_ctx_seed(m2d_ctx *c) { return c ? c->seed : 0; }

int m2d_ctx_synchronize(m2d_ctx *c)
{
    M2D_CHECK(c, M2D_ERR_BAD_ARG, "m2d_ctx_synchronize: ctx is NULL");
    M2D_HIP(hipStreamSynchronize(c->stream));
    return M2D_OK;
}

int m2d_ctx_status(m2d_ctx *c, uint32_t *flags, int clear)
{
    M2D_CHECK(c && flags, M2D_ERR_BAD_ARG, "m2d_ctx_status: NULL argument");
    M2D_TRY(read_counters(c));
    *flags = (uint32_t)c->h_counters[CNT_FLAGS];
    if (clear) {
        M2D_TRY(zero_words(c, c->counters + CNT_FLAGS, 1));
        M2D_HIP(hipStreamSynchronize(c->stream));
    }
    return M2D_OK;
}

int m2d_malloc(m2d_ctx *c, size_t bytes, void **out)
{
    M2D_CHECK(c && out, M2D_ERR_BAD_ARG, "m2d_malloc: NULL argument");
    *out = nullptr;
    if (bytes == 0) bytes = 1;
    hipError_t e = hipMalloc(out, bytes);
    if (e != hipSuccess) {
        *out = nullptr;
        return set_error(M2D_ERR_NO_MEMORY, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    }
    return M2D_OK;
}

int m2d_free(m2d_ctx *c, void *p)
{
    M2D_CHECK(c, M2D_ERR_BAD_ARG, "m2d_free: ctx is NULL");
    if (!p) return M2D_OK;
    M2D_HIP(hipStreamSynchronize(c->stream));
    M2D_HIP(hipFree(p));
    return M2D_OK;
}

int m2d_memcpy_h2d(m2d_ctx *c, void *dst, const void *src, size_t bytes)
{
    M2D_CHECK(c && (bytes == 0 || (dst && src)), M2D_ERR_BAD_ARG, "m2d_memcpy_h2d: NULL argument");
    if (!bytes) return M2D_OK;
    M2D_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    M2D_HIP(hipStreamSynchronize(c->stream));
    return M2D_OK;
}

int m2d_memcpy_d2h(m2d_ctx *c, void *dst, const void *src, size_t bytes)
{
    M2D_CHECK(c && (bytes == 0 || (dst && src)), M2D_ERR_BAD_ARG, "m2d_memcpy_d2h: NULL argument");
    if (!bytes) return M2D_OK;
    M2D_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    M2D_HIP(hipStreamSynchronize(c->stream));
    return M2D_OK;
}

int m2d_memcpy_d2d(m2d_ctx *c, void *dst, const void *src, size_t bytes)
{
    M2D_CHECK(c && (bytes == 0 || (dst && src)), M2D_ERR_BAD_ARG, "m2d_memcpy_d2d: NULL argument");
    if (!bytes) return M2D_OK;
    M2D_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c->stream));
    return M2D_OK;
}

int m2d_memset(m2d_ctx *c, void *dst, int value, size_t bytes)
{
    M2D_CHECK(c && (bytes == 0 || dst), M2D_ERR_BAD_ARG, "m2d_memset: NULL argument");
    return fill_bytes(c, dst, value, bytes);
}

int m2d_field_add(m2d_ctx *c, double *dst, const double *src, double scale, size_t n)
{
    M2D_CHECK(c && (n == 0 || (dst && src)), M2D_ERR_BAD_ARG, "m2d_field_add: NULL argument");
    if (n == 0) return M2D_OK;
    const unsigned g = (unsigned)std::min<size_t>((n + 255) / 256, 4096);
    k_axpy<<<g, 256, 0, c->stream>>>(dst, src, scale, n);
    M2D_HIP(hipGetLastError());
    return M2D_OK;
}

// ------------------------------------------------------------------ t_grid

static int check_grid(const m2d_grid2d *g)
{
    M2D_CHECK(g, M2D_ERR_BAD_ARG, "grid is NULL");
    M2D_CHECK(g->nx >= 3 && g->nz >= 3, M2D_ERR_BAD_ARG, "grid %dx%d: need at least 3x3 nodes",
              g->nx, g->nz);
    M2D_CHECK(g->dx > 0 && g->dz > 0, M2D_ERR_BAD_ARG, "grid spacing must be positive");
    return M2D_OK;
}

int m2d_grid_square_electrode_2d(const m2d_grid2d *g, uint8_t *mask, double *voltage, double rmin,
                                 double rmax, double zmin, double zmax, double v, int type)
{
    M2D_TRY(check_grid(g));
    M2D_CHECK(mask && voltage, M2D_ERR_BAD_ARG, "square_electrode: NULL array");
    for (int i = 0; i < g->nx; i++)
        for (int j = 0; j < g->nz; j++) {
            double r = i * g->dx, z = j * g->dz;
            if (r > rmin && r < rmax && z > zmin && z < zmax) {
                mask[(size_t)i * g->nz + j] = (uint8_t)type;
                voltage[(size_t)i * g->nz + j] = v;
            }
        }
    return M2D_OK;
}

int m2d_grid_circle_electrode_2d(const m2d_grid2d *g, uint8_t *mask, double *voltage, double xc,
                                 double zc, double radius, double v, int type)
{
    M2D_TRY(check_grid(g));
    M2D_CHECK(mask && voltage, M2D_ERR_BAD_ARG, "circle_electrode: NULL array");
    double r2 = radius * radius;
    for (int i = 0; i < g->nx; i++)
        for (int j = 0; j < g->nz; j++) {
            double a = i * g->dx - xc, b = j * g->dz - zc;
            if (a * a + b * b <= r2) {
                mask[(size_t)i * g->nz + j] = (uint8_t)type;
                voltage[(size_t)i * g->nz + j] = v;
            }
        }
    return M2D_OK;
}

int m2d_grid_mark_boundary_2d(const m2d_grid2d *g, uint8_t *mask)
{
    M2D_TRY(check_grid(g));
    M2D_CHECK(mask, M2D_ERR_BAD_ARG, "mark_boundary: NULL mask");
    const int M = g->nx, N = g->nz;
    auto at = [&](int i, int j) -> uint8_t & { return mask[(size_t)i * N + j]; };
    for (int i = 2; i < M - 2; i++)
        for (int j = 2; j < N - 2; j++)
            if ((at(i - 1, j) == M2D_FIXED || at(i + 1, j) == M2D_FIXED || at(i, j - 1) == M2D_FIXED ||
                 at(i, j + 1) == M2D_FIXED) && at(i, j) != M2D_FIXED)
                at(i, j) = M2D_BOUNDARY;
    return M2D_OK;
}

// outer frame of the builders: `all4` fixes i == 0 as well (the MAC and
// Penning builders leave it free); `linear` sets -extern_field*dz*(j - N/2)
static void build_frame(const m2d_grid2d *g, uint8_t *mask, double *voltage, bool all4, bool linear,
                        double ef)
{
    const int M = g->nx, N = g->nz;
    for (int i = 0; i < M; i++)
        for (int j = 0; j < N; j++) {
            size_t k = (size_t)i * N + j;
            bool edge = i == M - 1 || j == 0 || j == N - 1 || (all4 && i == 0);
            if (edge) {
                mask[k] = M2D_FIXED;
                voltage[k] = linear ? -ef * g->dz * (j - N / 2) : 0.0;
            } else {
                mask[k] = M2D_FREE;
            }
        }
}

static void build_multipole(const m2d_grid2d *g, uint8_t *mask, double *voltage, int npoles,
                            double r_pole, double r_rod)
{
    for (int i = 0; i < npoles; i++) {
        double ang = 2 * M_PI * (i + 1.0 / 32) / npoles;
        double x = 1e-2 + sin(ang) * r_pole, y = 1e-2 + cos(ang) * r_pole;
        m2d_grid_circle_electrode_2d(g, mask, voltage, x, y, r_rod, i % 2 == 0 ? -1.0 : 1.0,
                                     M2D_FIXED_RF);
    }
}

int m2d_grid_build_2d(const m2d_grid2d *g, const m2d_geom2d *p, uint8_t *mask, double *voltage)
{
    M2D_TRY(check_grid(g));
    M2D_CHECK(p && mask && voltage, M2D_ERR_BAD_ARG, "grid_build: NULL argument");
    const int M = g->nx, N = g->nz;
    std::memset(voltage, 0, sizeof(double) * (size_t)M * N);
    auto sq = [&](double a, double b, double c, double d, double v) {
        m2d_grid_square_electrode_2d(g, mask, voltage, a, b, c, d, v, M2D_FIXED);
    };
    switch (p->geometry) {
    case M2D_GEO_EMPTY:
        build_frame(g, mask, voltage, true, true, p->extern_field);
        return M2D_OK;  // empty() marks no BOUNDARY nodes (fields.cpp:417-433)
    case M2D_GEO_PROBE:
        build_frame(g, mask, voltage, true, true, p->extern_field);
        m2d_grid_circle_electrode_2d(g, mask, voltage, (M - 1) * g->dx / 2, (N - 1) * g->dz / 2,
                                     p->probe_radius, p->u_probe, M2D_FIXED);
        break;
    case M2D_GEO_RF_22PT:
        build_frame(g, mask, voltage, true, false, 0.0);
        build_multipole(g, mask, voltage, 22, 0.75e-2, 0.05e-2);
        break;
    case M2D_GEO_RF_8PT:
        build_frame(g, mask, voltage, true, true, p->extern_field);
        build_multipole(g, mask, voltage, 8, 0.3e-2 + 0.1e-2, 0.1e-2);
        break;
    case M2D_GEO_RF_HAITRAP:
        build_frame(g, mask, voltage, true, true, p->extern_field);
        build_multipole(g, mask, voltage, 8, 0.3e-2 + 0.01e-2, 0.01e-2);
        break;
    case M2D_GEO_TUBE: {
        build_frame(g, mask, voltage, true, false, 0.0);
        double c = p->x_max / 2.0, r2 = p->probe_radius * p->probe_radius;
        for (int i = 0; i < M; i++)
            for (int j = 0; j < N; j++) {
                double a = i * g->dx - c, b = j * g->dz - c;
                if (a * a + b * b >= r2) {
                    mask[(size_t)i * N + j] = M2D_FIXED;
                    voltage[(size_t)i * N + j] = 0.0;
                }
            }
        break;
    }
    case M2D_GEO_RF_QUAD:
        build_frame(g, mask, voltage, true, false, 0.0);
        m2d_grid_circle_electrode_2d(g, mask, voltage, 5e-3, 1e-2, 2e-3, 1.0, M2D_FIXED_RF);
        m2d_grid_circle_electrode_2d(g, mask, voltage, 15e-3, 1e-2, 2e-3, 1.0, M2D_FIXED_RF);
        m2d_grid_circle_electrode_2d(g, mask, voltage, 1e-2, 5e-3, 2e-3, -1.0, M2D_FIXED_RF);
        m2d_grid_circle_electrode_2d(g, mask, voltage, 1e-2, 15e-3, 2e-3, -1.0, M2D_FIXED_RF);
        break;
    case M2D_GEO_MAC: {
        build_frame(g, mask, voltage, false, false, 0.0);
        double th = p->u_probe, ofs = 3e-2;
        sq(5e-3, 4.5e-2, 1e-2, 1.5e-2, 0.0);
        sq(5e-3, 7e-3, 2e-2, 8e-2, 0.0);
        sq(5e-3, 4.5e-2, 8.5e-2, 9e-2, 0.0);
        sq(3e-2, 3.3e-2, 11e-2 + ofs, 14e-2 + ofs, 0.8 * th);
        sq(4.5e-2, 4.8e-2, 15e-2 + ofs, 25e-2 + ofs, th);
        sq(3e-2, 3.3e-2, 26e-2 + ofs, 29e-2 + ofs, th);
        sq(2.5e-2, 2.8e-2, 29e-2 + ofs, 30.5e-2 + ofs, th);
        sq(15e-3, 4.5e-2, 35e-2, 35.3e-2, 0.0);
        sq(0.0, 4.5e-2, 39.5e-2, 40e-2, 3e3);
        break;
    }
    case M2D_GEO_PENNING:
        build_frame(g, mask, voltage, false, false, 0.0);
        sq(1.57e-2 / 2, 1.67e-2 / 2, 1e-3, 25e-3, -5);
        sq(0, 1.46e-2 / 2, 15e-3, 16e-3, 10);
        sq(0, 7e-3 / 2, 12e-3, 19e-3, 10);
        sq(0, 1.9e-3, 1e-3, 12e-3, 10);
        sq(4e-3, 7e-3, 52e-3, 53e-3, -5);
        sq(2.5e-3, 7e-3, 46e-3, 47e-3, 0);
        break;
    case M2D_GEO_PENNING_SIMPLE: {
        build_frame(g, mask, voltage, false, false, 0.0);
        const double ri = 1e-2, ro = 1.1e-2;
        sq(ri, ro, 0, 1e-2, -0.5);
        sq(ri, ro, 1.1e-2, 2e-2, 0);
        sq(ri, ro, 2.1e-2, 6e-2, -1.0);  // default trap_voltage (fields.hpp:41)
        sq(ri, ro, 6.1e-2, 7.5e-2, -10);
        break;
    }
    default:
        return set_error(M2D_ERR_BAD_ARG, "grid_build: unknown geometry %d", p->geometry);
    }
    return m2d_grid_mark_boundary_2d(g, mask);
}

int m2d_grid_is_free_2d(const m2d_grid2d *g, const uint8_t *mask, double r, double z)
{
    if (!g || !mask) return 0;
    if (!(r >= 0.0) || !(z >= 0.0)) return 0;
    // Param::idx = 1.0/dx (param.cpp:129-130), products as in fields.hpp:96-97
    int i = (int)(r * (1.0 / g->dx)), j = (int)(z * (1.0 / g->dz));
    if (i > g->nx - 2 || j > g->nz - 2) return 0;
    const int N = g->nz;
    return mask[(size_t)i * N + j] == M2D_FREE || mask[(size_t)(i + 1) * N + j] == M2D_FREE ||
           mask[(size_t)i * N + j + 1] == M2D_FREE || mask[(size_t)(i + 1) * N + j + 1] == M2D_FREE;
}

}  // extern "C"
