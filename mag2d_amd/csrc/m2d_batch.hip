// m2d_batch.hip -- many independent 2D trap runs in one persistent pass.
//
// The pressure / RF-amplitude sweeps of the reference (drive_22PT_pressure.py,
// drive_8PT_pressure.py) start one process per run.  Here the runs share the
// field geometry (u, uRF, B, mask, box) and the species constants, and differ
// in particle count, RF amplitude and collision tables; their stores are
// concatenated.  Without space charge every particle moves independently, so
// one launch advances each particle through a whole segment of steps in
// registers (Species<CARTESIAN>::advance, particles.hpp:342-349: gather,
// Boris, MCC trial and scatter, boundary), reading only the L2-resident
// gather tables.  The result is bit for bit the one of m2d_advance_2d
// called step by step on each run: same gather, same FP32 push, same random
// streams (slot within the run, the run's stream id); the scatter is applied
// in line, which gives the same state because it only changes the velocity
// after the drift, and neither the boundary test nor anything else of the
// step reads that velocity.
//
// The launch shape (block size, particles per thread) is picked by a tuner
// that times the candidates on a copy of the store; a call's segments are
// captured in one hipGraph and replayed.
#include "m2d_objects.hpp"
#include "m2d_scatter.hpp"
#include "m2d_gather2d.hpp"
#include <algorithm>
#include <cmath>
#include <new>
#include <vector>

namespace m2d {

struct BatchRun {
    int64_t off;             // first slot of the run
    double amp;              // rf_amplitude
    uint64_t trial_thr;      // trial word threshold, 0: no collisions
    const McDev *mcc;
    uint32_t stream;
    uint32_t pad;
};

// argument block: field / grid members named as gather_E expects them
struct BatchArgs {
    float *x, *z, *vx, *vy, *vz;
    int64_t n;
    float *xl, *zl;          // position remainders (long-time positions), or null
    const float2 *gx, *gz;
    const uint8_t *cellfree;
    int nx, nz;
    double idx, idz, xmin, zmin;
    float idxh, idxl, idzh, idzl;  // idx, idz as FP32 pairs (the origin-0 gather)
    int econst;
    float ecx, ecz;
    float fx0, fx1, fz0, fz1;
    double xper, zper;
    int periodic, checkfree;
    float qmdt2, tx, ty, tz, sx, sy, sz, dt;
    float dtl;               // dt - (double)dt as FP32
    int rotate;
    double U0;
    uint64_t seed;
    const BatchRun *runs;
    const int64_t *offsets;  // n_runs + 1
    int n_runs;
    unsigned long long *counters;
};

__device__ __forceinline__ int run_of(const int64_t *off, int n_runs, int64_t p)
{
    int lo = 0, hi = n_runs;  // off[lo] <= p < off[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (off[mid] <= p) lo = mid; else hi = mid;
    }
    return lo;
}

// one particle state carried across the steps of a segment
struct BP {
    float x, z, vx, vy, vz;
    float xl, zl;   // position remainders (HIPOS)
    bool live;
    int r;
    uint32_t slot;  // slot within its run
};

// segment of `nsteps` steps from step niter0 for PPT particles per thread;
// cosv[s] = cos of the RF phase at step niter0 + s.  FASTG: the fields and
// species are those for which m2d_advance_2d runs its specialised pass (B =
// 0, no wrap, no constant field, origin 0, free-cell test on), whose gather
// is the FP32 origin-0 one; the same gather here keeps the runs bit for bit.
// HIPOS: long-time positions (FP32 pairs), as in m2d_advance_2d.
template <int PPT, bool FASTG, bool HIPOS>
__global__ void __launch_bounds__(256) k_batch2d(BatchArgs a, uint64_t niter0, int nsteps, const double *cosv)
{
    const int64_t nthreads = (int64_t)gridDim.x * blockDim.x;
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    BP q[PPT];
    int n_removed = 0, n_coll = 0;
    bool nan_any = false;
#pragma unroll
    for (int k = 0; k < PPT; k++) {
        const int64_t p = gid + k * nthreads;
        q[k].live = false;
        if (p < a.n) {
            q[k].x = a.x[p];
            q[k].live = !is_dead(q[k].x);
            if (q[k].live) {
                q[k].z = a.z[p]; q[k].vx = a.vx[p]; q[k].vy = a.vy[p]; q[k].vz = a.vz[p];
                q[k].xl = HIPOS ? a.xl[p] : 0.f;
                q[k].zl = HIPOS ? a.zl[p] : 0.f;
                q[k].r = run_of(a.offsets, a.n_runs, p);
                q[k].slot = (uint32_t)(p - a.runs[q[k].r].off);
            }
        }
    }
    for (int s = 0; s < nsteps; s++) {
        const uint64_t step = niter0 + (uint64_t)s;
        const double cs = cosv[s];
#pragma unroll
        for (int k = 0; k < PPT; k++) {
            if (!q[k].live) continue;
            const BatchRun &R = a.runs[q[k].r];
            // the per-step phase of m2d_advance_2d, rounded the same way (no contraction)
            const float ph = (float)__dadd_rn(__dmul_rn(R.amp, cs), a.U0);
            float x = q[k].x, z = q[k].z, vx = q[k].vx, vy = q[k].vy, vz = q[k].vz;
            float xl = q[k].xl, zl = q[k].zl;
            float ex, ez;
            if constexpr (FASTG) {
                if (gather_E_f(a, x, xl, z, zl, ph, ex, ez)) raise_flag(a.counters, M2D_FLAG_OUT_OF_RANGE);
            } else {
                gather_E(a, HIPOS ? pos64(x, xl) : (double)x, HIPOS ? pos64(z, zl) : (double)z, ph, ex, ez);
            }
            vx = fmaf(ex, a.qmdt2, vx);
            vz = fmaf(ez, a.qmdt2, vz);
            if (a.rotate) {
                float vpr = vx - vy * a.tz + vz * a.ty;
                float vpz = vz - vx * a.ty + vy * a.tx;
                float vpt = vy - vz * a.tx + vx * a.tz;
                vx = vx - vpt * a.sz + vpz * a.sy;
                vy = vy - vpz * a.sx + vpr * a.sz;
                vz = vz - vpr * a.sy + vpt * a.sx;
            }
            vx = fmaf(ex, a.qmdt2, vx);
            vz = fmaf(ez, a.qmdt2, vz);
            if constexpr (HIPOS) {
                drift_pair(x, xl, vx, a.dt, a.dtl);
                drift_pair(z, zl, vz, a.dt, a.dtl);
            } else {
                x = fmaf(vx, a.dt, x);
                z = fmaf(vz, a.dt, z);
            }
            if (R.trial_thr) {
                // the trial word of slot s: word s%4 of the Philox block (s/4, step, 0, purpose 2)
                uint32_t c0 = q[k].slot >> 2, c1 = (uint32_t)step, c2 = 0u;
                uint32_t c3 = (kPurposeTrial << 24) | ((uint32_t)(step >> 32) & 0xFFFFFFu);
                philox4x32_10(c0, c1, c2, c3, (uint32_t)a.seed, (uint32_t)(a.seed >> 32) ^ R.stream);
                const uint32_t w = (q[k].slot & 3) == 0 ? c0 : (q[k].slot & 3) == 1 ? c1 : (q[k].slot & 3) == 2 ? c2 : c3;
                if ((uint64_t)w < R.trial_thr) {
                    Rng rg;
                    rg.begin(a.seed, R.stream, q[k].slot, step, kPurposeMcc);
                    scatter(*R.mcc, rg, vx, vy, vz, a.counters);
                    n_coll++;
                }
            }
            // advance_boundary (particles.hpp:376-406), no deposit
            bool dead = false;
            const double xd = HIPOS ? pos64(x, xl) : (double)x, zd = HIPOS ? pos64(z, zl) : (double)z;
            const bool inside = HIPOS ? xd >= a.xmin && xd <= a.xper && zd >= a.zmin && zd <= a.zper
                                      : x >= a.fx0 && x <= a.fx1 && z >= a.fz0 && z <= a.fz1;
            if (!inside) {
                if (x != x || z != z) { dead = true; nan_any = true; }
                else if (!a.periodic) dead = true;
                else {
                    double xm = fmod(xd, a.xper), zm = fmod(zd, a.zper);
                    if (xm < 0) xm += a.xper;
                    if (zm < 0) zm += a.zper;
                    if constexpr (HIPOS) {
                        pos_split(xm, x, xl);
                        pos_split(zm, z, zl);
                    } else {
                        x = (float)xm; z = (float)zm;
                    }
                }
            }
            if (!dead && a.checkfree) {
                bool fr = false;
                if constexpr (FASTG) {
                    // the cell of the specialised single-run pass (the box test gave x, z >= 0)
                    int ci, cj;
                    float u, ou;
                    cell_f(x, xl, a.idxh, a.idxl, ci, u, ou);
                    cell_f(z, zl, a.idzh, a.idzl, cj, u, ou);
                    if ((unsigned)ci <= (unsigned)(a.nx - 1) && (unsigned)cj <= (unsigned)(a.nz - 1))
                        fr = a.cellfree[ci * a.nz + cj] != 0;
                } else {
                    const double X = ((HIPOS ? pos64(x, xl) : (double)x) - a.xmin) * a.idx;
                    const double Z = ((HIPOS ? pos64(z, zl) : (double)z) - a.zmin) * a.idz;
                    if (X >= 0.0 && X < (double)a.nx && Z >= 0.0 && Z < (double)a.nz)
                        fr = a.cellfree[(int)X * a.nz + (int)Z] != 0;
                }
                dead = !fr;
            }
            if (dead) { q[k].live = false; n_removed++; }
            q[k].x = x; q[k].z = z; q[k].vx = vx; q[k].vy = vy; q[k].vz = vz;
            q[k].xl = xl; q[k].zl = zl;
        }
    }
#pragma unroll
    for (int k = 0; k < PPT; k++) {
        const int64_t p = gid + k * nthreads;
        if (p >= a.n || is_dead(a.x[p])) continue;
        if (!q[k].live) { a.x[p] = dead_value(); continue; }
        a.x[p] = q[k].x; a.z[p] = q[k].z; a.vx[p] = q[k].vx; a.vy[p] = q[k].vy; a.vz[p] = q[k].vz;
        if (HIPOS) { a.xl[p] = q[k].xl; a.zl[p] = q[k].zl; }
    }
    int sum = n_removed;
    for (int o = 32; o > 0; o >>= 1) sum += __shfl_xor(sum, o);
    if (__lane_id() == 0 && sum) atomicAdd(&a.counters[CNT_REMOVED], (unsigned long long)sum);
    int c = n_coll;
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
    if (__lane_id() == 0 && c) atomicAdd(&a.counters[CNT_COLL], (unsigned long long)c);
    if (__ballot(nan_any) && __lane_id() == 0) raise_flag(a.counters, M2D_FLAG_NAN);
}

// per-run sums of the kinetic energy (J) and live counts; runs are contiguous
// slot ranges, so the lanes of a wavefront share a run and are summed first
__global__ void __launch_bounds__(256) k_batch_energy(const float *x, const float *vx, const float *vy,
                                                      const float *vz, int64_t n, const int64_t *offsets,
                                                      int n_runs, double half_m, double *esum,
                                                      unsigned long long *count)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double e = 0.0;
    int live = 0, r = -1;
    if (p < n) {
        r = run_of(offsets, n_runs, p);
        if (!is_dead(x[p])) {
            const double a = vx[p], b = vy[p], c = vz[p];
            e = half_m * (a * a + b * b + c * c);
            live = 1;
        }
    }
    // segmented sum keyed by run over the wavefront, tails add
    const int lane = __lane_id();
    const int key = r >= 0 ? r : -1 - lane;
    const int prev = __shfl_up(key, 1);
    const unsigned long long heads = __ballot(lane == 0 || prev != key);
    const int seg = 63 - __clzll(heads & ((2ull << lane) - 1ull));
    for (int off = 1; off < 64; off <<= 1) {
        const double te = __shfl_up(e, off);
        const int tl = __shfl_up(live, off);
        if (lane - off >= seg) { e += te; live += tl; }
    }
    const bool tail = lane == 63 || ((heads >> (lane + 1)) & 1ull);
    if (r >= 0 && tail) {
        if (e != 0.0) atomicAdd(esum + r, e);
        if (live) atomicAdd(count + r, (unsigned long long)live);
    }
}

__global__ void k_batch_zero(double *e, unsigned long long *c, int n)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) { e[k] = 0.0; c[k] = 0ull; }
}

}  // namespace m2d

struct m2d_batch {
    m2d_ctx *ctx = nullptr;
    m2d_fields2d *f = nullptr;
    m2d_species2d s{};
    m2d_particles2d p{};
    int n_runs = 0;
    m2d::BatchRun *d_runs = nullptr;
    int64_t *d_off = nullptr;
    std::vector<int64_t> off;
    double *cosv = nullptr;
    size_t cosv_n = 0;
    double *esum = nullptr;
    unsigned long long *ecount = nullptr;
    m2d_batch_shape shape{0, 0, 0};
    bool tuned = false;
    bool retained = false;
};

namespace m2d {

static float f32_ge(double v)
{
    float f = (float)v;
    if ((double)f < v) f = nextafterf(f, INFINITY);
    return f;
}
static float f32_le(double v)
{
    float f = (float)v;
    if ((double)f > v) f = nextafterf(f, -INFINITY);
    return f;
}

static void batch_args(const m2d_batch *b, const m2d_particles2d &p, BatchArgs &a)
{
    const m2d_fields2d *f = b->f;
    const m2d_fields2d_params &P = f->p;
    const m2d_species2d *s = &b->s;
    a.x = p.x; a.z = p.z; a.vx = p.vx; a.vy = p.vy; a.vz = p.vz; a.n = p.n;
    a.xl = p.xl; a.zl = p.zl;
    a.gx = f->gx; a.gz = f->gz; a.cellfree = f->cellfree;
    a.nx = f->nx; a.nz = f->nz; a.idx = f->idx; a.idz = f->idz;
    split_f32(f->idx, a.idxh, a.idxl);
    split_f32(f->idz, a.idzh, a.idzl);
    a.xmin = P.grid.xmin; a.zmin = P.grid.zmin;
    a.econst = P.geometry_empty && !P.selfconsistent;
    a.ecx = 0.f;
    a.ecz = (float)P.extern_field;
    a.fx0 = f32_ge(P.grid.xmin); a.fx1 = f32_le(P.x_max);
    a.fz0 = f32_ge(P.grid.zmin); a.fz1 = f32_le(P.z_max);
    a.xper = P.x_max; a.zper = P.z_max;
    a.periodic = P.boundary == M2D_BC_PERIODIC;
    a.checkfree = !P.electric_field_from_file;
    const double tmp = s->charge * s->dt / (2.0 * s->mass);
    const double tx = P.Br * tmp, ty = P.Bt * tmp, tz = P.Bz * tmp;
    const double sc = 2.0 / (1 + tx * tx + ty * ty + tz * tz);
    a.tx = (float)tx; a.ty = (float)ty; a.tz = (float)tz;
    a.sx = (float)(tx * sc); a.sy = (float)(ty * sc); a.sz = (float)(tz * sc);
    a.rotate = tx != 0.0 || ty != 0.0 || tz != 0.0;
    a.qmdt2 = (float)(s->charge / s->mass * s->dt / 2.0);
    a.dt = (float)s->dt;
    a.dtl = (float)(s->dt - (double)a.dt);
    a.U0 = P.rf_U0;
    a.seed = b->ctx->seed;
    a.runs = b->d_runs;
    a.offsets = b->d_off;
    a.n_runs = b->n_runs;
    a.counters = b->ctx->counters;
}

static int ensure_cos(m2d_batch *b, uint64_t niter0, int steps)
{
    if ((size_t)steps > b->cosv_n) {
        if (b->cosv) (void)hipFree(b->cosv);
        b->cosv = nullptr;
        b->cosv_n = 0;
        M2D_HIP(hipMalloc((void **)&b->cosv, sizeof(double) * (size_t)steps));
        b->cosv_n = (size_t)steps;
    }
    std::vector<double> c((size_t)steps);
    const m2d_fields2d_params &P = b->f->p;
    for (int s = 0; s < steps; s++) c[(size_t)s] = P.rf ? rf_cos(niter0 + (uint64_t)s, b->s.dt, P.rf_omega) : 0.0;
    M2D_HIP(hipMemcpyAsync(b->cosv, c.data(), sizeof(double) * (size_t)steps, hipMemcpyHostToDevice, b->ctx->stream));
    M2D_HIP(hipStreamSynchronize(b->ctx->stream));
    return M2D_OK;
}

template <bool HIPOS>
static void launch_batch(const BatchArgs &a, const m2d_batch_shape &sh, unsigned grid, bool fastg, uint64_t niter0,
                         int nsteps, const double *cosv, hipStream_t st)
{
    if (fastg) {
        switch (sh.ppt) {
        case 1: k_batch2d<1, true, HIPOS><<<grid, sh.block, 0, st>>>(a, niter0, nsteps, cosv); break;
        case 2: k_batch2d<2, true, HIPOS><<<grid, sh.block, 0, st>>>(a, niter0, nsteps, cosv); break;
        default: k_batch2d<4, true, HIPOS><<<grid, sh.block, 0, st>>>(a, niter0, nsteps, cosv); break;
        }
    } else {
        switch (sh.ppt) {
        case 1: k_batch2d<1, false, HIPOS><<<grid, sh.block, 0, st>>>(a, niter0, nsteps, cosv); break;
        case 2: k_batch2d<2, false, HIPOS><<<grid, sh.block, 0, st>>>(a, niter0, nsteps, cosv); break;
        default: k_batch2d<4, false, HIPOS><<<grid, sh.block, 0, st>>>(a, niter0, nsteps, cosv); break;
        }
    }
}

static int launch_segment(m2d_batch *b, const m2d_particles2d &p, const m2d_batch_shape &sh, uint64_t niter0,
                          int nsteps, const double *cosv)
{
    BatchArgs a;
    batch_args(b, p, a);
    const int64_t threads = (p.n + sh.ppt - 1) / sh.ppt;
    const unsigned grid = (unsigned)std::max<int64_t>(1, (threads + sh.block - 1) / sh.block);
    hipStream_t st = b->ctx->stream;
    const bool fastg = a.checkfree && !a.rotate && !a.periodic && !a.econst && a.xmin == 0.0 && a.zmin == 0.0;
    if (a.xl) launch_batch<true>(a, sh, grid, fastg, niter0, nsteps, cosv, st);
    else launch_batch<false>(a, sh, grid, fastg, niter0, nsteps, cosv, st);
    M2D_HIP(hipGetLastError());
    return M2D_OK;
}

static int check_shape(const m2d_batch_shape *sh)
{
    M2D_CHECK(sh->block == 64 || sh->block == 128 || sh->block == 256, M2D_ERR_BAD_ARG,
              "batch shape: block %d (64, 128 or 256)", sh->block);
    M2D_CHECK(sh->ppt == 1 || sh->ppt == 2 || sh->ppt == 4, M2D_ERR_BAD_ARG, "batch shape: %d particles per thread",
              sh->ppt);
    M2D_CHECK(sh->seg_steps >= 1, M2D_ERR_BAD_ARG, "batch shape: steps per launch %d", sh->seg_steps);
    return M2D_OK;
}

}  // namespace m2d

using namespace m2d;

extern "C" {

int m2d_batch_create(m2d_ctx *ctx, m2d_fields2d *f, const m2d_species2d *s, const m2d_particles2d *p,
                     const m2d_batch_desc *d, m2d_batch **out)
{
    M2D_CHECK(ctx && f && s && p && d && out && d->offset && d->rf_amplitude, M2D_ERR_BAD_ARG,
              "m2d_batch_create: NULL argument");
    *out = nullptr;
    M2D_CHECK(f->ctx == ctx, M2D_ERR_BAD_ARG, "m2d_batch_create: fields belong to another context");
    M2D_CHECK(!f->bt, M2D_ERR_UNSUPPORTED, "m2d_batch_create: gridded B is not supported by the batched pass");
    M2D_CHECK(!f->p.selfconsistent, M2D_ERR_UNSUPPORTED,
              "m2d_batch_create: runs of a batch carry no space charge (selfconsistent = 0)");
    M2D_CHECK(d->n_runs >= 1, M2D_ERR_BAD_ARG, "m2d_batch_create: no runs");
    M2D_CHECK(s->mass > 0 && s->dt > 0, M2D_ERR_BAD_ARG, "m2d_batch_create: mass and dt must be > 0");
    M2D_CHECK(d->offset[0] == 0 && d->offset[d->n_runs] == p->n, M2D_ERR_BAD_ARG,
              "m2d_batch_create: offsets must run from 0 to the store's %lld slots", (long long)p->n);
    for (int r = 0; r < d->n_runs; r++)
        M2D_CHECK(d->offset[r + 1] >= d->offset[r], M2D_ERR_BAD_ARG, "m2d_batch_create: offsets decrease at %d", r);
    M2D_CHECK(p->n < (1ll << 31) && (p->n == 0 || (p->x && p->z && p->vx && p->vy && p->vz)), M2D_ERR_BAD_ARG,
              "m2d_batch_create: particle store");
    M2D_CHECK((p->xl == nullptr) == (p->zl == nullptr), M2D_ERR_BAD_ARG,
              "m2d_batch_create: position remainders xl and zl must both be set or both be NULL");
    m2d_batch *b = new (std::nothrow) m2d_batch();
    M2D_CHECK(b, M2D_ERR_NO_MEMORY, "m2d_batch_create: host allocation");
    b->ctx = ctx;
    b->f = f;
    b->s = *s;
    b->p = *p;
    b->n_runs = d->n_runs;
    b->off.assign(d->offset, d->offset + d->n_runs + 1);
    std::vector<BatchRun> runs((size_t)d->n_runs);
    for (int r = 0; r < d->n_runs; r++) {
        BatchRun &R = runs[(size_t)r];
        R.off = d->offset[r];
        R.amp = f->p.rf ? d->rf_amplitude[r] : 0.0;
        const m2d_mcc *m = d->mcc ? d->mcc[r] : nullptr;
        if (m && m->h.partner_update) {
            m2d_batch_destroy(b);
            return set_error(M2D_ERR_UNSUPPORTED, "m2d_batch_create: COULOMB (partner update) runs in "
                             "m2d_advance_2d only");
        }
        if (m && mcc_partners_alias(m, p->x)) {
            m2d_batch_destroy(b);
            return set_error(M2D_ERR_UNSUPPORTED, "m2d_batch_create: partners from the batch's own store "
                             "run in m2d_advance_2d only");
        }
        R.mcc = m ? m->d : nullptr;
        R.trial_thr = 0;
        if (m && std::isfinite(m->h.lifetime) && m->h.lifetime > 0) {
            const double prob = 1.0 - exp(-s->dt / m->h.lifetime);  // particles.cpp:937
            R.trial_thr = prob >= 1.0 ? (1ull << 32) : (uint64_t)ceil(prob * 4294967296.0 - 0.5);
        }
        R.stream = d->stream ? d->stream[r] : (uint32_t)r;
        R.pad = 0;
    }
    auto fail = [&](int code) { m2d_batch_destroy(b); return code; };
    if (hipMalloc((void **)&b->d_runs, sizeof(BatchRun) * runs.size()) != hipSuccess ||
        hipMalloc((void **)&b->d_off, sizeof(int64_t) * b->off.size()) != hipSuccess ||
        hipMalloc((void **)&b->esum, sizeof(double) * (size_t)d->n_runs) != hipSuccess ||
        hipMalloc((void **)&b->ecount, sizeof(unsigned long long) * (size_t)d->n_runs) != hipSuccess)
        return fail(set_error(M2D_ERR_NO_MEMORY, "m2d_batch_create: device allocation"));
    if (hipMemcpy(b->d_runs, runs.data(), sizeof(BatchRun) * runs.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(b->d_off, b->off.data(), sizeof(int64_t) * b->off.size(), hipMemcpyHostToDevice) != hipSuccess)
        return fail(set_error(M2D_ERR_HIP, "m2d_batch_create: upload"));
    ctx_retain(ctx);
    b->retained = true;
    b->shape = m2d_batch_shape{256, 2, 500};
    *out = b;
    return M2D_OK;
}

int m2d_batch_destroy(m2d_batch *b)
{
    if (!b) return M2D_OK;
    if (b->ctx) (void)hipStreamSynchronize(b->ctx->stream);
    for (void *q : {(void *)b->d_runs, (void *)b->d_off, (void *)b->cosv, (void *)b->esum, (void *)b->ecount})
        if (q) (void)hipFree(q);
    if (b->retained) ctx_release(b->ctx);
    delete b;
    return M2D_OK;
}

int m2d_batch_set_shape(m2d_batch *b, const m2d_batch_shape *sh)
{
    M2D_CHECK(b && sh, M2D_ERR_BAD_ARG, "m2d_batch_set_shape: NULL argument");
    M2D_TRY(check_shape(sh));
    b->shape = *sh;
    b->tuned = true;
    return M2D_OK;
}

int m2d_batch_tune(m2d_batch *b, uint64_t niter0, int32_t probe_steps, int32_t seg_steps, m2d_batch_shape *best)
{
    M2D_CHECK(b, M2D_ERR_BAD_ARG, "m2d_batch_tune: NULL argument");
    M2D_CHECK(probe_steps >= 1 && seg_steps >= 1, M2D_ERR_BAD_ARG, "m2d_batch_tune: steps");
    m2d_ctx *ctx = b->ctx;
    hipStream_t st = ctx->stream;
    const size_t np = (size_t)b->p.n;
    // candidates run on a copy of the store: the batch's own state is not touched
    const int nc = b->p.xl ? 7 : 5;  // state components (with position remainders: 7)
    const size_t stride = (std::max<size_t>(np, 1) + 3) / 4 * 4;  // 16-byte aligned arrays
    M2D_TRY(ensure(ctx, ctx->tmp, sizeof(float) * (size_t)nc * stride + 256));
    float *t = (float *)ctx->tmp.ptr;
    m2d_particles2d copy{t, t + stride, t + 2 * stride, t + 3 * stride, t + 4 * stride, b->p.n};
    if (b->p.xl) { copy.xl = t + 5 * stride; copy.zl = t + 6 * stride; }
    float *const src[7] = {b->p.x, b->p.z, b->p.vx, b->p.vy, b->p.vz, b->p.xl, b->p.zl};
    float *const dst[7] = {copy.x, copy.z, copy.vx, copy.vy, copy.vz, copy.xl, copy.zl};
    M2D_TRY(ensure_cos(b, niter0, probe_steps));
    hipEvent_t e0, e1;
    M2D_HIP(hipEventCreate(&e0));
    M2D_HIP(hipEventCreate(&e1));
    float best_ms = INFINITY;
    m2d_batch_shape pick = b->shape;
    for (int block : {64, 128, 256})
        for (int ppt : {1, 2, 4}) {
            const m2d_batch_shape sh{block, ppt, seg_steps};
            for (int c = 0; c < nc; c++)
                M2D_HIP(hipMemcpyAsync(dst[c], src[c], sizeof(float) * np, hipMemcpyDeviceToDevice, st));
            M2D_HIP(hipEventRecord(e0, st));
            M2D_TRY(launch_segment(b, copy, sh, niter0, probe_steps, b->cosv));
            M2D_HIP(hipEventRecord(e1, st));
            M2D_HIP(hipEventSynchronize(e1));
            float ms = 0.f;
            M2D_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (ms < best_ms) { best_ms = ms; pick = sh; }
        }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    b->shape = pick;
    b->tuned = true;
    if (best) *best = pick;
    return M2D_OK;
}

int m2d_batch_advance(m2d_batch *b, uint64_t niter0, int32_t steps)
{
    M2D_CHECK(b, M2D_ERR_BAD_ARG, "m2d_batch_advance: NULL argument");
    M2D_CHECK(steps >= 0, M2D_ERR_BAD_ARG, "m2d_batch_advance: steps %d", steps);
    if (steps == 0 || b->p.n == 0) return M2D_OK;
    M2D_TRY(ensure_cos(b, niter0, steps));
    hipStream_t st = b->ctx->stream;
    // the call's segments as one graph: kernel nodes only
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    M2D_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    int rc = M2D_OK;
    for (int s0 = 0; s0 < steps && rc == M2D_OK; s0 += b->shape.seg_steps) {
        const int ns = std::min(b->shape.seg_steps, steps - s0);
        rc = launch_segment(b, b->p, b->shape, niter0 + (uint64_t)s0, ns, b->cosv + s0);
    }
    const hipError_t ec = hipStreamEndCapture(st, &graph);
    if (rc != M2D_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    M2D_CHECK(ec == hipSuccess, M2D_ERR_HIP, "m2d_batch_advance: capture: %s", hipGetErrorString(ec));
    hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (ei != hipSuccess) {
        (void)hipGraphDestroy(graph);
        return set_error(M2D_ERR_HIP, "m2d_batch_advance: instantiate: %s", hipGetErrorString(ei));
    }
    const hipError_t el = hipGraphLaunch(exec, st);
    const hipError_t es = hipStreamSynchronize(st);
    (void)hipGraphExecDestroy(exec);
    (void)hipGraphDestroy(graph);
    M2D_CHECK(el == hipSuccess && es == hipSuccess, M2D_ERR_HIP, "m2d_batch_advance: graph launch: %s",
              hipGetErrorString(el != hipSuccess ? el : es));
    return M2D_OK;
}

int m2d_batch_energy(m2d_batch *b, double *mean_eV, int64_t *n_live)
{
    M2D_CHECK(b && mean_eV, M2D_ERR_BAD_ARG, "m2d_batch_energy: NULL argument");
    hipStream_t st = b->ctx->stream;
    const int R = b->n_runs;
    k_batch_zero<<<(R + 255) / 256, 256, 0, st>>>(b->esum, b->ecount, R);
    if (b->p.n > 0)
        k_batch_energy<<<(unsigned)((b->p.n + 255) / 256), 256, 0, st>>>(
            b->p.x, b->p.vx, b->p.vy, b->p.vz, b->p.n, b->d_off, R, 0.5 * b->s.mass / kQE, b->esum, b->ecount);
    M2D_HIP(hipGetLastError());
    std::vector<double> e((size_t)R);
    std::vector<unsigned long long> c((size_t)R);
    M2D_HIP(hipMemcpyAsync(e.data(), b->esum, sizeof(double) * (size_t)R, hipMemcpyDeviceToHost, st));
    M2D_HIP(hipMemcpyAsync(c.data(), b->ecount, sizeof(unsigned long long) * (size_t)R, hipMemcpyDeviceToHost, st));
    M2D_HIP(hipStreamSynchronize(st));
    for (int r = 0; r < R; r++) {
        mean_eV[r] = c[(size_t)r] ? e[(size_t)r] / (double)c[(size_t)r] : 0.0;
        if (n_live) n_live[r] = (int64_t)c[(size_t)r];
    }
    return M2D_OK;
}

}  // extern "C"
