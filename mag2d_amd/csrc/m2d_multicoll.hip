// m2d_multicoll.hip -- the multicoll mover of the 2D species:
// Species<CARTESIAN>::advance_multicoll (particles.cpp:813-861) and the
// time_to_death draw of BaseSpecies::lifetime_init (particles.cpp:163-169).
//
// One thread per slot.  A particle's events inside a step form a sequential
// chain (each flight ends where the scatter draws the next velocity), so the
// chain stays in one thread, in FP64 registers, and the FP32 state is read
// and written once.  Every slot draws from its own Philox stream keyed by
// (slot, step), consumed in event order: scatter draws, then the next
// time_to_death, as the reference consumes its shared stream per particle.
#include "m2d_objects.hpp"
#include "m2d_scatter.hpp"

namespace m2d {

constexpr int kMaxEvents = 4096;        // events of one particle in one step
constexpr double kMaxEventRate = 256.0; // dt / lifetime accepted by the mover

struct MulticollArgs {
    float *x, *z, *vx, *vy, *vz;
    float *xl, *zl;          // position remainders (long-time positions), or null
    double *ttd;
    int64_t n;
    double fx, fz, qm;       // Fields::E(0, 0) (un-negated constant field) and charge / mass
    double dt, lifetime;
    const McDev *mcc;        // null: every event is a null collision
    uint64_t seed;
    uint32_t stream;
    uint64_t step;
    unsigned long long *counters;
};

// flight of duration t in the constant field, in the reference's order:
// kick, then x += (v + a t / 2) t with the kicked v (particles.cpp:841-844)
__device__ __forceinline__ void fly(const MulticollArgs &a, double t, double &x, double &z, double &vx,
                                    double &vz)
{
    vx += a.fx * a.qm * t;
    vz += a.fz * a.qm * t;
    x += (vx + 0.5 * a.fx * a.qm * t) * t;
    z += (vz + 0.5 * a.fz * a.qm * t) * t;
}

__global__ void __launch_bounds__(kBlock) k_multicoll2d(MulticollArgs a)
{
    const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    int events = 0;
    bool capped = false;
    if (p < a.n && !is_dead(a.x[p])) {
        double x = a.xl ? pos64(a.x[p], a.xl[p]) : a.x[p], z = a.zl ? pos64(a.z[p], a.zl[p]) : a.z[p];
        double vx = a.vx[p], vy = a.vy[p], vz = a.vz[p];
        double ttd = a.ttd[p];
        double local = 0.0;
        Rng rg;
        rg.begin(a.seed, a.stream, (uint32_t)p, a.step, kPurposeMulticoll);
        // particles.cpp:838-849
        while (local + ttd < a.dt) {
            if (events == kMaxEvents) {
                capped = true;
                break;
            }
            fly(a, ttd, x, z, vx, vz);
            local += ttd;
            if (a.mcc) scatter_d(*a.mcc, rg, vx, vy, vz, a.counters);
            ttd = a.lifetime * rg.exponential();
            events++;
        }
        // particles.cpp:850-855
        const double rest = a.dt - local;
        fly(a, rest, x, z, vx, vz);
        ttd -= rest;
        if (a.xl) {
            pos_split(x, a.x[p], a.xl[p]);
            pos_split(z, a.z[p], a.zl[p]);
        } else {
            a.x[p] = (float)x; a.z[p] = (float)z;
        }
        a.vx[p] = (float)vx; a.vy[p] = (float)vy; a.vz[p] = (float)vz;
        a.ttd[p] = ttd;
    }
    // events per wavefront into the collision counter, one atomic
    int sum = events;
    for (int o = 32; o > 0; o >>= 1) sum += __shfl_xor(sum, o);
    if (__lane_id() == 0 && sum) atomicAdd(&a.counters[CNT_COLL], (unsigned long long)sum);
    if (__ballot(capped) && __lane_id() == 0) raise_flag(a.counters, M2D_FLAG_EVENTS);
}

// BaseSpecies::lifetime_init (particles.cpp:163-169): time_to_death of every
// live particle
__global__ void __launch_bounds__(kBlock) k_ttd_init2d(const float *x, double *ttd, int64_t n, double lifetime,
                                                      uint64_t seed, uint32_t stream, uint64_t step)
{
    const int64_t p = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (p >= n || is_dead(x[p])) return;
    Rng rg;
    rg.begin(seed, stream, (uint32_t)p, step, kPurposeTtd);
    ttd[p] = lifetime * rg.exponential();
}

}  // namespace m2d

using namespace m2d;

extern "C" {

int m2d_advance_multicoll_2d(m2d_ctx *ctx, m2d_fields2d *f, const m2d_species2d *s, const m2d_particles2d *p,
                             double *time_to_death, const m2d_mcc *mcc, m2d_step_stats *stats)
{
    M2D_CHECK(ctx && f && s, M2D_ERR_BAD_ARG, "m2d_advance_multicoll_2d: NULL argument");
    M2D_CHECK(f->ctx == ctx, M2D_ERR_BAD_ARG, "m2d_advance_multicoll_2d: fields belong to another context");
    M2D_TRY(check_parts(p));
    M2D_CHECK(p->n == 0 || time_to_death, M2D_ERR_BAD_ARG, "m2d_advance_multicoll_2d: time_to_death is NULL");
    M2D_CHECK(s->mass > 0 && s->dt > 0, M2D_ERR_BAD_ARG, "m2d_advance_multicoll_2d: mass and dt must be > 0");
    M2D_CHECK(!mcc || mcc->ctx == ctx, M2D_ERR_BAD_ARG, "m2d_advance_multicoll_2d: mcc belongs to another context");
    M2D_CHECK(!mcc || !mcc->h.partner_update, M2D_ERR_UNSUPPORTED,
              "m2d_advance_multicoll_2d: COULOMB (partner update) runs in m2d_advance_2d only");
    M2D_CHECK(!mcc || !mcc_partners_alias(mcc, p->x), M2D_ERR_UNSUPPORTED,
              "m2d_advance_multicoll_2d: partners from the colliding species' own store run in "
              "m2d_advance_2d only");
    const m2d_fields2d_params &P = f->p;
    // the reference's own restrictions (particles.cpp:818-823)
    M2D_CHECK(!f->bt && P.Br == 0.0 && P.Bz == 0.0 && P.Bt == 0.0, M2D_ERR_UNSUPPORTED,
              "Species<CARTESIAN>::advance_multicoll() implemented for zero B only");
    M2D_CHECK(!P.selfconsistent && P.geometry_empty, M2D_ERR_UNSUPPORTED,
              "Species<CARTESIAN>::advance_multicoll() implemented for const extern field only: "
              "set selfconsistent=0 and geometry=EMPTY");
    M2D_CHECK(s->lifetime > 0 && s->dt / s->lifetime <= kMaxEventRate, M2D_ERR_BAD_ARG,
              "m2d_advance_multicoll_2d: dt / lifetime = %g (at most %g events per step expected)",
              s->dt / s->lifetime, kMaxEventRate);
    if (stats) {
        M2D_TRY(zero_words(ctx, ctx->counters + CNT_REMOVED, 2));
    }
    if (p->n > 0) {
        MulticollArgs a;
        a.x = p->x; a.z = p->z; a.vx = p->vx; a.vy = p->vy; a.vz = p->vz;
        a.xl = p->xl; a.zl = p->zl;
        a.ttd = time_to_death;
        a.n = p->n;
        // Fields::E(0, 0): EMPTY and not selfconsistent (fields.hpp:126-131)
        a.fx = 0.0;
        a.fz = P.extern_field;
        a.qm = s->charge / s->mass;
        a.dt = s->dt;
        a.lifetime = s->lifetime;
        a.mcc = mcc ? mcc->d : nullptr;
        a.seed = ctx->seed;
        a.stream = s->stream;
        a.step = s->niter;
        a.counters = ctx->counters;
        k_multicoll2d<<<grid_for(p->n, 1), kBlock, 0, ctx->stream>>>(a);
        M2D_HIP(hipGetLastError());
    }
    if (stats) {
        M2D_TRY(read_counters(ctx));
        stats->removed = 0;
        stats->spilled = 0;
        stats->collisions = (int64_t)ctx->h_counters[CNT_COLL];
    }
    return M2D_OK;
}

int m2d_time_to_death_init_2d(m2d_ctx *ctx, const m2d_species2d *s, const m2d_particles2d *p,
                              double *time_to_death)
{
    M2D_CHECK(ctx && s, M2D_ERR_BAD_ARG, "m2d_time_to_death_init_2d: NULL argument");
    M2D_TRY(check_parts(p));
    M2D_CHECK(p->n == 0 || time_to_death, M2D_ERR_BAD_ARG, "m2d_time_to_death_init_2d: time_to_death is NULL");
    M2D_CHECK(s->lifetime > 0, M2D_ERR_BAD_ARG, "m2d_time_to_death_init_2d: lifetime must be > 0");
    if (p->n > 0) {
        k_ttd_init2d<<<grid_for(p->n, 1), kBlock, 0, ctx->stream>>>(p->x, time_to_death, p->n, s->lifetime,
                                                                    ctx->seed, s->stream, s->niter);
        M2D_HIP(hipGetLastError());
    }
    return M2D_OK;
}

}  // extern "C"
