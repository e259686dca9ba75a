/*
 * mag2d_amd.h -- C ABI of the MI355X-native PIC-MCC inner loop of mag2d.
 *
 * Plain C: opaque handles, plain pointers and sizes, int status codes.  No
 * call throws.  Every entry point below names the reference interface it
 * replaces (rouckas/mag2d, file:line).  The C++ host layer
 * (mag2d_amd/host/mag2d.hpp) rebuilds the reference class API
 * (Species / Field2D / Fields / t_grid) on top of these calls.
 *
 * Conventions
 *  - Units are SI: positions in m, velocities in m/s, charge density in C per
 *    node (Field2D::accumulate), potentials in V.
 *  - Grids are row-major double arrays data[i*nz + j], z fastest, like
 *    Array2D (Array.hpp:7-23).
 *  - Particles are structure-of-arrays FP32 (x, z, vx, vy, vz); an empty slot
 *    has x == M2D_DEAD (a NaN payload), the counterpart of t_particle::empty
 *    (particles.hpp:32).
 *  - All device work is queued on the context's stream; calls that return a
 *    host value synchronise that stream.  One host thread per context.
 *  - Device pointers in structs are device memory from m2d_malloc (or any
 *    allocation of the same device, e.g. a torch tensor's data_ptr()).
 */
#ifndef MAG2D_AMD_H
#define MAG2D_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define M2D_ABI_VERSION 2

/* ---- status codes (the reference throws std::runtime_error instead) ---- */
#define M2D_OK                 0
#define M2D_ERR_BAD_ARG       -1
#define M2D_ERR_OUT_OF_RANGE  -2   /* Field2D::accumulate "outside of range" (Field2D.hpp:55-56) */
#define M2D_ERR_NAN           -3
#define M2D_ERR_HIP           -4
#define M2D_ERR_NOT_CONVERGED -5
#define M2D_ERR_UNSUPPORTED   -6
#define M2D_ERR_NO_MEMORY     -7

/* device status flags accumulated by kernels (m2d_ctx_status) */
#define M2D_FLAG_OUT_OF_RANGE  1u  /* a gather/deposit index left the grid (value clamped / skipped) */
#define M2D_FLAG_NAN           2u  /* a live particle had a NaN position and was removed */
#define M2D_FLAG_PARTNER       4u  /* MCC found no live partner particle in the retry budget */
#define M2D_FLAG_EVENTS        8u  /* the multicoll mover hit its per-particle event budget in a step */

/* mask values (fields.hpp:20) */
#define M2D_FIXED    0
#define M2D_FIXED_RF 1
#define M2D_FREE     2
#define M2D_BOUNDARY 3

/* empty-slot marker stored in x (bit pattern of a quiet NaN) */
#define M2D_DEAD_BITS 0xFFC0DEADu

typedef struct m2d_ctx m2d_ctx;
typedef struct m2d_fields2d m2d_fields2d;
typedef struct m2d_mcc m2d_mcc;

/* ======================= context, errors, memory ======================== */

/* Message of the last failing call on this thread ("" if none). */
const char *m2d_last_error(void);
int m2d_abi_version(void);

/* Opens `device` and creates a stream of default priority; `seed` keys the
 * counter-based RNG (replaces the shared t_random of pic.cpp:124). */
int m2d_ctx_create(int device, uint64_t seed, m2d_ctx **out);
int m2d_ctx_destroy(m2d_ctx *ctx);
/* Queue all later work on `stream` (a hipStream_t of the same device, e.g.
 * torch.cuda.current_stream().cuda_stream); NULL restores the own stream. */
int m2d_ctx_set_stream(m2d_ctx *ctx, void *stream);
void *m2d_ctx_stream(m2d_ctx *ctx);
uint64_t m2d_ctx_seed(m2d_ctx *ctx);
int m2d_ctx_synchronize(m2d_ctx *ctx);
/* OR of M2D_FLAG_* raised since the last clear; synchronises. */
int m2d_ctx_status(m2d_ctx *ctx, uint32_t *flags, int clear);
/* Kernel timing (the counterpart of t_timer around Pic::advance, pic.cpp:332,355):
 * while enabled, every launch of the particle pass is bracketed by HIP events
 * on the context's stream.  _read synchronises, returns the summed kernel
 * time in ms and the launch count, and clears them. */
int m2d_ctx_profile(m2d_ctx *ctx, int enable);
int m2d_ctx_profile_read(m2d_ctx *ctx, double *ms, int64_t *launches);
/* Which compile-time form of the 2D particle pass the last m2d_advance_2d
 * launched (0 if none): an OR of the bits below.  A diagnostic for tests
 * that must reach a given form; no counterpart in the reference. */
#define M2D_PASS_DEPOSIT 1u   /* deposit fused */
#define M2D_PASS_MCC     2u   /* MCC trial and collision queue */
#define M2D_PASS_TILED   4u   /* sorted store: per-tile LDS deposit windows */
#define M2D_PASS_FAST    8u   /* B = 0, no wrap, origin 0, box and free-cell tests */
#define M2D_PASS_BGRID  16u   /* B gathered from a B grid per particle */
#define M2D_PASS_PHT    32u   /* gather from the phase-applied tables */
#define M2D_PASS_HIPOS  64u   /* long-time positions (FP32 pairs) */
int m2d_ctx_last_pass(m2d_ctx *ctx, uint32_t *variant);

int m2d_malloc(m2d_ctx *ctx, size_t bytes, void **out);
int m2d_free(m2d_ctx *ctx, void *p);
int m2d_memcpy_h2d(m2d_ctx *ctx, void *dst, const void *src, size_t bytes);   /* synchronous */
int m2d_memcpy_d2h(m2d_ctx *ctx, void *dst, const void *src, size_t bytes);   /* synchronous */
int m2d_memcpy_d2d(m2d_ctx *ctx, void *dst, const void *src, size_t bytes);   /* stream-ordered */
/* Fill with a byte value by a kernel (no hipMemsetAsync: safe under capture). */
int m2d_memset(m2d_ctx *ctx, void *dst, int value, size_t bytes);
/* dst[k] += scale * src[k] for k < n, device arrays (Array2D::add with
 * scale 1, Array.hpp:51-58; Pic::advance sums species rho this way, pic.cpp:351-355) */
int m2d_field_add(m2d_ctx *ctx, double *dst, const double *src, double scale, size_t n);

/* ======================= grid and geometry (host) ======================= */

/* Field2D geometry (Field2D.hpp:16-19); dx = x_max/(x_sampl-1) (param.cpp:126-127) */
typedef struct {
    int32_t nx, nz;        /* x_sampl, z_sampl (Array2D jmax, lmax) */
    double dx, dz;
    double xmin, zmin;
} m2d_grid2d;

/* geometries of Param::geometry (param.hpp:21-22) */
enum { M2D_GEO_EMPTY = 0, M2D_GEO_PROBE, M2D_GEO_RF_22PT, M2D_GEO_RF_8PT, M2D_GEO_RF_HAITRAP,
       M2D_GEO_RF_QUAD, M2D_GEO_MAC, M2D_GEO_PENNING, M2D_GEO_PENNING_SIMPLE, M2D_GEO_TUBE };

typedef struct {
    int32_t geometry;
    double x_max, z_max;     /* Param::x_max, z_max */
    double probe_radius;     /* param.cpp:31 */
    double u_probe;          /* param.cpp:33 */
    double extern_field;     /* param.cpp:34 */
} m2d_geom2d;

/* t_grid::t_grid + builders (fields.cpp:374-833): fills mask[nx*nz] and
 * voltage[nx*nz] (host arrays). */
int m2d_grid_build_2d(const m2d_grid2d *g, const m2d_geom2d *p, uint8_t *mask, double *voltage);
/* t_grid::square_electrode (fields.cpp:835-850) */
int m2d_grid_square_electrode_2d(const m2d_grid2d *g, uint8_t *mask, double *voltage,
                                 double rmin, double rmax, double zmin, double zmax,
                                 double v, int mask_type);
/* t_grid::circle_electrode (fields.cpp:852-868) */
int m2d_grid_circle_electrode_2d(const m2d_grid2d *g, uint8_t *mask, double *voltage,
                                 double xc, double zc, double radius, double v, int mask_type);
/* BOUNDARY marking of the builders (e.g. fields.cpp:507-518) */
int m2d_grid_mark_boundary_2d(const m2d_grid2d *g, uint8_t *mask);
/* t_grid::is_free (fields.hpp:94-101), host; cells on the last row/column are not free */
int m2d_grid_is_free_2d(const m2d_grid2d *g, const uint8_t *mask, double x, double z);

/* ======================= particles ====================================== */

/* Species particle store as SoA (replaces vector<t_particle>, particles.hpp:26-33,113).
 * Long-time positions: with xl and zl set, a position is the exact FP64 sum
 * x + xl (z + zl) of an FP32 value and its FP32 remainder, about 48
 * significant bits, and every pass, sampler, sort and compaction reads and
 * writes it as such (the reference's FP64 t_particle::x, z drift,
 * particles.cpp:986-987).  Without them (NULL) positions are FP32. */
typedef struct {
    float *x, *z;          /* in-plane position (reference x, z) */
    float *vx, *vy, *vz;   /* vy is the out-of-plane component */
    int64_t n;             /* slots, live + empty */
    float *xl, *zl;        /* FP32 remainders of x and z, or both NULL */
} m2d_particles2d;

/* per-species constants of one step (BaseSpecies members, particles.hpp:118-131) */
typedef struct {
    double charge, mass, dt;
    double lifetime;       /* 1/sum(rates); INFINITY disables MCC */
    uint64_t niter;        /* Species::niter: E at niter*dt, RNG counter */
    uint32_t stream;       /* RNG stream id of this species */
    uint32_t reserved;
} m2d_species2d;

/* ======================= fields ========================================== */

enum { M2D_BC_FREE = 0, M2D_BC_PERIODIC = 1 };

/* Param members read by Fields and Species::advance_boundary (param.hpp) */
typedef struct {
    m2d_grid2d grid;
    double x_max, z_max;             /* box is [xmin, x_max] x [zmin, z_max] (param.cpp:23) */
    int32_t geometry_empty;          /* Param::geometry == EMPTY (fields.hpp:126) */
    int32_t selfconsistent;
    int32_t electric_field_from_file;
    int32_t boundary;                /* M2D_BC_FREE / M2D_BC_PERIODIC */
    double extern_field;
    int32_t rf;
    double rf_amplitude, rf_omega, rf_U0;
    double Br, Bz, Bt;               /* magnetic_field_const (fields.hpp:154-169) */
    double dV;                       /* Param::dV (param.cpp:134-136) */
} m2d_fields2d_params;

enum { M2D_FIELD_U = 0, M2D_FIELD_URF = 1, M2D_FIELD_RHO = 2, M2D_FIELD_VOLTAGE = 3 };

/* Fields::Fields (fields.cpp:115-276): uploads mask and voltage, builds the
 * multigrid hierarchy that replaces the UMFPACK factorisation.  Grids need
 * nx, nz >= 3 and (nx+1) (nz+1) < 2^30 (M2D_ERR_BAD_ARG otherwise). */
int m2d_fields2d_create(m2d_ctx *ctx, const m2d_fields2d_params *p, const uint8_t *mask,
                        const double *voltage, m2d_fields2d **out);
int m2d_fields2d_destroy(m2d_fields2d *f);
/* device pointer of Fields::u / uRF / rho / grid.voltage (double[nx*nz]) */
double *m2d_fields2d_data(m2d_fields2d *f, int which);
int m2d_fields2d_upload(m2d_fields2d *f, int which, const double *host);
int m2d_fields2d_download(m2d_fields2d *f, int which, double *host);
/* Rebuild the gather tables after u or uRF changed outside a solve
 * (e.g. Field2D::load for electric_field_from_file, pic.cpp:154-158). */
int m2d_fields2d_update(m2d_fields2d *f);
/* Fields::u_smooth (fields.cpp:28-113) of u (which = M2D_FIELD_U) or uRF:
 * with `symmetry` (square grids only, M2D_ERR_UNSUPPORTED otherwise) first
 * the average of the 8 mirror images of every node; then the 2dx x 2dx
 * square-particle filter (u + edge neighbours / 2 + diagonals / 4) / 4 at
 * every interior node within `radius` (m) of the grid centre (<= 0: all),
 * electrode nodes included as in the reference.  The reference's loop also
 * writes the last column from reads past the row end; here it stays as it
 * is.  Rebuilds the gather tables.  Bits equal to the FP64 formula. */
int m2d_fields2d_u_smooth(m2d_fields2d *f, int which, int symmetry, double radius);
/* Fields::reset (fields.cpp:355-360) */
int m2d_fields2d_reset(m2d_fields2d *f);

enum { M2D_SOLVE_STATIC = 0, M2D_SOLVE_RF = 1, M2D_SOLVE_RAW = 2 };
typedef struct {
    int32_t outer_iters;      /* FP64 refinement steps */
    int32_t inner_iters;      /* FP32 MG-PCG iterations in total */
    double rel_residual;      /* ||b - A u||_2 / ||b_eff||_2 in FP64 over free nodes */
} m2d_solve_info;
/* Fields::boundary_solve (STATIC -> u, fields.cpp:278-312), boundary_solve_rf
 * (RF -> uRF, fields.cpp:314-348), solve (RAW -> u, fields.cpp:350-353).
 * Builds the rhs in place on Fields::rho like the reference, then solves to
 * FP64 relative residual `tol` (<= 0: 1e-10). */
int m2d_fields2d_solve(m2d_fields2d *f, int mode, double tol, m2d_solve_info *info);
/* Fields::E at host-given points (fields.hpp:124-150), FP64 inputs, for tests */
int m2d_fields2d_E(m2d_fields2d *f, const double *x, const double *z, int64_t n, double t,
                   double *ex, double *ez);
/* Fields::load_magnetic_field (fields.cpp:870-959) as data, and Fields::B
 * with magnetic_field_const = 0 (fields.hpp:171-176): Br and Bz on their own
 * grid g (host arrays, row-major [nx*nz]), Bt = 0.  The particle passes then
 * rotate every particle by the field interpolated at its position
 * (Field2D::interpolate, Field2D.hpp:64-78), and the params' constant B is
 * not used; a position outside the B grid is clamped onto it and raises
 * M2D_FLAG_OUT_OF_RANGE (the reference throws).  NaN values are rejected
 * ("garbage loaded").  Br == NULL goes back to the constant B.  The batched,
 * adjoint and multicoll paths reject fields with a B grid. */
int m2d_fields2d_set_magnetic_field(m2d_fields2d *f, const m2d_grid2d *g, const double *Br,
                                    const double *Bz);
/* Fields::B at n points (fields.hpp:152-177), device arrays, through the
 * tables the passes use (FP32 values of a B grid) */
int m2d_fields2d_B(m2d_fields2d *f, const double *x, const double *z, int64_t n, double *Br,
                   double *Bz, double *Bt);

/* ======================= Field2D kernels ================================= */

/* Field2D::accumulate for all live particles (Species::accumulate,
 * particles.hpp:358-367): rho[nx*nz] += charge * CIC weights.  Deterministic:
 * weights are summed as 64-bit fixed point (2^-31) and converted once. */
int m2d_accumulate_2d(m2d_ctx *ctx, const m2d_grid2d *g, const m2d_particles2d *p,
                      double charge, double *rho);
/* No counterpart in the reference: m2d_accumulate_2d for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same kernels and the same bits; the
 * fixed-point grid lives in a static per-GPU device array of 2^20 nodes
 * (8 MiB), so nx * nz must be at most 1048576 (M2D_ERR_BAD_ARG otherwise).
 * No allocation of any kind.  Calls on one GPU share that array: calls that
 * overlap on one GPU (other streams, other threads) race. */
int m2d_accumulate_2d_noalloc(m2d_ctx *ctx, const m2d_grid2d *g, const m2d_particles2d *p,
                              double charge, double *rho);
/* Field2D::interpolate / grad (Field2D.hpp:64-168) at n FP64 points, device arrays */
int m2d_interpolate_2d(m2d_ctx *ctx, const m2d_grid2d *g, const double *data, const double *x,
                       const double *z, int64_t n, double *out);
int m2d_grad_2d(m2d_ctx *ctx, const m2d_grid2d *g, const double *data, const double *x,
                const double *z, int64_t n, double *gx, double *gz);

/* ======================= MCC ============================================= */

/* interaction types (parser.hpp:15) */
enum { M2D_ELASTIC = 0, M2D_LANGEVIN = 1, M2D_CX = 2, M2D_COULOMB = 3, M2D_SUPERELASTIC = 4 };

/* one INTERACTION block of the species conf (parser.cpp:97-143), values as
 * written there: DE in eV, RATE raw (the library applies the LANGEVIN
 * cutoff^2 factor of particles.hpp:81-82) */
typedef struct {
    int32_t type;
    int32_t target;           /* index into m2d_mcc_desc.targets */
    double DE;
    double rate;
    double cutoff;
    int32_t tab_offset, tab_len;  /* CROSS_SECTION rows in tab_E / tab_sigma */
} m2d_interaction;

/* one secondary species (speclist entry, particles.hpp:113-131) */
typedef struct {
    double mass, density, temperature, charge;
    /* live ensemble of a particle species (random_particle, particles.hpp:201-211);
     * x == NULL for a continuum sampled by rndv (particles.hpp:188-193) */
    m2d_particles2d partner;
} m2d_target;

typedef struct {
    double mass;              /* primary species */
    double charge, temperature, density;  /* primary (COULOMB, particles.cpp:19-26) */
    double E_max;             /* eV; svmax is searched up to veV(E_max) (particles.cpp:163) */
    int32_t n_targets;
    const m2d_target *targets;
    int32_t n_inter;
    const m2d_interaction *inter;
    int32_t n_tab;
    const double *tab_E;      /* eV */
    const double *tab_sigma;  /* m^2 */
} m2d_mcc_desc;

/* BaseSpecies::lifetime_init + svmax_find (particles.cpp:142-206) and the
 * device tables for BaseSpecies::scatter (particles.cpp:208-365).  COULOMB
 * (Interaction::coulomb_sigma, particles.cpp:19-26) needs the primary's
 * charge, temperature and density and the target's charge; its partner
 * update (particles.cpp:309-314) writes the partner store's velocities, so
 * m2d_advance_2d applies a COULOMB table's hits one after another in slot
 * order, as the reference's loop does (one device thread, after a host read
 * of the hit count), and the batched, 3D and multicoll passes, which scatter
 * their hits in parallel, refuse such a table (M2D_ERR_UNSUPPORTED). */
int m2d_mcc_create(m2d_ctx *ctx, const m2d_mcc_desc *d, m2d_mcc **out);
int m2d_mcc_destroy(m2d_mcc *m);
double m2d_mcc_lifetime(const m2d_mcc *m);
int m2d_mcc_rates(const m2d_mcc *m, double *rates_by_species);
/* point a particle target at its current store (after insert/compact/sort) */
int m2d_mcc_set_partner(m2d_mcc *m, int32_t target, const m2d_particles2d *p);

/* ======================= species step ==================================== */

#define M2D_STEP_PUSH      1u   /* advance_boris incl. the MCC trial (particles.cpp:924-995);
                                   without it the pass is advance_boundary alone */
#define M2D_STEP_BOUNDARY  2u   /* advance_boundary (particles.hpp:370-406) */
#define M2D_STEP_DEPOSIT   4u   /* rho.accumulate in advance_boundary (particles.hpp:408-409) */

typedef struct {
    int64_t removed;          /* particles removed by the boundary pass */
    int64_t collisions;       /* MCC trials that called scatter */
    int64_t spilled;          /* m2d_advance_3d over a store sorted by
                                 m2d_sort_3d: deposits that fell outside their
                                 tile's LDS window and went to the grid one by
                                 one (particles that moved more than 2 cells
                                 past their sort tile).  A rising count says
                                 the store wants sorting again.  0 for an
                                 unsorted store, whose pass has no windows,
                                 and for the 2D passes, which do not count
                                 (it cost the 2D pass 1.5% of its time) */
} m2d_step_stats;

/* Species<CARTESIAN>::advance (particles.hpp:342-349) fused into one pass:
 * E/B gather + Boris + drift + MCC trial/scatter, then box test, is_free and
 * (flag DEPOSIT) CIC deposit of the survivors into rho (C/node, added).
 * Does not increment niter (the caller owns Species::niter).  `mcc` may be
 * NULL.  `stats` may be NULL (no synchronisation then).  The trial's hits
 * are queued and scattered after the pass (a hit changes only the velocity
 * after the drift).  When a target draws its partners from this very store,
 * every scatter reads its partner as the push left it and the new velocities
 * are stored after all of them (the reference's loop would show a later
 * partner its previous-step velocity). */
int m2d_advance_2d(m2d_ctx *ctx, m2d_fields2d *f, const m2d_species2d *s,
                   const m2d_particles2d *p, const m2d_mcc *mcc, double *rho,
                   uint32_t flags, m2d_step_stats *stats);
/* Species<CARTESIAN>::advance_boris_init (particles.cpp:997-1050) */
int m2d_advance_init_2d(m2d_ctx *ctx, m2d_fields2d *f, const m2d_species2d *s,
                        const m2d_particles2d *p);
/* Species<CARTESIAN>::advance_multicoll (particles.cpp:813-861), the mover
 * of Param::ADVANCE_MULTICOLL: every live particle flies in the constant
 * field of Fields::E(0, 0) (fields.hpp:126-131) from one collision event to
 * the next inside the step; each event calls BaseSpecies::scatter (`mcc`
 * NULL: a null collision) and draws time_to_death = lifetime * Exp(1).
 * time_to_death[p->n] (FP64, device) is t_particle::time_to_death
 * (particles.hpp:31), per slot.  Like the reference it requires B = 0, the
 * EMPTY geometry and selfconsistent = 0 (M2D_ERR_UNSUPPORTED otherwise), and
 * reproduces its position update x += (v + a t/2) t with v already kicked.
 * dt / lifetime must be at most 256; a particle with more than 4096 events
 * in one step raises M2D_FLAG_EVENTS and flies the rest of the step freely.
 * advance_boundary is a separate call (m2d_advance_2d with BOUNDARY only).
 * stats->collisions counts events.  Does not increment niter. */
int m2d_advance_multicoll_2d(m2d_ctx *ctx, m2d_fields2d *f, const m2d_species2d *s,
                             const m2d_particles2d *p, double *time_to_death, const m2d_mcc *mcc,
                             m2d_step_stats *stats);
/* The time_to_death draw of BaseSpecies::lifetime_init (particles.cpp:163-169):
 * time_to_death = s->lifetime * Exp(1) for the live slots of p (empty slots
 * keep theirs).  Draws are keyed by (slot, s->niter, s->stream). */
int m2d_time_to_death_init_2d(m2d_ctx *ctx, const m2d_species2d *s, const m2d_particles2d *p,
                              double *time_to_death);
/* energy_dist.mean_tot() after energy_dist_compute (particles.cpp:408-414,
 * histogram.cpp:68-71): mean kinetic energy in eV over live particles */
int m2d_mean_energy_2d(m2d_ctx *ctx, const m2d_particles2d *p, double mass, double *mean_eV,
                       int64_t *n_live);
/* n_particles() (particles.hpp:248) */
int m2d_count_live_2d(m2d_ctx *ctx, const m2d_particles2d *p, int64_t *n_live);
/* Stable stream compaction src -> dst (dst has room for src->n), replacing
 * the free list of BaseSpecies::remove/insert (particles.hpp:223-247). */
int m2d_compact_2d(m2d_ctx *ctx, const m2d_particles2d *src, const m2d_particles2d *dst,
                   int64_t *n_live);
/* Rebuild the store in tile order (8 x 8-cell tiles, row-major; inside a
 * tile the particles keep their slot order) into dst, empty slots dropped
 * to the end; *n_live (may be NULL) gets the live count.
 * Replaces the insertion-order vector<t_particle> with holes
 * (particles.hpp:223-249); the fused pass is fastest on a recently sorted
 * store.  Slot order keys the MCC random streams, so sorting changes which
 * draws a particle gets, never the statistics. */
int m2d_sort_2d(m2d_ctx *ctx, m2d_fields2d *f, const m2d_particles2d *src,
                const m2d_particles2d *dst, int64_t *n_live);
/* The sort also records the tile schedule of dst (same x pointer and slot
 * count) that the deposit of m2d_advance_2d uses.  Call this after writing
 * the store by any other means (uploads, Species::insert): the schedule is
 * dropped and the pass falls back to its unsorted deposit.  m2d_compact_2d
 * and the samplers drop it for their destination themselves. */
int m2d_sort_forget_2d(m2d_ctx *ctx, const m2d_particles2d *p);
/* Species<CARTESIAN>::add_particles_on_disk (particles.cpp:728-749) into the
 * n slots of p: uniform on the disk by rejection, Maxwellian velocities of
 * thermal speed vth = v_max/sqrt(2) (rndv, particles.hpp:188-193).  Samples
 * outside the box become empty slots, as the reference skips them. */
int m2d_add_particles_on_disk_2d(m2d_ctx *ctx, m2d_fields2d *f, const m2d_particles2d *p,
                                 double cx, double cz, double radius, double vth,
                                 uint32_t stream);
/* Species<CARTESIAN>::add_particles_bessel (particles.cpp:701-726) into the
 * n slots of p: density J0(2.405 r / radius) on the disk, by rejection from
 * the square with acceptance uni() <= J0; velocities as rndv.  Samples
 * outside the box become empty slots, as the reference skips them. */
int m2d_add_particles_bessel_2d(m2d_ctx *ctx, m2d_fields2d *f, const m2d_particles2d *p,
                                double cx, double cz, double radius, double vth, uint32_t stream);
/* Species<CARTESIAN>::add_particles_everywhere (particles.cpp:686-699) into the
 * n slots of p: uniform in the box [xmin, x_max] x [zmin, z_max], Maxwellian
 * velocities of thermal speed vth, and (time_to_death not NULL, FP64 device
 * array of n) time_to_death = lifetime * Exp(1). */
int m2d_add_particles_everywhere_2d(m2d_ctx *ctx, m2d_fields2d *f, const m2d_particles2d *p,
                                    double vth, double *time_to_death, double lifetime,
                                    uint32_t stream);

/* ======================= adjoint ========================================= */

typedef struct {
    int32_t steps;               /* self-consistent steps of the run */
    int32_t checkpoint_every;    /* particle checkpoints every K steps; the reverse sweep replays
                                    the pushes of one K-step segment from its checkpoint */
    int32_t n_groups;            /* electrode voltages differentiated, <= 64 */
    const int32_t *group;        /* host, nx*nz: group of each FIXED node, -1 for none */
    double adjoint_tol;          /* FP64 relative residual of the adjoint solves; 0: 1e-10, the
                                    forward solves' (the forward run always solves to 1e-10) */
    m2d_mcc *mcc;                /* NULL: no collisions; else the species' MCC (continuum targets
                                    only), stepped with the species' lifetime */
} m2d_adjoint_cfg;

typedef struct {
    int64_t n_live;              /* live particles at the end (the mean's denominator) */
    double forward_ms;           /* the forward run alone */
    double backward_ms;          /* replay + reverse sweep */
} m2d_adjoint_info;

/* J = energy_dist.mean_tot() (histogram.cpp:68-71) after cfg->steps iterations
 * of Pic::advance with a solve every step (pic.cpp:330-358; one species, no
 * RF), starting from the store p already started by advance_boris_init,
 * whose charge is deposited first (pic.cpp:362-375).  With cfg->mcc the
 * steps collide as advance_boris does (particles.cpp:937-993,208-365).
 * grad (may be NULL: forward only) receives dJ/dV_g for each group of FIXED
 * nodes whose voltages move together, by reverse-mode differentiation of
 * the push, the deposit and the Poisson solve (self-adjoint on the free
 * nodes); particles removed by the boundary are frozen out.  Collisions are
 * differentiated under common random numbers: the same trials fire and each
 * keeps its draws and its interaction, and the scatter's Jacobian in the
 * particle's velocity enters the reverse sweep.  The store, Fields::u and
 * rho end as after a plain forward run.  No counterpart in the reference,
 * which has no derivatives. */
int m2d_adjoint_energy_2d(m2d_ctx *ctx, m2d_fields2d *f, const m2d_species2d *s, const m2d_particles2d *p,
                          const m2d_adjoint_cfg *cfg, double *J, double *grad, m2d_adjoint_info *info);

/* ======================= batched ensembles =============================== */

typedef struct m2d_batch m2d_batch;

/* Independent runs sharing a field geometry and species constants (the
 * one-process-per-run sweeps of drive_22PT_pressure.py:145-160 and
 * drive_8PT_pressure.py:151-166): run r owns slots [offset[r], offset[r+1])
 * of one store, with its own RF amplitude, collision tables and random
 * stream. */
typedef struct {
    int32_t n_runs;
    const int64_t *offset;        /* host, n_runs + 1, offset[0] = 0, offset[n_runs] = store slots */
    const double *rf_amplitude;   /* host, per run V_rf (ignored when the fields have rf = 0) */
    const m2d_mcc *const *mcc;    /* host, per run collision tables; NULL or NULL entries: none */
    const uint32_t *stream;       /* host, per run RNG stream; NULL: the run index */
} m2d_batch_desc;

/* launch shape of the persistent pass */
typedef struct {
    int32_t block;                /* threads per block: 64, 128 or 256 */
    int32_t ppt;                  /* particles per thread: 1, 2 or 4 */
    int32_t seg_steps;            /* steps per launch */
} m2d_batch_shape;

/* The batch over fields f (no space charge) and species constants s
 * (charge, mass, dt); the store p is advanced in place by m2d_batch_advance. */
int m2d_batch_create(m2d_ctx *ctx, m2d_fields2d *f, const m2d_species2d *s, const m2d_particles2d *p,
                     const m2d_batch_desc *d, m2d_batch **out);
int m2d_batch_destroy(m2d_batch *b);
/* time every shape on a copy of the store for probe_steps steps from niter0
 * and keep the fastest (seg_steps as given); the batch's state is unchanged */
int m2d_batch_tune(m2d_batch *b, uint64_t niter0, int32_t probe_steps, int32_t seg_steps, m2d_batch_shape *best);
int m2d_batch_set_shape(m2d_batch *b, const m2d_batch_shape *sh);
/* Species<CARTESIAN>::advance of every run for steps niter0 .. niter0+steps-1
 * (particles.hpp:342-349, no deposit): one persistent launch per seg_steps
 * steps, the call's launches captured and replayed as one hipGraph.  Bit for
 * bit the states m2d_advance_2d gives each run step by step. */
int m2d_batch_advance(m2d_batch *b, uint64_t niter0, int32_t steps);
/* per run: energy_dist.mean_tot() (eV) and live count */
int m2d_batch_energy(m2d_batch *b, double *mean_eV, int64_t *n_live);

/* ======================= 3D (Field3D, Species<CARTESIAN3D>) ============== */

typedef struct m2d_fields3d m2d_fields3d;

/* Field3D geometry (Field3D.hpp:14-18) with the origin at 0; data[(i*ny + j)*nz + k]
 * like Array3D (Array.hpp:148-151) */
typedef struct {
    int32_t nx, ny, nz;
    double dx, dy, dz;
} m2d_grid3d;

/* 3D particle store: SoA FP32, x == M2D_DEAD bits marks an empty slot */
typedef struct {
    float *x, *y, *z;
    float *vx, *vy, *vz;
    int64_t n;
} m2d_particles3d;

typedef struct {
    m2d_grid3d grid;
    double x_max, y_max, z_max;      /* box [0, max] per axis (species3d.cpp:63-66) */
    int32_t boundary;                /* M2D_BC_FREE removes; M2D_BC_PERIODIC wraps with mod */
    double Bx, By, Bz;               /* uniform B (the reference's 3D mover sets B = 0) */
    double rhs_scale;                /* free rows: rhs = -rho * rhs_scale; the reference uses
                                        macroparticle_factor / eps_0 (fields3d.cpp:92) */
} m2d_fields3d_params;

/* ElMag3D / Geometry / Solver (fields3d.hpp:38-115, fields3d.cpp:12-101):
 * mask (M2D_FIXED / M2D_FIXED_RF / M2D_FREE per node, Dirichlet on the
 * grid faces) and voltage on the host; builds the multigrid hierarchy that
 * replaces the UMFPACK factorisation of Solver::matrix_init / umfpack_init. */
int m2d_fields3d_create(m2d_ctx *ctx, const m2d_fields3d_params *p, const uint8_t *mask,
                        const double *voltage, m2d_fields3d **out);
int m2d_fields3d_destroy(m2d_fields3d *f);
/* device pointer of u / rho / voltage (M2D_FIELD_U, M2D_FIELD_RHO, M2D_FIELD_VOLTAGE) */
double *m2d_fields3d_data(m2d_fields3d *f, int which);
int m2d_fields3d_upload(m2d_fields3d *f, int which, const double *host);
int m2d_fields3d_download(m2d_fields3d *f, int which, double *host);
int m2d_fields3d_reset(m2d_fields3d *f);
/* Solver::solve (fields3d.cpp:84-96): rhs built in place on rho, 7-point
 * system solved into u to FP64 relative residual tol (<= 0: 1e-10) */
int m2d_fields3d_solve(m2d_fields3d *f, double tol, m2d_solve_info *info);

/* Field3D::accumulate over the live particles (Field3D.hpp:40-66), 2^-31 fixed point */
int m2d_accumulate_3d(m2d_ctx *ctx, const m2d_grid3d *g, const m2d_particles3d *p, double charge,
                      double *rho);
/* Field3D::interpolate / grad (Field3D.hpp:68-163) at n FP64 points, device arrays */
int m2d_interpolate_3d(m2d_ctx *ctx, const m2d_grid3d *g, const double *data, const double *x,
                       const double *y, const double *z, int64_t n, double *out);
int m2d_grad_3d(m2d_ctx *ctx, const m2d_grid3d *g, const double *data, const double *x,
                const double *y, const double *z, int64_t n, double *gx, double *gy, double *gz);
/* Species<CARTESIAN3D>::advance (species3d.cpp:3-94) in one pass: E = -grad u,
 * Boris in (x, y, z), drift, MCC trial (scatter applied after the pass),
 * box / mod wrap, Geometry::is_free, CIC deposit (flag DEPOSIT) into rho. */
int m2d_advance_3d(m2d_ctx *ctx, m2d_fields3d *f, const m2d_species2d *s, const m2d_particles3d *p,
                   const m2d_mcc *mcc, double *rho, uint32_t flags, m2d_step_stats *stats);
/* leapfrog start of the 3D mover (the 2D advance_boris_init, particles.cpp:997-1050) */
int m2d_advance_init_3d(m2d_ctx *ctx, m2d_fields3d *f, const m2d_species2d *s,
                        const m2d_particles3d *p);
/* stable re-order by cell, tile-major (4 x 4 x 4-cell tiles), into dst (same slot count), empty slots
 * dropped to the end */
int m2d_sort_3d(m2d_ctx *ctx, m2d_fields3d *f, const m2d_particles3d *src, const m2d_particles3d *dst,
                int64_t *n_live);
/* drop the tile schedule of a 3D store written by other means than m2d_sort_3d */
int m2d_sort_forget_3d(m2d_ctx *ctx, const m2d_particles3d *p);
/* n_particles() and energy_dist.mean_tot() of a 3D store */
int m2d_count_live_3d(m2d_ctx *ctx, const m2d_particles3d *p, int64_t *n_live);
int m2d_mean_energy_3d(m2d_ctx *ctx, const m2d_particles3d *p, double mass, double *mean_eV,
                       int64_t *n_live);

#ifdef __cplusplus
}
#endif
#endif
