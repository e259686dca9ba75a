/*
 * m2d_oracle.h -- CPU oracle for the mag2d PIC-MCC inner loop.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C, FP64 restatement of the
 * rouckas/mag2d algorithms on the PIC-MCC path, written to check the HIP
 * implementation.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it, and only as the checker; the product library
 * (mag2d_amd/libmag2d_amd.so) never links or calls it.
 *
 * Every function cites the reference file:line it restates.  Layouts follow
 * the reference: row-major grids data[i*nz + j] (Array.hpp:7-23, z fastest),
 * particles as the 64-byte AoS t_particle (particles.hpp:26-33).
 *
 * Pinning status of each function is recorded in the comments below and in
 * the tests that check it (tests/test_oracle_*.py).
 */
#ifndef M2D_ORACLE_H
#define M2D_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* physical constants of the reference (param.cpp:7-11) */
#define O_EPS0 8.854187817e-12
#define O_KB   1.380662e-23
#define O_QE   1.602189e-19

/* mask values (fields.hpp:20) */
enum { O_FIXED = 0, O_FIXED_RF = 1, O_FREE = 2, O_BOUNDARY = 3 };

/* geometries (param.hpp:21-22) */
enum { O_GEO_EMPTY = 0, O_GEO_PROBE, O_GEO_RF_22PT, O_GEO_RF_8PT, O_GEO_RF_HAITRAP,
       O_GEO_RF_QUAD, O_GEO_MAC, O_GEO_PENNING, O_GEO_PENNING_SIMPLE, O_GEO_TUBE };

/* interaction types (parser.hpp:15) */
enum { O_ELASTIC = 0, O_LANGEVIN, O_CX, O_COULOMB, O_SUPERELASTIC };

/* boundary conditions (param.hpp:19) */
enum { O_BC_FREE = 0, O_BC_PERIODIC = 1 };

/* 2D grid geometry: Field2D(x_sampl, z_sampl, dx, dz, xmin, zmin) (Field2D.hpp:16-19) */
typedef struct {
    int nx, nz;          /* jmax, lmax */
    double dx, dz;
    double idx, idz;
    double xmin, zmin;
} o_grid2;

/* t_particle (particles.hpp:26-33): 7 doubles + bool = 64 B */
typedef struct {
    double x, y, z;
    double vx, vy, vz;
    double time_to_death;
    uint8_t empty;
    uint8_t pad_[7];
} o_particle;

/* ---------------- random numbers ------------------------------------------
 * Two sources of uniforms drive the same sampling code:
 *  - O_RNG_SHR3:   the reference's shared SHR3 stream, uni() rounded to float
 *                  (random.cpp:33-42, seeding random.cpp:192-200);
 *  - O_RNG_PHILOX: Philox4x32-10 (Salmon et al., SC'11) keyed per
 *                  (seed, stream) with counter (slot, step, block, 0); this is
 *                  the device library's generator, so decisions can be
 *                  compared draw for draw.
 */
enum { O_RNG_SHR3 = 0, O_RNG_PHILOX = 1 };
typedef struct {
    int kind;
    /* SHR3 */
    uint32_t jsr;
    /* Philox */
    uint32_t key[2];
    uint32_t ctr[4];
    uint32_t buf[4];
    int nbuf;          /* words left in buf */
} o_rng;

void     o_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void     o_rng_shr3_seed(o_rng *r, uint32_t seed);             /* random.cpp:192-200 */
/* counter = (slot, step lo, block, purpose << 24 | step hi), key = (seed lo,
 * seed hi ^ stream); purpose 0 = MCC draws of a step, 1 = initial sampling */
void     o_rng_philox_begin(o_rng *r, uint64_t seed, uint32_t stream,
                            uint32_t slot, uint64_t step, uint32_t purpose);
uint32_t o_rng_u32(o_rng *r);
/* MCC trial uniform of slot `slot` at `step` for the Philox source: four
 * slots share one block (counter (slot/4, step, 0, 2<<24 | step hi), word
 * slot%4); a scatter then draws from the slot's own purpose-0 stream. */
double   o_trial_uniform(uint64_t seed, uint32_t stream, uint32_t slot, uint64_t step);
double   o_rng_uni(o_rng *r);     /* SHR3: random.cpp:42 ; Philox: (w+0.5)*2^-32 */
double   o_rng_normal(o_rng *r);  /* Box-Muller, both sources (statistical parity only) */
void     o_rot_len(o_rng *r, double len, double *x, double *y, double *z);   /* random.cpp:103-115 */
void     o_rot(o_rng *r, double *x, double *y, double *z);                   /* random.cpp:117-131 */
void     o_deflect(o_rng *r, double angle, double *x, double *y, double *z); /* random.cpp:133-175 */
double   o_radius(o_rng *r);                                               /* random.cpp:177 */

/* ---------------- math helpers ------------------------------------------ */
double o_mod(double x, double y);                     /* mymath.cpp:62-67 */
double o_vec_interpolate(const double *xs, const double *ys, int n, double x); /* tabulate.cpp:124-141 */
double o_ellint_k(double k);                          /* complete elliptic K(modulus k), AGM */
double o_langevin_chi(double beta);                   /* particles.cpp:337-345 */

/* ---------------- Field2D (Field2D.hpp:45-168) --------------------------- */
int    o_f2_accumulate(const o_grid2 *g, double *data, double q, double x, double z);
int    o_f2_interpolate(const o_grid2 *g, const double *data, double x, double z, double *out);
void   o_f2_grad(const o_grid2 *g, const double *data, double x, double z, double *gx, double *gz);

/* ---------------- t_grid (fields.cpp:374-868, fields.hpp:94-101) -------- */
typedef struct {
    int geometry;
    double x_max, z_max;      /* Param x_max, z_max */
    double probe_radius;      /* param.cpp:31 */
    double u_probe;           /* param.cpp:33 */
    double extern_field;      /* param.cpp:34 */
} o_geom;
void o_grid_build(const o_grid2 *g, const o_geom *p, uint8_t *mask, double *voltage);
void o_square_electrode(const o_grid2 *g, uint8_t *mask, double *voltage,
                        double rmin, double rmax, double zmin, double zmax,
                        double v, int mask_type);                        /* fields.cpp:835-850 */
void o_circle_electrode(const o_grid2 *g, uint8_t *mask, double *voltage,
                        double xc, double zc, double radius, double v,
                        int mask_type);                                  /* fields.cpp:852-868 */
void o_mark_boundary(const o_grid2 *g, uint8_t *mask);                   /* e.g. fields.cpp:507-518 */
int  o_is_free(const o_grid2 *g, const uint8_t *mask, double x, double z); /* fields.hpp:94-101 */

/* ---------------- Poisson (fields.cpp:211-360) --------------------------- */
enum { O_SOLVE_STATIC = 0, O_SOLVE_RF = 1, O_SOLVE_RAW = 2 };
/* rhs assembly in place on rho (fields.cpp:297-310, 333-346) */
void o_poisson_rhs(const o_grid2 *g, const uint8_t *mask, const double *voltage,
                   double *rho, int mode, double dV);
/* y = A x with the reference's rows: identity on FIXED/FIXED_RF, [1,1,-4,1,1] else
 * (fields.cpp:211-259, solved as A^T with UMFPACK_At at fields.cpp:311) */
int  o_poisson_apply(const o_grid2 *g, const uint8_t *mask, const double *x, double *y);
/* FP64 direct solve of A u = b (banded Cholesky of the Dirichlet-eliminated SPD
 * system); stands in for UMFPACK's LU (fields.cpp:273-274,311). */
int  o_poisson_solve(const o_grid2 *g, const uint8_t *mask, const double *b, double *u);

/* ---------------- MCC tables (particles.hpp:49-92, particles.cpp:142-206) */
typedef struct {
    int type;
    double DE;          /* J (parser value * q_e, particles.hpp:73) */
    double rate;        /* constant sigma*v, already * cutoff^2 for LANGEVIN (particles.hpp:81-82) */
    double cutoff;
    double m1, m2;      /* primary, secondary masses */
    /* COULOMB needs the primary's temperature, density and both charges (particles.cpp:19-26) */
    double t1, n1, q1, q2;
    int ntab;
    const double *tabE; /* eV */
    const double *tabS; /* m^2 */
} o_inter;

typedef struct {
    double rate;        /* rates_by_species[k] */
    double density;
    double mass;
    double vth;         /* v_max * sqrt(1/2) of the target (particles.hpp:180,190-192) */
    int n_inter;
    const o_inter *inter;
    o_particle *parts;  /* particle ensemble of the target, NULL for a continuum */
    long n_parts;
} o_target;

typedef struct {
    int n_targets;
    const o_target *targets;
    double lifetime;
    double mass;        /* primary species mass */
} o_mcc;

double o_sigma_v(const o_inter *it, double v);                       /* particles.hpp:61-70 */
double o_svmax_find(const o_inter *its, int n, double vmax, int samples); /* particles.cpp:190-206 */
/* scatter one particle (particles.cpp:208-365). Returns the interaction index
 * chosen, -1 for a null collision, -2 for an unsupported type. */
int o_scatter(const o_mcc *m, o_particle *p, o_rng *r);

/* ---------------- one species step (particles.cpp:924-1050, particles.hpp:342-411) */
typedef struct {
    o_grid2 g;
    double x_min, x_max, z_min, z_max;
    int geometry_empty;          /* Param::geometry == EMPTY (fields.hpp:126) */
    int selfconsistent;
    int electric_field_from_file;
    double extern_field;
    int rf;
    double rf_amplitude, rf_omega, rf_U0;
    double Br, Bz, Bt;           /* magnetic_field_const (fields.hpp:154-169) */
    int boundary;                /* O_BC_FREE / O_BC_PERIODIC */
    const double *u, *uRF;
    const uint8_t *mask;
    /* magnetic_field_const = 0 (fields.hpp:171-176): Br, Bz on the grid gB */
    int b_grid;
    o_grid2 gB;
    const double *BgR, *BgZ;
} o_fields;

typedef struct {
    double charge, mass, dt, lifetime;
    unsigned long niter;
    double *rho;                 /* species rho (Field2D), may be NULL when not selfconsistent */
    const o_mcc *mcc;            /* NULL: no collisions */
    int rng_kind;
    uint64_t seed;
    uint32_t stream;
    long slot0;                  /* Philox slot of p[0] (sub-array offset) */
    o_rng *shr3;                 /* shared SHR3 stream when rng_kind == O_RNG_SHR3 */
} o_species;

void o_fields_E(const o_fields *f, double x, double z, double t, double *ex, double *ez); /* fields.hpp:124-150 */
/* Fields::u_smooth (fields.cpp:28-113); -1 for symmetry on a non-square grid */
int  o_u_smooth(const o_grid2 *g, double *u, int symmetry, double radius);
/* Fields::B (fields.hpp:152-177); -2 for a gridded B outside its grid (the reference throws) */
int  o_fields_B(const o_fields *f, double x, double z, double *Br, double *Bz, double *Bt);
void o_advance_boris(const o_fields *f, o_species *s, o_particle *p, long n, long *n_coll);
void o_advance_boris_init(const o_fields *f, o_species *s, o_particle *p, long n);
double o_rng_rexp(o_rng *r);   /* Exp(1) by inversion (the reference's ziggurat: statistical parity) */
/* BaseSpecies::lifetime_init's time_to_death draw (particles.cpp:163-169) */
void o_ttd_init(const o_species *s, o_particle *p, long n);
/* Species<CARTESIAN>::advance_multicoll (particles.cpp:813-861); 0, or -1 (B != 0) / -2 (not EMPTY
 * or selfconsistent) where the reference throws */
int  o_advance_multicoll(const o_fields *f, o_species *s, o_particle *p, long n, long *n_events);
/* advance_boundary incl. deposit; returns particles removed; *n_bad counts
 * deposits outside the grid (Field2D::accumulate throws there). */
long o_advance_boundary(const o_fields *f, o_species *s, o_particle *p, long n, long *n_bad);
long o_accumulate(const o_fields *f, o_species *s, const o_particle *p, long n);
/* energy_dist.mean_tot() after energy_dist_compute (particles.cpp:408-414, histogram.cpp:68-71) */
double o_mean_energy_eV(const o_particle *p, long n, double mass);
/* add_particles_on_disk (particles.cpp:728-749) with rndv (particles.hpp:188-193) */
long o_add_particles_on_disk(o_particle *p, long cap, long n_req, double cx, double cz,
                             double radius, double vth, const o_fields *f, o_rng *r);
void o_add_particles_on_disk_philox(o_particle *p, long n, double cx, double cz, double radius,
                                    double vth, const o_fields *f, uint64_t seed, uint32_t stream);
/* add_particles_bessel (particles.cpp:701-726), add_particles_everywhere (particles.cpp:686-699),
 * on the device sampler's Philox streams */
void o_add_particles_bessel_philox(o_particle *p, long n, double cx, double cz, double radius,
                                   double vth, const o_fields *f, uint64_t seed, uint32_t stream);
void o_add_particles_everywhere_philox(o_particle *p, long n, double vth, double lifetime,
                                       const o_fields *f, uint64_t seed, uint32_t stream);

/* Pic<CARTESIAN>::advance_init / advance with a solve every step (pic.cpp:330-384),
 * one species, no collisions, from a store already started by advance_boris_init;
 * f->u is the potential array it writes.  Returns the final mean kinetic energy (eV). */
double o_selfconsistent_run(o_fields *f, o_species *s, o_particle *p, long n, const double *voltage,
                            double dV, int steps);

/* ---------------- Field3D (Field3D.hpp:40-163) --------------------------- */
typedef struct {
    int nx, ny, nz;
    double dx, dy, dz, idx, idy, idz;
    double xmin, ymin, zmin;
    double xmax, ymax, zmax;     /* extent from the origin; the reference never sets these */
} o_grid3;
int    o_f3_accumulate(const o_grid3 *g, double *data, double q, double x, double y, double z);
int    o_f3_interpolate(const o_grid3 *g, const double *data, double x, double y, double z, double *out);
void   o_f3_grad(const o_grid3 *g, const double *data, double x, double y, double z,
                 double *gx, double *gy, double *gz);
int    o_poisson3_apply(const o_grid3 *g, const uint8_t *mask, const double *x, double *y);
/* Solver::solve rhs (fields3d.cpp:84-93): FIXED rows take the voltage, free
 * rows rho * -scale (the reference's scale is macroparticle_factor/eps_0) */
void   o_poisson3_rhs(const o_grid3 *g, const uint8_t *mask, const double *voltage, double *rho,
                      double scale);

/* 3D fields of the 3D step (ElMag3D, fields3d.hpp:83-115): E = -grad u,
 * B constant (the reference's species3d.cpp:17-18 sets it to zero; the
 * magnetised 3D trap needs a uniform B, taken as a parameter here) */
typedef struct {
    o_grid3 g;                   /* origin at 0 */
    double x_max, y_max, z_max;  /* box [0, max] (species3d.cpp:63-66) */
    int boundary;                /* O_BC_FREE / O_BC_PERIODIC */
    double Bx, By, Bz;
    const double *u;
    const uint8_t *mask;
} o_fields3;
/* Geometry::is_free (fields3d.hpp:48-55): any of the 8 corners FREE; cells
 * outside [0, n-2] are not free (the reference reads past the mask there) */
int  o_is_free3(const o_grid3 *g, const uint8_t *mask, double x, double y, double z);
/* Species<CARTESIAN3D>::advance (species3d.cpp:3-94): E gather, Boris kick-
 * rotate-kick in (x, y, z), drift, MCC trial + scatter, box (FREE removes,
 * PERIODIC wraps with mod), is_free, deposit into rho (may be NULL). Returns
 * removed particles. */
long o_advance3d(const o_fields3 *f, o_species *s, o_particle *p, long n, long *n_coll, long *n_bad);
/* leapfrog start of the 3D mover: rotation by -dt/2 and a backward half
 * kick, the 2D advance_boris_init (particles.cpp:997-1050) on three axes */
void o_advance3d_init(const o_fields3 *f, o_species *s, o_particle *p, long n);

/* timing helper for the CPU baseline: n_steps of boris + boundary + deposit
 * with the SHR3 stream, returns wall seconds */
double o_bench_steps(const o_fields *f, o_species *s, o_particle *p, long n, int n_steps);
double o_bench_steps_omp(const o_fields *f, o_species *s, o_particle *p, long n, int n_steps,
                         int threads);

#ifdef __cplusplus
}
#endif
#endif
