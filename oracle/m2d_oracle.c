/*
 * m2d_oracle.c -- CPU oracle for the mag2d PIC-MCC inner loop.
 *
 * TEST INFRASTRUCTURE ONLY (see m2d_oracle.h).  Plain C99, FP64, serial
 * unless a function says otherwise.  Each function restates the reference
 * routine it cites; deviations from the reference are listed where they
 * occur (all of them replace undefined behaviour of the reference by a
 * defined result).
 */
#define _GNU_SOURCE
#include "m2d_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define SQR(x) ((x) * (x))

/* ======================================================================= */
/* random numbers                                                          */
/* ======================================================================= */

/* Philox4x32-10, Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
 * easy as 1, 2, 3", SC'11: ten rounds of two 32x32->64 multiplies with the
 * Weyl key schedule. */
void o_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* t_random::initialize_seed (random.cpp:192-200): the seeding itself draws
 * three SHR3 words (for the unused z, w, jcong generators). */
void o_rng_shr3_seed(o_rng *r, uint32_t seed)
{
    memset(r, 0, sizeof(*r));
    r->kind = O_RNG_SHR3;
    r->jsr = 123456789u ^ seed;
    for (int i = 0; i < 3; i++) (void)o_rng_u32(r);
}

void o_rng_philox_begin(o_rng *r, uint64_t seed, uint32_t stream, uint32_t slot, uint64_t step,
                        uint32_t purpose)
{
    r->kind = O_RNG_PHILOX;
    r->key[0] = (uint32_t)seed;
    r->key[1] = (uint32_t)(seed >> 32) ^ stream;
    r->ctr[0] = slot;
    r->ctr[1] = (uint32_t)step;
    r->ctr[2] = 0;                       /* block index within this (slot, step) */
    r->ctr[3] = (purpose << 24) | ((uint32_t)(step >> 32) & 0xFFFFFFu);
    r->nbuf = 0;
}

/* shr3() (random.cpp:33-37): jz = jsr; xorshift 13/17/5; return jz + jsr */
uint32_t o_rng_u32(o_rng *r)
{
    if (r->kind == O_RNG_SHR3) {
        uint32_t jz = r->jsr;
        uint32_t j = r->jsr;
        j ^= j << 13; j ^= j >> 17; j ^= j << 5;
        r->jsr = j;
        return jz + j;
    }
    if (r->nbuf == 0) {
        o_philox4x32_10(r->ctr, r->key, r->buf);
        r->ctr[2]++;
        r->nbuf = 4;
    }
    return r->buf[4 - r->nbuf--];
}

/* uni() (random.cpp:42): float(.5 + (int32)iuni() * 2^-32).  The Philox
 * source uses (w + 0.5) * 2^-32, exact in FP64 and strictly inside (0,1). */
double o_rng_uni(o_rng *r)
{
    uint32_t w = o_rng_u32(r);
    if (r->kind == O_RNG_SHR3)
        return (double)(float)(.5 + (double)(int32_t)w * 2.3283064365386963e-10);
    return ((double)w + 0.5) * 2.3283064365386963e-10;
}

/* Gaussian variate.  The reference uses the ziggurat rnor (random.cpp:49-52);
 * only the distribution matters on this path (statistical parity).  Both
 * sources here use Box-Muller with two uniforms, cosine branch. */
double o_rng_normal(o_rng *r)
{
    double u1 = o_rng_uni(r);
    double u2 = o_rng_uni(r);
    if (u1 <= 0.0) u1 = 1e-300;
    return sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
}

/* 1 - 2u of rot/deflect (random.cpp:106,137): uni() is a float there, so the
 * expression is evaluated in float for the SHR3 source */
static double one_minus_2u(o_rng *r)
{
    double u = o_rng_uni(r);
    if (r->kind == O_RNG_SHR3) return (double)(1.0f - 2.0f * (float)u);
    return 1.0 - 2.0 * u;
}

/* rot(len, x, y, z) (random.cpp:103-115): cos(theta) = 1 - 2u, phi = 2 pi u */
void o_rot_len(o_rng *r, double len, double *x, double *y, double *z)
{
    double cs = one_minus_2u(r);
    *x = len * cs;
    cs = sqrt(1.0 - cs * cs);
    double ph = 2.0 * M_PI * o_rng_uni(r);
    double sp = sin(ph), cp = cos(ph);
    *y = len * cs * sp;
    *z = len * cs * cp;
}

/* rot(x, y, z) (random.cpp:117-131): isotropic direction, length kept */
void o_rot(o_rng *r, double *x, double *y, double *z)
{
    double len = sqrt(SQR(*x) + SQR(*y) + SQR(*z));
    o_rot_len(r, len, x, y, z);
}

/* deflect(angle, x, y, z) (random.cpp:133-175): Boris-type rotation of v by
 * `angle` about an axis perpendicular to v chosen through a random vector. */
void o_deflect(o_rng *r, double angle, double *x, double *y, double *z)
{
    double len = tan(angle / 2.0);
    double x1, y1, z1;
    double t = one_minus_2u(r);
    x1 = len * t;
    t = sqrt(1.0 - t * t);
    double ph = 2.0 * M_PI * o_rng_uni(r);
    double sp = sin(ph), cp = cos(ph);
    y1 = len * t * sp;
    z1 = len * t * cp;

    double tx = *y * z1 - *z * y1;
    double ty = *z * x1 - *x * z1;
    double tz = *x * y1 - *y * x1;
    t = len / sqrt(tx * tx + ty * ty + tz * tz);
    tx *= t; ty *= t; tz *= t;

    double xp = *x - *y * tz + *z * ty;
    double yp = *y - *z * tx + *x * tz;
    double zp = *z - *x * ty + *y * tx;

    t = 2.0 / (1.0 + len * len);
    x1 = tx * t; y1 = ty * t; z1 = tz * t;

    *x += -yp * z1 + zp * y1;
    *y += -zp * x1 + xp * z1;
    *z += -xp * y1 + yp * x1;
}

/* radius() (random.cpp:177): std::sqrt of the float uni() is a float sqrt */
double o_radius(o_rng *r)
{
    double u = o_rng_uni(r);
    if (r->kind == O_RNG_SHR3) return (double)sqrtf((float)u);
    return sqrt(u);
}

/* ======================================================================= */
/* math helpers                                                            */
/* ======================================================================= */

/* mod() (mymath.cpp:62-67) */
double o_mod(double x, double y)
{
    if (x >= 0.0 && x <= y) return x;
    return x - y * (int)(x / y) + (x < 0 ? y : 0);
}

/* vec_interpolate::operator() (tabulate.cpp:124-141): clamp at the ends,
 * bisection, linear interpolation */
double o_vec_interpolate(const double *xs, const double *ys, int n, double x)
{
    if (x >= xs[n - 1]) return ys[n - 1];
    if (x <= xs[0]) return ys[0];
    int j1 = 0, j2 = n - 1;
    while (j2 - j1 > 1) {
        int j3 = (j1 + j2) / 2;
        if (x < xs[j3]) j2 = j3; else j1 = j3;
    }
    double w = (x - xs[j1]) / (xs[j2] - xs[j1]);
    return ys[j1] * (1 - w) + ys[j2] * w;
}

/* Complete elliptic integral of the first kind K(k), modulus k, as
 * std::tr1::comp_ellint_1 takes it: K = pi / (2 AGM(1, sqrt(1-k^2)))
 * (Abramowitz & Stegun 17.6). */
double o_ellint_k(double k)
{
    double a = 1.0, b = sqrt(1.0 - k * k);
    for (int i = 0; i < 40 && fabs(a - b) > 1e-16 * a; i++) {
        double an = 0.5 * (a + b);
        b = sqrt(a * b);
        a = an;
    }
    return M_PI / (a + b);
}

/* Nanbu-Kitatani small-angle Langevin deflection (particles.cpp:337-345,
 * tests/test_langevin.cpp:9-17) */
double o_langevin_chi(double beta)
{
    double t = sqrt(beta * beta * beta * beta - 1.0);
    double xi0 = sqrt(beta * beta - t);
    double xi1 = sqrt(beta * beta + t);
    double zeta = xi0 / xi1;
    double theta = o_ellint_k(zeta) * M_SQRT2 * beta / xi1;
    return M_PI - 2 * theta;
}

/* ======================================================================= */
/* Field2D                                                                 */
/* ======================================================================= */

/* Field2D::accumulate (Field2D.hpp:45-62).  The reference accepts
 * i == jmax-1 and then writes row jmax (out of bounds); here any deposit
 * whose 2x2 stencil leaves the grid returns -2 without writing. */
int o_f2_accumulate(const o_grid2 *g, double *d, double q, double x, double z)
{
    x -= g->xmin; z -= g->zmin;
    if (!(x >= 0.0) || !(z >= 0.0)) return -2;
    int i = (int)(x * g->idx), j = (int)(z * g->idz);
    double u = x * g->idx - i, v = z * g->idz - j;
    if (i < 0 || i > g->nx - 2 || j < 0 || j > g->nz - 2) return -2;
    int nz = g->nz;
    d[i * nz + j] += (1 - u) * (1 - v) * q;
    d[(i + 1) * nz + j] += u * (1 - v) * q;
    d[i * nz + j + 1] += (1 - u) * v * q;
    d[(i + 1) * nz + j + 1] += u * v * q;
    return 0;
}

/* Field2D::interpolate (Field2D.hpp:64-78), same range rule as accumulate */
int o_f2_interpolate(const o_grid2 *g, const double *d, double x, double z, double *out)
{
    x -= g->xmin; z -= g->zmin;
    if (!(x >= 0.0) || !(z >= 0.0)) return -2;
    int i = (int)(x * g->idx), j = (int)(z * g->idz);
    double u = x * g->idx - i, v = z * g->idz - j;
    if (i < 0 || i > g->nx - 2 || j < 0 || j > g->nz - 2) return -2;
    int nz = g->nz;
    *out = (1 - u) * (1 - v) * d[i * nz + j] + u * (1 - v) * d[(i + 1) * nz + j]
         + (1 - u) * v * d[i * nz + j + 1] + u * v * d[(i + 1) * nz + j + 1];
    return 0;
}

/* Field2D::grad (Field2D.hpp:80-168): staggered edge differences, bilinear
 * weights, 1-D fallbacks on the first and last row / column.  When the
 * index is out of range the reference leaves the output untouched; so does
 * this function. */
void o_f2_grad(const o_grid2 *g, const double *d, double x, double z, double *gx, double *gz)
{
    const int jmax = g->nx, lmax = g->nz, nz = g->nz;
    const double idx = g->idx, idy = g->idz;
#define D(a, b) d[(a) * nz + (b)]
    double g1, g2, g3, g4, fx, fy;
    x -= g->xmin; z -= g->zmin;
    /* x component */
    int i = (int)(x * idx + 0.5);
    int j = (int)(z * idy);
    if (j > lmax - 2) j = lmax - 2;
    if (i > 0 && i < jmax - 1) {
        g1 = (D(i, j) - D(i - 1, j)) * idx;
        g2 = (D(i, j + 1) - D(i - 1, j + 1)) * idx;
        g3 = (D(i + 1, j + 1) - D(i, j + 1)) * idx;
        g4 = (D(i + 1, j) - D(i, j)) * idx;
        fx = x * idx - i + .5;
        fy = z * idy - j;
        *gx = g1 * (1 - fx) * (1 - fy) + g2 * (1 - fx) * fy + g3 * fx * fy + g4 * fx * (1 - fy);
    } else if (i == jmax - 1) {
        g1 = (D(i, j) - D(i - 1, j)) * idx;
        g2 = (D(i, j + 1) - D(i - 1, j + 1)) * idx;
        fy = z * idy - j;
        *gx = g1 * (1 - fy) + g2 * fy;
    } else if (i == 0) {
        g3 = (D(i + 1, j + 1) - D(i, j + 1)) * idx;
        g4 = (D(i + 1, j) - D(i, j)) * idx;
        fy = z * idy - j;
        *gx = g3 * fy + g4 * (1 - fy);
    }
    /* z component */
    i = (int)(x * idx);
    j = (int)(z * idy + 0.5);
    if (i > jmax - 2) i = jmax - 2;
    if (j > 0 && j < lmax - 1) {
        g1 = (D(i, j) - D(i, j - 1)) * idy;
        g2 = (D(i + 1, j) - D(i + 1, j - 1)) * idy;
        g3 = (D(i + 1, j + 1) - D(i + 1, j)) * idy;
        g4 = (D(i, j + 1) - D(i, j)) * idy;
        fx = x * idx - i;
        fy = z * idy - j + 0.5;
        *gz = g1 * (1 - fx) * (1 - fy) + g2 * (1 - fy) * fx + g3 * fx * fy + g4 * fy * (1 - fx);
    } else if (j == lmax - 1) {
        g1 = (D(i, j) - D(i, j - 1)) * idy;
        g2 = (D(i + 1, j) - D(i + 1, j - 1)) * idy;
        fx = x * idx - i;
        *gz = g1 * (1 - fx) + g2 * fx;
    } else if (j == 0) {
        g3 = (D(i + 1, j + 1) - D(i + 1, j)) * idy;
        g4 = (D(i, j + 1) - D(i, j)) * idy;
        fx = x * idx - i;
        *gz = g3 * fx + g4 * (1 - fx);
    }
#undef D
}

/* ======================================================================= */
/* t_grid                                                                  */
/* ======================================================================= */

/* square_electrode (fields.cpp:835-850): strict inequalities on node coords */
void o_square_electrode(const o_grid2 *g, uint8_t *mask, double *voltage,
                        double rmin, double rmax, double zmin, double zmax, double v, int type)
{
    for (int i = 0; i < g->nx; i++)
        for (int j = 0; j < g->nz; j++) {
            double r = i * g->dx, z = j * g->dz;
            if (r > rmin && r < rmax && z > zmin && z < zmax) {
                mask[i * g->nz + j] = (uint8_t)type;
                voltage[i * g->nz + j] = v;
            }
        }
}

/* circle_electrode (fields.cpp:852-868): closed disk on node coords */
void o_circle_electrode(const o_grid2 *g, uint8_t *mask, double *voltage,
                        double xc, double zc, double radius, double v, int type)
{
    double r2 = SQR(radius);
    for (int i = 0; i < g->nx; i++)
        for (int j = 0; j < g->nz; j++) {
            double r = i * g->dx - xc, z = j * g->dz - zc;
            if (SQR(r) + SQR(z) <= r2) {
                mask[i * g->nz + j] = (uint8_t)type;
                voltage[i * g->nz + j] = v;
            }
        }
}

/* BOUNDARY marking shared by every builder (e.g. fields.cpp:507-518): an
 * interior node (2 <= i,j < M-2) that is not FIXED and has a FIXED 4-neighbour
 * becomes BOUNDARY (FIXED_RF nodes included). */
void o_mark_boundary(const o_grid2 *g, uint8_t *mask)
{
    const int M = g->nx, N = g->nz;
#define MK(a, b) mask[(a) * N + (b)]
    for (int i = 2; i < M - 2; i++)
        for (int j = 2; j < N - 2; j++)
            if ((MK(i - 1, j) == O_FIXED || MK(i + 1, j) == O_FIXED ||
                 MK(i, j - 1) == O_FIXED || MK(i, j + 1) == O_FIXED) && MK(i, j) != O_FIXED)
                MK(i, j) = O_BOUNDARY;
#undef MK
}

/* frame: all four sides FIXED with voltage f(j) (fields.cpp:421-433 pattern);
 * `sides` = 4 for all sides, 3 leaves i == 0 FREE (MAC, PENNING*) */
static void frame(const o_grid2 *g, uint8_t *mask, double *voltage, int sides, int linear,
                  double extern_field)
{
    const int M = g->nx, N = g->nz;
    for (int i = 0; i < M; i++)
        for (int j = 0; j < N; j++) {
            int edge = (i == M - 1 || j == 0 || j == N - 1 || (sides == 4 && i == 0));
            if (edge) {
                mask[i * N + j] = O_FIXED;
                voltage[i * N + j] = linear ? -extern_field * g->dz * (j - N / 2) : 0.0;
            } else {
                mask[i * N + j] = O_FREE;
            }
        }
}

static void multipole(const o_grid2 *g, uint8_t *mask, double *voltage, int npoles,
                      double r_pole, double r_rod)
{
    /* fields.cpp:494-505 (22PT), 543-552 (8PT), 584-593 (HAITRAP) */
    double xc = 1e-2, zc = 1e-2;
    for (int i = 0; i < npoles; i++) {
        double x = xc + sin(2 * M_PI * (i + 1.0 / 32) / npoles) * r_pole;
        double y = zc + cos(2 * M_PI * (i + 1.0 / 32) / npoles) * r_pole;
        int sign = i % 2 == 0 ? -1 : 1;
        o_circle_electrode(g, mask, voltage, x, y, r_rod, (double)sign, O_FIXED_RF);
    }
}

/* t_grid::t_grid dispatch (fields.cpp:374-415) and the builders
 * fields.cpp:417-833.  `voltage` entries of non-electrode nodes are set to 0
 * (the reference leaves them uninitialised; they are never read). */
void o_grid_build(const o_grid2 *g, const o_geom *p, uint8_t *mask, double *voltage)
{
    const int M = g->nx, N = g->nz;
    for (int k = 0; k < M * N; k++) voltage[k] = 0.0;
    switch (p->geometry) {
    case O_GEO_PROBE: /* fields.cpp:435-473 */
        frame(g, mask, voltage, 4, 1, p->extern_field);
        o_circle_electrode(g, mask, voltage, (M - 1) * g->dx / 2, (N - 1) * g->dz / 2,
                           p->probe_radius, p->u_probe, O_FIXED);
        break;
    case O_GEO_RF_22PT: /* fields.cpp:475-519 */
        frame(g, mask, voltage, 4, 0, 0.0);
        multipole(g, mask, voltage, 22, 0.75e-2, 0.05e-2);
        break;
    case O_GEO_RF_8PT: /* fields.cpp:521-565 */
        frame(g, mask, voltage, 4, 1, p->extern_field);
        multipole(g, mask, voltage, 8, 0.3e-2 + 0.1e-2, 0.1e-2);
        break;
    case O_GEO_RF_HAITRAP: /* fields.cpp:566-610 */
        frame(g, mask, voltage, 4, 1, p->extern_field);
        multipole(g, mask, voltage, 8, 0.3e-2 + 0.01e-2, 0.01e-2);
        break;
    case O_GEO_TUBE: { /* fields.cpp:612-661 */
        frame(g, mask, voltage, 4, 0, 0.0);
        double rc = p->x_max / 2.0, zc = p->x_max / 2.0, r2 = SQR(p->probe_radius);
        for (int i = 0; i < M; i++)
            for (int j = 0; j < N; j++) {
                double r = i * g->dx - rc, z = j * g->dz - zc;
                if (SQR(r) + SQR(z) >= r2) { mask[i * N + j] = O_FIXED; voltage[i * N + j] = 0.0; }
            }
        break;
    }
    case O_GEO_RF_QUAD: /* fields.cpp:663-699 */
        frame(g, mask, voltage, 4, 0, 0.0);
        o_circle_electrode(g, mask, voltage, 5e-3, 1e-2, 2e-3, 1.0, O_FIXED_RF);
        o_circle_electrode(g, mask, voltage, 15e-3, 1e-2, 2e-3, 1.0, O_FIXED_RF);
        o_circle_electrode(g, mask, voltage, 1e-2, 5e-3, 2e-3, -1.0, O_FIXED_RF);
        o_circle_electrode(g, mask, voltage, 1e-2, 15e-3, 2e-3, -1.0, O_FIXED_RF);
        break;
    case O_GEO_MAC: { /* fields.cpp:701-745 */
        frame(g, mask, voltage, 3, 0, 0.0);
        o_square_electrode(g, mask, voltage, 5e-3, 4.5e-2, 1e-2, 1.5e-2, -.00, O_FIXED);
        o_square_electrode(g, mask, voltage, 5e-3, 7e-3, 2e-2, 8e-2, 0.0, O_FIXED);
        o_square_electrode(g, mask, voltage, 5e-3, 4.5e-2, 8.5e-2, 9e-2, -.00, O_FIXED);
        double th = p->u_probe, ofs = 3e-2;
        o_square_electrode(g, mask, voltage, 3e-2, 3.3e-2, 11e-2 + ofs, 14e-2 + ofs, 0.8 * th, O_FIXED);
        o_square_electrode(g, mask, voltage, 4.5e-2, 4.8e-2, 15e-2 + ofs, 25e-2 + ofs, th, O_FIXED);
        o_square_electrode(g, mask, voltage, 3e-2, 3.3e-2, 26e-2 + ofs, 29e-2 + ofs, 1.0 * th, O_FIXED);
        o_square_electrode(g, mask, voltage, 2.5e-2, 2.8e-2, 29e-2 + ofs, 30.5e-2 + ofs, 1.0 * th, O_FIXED);
        o_square_electrode(g, mask, voltage, 15e-3, 4.5e-2, 35e-2, 35.3e-2, .0, O_FIXED);
        o_square_electrode(g, mask, voltage, 0.0, 4.5e-2, 39.5e-2, 40e-2, 3e3, O_FIXED);
        break;
    }
    case O_GEO_PENNING: /* fields.cpp:747-786 */
        frame(g, mask, voltage, 3, 0, 0.0);
        o_square_electrode(g, mask, voltage, 1.57e-2 / 2, 1.67e-2 / 2, 1e-3, 25e-3, -5, O_FIXED);
        o_square_electrode(g, mask, voltage, 0, 1.46e-2 / 2, 15e-3, 16e-3, 10, O_FIXED);
        o_square_electrode(g, mask, voltage, 0, 7e-3 / 2, 12e-3, 19e-3, 10, O_FIXED);
        o_square_electrode(g, mask, voltage, 0, 1.9e-3, 1e-3, 12e-3, 10, O_FIXED);
        o_square_electrode(g, mask, voltage, 4e-3, 7e-3, 52e-3, 53e-3, -5, O_FIXED);
        o_square_electrode(g, mask, voltage, 2.5e-3, 7e-3, 46e-3, 47e-3, 0, O_FIXED);
        break;
    case O_GEO_PENNING_SIMPLE: { /* fields.cpp:788-833, default trap_voltage -1 */
        frame(g, mask, voltage, 3, 0, 0.0);
        double ri = 1e-2, ro = 1.1e-2, tv = -1.0;
        o_square_electrode(g, mask, voltage, ri, ro, 0, 1e-2, -0.5, O_FIXED);
        o_square_electrode(g, mask, voltage, ri, ro, 1.1e-2, 2e-2, 0, O_FIXED);
        o_square_electrode(g, mask, voltage, ri, ro, 2.1e-2, 6e-2, tv, O_FIXED);
        o_square_electrode(g, mask, voltage, ri, ro, 6.1e-2, 7.5e-2, -10, O_FIXED);
        break;
    }
    case O_GEO_EMPTY: /* fields.cpp:417-433 */
    default:
        frame(g, mask, voltage, 4, 1, p->extern_field);
        return; /* empty() does not mark BOUNDARY */
    }
    o_mark_boundary(g, mask);
}

/* t_grid::is_free (fields.hpp:94-101): any of the 4 corners of the cell is
 * FREE.  The reference reads past the mask for i == M-1 or j == N-1 (a
 * particle exactly on x_max / z_max); here such a cell is not free. */
int o_is_free(const o_grid2 *g, const uint8_t *mask, double r, double z)
{
    int i = (int)(r * g->idx), j = (int)(z * g->idz);
    if (!(r >= 0.0) || !(z >= 0.0)) return 0;
    if (i < 0 || j < 0 || i > g->nx - 2 || j > g->nz - 2) return 0;
    const int N = g->nz;
    return mask[i * N + j] == O_FREE || mask[(i + 1) * N + j] == O_FREE ||
           mask[i * N + j + 1] == O_FREE || mask[(i + 1) * N + j + 1] == O_FREE;
}

/* ======================================================================= */
/* Poisson                                                                 */
/* ======================================================================= */

/* Fields::boundary_solve / boundary_solve_rf rhs (fields.cpp:297-310,
 * 333-346), Cartesian rows: FIXED -> voltage (static) or 0 (rf), FIXED_RF ->
 * 0 (static) or voltage (rf), other nodes -> rho * (-dx^2 / (eps0 dV)).
 * O_SOLVE_RAW leaves rho as is (Fields::solve, fields.cpp:350-353). */
void o_poisson_rhs(const o_grid2 *g, const uint8_t *mask, const double *voltage,
                   double *rho, int mode, double dV)
{
    if (mode == O_SOLVE_RAW) return;
    const int n = g->nx * g->nz;
    const double s = -SQR(g->dx) / O_EPS0 / dV;
    for (int k = 0; k < n; k++) {
        if (mask[k] == O_FIXED) rho[k] = (mode == O_SOLVE_STATIC) ? voltage[k] : 0.0;
        else if (mask[k] == O_FIXED_RF) rho[k] = (mode == O_SOLVE_RF) ? voltage[k] : 0.0;
        else rho[k] *= s;
    }
}

/* Matrix rows of Fields::Fields, Cartesian (fields.cpp:211-259): the CSC
 * columns are solved transposed (UMFPACK_At, fields.cpp:311), so row k is
 * identity for FIXED/FIXED_RF nodes and u[k-N] + u[k-1] - 4u[k] + u[k+1] +
 * u[k+N] otherwise.  Returns -1 if a stencil row leaves the grid (the
 * reference would index outside its arrays). */
int o_poisson_apply(const o_grid2 *g, const uint8_t *mask, const double *x, double *y)
{
    const int M = g->nx, N = g->nz;
    for (int i = 0; i < M; i++)
        for (int j = 0; j < N; j++) {
            int k = i * N + j;
            if (mask[k] == O_FIXED || mask[k] == O_FIXED_RF) { y[k] = x[k]; continue; }
            if (i == 0 || j == 0 || i == M - 1 || j == N - 1) return -1;
            y[k] = x[k - N] + x[k - 1] - 4.0 * x[k] + x[k + 1] + x[k + N];
        }
    return 0;
}

/* Direct FP64 solve of A u = b with the rows above.  Dirichlet unknowns are
 * eliminated, leaving the SPD system 4u_k - sum_free u_n = -(b_k - sum_dir b_n),
 * factored as a band Cholesky of half-bandwidth N over the row-major order. */
int o_poisson_solve(const o_grid2 *g, const uint8_t *mask, const double *b, double *u)
{
    const int M = g->nx, N = g->nz;
    const long n = (long)M * N;
    const int bw = N;
    double *L = (double *)calloc((size_t)n * (bw + 1), sizeof(double));
    double *c = (double *)malloc(sizeof(double) * n);
    if (!L || !c) { free(L); free(c); return -3; }
#define LB(k, d) L[(size_t)(k) * (bw + 1) + (d)]
#define DIR(k) (mask[k] == O_FIXED || mask[k] == O_FIXED_RF)
    for (int i = 0; i < M; i++)
        for (int j = 0; j < N; j++) {
            long k = (long)i * N + j;
            if (DIR(k)) { LB(k, 0) = 1.0; c[k] = b[k]; continue; }
            if (i == 0 || j == 0 || i == M - 1 || j == N - 1) { free(L); free(c); return -1; }
            LB(k, 0) = 4.0;
            double rhs = b[k];
            long nb[2] = { k - N, k - 1 };
            for (int q = 0; q < 2; q++) {
                if (DIR(nb[q])) rhs -= b[nb[q]];
                else LB(k, k - nb[q]) = -1.0;
            }
            if (DIR(k + 1)) rhs -= b[k + 1];
            if (DIR(k + N)) rhs -= b[k + N];
            c[k] = -rhs;
        }
    /* band Cholesky, lower factor stored in place: LB(k,d) = L[k][k-d] */
    for (long k = 0; k < n; k++) {
        long dmax = k < bw ? k : bw;
        for (long d = dmax; d >= 1; d--) {
            long j = k - d;
            double s = LB(k, d);
            long m0 = k - bw > j - bw ? k - bw : j - bw;
            if (m0 < 0) m0 = 0;
            for (long m = m0; m < j; m++) s -= LB(k, k - m) * LB(j, j - m);
            LB(k, d) = s / LB(j, 0);
        }
        double s = LB(k, 0);
        for (long m = (k - bw > 0 ? k - bw : 0); m < k; m++) s -= SQR(LB(k, k - m));
        if (!(s > 0.0)) { free(L); free(c); return -2; }
        LB(k, 0) = sqrt(s);
    }
    for (long k = 0; k < n; k++) { /* L y = c */
        double s = c[k];
        for (long m = (k - bw > 0 ? k - bw : 0); m < k; m++) s -= LB(k, k - m) * c[m];
        c[k] = s / LB(k, 0);
    }
    for (long k = n - 1; k >= 0; k--) { /* L^T u = y */
        double s = c[k];
        long mmax = k + bw < n - 1 ? k + bw : n - 1;
        for (long m = k + 1; m <= mmax; m++) s -= LB(m, m - k) * c[m];
        c[k] = s / LB(k, 0);
    }
    memcpy(u, c, sizeof(double) * n);
    free(L); free(c);
#undef LB
#undef DIR
    return 0;
}

/* ======================================================================= */
/* MCC                                                                     */
/* ======================================================================= */

static double inter_EeV(const o_inter *it, double v)   /* particles.hpp:415-420 */
{
    double mu = it->m1 * it->m2 / (it->m1 + it->m2);
    return 0.5 * mu * v * v / O_QE;
}
static double inter_E(const o_inter *it, double v)     /* particles.hpp:421-426 */
{
    double mu = it->m1 * it->m2 / (it->m1 + it->m2);
    return 0.5 * mu * v * v;
}
static double inter_vrel(const o_inter *it, double E)  /* particles.hpp:428-433 */
{
    double mu = it->m1 * it->m2 / (it->m1 + it->m2);
    return sqrt(2 * E / mu);
}
static double coulomb_sigma(const o_inter *it, double EeV) /* particles.cpp:19-26 */
{
    double E = EeV * O_QE;
    double lD = sqrt(O_EPS0 * O_KB * it->t1 / (it->n1 * it->q1 * it->q1));
    double L = it->q1 * it->q2 / (4 * M_PI * O_EPS0 * E);
    return M_PI * L * L * log(lD / L);
}

/* Interaction::sigma_v (particles.hpp:61-70) */
double o_sigma_v(const o_inter *it, double v)
{
    if (it->type == O_COULOMB) return coulomb_sigma(it, inter_EeV(it, v)) * v;
    if (it->ntab > 0) return o_vec_interpolate(it->tabE, it->tabS, it->ntab, inter_EeV(it, v)) * v;
    return it->rate;
}

/* BaseSpecies::svmax_find (particles.cpp:190-206): max of sum sigma_v over
 * v = 0, dv, 2dv, ... < vmax (accumulated loop variable, NaNs skipped) */
double o_svmax_find(const o_inter *its, int n, double vmax, int samples)
{
    double dv = vmax / samples, svmax = 0.0;
    for (double v = 0; v < vmax; v += dv) {
        double sv = 0;
        for (int i = 0; i < n; i++) sv += o_sigma_v(&its[i], v);
        if (isnan(sv)) continue;
        if (sv > svmax) svmax = sv;
    }
    return svmax;
}

/* BaseSpecies::scatter (particles.cpp:208-365) */
int o_scatter(const o_mcc *m, o_particle *p, o_rng *r)
{
    double gamma = o_rng_uni(r) / m->lifetime;
    double tmp = 0.0;
    int sid;
    for (sid = 0; sid < m->n_targets - 1; sid++) {
        tmp += m->targets[sid].rate;
        if (tmp > gamma) break;
    }
    const o_target *T = &m->targets[sid];
    double vr2, vz2, vt2;
    o_particle *partner = NULL;
    if (T->parts == NULL || T->n_parts == 0) {
        /* rndv(vr, vz, vt) (particles.hpp:188-193) */
        vr2 = o_rng_normal(r) * T->vth;
        vz2 = o_rng_normal(r) * T->vth;
        vt2 = o_rng_normal(r) * T->vth;
    } else {
        /* random_particle() (particles.hpp:201-211); bounded here */
        long i = (long)(o_rng_u32(r) % (uint32_t)T->n_parts);
        int tries = 0;
        while (T->parts[i].empty && tries++ < 1000) i = (long)(o_rng_u32(r) % (uint32_t)T->n_parts);
        partner = &T->parts[i];
        vr2 = partner->vx; vz2 = partner->vz; vt2 = partner->vy;
    }
    double v_rel = sqrt(SQR(p->vx - vr2) + SQR(p->vz - vz2) + SQR(p->vy - vt2));
    double m2 = T->mass, mass = m->mass;

    gamma = o_rng_uni(r) * T->rate;
    tmp = 0.0;
    int iid;
    for (iid = 0; iid < T->n_inter; iid++) {
        tmp += o_sigma_v(&T->inter[iid], v_rel) * T->density;
        if (tmp > gamma) break;
    }
    if (iid == T->n_inter) return -1; /* null collision */
    const o_inter *it = &T->inter[iid];
    tmp = 1.0 / (mass + m2);
    switch (it->type) {
    case O_SUPERELASTIC: { /* particles.cpp:268-288 */
        double E = inter_E(it, v_rel) + it->DE;
        double vr = inter_vrel(it, E);
        double rx, rz, ry;
        o_rot_len(r, vr, &rx, &rz, &ry);
        p->vx = (rx * m2 + p->vx * mass + vr2 * m2) * tmp;
        p->vz = (ry * m2 + p->vz * mass + vz2 * m2) * tmp;
        p->vy = (rz * m2 + p->vy * mass + vt2 * m2) * tmp;
        break;
    }
    case O_COULOMB:
    case O_ELASTIC: { /* particles.cpp:290-316 */
        double cx = (p->vx * mass + vr2 * m2) * tmp;
        double cz = (p->vz * mass + vz2 * m2) * tmp;
        double cy = (p->vy * mass + vt2 * m2) * tmp;
        double rx, ry, rz;
        o_rot_len(r, v_rel, &rx, &ry, &rz);
        p->vx = rx * m2 * tmp + cx;
        p->vz = rz * m2 * tmp + cz;
        p->vy = ry * m2 * tmp + cy;
        if (it->type == O_COULOMB && partner) {
            partner->vx = -rx * mass * tmp + cx;
            partner->vz = -rz * mass * tmp + cz;
            partner->vy = -ry * mass * tmp + cy;
        }
        break;
    }
    case O_CX: /* particles.cpp:318-325 */
        p->vx = vr2; p->vz = vz2; p->vy = vt2;
        break;
    case O_LANGEVIN: { /* particles.cpp:327-360 */
        double ax = (p->vx - vr2) * m2 * tmp;
        double ay = (p->vz - vz2) * m2 * tmp;
        double az = (p->vy - vt2) * m2 * tmp;
        double beta = o_radius(r) * it->cutoff;
        if (beta > 1.0) o_deflect(r, o_langevin_chi(beta), &ax, &ay, &az);
        else o_rot(r, &ax, &ay, &az);
        p->vx = ax + (p->vx * mass + vr2 * m2) * tmp;
        p->vz = ay + (p->vz * mass + vz2 * m2) * tmp;
        p->vy = az + (p->vy * mass + vt2 * m2) * tmp;
        break;
    }
    default:
        return -2;
    }
    return iid;
}

/* ======================================================================= */
/* species step                                                            */
/* ======================================================================= */

/* Fields::E (fields.hpp:124-150) */
void o_fields_E(const o_fields *f, double x, double z, double t, double *ex, double *ez)
{
    if (f->geometry_empty && !f->selfconsistent) { *ex = 0.; *ez = f->extern_field; return; }
    double gx = 0, gz = 0;
    o_f2_grad(&f->g, f->u, x, z, &gx, &gz);
    if (f->rf) {
        double rx = 0, rz = 0;
        o_f2_grad(&f->g, f->uRF, x, z, &rx, &rz);
        double ph = o_mod(f->rf_omega * t, 10000000 * M_PI);
        ph = f->rf_amplitude * cos(ph) + f->rf_U0;
        gx += rx * ph; gz += rz * ph;
    }
    *ex = -gx; *ez = -gz;
}

/* Trial uniform of the Philox source: slots 4m..4m+3 share one block,
 * counter (m, step lo, 0, 2 << 24 | step hi), word slot & 3. */
double o_trial_uniform(uint64_t seed, uint32_t stream, uint32_t slot, uint64_t step)
{
    uint32_t key[2] = { (uint32_t)seed, (uint32_t)(seed >> 32) ^ stream };
    uint32_t ctr[4] = { slot >> 2, (uint32_t)step, 0, (2u << 24) | ((uint32_t)(step >> 32) & 0xFFFFFFu) };
    uint32_t out[4];
    o_philox4x32_10(ctr, key, out);
    return ((double)out[slot & 3] + 0.5) * 2.3283064365386963e-10;
}

static o_rng *species_rng(o_species *s, o_rng *local, long slot)
{
    if (s->rng_kind == O_RNG_SHR3) return s->shr3;
    o_rng_philox_begin(local, s->seed, s->stream, (uint32_t)(slot + s->slot0), s->niter, 0);
    return local;
}

/* Fields::B (fields.hpp:152-177) */
int o_fields_B(const o_fields *f, double x, double z, double *Br, double *Bz, double *Bt)
{
    if (!f->b_grid) {
        *Br = f->Br; *Bz = f->Bz; *Bt = f->Bt;
        return 0;
    }
    *Bt = 0.0;
    if (o_f2_interpolate(&f->gB, f->BgR, x, z, Br)) return -2;
    return o_f2_interpolate(&f->gB, f->BgZ, x, z, Bz);
}

/* Fields::u_smooth (fields.cpp:28-113): the 8-fold mirror average on a
 * square grid (symmetry), then the 2dx x 2dx filter at interior nodes within
 * `radius` of the centre.  The reference's filter loop also runs over the
 * last column, reading past the row end; here that column is left alone, as
 * in the device library.  -1 for symmetry on a non-square grid (the
 * reference throws). */
int o_u_smooth(const o_grid2 *g, double *u, int symmetry, double radius)
{
    const int jmax = g->nx, lmax = g->nz, ic = jmax - 1, jc = lmax - 1;
#define U(a, b) u[(size_t)(a) * lmax + (b)]
    if (symmetry) {
        if (ic != jc) return -1;
        for (int i = 0; i <= ic / 2; i++)
            for (int j = 0; j <= i; j++) {
                double sum = 0;
                sum += U(i, j);
                sum += U(ic - i, j);
                sum += U(i, jc - j);
                sum += U(ic - i, jc - j);
                sum += U(jc - j, ic - i);
                sum += U(j, i);
                sum += U(jc - j, i);
                sum += U(j, ic - i);
                sum /= 8.0;
                U(i, j) = sum;
                U(ic - i, j) = sum;
                U(i, jc - j) = sum;
                U(ic - i, jc - j) = sum;
                U(jc - j, ic - i) = sum;
                U(j, i) = sum;
                U(jc - j, i) = sum;
                U(j, ic - i) = sum;
            }
    }
    if (radius > 0) radius = SQR(radius / g->dx);
    double *t = malloc(sizeof(double) * (size_t)jmax * lmax);
    memcpy(t, u, sizeof(double) * (size_t)jmax * lmax);
#define T(a, b) t[(size_t)(a) * lmax + (b)]
    const double icf = ic / 2.0, jcf = jc / 2.0;
    for (int i = 1; i < jmax - 1; i++)
        for (int j = 1; j < lmax - 1; j++) {
            double r = SQR(i - icf) + SQR(j - jcf);
            if (radius > 0 && r > radius) continue;
            double sum = T(i, j) + T(i - 1, j) * 0.5 + T(i + 1, j) * 0.5 + T(i, j - 1) * 0.5 + T(i, j + 1) * 0.5 +
                         T(i - 1, j - 1) * 0.25 + T(i + 1, j - 1) * 0.25 + T(i - 1, j + 1) * 0.25 +
                         T(i + 1, j + 1) * 0.25;
            U(i, j) = sum / 4.0;
        }
#undef T
#undef U
    free(t);
    return 0;
}

/* Species<CARTESIAN>::advance_boris (particles.cpp:924-995) */
void o_advance_boris(const o_fields *f, o_species *s, o_particle *P, long n, long *n_coll)
{
    const double qmdt = s->charge / s->mass * s->dt;
    const double prob = 1.0 - exp(-s->dt / s->lifetime);
    long nc = 0;
    for (long k = 0; k < n; k++) {
        o_particle *I = &P[k];
        if (I->empty) continue;
        double fx, fz;
        o_fields_E(f, I->x, I->z, s->niter * s->dt, &fx, &fz);
        /* field->B(I->x, I->z, Bx, Bz, By) (particles.cpp:946-949) */
        double Bx, Bz, By;
        o_fields_B(f, I->x, I->z, &Bx, &Bz, &By);
        I->vx += fx * qmdt / 2.0;
        I->vz += fz * qmdt / 2.0;
        double tmp = s->charge * s->dt / (2.0 * s->mass);
        double tx = Bx * tmp, ty = By * tmp, tz = Bz * tmp;
        double vpr = I->vx - I->vy * tz + I->vz * ty;
        double vpz = I->vz - I->vx * ty + I->vy * tx;
        double vpt = I->vy - I->vz * tx + I->vx * tz;
        tmp = 2.0 / (1 + SQR(tx) + SQR(ty) + SQR(tz));
        double sx = tx * tmp, sy = ty * tmp, sz = tz * tmp;
        I->vx = I->vx - vpt * sz + vpz * sy;
        I->vy = I->vy - vpz * sx + vpr * sz;
        I->vz = I->vz - vpr * sy + vpt * sx;
        I->vx += fx * qmdt / 2.0;
        I->vz += fz * qmdt / 2.0;
        I->x += I->vx * s->dt;
        I->z += I->vz * s->dt;
        if (s->mcc) {
            o_rng local;
            o_rng *r = species_rng(s, &local, k);
            double u = s->rng_kind == O_RNG_SHR3 ? o_rng_uni(r)
                                                 : o_trial_uniform(s->seed, s->stream,
                                                                   (uint32_t)(k + s->slot0), s->niter);
            if (u < prob) { o_scatter(s->mcc, I, r); nc++; }
        }
    }
    if (n_coll) *n_coll = nc;
}

/* Species<CARTESIAN>::advance_boris_init (particles.cpp:997-1050) */
void o_advance_boris_init(const o_fields *f, o_species *s, o_particle *P, long n)
{
    const double qmdt = -s->charge / s->mass * s->dt;
    for (long k = 0; k < n; k++) {
        o_particle *I = &P[k];
        if (I->empty) continue;
        double fx, fz;
        o_fields_E(f, I->x, I->z, s->niter * s->dt, &fx, &fz);
        double Bx, Bz, By;
        o_fields_B(f, I->x, I->z, &Bx, &Bz, &By);
        double tmp = -0.5 * s->charge * s->dt / (2.0 * s->mass);
        double tx = Bx * tmp, ty = By * tmp, tz = Bz * tmp;
        double vpr = I->vx - I->vy * tz + I->vz * ty;
        double vpz = I->vz - I->vx * ty + I->vy * tx;
        double vpt = I->vy - I->vz * tx + I->vx * tz;
        tmp = 2.0 / (1 + SQR(tx) + SQR(ty) + SQR(tz));
        double sx = tx * tmp, sy = ty * tmp, sz = tz * tmp;
        I->vx = I->vx - vpt * sz + vpz * sy;
        I->vy = I->vy - vpz * sx + vpr * sz;
        I->vz = I->vz - vpr * sy + vpt * sx;
        I->vx += fx * qmdt / 2.0;
        I->vz += fz * qmdt / 2.0;
    }
}

/* Exp(1) variate.  The reference draws it by the ziggurat (t_random::rexp,
 * random.cpp:54); only the distribution matters (statistical parity).  Both
 * sources here invert a uniform, as the device library does. */
double o_rng_rexp(o_rng *r)
{
    double u = o_rng_uni(r);
    if (u <= 0.0) u = 1e-300;
    return -log(u);
}

/* BaseSpecies::lifetime_init (particles.cpp:163-169): time_to_death of the
 * live particles.  Philox: counter (slot, niter, block, 4 << 24). */
void o_ttd_init(const o_species *s, o_particle *P, long n)
{
    for (long k = 0; k < n; k++) {
        o_particle *I = &P[k];
        if (I->empty) continue;
        o_rng local, *r = s->shr3;
        if (s->rng_kind != O_RNG_SHR3) {
            o_rng_philox_begin(&local, s->seed, s->stream, (uint32_t)(k + s->slot0), s->niter, 4);
            r = &local;
        }
        I->time_to_death = o_rng_rexp(r) * s->lifetime;
    }
}

/* Species<CARTESIAN>::advance_multicoll (particles.cpp:813-861).  Returns 0,
 * -1 for B != 0, -2 for a selfconsistent or non-EMPTY run (the reference
 * throws there).  Philox: one stream per slot and step, counter (slot,
 * niter, block, 3 << 24), drawn in event order (scatter, then the next
 * time_to_death).  s->mcc NULL: every event is a null collision.  niter is
 * left to the caller. */
int o_advance_multicoll(const o_fields *f, o_species *s, o_particle *P, long n, long *n_events)
{
    if (f->Br != 0. || f->Bz != 0. || f->Bt != 0.) return -1;
    if (f->selfconsistent || !f->geometry_empty) return -2;
    double fx, fz;
    o_fields_E(f, 0., 0., 0., &fx, &fz);
    const double qm = s->charge / s->mass;
    const double dt = s->dt;
    long ne = 0;
    for (long k = 0; k < n; k++) {
        o_particle *I = &P[k];
        if (I->empty) continue;
        o_rng local, *r = s->shr3;
        if (s->rng_kind != O_RNG_SHR3) {
            o_rng_philox_begin(&local, s->seed, s->stream, (uint32_t)(k + s->slot0), s->niter, 3);
            r = &local;
        }
        double local_time = 0., ttd;
        while (local_time + I->time_to_death < dt) {
            ttd = I->time_to_death;
            I->vx += fx * qm * ttd;
            I->vz += fz * qm * ttd;
            I->x += (I->vx + 0.5 * fx * qm * ttd) * ttd;
            I->z += (I->vz + 0.5 * fz * qm * ttd) * ttd;
            local_time += I->time_to_death;
            if (s->mcc) o_scatter(s->mcc, I, r);
            I->time_to_death = s->lifetime * o_rng_rexp(r);
            ne++;
        }
        ttd = dt - local_time;
        I->vx += fx * qm * ttd;
        I->vz += fz * qm * ttd;
        I->x += (I->vx + 0.5 * fx * qm * ttd) * ttd;
        I->z += (I->vz + 0.5 * fz * qm * ttd) * ttd;
        I->time_to_death -= ttd;
    }
    if (n_events) *n_events = ne;
    return 0;
}

/* Species<D>::advance_boundary (particles.hpp:370-411) */
long o_advance_boundary(const o_fields *f, o_species *s, o_particle *P, long n, long *n_bad)
{
    long removed = 0, bad = 0;
    for (long k = 0; k < n; k++) {
        o_particle *I = &P[k];
        if (I->empty) continue;
        if (I->x < f->x_min || I->x > f->x_max || I->z < f->z_min || I->z > f->z_max) {
            if (f->boundary == O_BC_FREE) { I->empty = 1; removed++; continue; }
            I->x = fmod(I->x, f->x_max);
            if (I->x < 0) I->x += f->x_max;
            I->z = fmod(I->z, f->z_max);
            if (I->z < 0) I->z += f->z_max;
        }
        if (!f->electric_field_from_file && !o_is_free(&f->g, f->mask, I->x, I->z)) {
            I->empty = 1; removed++; continue;
        }
        if (f->selfconsistent && s->rho)
            if (o_f2_accumulate(&f->g, s->rho, s->charge, I->x, I->z)) bad++;
    }
    if (n_bad) *n_bad = bad;
    return removed;
}

/* Species<D>::accumulate (particles.hpp:358-367) */
long o_accumulate(const o_fields *f, o_species *s, const o_particle *P, long n)
{
    long bad = 0;
    for (long k = 0; k < n; k++)
        if (!P[k].empty && o_f2_accumulate(&f->g, s->rho, s->charge, P[k].x, P[k].z)) bad++;
    return bad;
}

/* energy_dist_compute + Histogram::mean_tot (particles.cpp:408-414,
 * histogram.cpp:22-35,68-71): mean of |v|^2 m / 2 / q_e over live particles */
double o_mean_energy_eV(const o_particle *P, long n, double mass)
{
    double sum = 0; long cnt = 0;
    for (long k = 0; k < n; k++)
        if (!P[k].empty) { sum += (SQR(P[k].vx) + SQR(P[k].vz) + SQR(P[k].vy)) * mass * 0.5 / O_QE; cnt++; }
    return cnt ? sum / cnt : 0.0;
}

/* Species<CARTESIAN>::add_particles_on_disk (particles.cpp:728-749) */
long o_add_particles_on_disk(o_particle *P, long cap, long n_req, double cx, double cz,
                             double radius, double vth, const o_fields *f, o_rng *r)
{
    long n = 0;
    for (long i = 0; i < n_req && n < cap; i++) {
        double x, y;
        do { x = o_rng_uni(r) - 0.5; y = o_rng_uni(r) - 0.5; } while (SQR(x) + SQR(y) > 0.25);
        x = x * 2 * radius + cx;
        y = y * 2 * radius + cz;
        if (x < f->x_min || x > f->x_max || y < f->z_min || y > f->z_max) continue;
        o_particle *p = &P[n++];
        memset(p, 0, sizeof(*p));
        p->x = x; p->z = y;
        p->vx = o_rng_normal(r) * vth;
        p->vz = o_rng_normal(r) * vth;
        p->vy = o_rng_normal(r) * vth;
    }
    return n;
}

/* The same sampler with one Philox stream per slot (purpose 1, step 0) and
 * samples outside the box left as empty slots instead of skipped: the
 * slot-parallel form the device library uses. */
void o_add_particles_on_disk_philox(o_particle *P, long n, double cx, double cz, double radius,
                                    double vth, const o_fields *f, uint64_t seed, uint32_t stream)
{
    for (long k = 0; k < n; k++) {
        o_rng r;
        o_rng_philox_begin(&r, seed, stream, (uint32_t)k, 0, 1);
        double x = 1, y = 1;
        for (int t = 0; t < 64 && SQR(x) + SQR(y) > 0.25; t++) {
            x = o_rng_uni(&r) - 0.5;
            y = o_rng_uni(&r) - 0.5;
        }
        x = x * 2 * radius + cx;
        y = y * 2 * radius + cz;
        o_particle *p = &P[k];
        memset(p, 0, sizeof(*p));
        if (x < f->x_min || x > f->x_max || y < f->z_min || y > f->z_max) { p->empty = 1; continue; }
        p->x = x; p->z = y;
        p->vx = o_rng_normal(&r) * vth;
        p->vz = o_rng_normal(&r) * vth;
        p->vy = o_rng_normal(&r) * vth;
    }
}

/* add_particles_bessel (particles.cpp:701-726) on the Philox streams of the
 * device sampler (slot k: counter (k, 0, block, 1 << 24)); J0 is the C
 * library's j0, where the reference calls std::tr1::cyl_bessel_j(0, .) */
void o_add_particles_bessel_philox(o_particle *P, long n, double cx, double cz, double radius,
                                   double vth, const o_fields *f, uint64_t seed, uint32_t stream)
{
    const double bessel_root = 2.404825557695773;
    for (long k = 0; k < n; k++) {
        o_rng r;
        o_rng_philox_begin(&r, seed, stream, (uint32_t)k, 0, 1);
        double x = 0, y = 0;
        for (int t = 0; t < 256; t++) {
            x = o_rng_uni(&r) * 2 - 1.0;
            y = o_rng_uni(&r) * 2 - 1.0;
            double rr = sqrt(SQR(x) + SQR(y));
            if (rr > 1.0) continue;
            if (!(o_rng_uni(&r) > j0(rr * bessel_root))) break;
        }
        x = x * radius + cx;
        y = y * radius + cz;
        o_particle *p = &P[k];
        memset(p, 0, sizeof(*p));
        if (x < f->x_min || x > f->x_max || y < f->z_min || y > f->z_max) { p->empty = 1; continue; }
        p->x = x; p->z = y;
        p->vx = o_rng_normal(&r) * vth;
        p->vz = o_rng_normal(&r) * vth;
        p->vy = o_rng_normal(&r) * vth;
    }
}

/* add_particles_everywhere (particles.cpp:686-699) on the same streams: x, y,
 * z uniform in the box (y is drawn and dropped: the Cartesian store has no
 * y), rnor velocities vx, vy, vz, then time_to_death = rexp() * lifetime */
void o_add_particles_everywhere_philox(o_particle *P, long n, double vth, double lifetime,
                                       const o_fields *f, uint64_t seed, uint32_t stream)
{
    for (long k = 0; k < n; k++) {
        o_rng r;
        o_rng_philox_begin(&r, seed, stream, (uint32_t)k, 0, 1);
        o_particle *p = &P[k];
        memset(p, 0, sizeof(*p));
        p->x = (f->x_max - f->x_min) * o_rng_uni(&r) + f->x_min;
        (void)o_rng_uni(&r);
        p->z = (f->z_max - f->z_min) * o_rng_uni(&r) + f->z_min;
        p->vx = o_rng_normal(&r) * vth;
        p->vy = o_rng_normal(&r) * vth;
        p->vz = o_rng_normal(&r) * vth;
        p->time_to_death = o_rng_rexp(&r) * lifetime;
    }
}

/* ======================================================================= */
/* Field3D                                                                 */
/* ======================================================================= */

#define F3(d, g, a, b, c) (d)[((size_t)(a) * (g)->ny + (b)) * (g)->nz + (c)]

/* Field3D::accumulate (Field3D.hpp:40-66), 8-point CIC; a stencil leaving
 * the grid returns -2 (the reference accepts i == imax-1 and writes past it) */
int o_f3_accumulate(const o_grid3 *g, double *d, double q, double x, double y, double z)
{
    x -= g->xmin; y -= g->ymin; z -= g->zmin;
    if (!(x >= 0) || !(y >= 0) || !(z >= 0)) return -2;
    int i = (int)(x * g->idx), j = (int)(y * g->idy), k = (int)(z * g->idz);
    double u = x * g->idx - i, v = y * g->idy - j, w = z * g->idz - k;
    if (i > g->nx - 2 || j > g->ny - 2 || k > g->nz - 2) return -2;
    F3(d, g, i, j, k) += (1 - u) * (1 - v) * (1 - w) * q;
    F3(d, g, i + 1, j, k) += u * (1 - v) * (1 - w) * q;
    F3(d, g, i, j + 1, k) += (1 - u) * v * (1 - w) * q;
    F3(d, g, i + 1, j + 1, k) += u * v * (1 - w) * q;
    F3(d, g, i, j, k + 1) += (1 - u) * (1 - v) * w * q;
    F3(d, g, i + 1, j, k + 1) += u * (1 - v) * w * q;
    F3(d, g, i, j + 1, k + 1) += (1 - u) * v * w * q;
    F3(d, g, i + 1, j + 1, k + 1) += u * v * w * q;
    return 0;
}

/* Field3D::interpolate (Field3D.hpp:68-96) */
int o_f3_interpolate(const o_grid3 *g, const double *d, double x, double y, double z, double *out)
{
    x -= g->xmin; y -= g->ymin; z -= g->zmin;
    if (!(x >= 0) || !(y >= 0) || !(z >= 0)) return -2;
    int i = (int)(x * g->idx), j = (int)(y * g->idy), k = (int)(z * g->idz);
    double u = x * g->idx - i, v = y * g->idy - j, w = z * g->idz - k;
    if (i > g->nx - 2 || j > g->ny - 2 || k > g->nz - 2) return -2;
    double r = 0;
    r += (1 - u) * (1 - v) * (1 - w) * F3(d, g, i, j, k);
    r += u * (1 - v) * (1 - w) * F3(d, g, i + 1, j, k);
    r += (1 - u) * v * (1 - w) * F3(d, g, i, j + 1, k);
    r += u * v * (1 - w) * F3(d, g, i + 1, j + 1, k);
    r += (1 - u) * (1 - v) * w * F3(d, g, i, j, k + 1);
    r += u * (1 - v) * w * F3(d, g, i + 1, j, k + 1);
    r += (1 - u) * v * w * F3(d, g, i, j + 1, k + 1);
    r += u * v * w * F3(d, g, i + 1, j + 1, k + 1);
    *out = r;
    return 0;
}

static double clampd(double x, double lo, double hi) { return x > hi ? hi : (x < lo ? lo : x); }

/* Field3D::grad_component (Field3D.hpp:112-141): coordinates clamped half a
 * cell inside along the differenced axis, staggered differences on the 8
 * corners, trilinear weights.  xmax/ymax/zmax are the grid extent (the
 * reference never initialises them, Field3D.hpp:37), and the lower corner
 * index is capped at n-2 so the upper corner stays inside the grid (the
 * reference can read one past the end when the clamp lands on the edge). */
static double f3_grad_component(const o_grid3 *g, const double *d, double x, double y, double z,
                                int dx_, int dy_, int dz_)
{
    double X = clampd(x * g->idx, 0.5 * dx_, g->xmax * g->idx - 0.5 * dx_) + 0.5 * dx_;
    double Y = clampd(y * g->idy, 0.5 * dy_, g->ymax * g->idy - 0.5 * dy_) + 0.5 * dy_;
    double Z = clampd(z * g->idz, 0.5 * dz_, g->zmax * g->idz - 0.5 * dz_) + 0.5 * dz_;
    int i = (int)X, j = (int)Y, k = (int)Z;
    if (i > g->nx - 2) i = g->nx - 2;
    if (j > g->ny - 2) j = g->ny - 2;
    if (k > g->nz - 2) k = g->nz - 2;
    double u = X - i, v = Y - j, w = Z - k;
    double gg[8];
    for (int c = 0; c < 8; c++) {
        int a = i + (c & 1), b = j + ((c >> 1) & 1), e = k + ((c >> 2) & 1);
        gg[c] = F3(d, g, a, b, e) - F3(d, g, a - dx_, b - dy_, e - dz_);
    }
    double r = (1 - u) * (1 - v) * (1 - w) * gg[0] + u * (1 - v) * (1 - w) * gg[1]
             + (1 - u) * v * (1 - w) * gg[2] + u * v * (1 - w) * gg[3]
             + (1 - u) * (1 - v) * w * gg[4] + u * (1 - v) * w * gg[5]
             + (1 - u) * v * w * gg[6] + u * v * w * gg[7];
    return r * (g->idx * dx_ + g->idy * dy_ + g->idz * dz_);
}

/* Field3D::grad (Field3D.hpp:157-162) */
void o_f3_grad(const o_grid3 *g, const double *d, double x, double y, double z,
               double *gx, double *gy, double *gz)
{
    *gx = f3_grad_component(g, d, x, y, z, 1, 0, 0);
    *gy = f3_grad_component(g, d, x, y, z, 0, 1, 0);
    *gz = f3_grad_component(g, d, x, y, z, 0, 0, 1);
}

/* Solver::matrix_init rows (fields3d.cpp:35-76), transposed solve as in 2D:
 * identity on FIXED (mask 0) and electrode (mask < 0, here mask values 0/1),
 * 7-point [1 x6, -6] otherwise */
int o_poisson3_apply(const o_grid3 *g, const uint8_t *mask, const double *x, double *y)
{
    const int X = g->nx, Y = g->ny, Z = g->nz;
    const long sy = Z, sx = (long)Y * Z;
    for (int i = 0; i < X; i++)
        for (int j = 0; j < Y; j++)
            for (int k = 0; k < Z; k++) {
                long m = i * sx + j * sy + k;
                if (mask[m] == O_FIXED || mask[m] == O_FIXED_RF) { y[m] = x[m]; continue; }
                if (i == 0 || j == 0 || k == 0 || i == X - 1 || j == Y - 1 || k == Z - 1) return -1;
                y[m] = x[m - sx] + x[m - sy] + x[m - 1] - 6.0 * x[m] + x[m + 1] + x[m + sy] + x[m + sx];
            }
    return 0;
}

/* Solver::solve rhs (fields3d.cpp:84-93) */
void o_poisson3_rhs(const o_grid3 *g, const uint8_t *mask, const double *voltage, double *rho,
                    double scale)
{
    const long n = (long)g->nx * g->ny * g->nz;
    for (long m = 0; m < n; m++)
        rho[m] = (mask[m] == O_FIXED || mask[m] == O_FIXED_RF) ? voltage[m] : rho[m] * -scale;
}

/* Geometry::is_free (fields3d.hpp:48-55) */
int o_is_free3(const o_grid3 *g, const uint8_t *mask, double x, double y, double z)
{
    if (!(x >= 0.0) || !(y >= 0.0) || !(z >= 0.0)) return 0;
    int i = (int)(x * g->idx), j = (int)(y * g->idy), k = (int)(z * g->idz);
    if (i > g->nx - 2 || j > g->ny - 2 || k > g->nz - 2) return 0;
    for (int c = 0; c < 8; c++)
        if (F3(mask, g, i + (c & 1), j + ((c >> 1) & 1), k + ((c >> 2) & 1)) == O_FREE) return 1;
    return 0;
}

static void boris3(double qmdt2, double tx, double ty, double tz, double ex, double ey, double ez,
                   o_particle *I)
{
    I->vx += ex * qmdt2;
    I->vy += ey * qmdt2;
    I->vz += ez * qmdt2;
    /* species3d.cpp:30-50 */
    double vpx = I->vx - I->vy * tz + I->vz * ty;
    double vpy = I->vy - I->vz * tx + I->vx * tz;
    double vpz = I->vz - I->vx * ty + I->vy * tx;
    double tmp = 2.0 / (1 + SQR(tx) + SQR(ty) + SQR(tz));
    double sx = tx * tmp, sy = ty * tmp, sz = tz * tmp;
    I->vx = I->vx - vpy * sz + vpz * sy;
    I->vy = I->vy - vpz * sx + vpx * sz;
    I->vz = I->vz - vpx * sy + vpy * sx;
}

/* Species<CARTESIAN3D>::advance (species3d.cpp:3-94) */
long o_advance3d(const o_fields3 *f, o_species *s, o_particle *P, long n, long *n_coll, long *n_bad)
{
    const double qmdt = s->charge / s->mass * s->dt;
    const double prob = 1.0 - exp(-s->dt / s->lifetime);
    const double tmp = s->charge * s->dt / (2.0 * s->mass);
    long nc = 0, removed = 0, bad = 0;
    for (long k = 0; k < n; k++) {
        o_particle *I = &P[k];
        if (I->empty) continue;
        double ex, ey, ez;
        o_f3_grad(&f->g, f->u, I->x, I->y, I->z, &ex, &ey, &ez);  /* ElMag3D::E */
        ex = -ex; ey = -ey; ez = -ez;
        /* half kick, rotation, half kick (species3d.cpp:24-56) */
        boris3(qmdt / 2.0, f->Bx * tmp, f->By * tmp, f->Bz * tmp, ex, ey, ez, I);
        I->vx += ex * qmdt / 2.0;
        I->vy += ey * qmdt / 2.0;
        I->vz += ez * qmdt / 2.0;
        I->x += I->vx * s->dt;
        I->y += I->vy * s->dt;
        I->z += I->vz * s->dt;
        if (s->mcc) {
            o_rng local;
            o_rng *r = species_rng(s, &local, k);
            double u = s->rng_kind == O_RNG_SHR3 ? o_rng_uni(r)
                                                 : o_trial_uniform(s->seed, s->stream,
                                                                   (uint32_t)(k + s->slot0), s->niter);
            if (u < prob) { o_scatter(s->mcc, I, r); nc++; }
        }
        if (I->x > f->x_max || I->x < 0 || I->y > f->y_max || I->y < 0 || I->z > f->z_max || I->z < 0) {
            if (f->boundary == O_BC_FREE) { I->empty = 1; removed++; continue; }
            I->x = o_mod(I->x, f->x_max);
            I->y = o_mod(I->y, f->y_max);
            I->z = o_mod(I->z, f->z_max);
        }
        if (!o_is_free3(&f->g, f->mask, I->x, I->y, I->z)) { I->empty = 1; removed++; continue; }
        if (s->rho && o_f3_accumulate(&f->g, s->rho, s->charge, I->x, I->y, I->z)) bad++;
    }
    if (n_coll) *n_coll = nc;
    if (n_bad) *n_bad = bad;
    return removed;
}

void o_advance3d_init(const o_fields3 *f, o_species *s, o_particle *P, long n)
{
    const double qmdt = -s->charge / s->mass * s->dt;
    const double tmp = -0.5 * s->charge * s->dt / (2.0 * s->mass);
    for (long k = 0; k < n; k++) {
        o_particle *I = &P[k];
        if (I->empty) continue;
        double ex, ey, ez;
        o_f3_grad(&f->g, f->u, I->x, I->y, I->z, &ex, &ey, &ez);
        ex = -ex; ey = -ey; ez = -ez;
        boris3(0.0, f->Bx * tmp, f->By * tmp, f->Bz * tmp, 0, 0, 0, I);
        I->vx += ex * qmdt / 2.0;
        I->vy += ey * qmdt / 2.0;
        I->vz += ez * qmdt / 2.0;
    }
}

/* Pic<CARTESIAN>::advance_init / advance with a solve every step
 * (pic.cpp:330-384) for one species and no collisions, from a store already
 * started by advance_boris_init: rho^0 = accumulate(p); then per step
 * u = boundary_solve(rho), reset, advance_boris, advance_boundary (deposit).
 * The band Cholesky factor of the Dirichlet-eliminated matrix is built once.
 * Returns the mean kinetic energy (eV) of the live particles at the end, or
 * NaN when the matrix cannot be factored. */
double o_selfconsistent_run(o_fields *f, o_species *s, o_particle *P, long np, const double *voltage,
                            double dV, int steps)
{
    const o_grid2 *g = &f->g;
    const int M = g->nx, N = g->nz;
    const long n = (long)M * N;
    const int bw = N;
    const uint8_t *mask = f->mask;
    double *L = (double *)calloc((size_t)n * (bw + 1), sizeof(double));
    double *c = (double *)malloc(sizeof(double) * n);
    double *rho = (double *)calloc((size_t)n, sizeof(double));
    double *u = (double *)f->u;
    if (!L || !c || !rho) { free(L); free(c); free(rho); return NAN; }
#define LB(k, d) L[(size_t)(k) * (bw + 1) + (d)]
#define DIR(k) (mask[k] == O_FIXED || mask[k] == O_FIXED_RF)
    for (long k = 0; k < n; k++) {
        if (DIR(k)) { LB(k, 0) = 1.0; continue; }
        LB(k, 0) = 4.0;
        if (!DIR(k - N)) LB(k, N) = -1.0;
        if (!DIR(k - 1)) LB(k, 1) = -1.0;
    }
    for (long k = 0; k < n; k++) {
        long dmax = k < bw ? k : bw;
        for (long d = dmax; d >= 1; d--) {
            long j = k - d;
            double sm = LB(k, d);
            long m0 = k - bw > j - bw ? k - bw : j - bw;
            if (m0 < 0) m0 = 0;
            for (long m = m0; m < j; m++) sm -= LB(k, k - m) * LB(j, j - m);
            LB(k, d) = sm / LB(j, 0);
        }
        double sm = LB(k, 0);
        for (long m = (k - bw > 0 ? k - bw : 0); m < k; m++) sm -= SQR(LB(k, k - m));
        if (!(sm > 0.0)) { free(L); free(c); free(rho); return NAN; }
        LB(k, 0) = sqrt(sm);
    }
    double *sr = s->rho;
    s->rho = rho;
    o_accumulate(f, s, P, np);
    const double scale = -(g->dx * g->dx) / O_EPS0 / dV;
    for (int st = 0; st < steps; st++) {
        /* boundary_solve (fields.cpp:278-312) on the factor */
        for (long k = 0; k < n; k++) {
            /* STATIC rows (fields.cpp:297-310): FIXED -> voltage, FIXED_RF -> 0 */
            if (DIR(k)) { c[k] = mask[k] == O_FIXED ? voltage[k] : 0.0; continue; }
            double rhs = rho[k] * scale;
            const long nb[4] = {k - N, k - 1, k + 1, k + N};
            for (int q = 0; q < 4; q++)
                if (DIR(nb[q])) rhs -= mask[nb[q]] == O_FIXED ? voltage[nb[q]] : 0.0;
            c[k] = -rhs;
        }
        for (long k = 0; k < n; k++) {
            double sm = c[k];
            for (long m = (k - bw > 0 ? k - bw : 0); m < k; m++) sm -= LB(k, k - m) * c[m];
            c[k] = sm / LB(k, 0);
        }
        for (long k = n - 1; k >= 0; k--) {
            double sm = c[k];
            long mmax = k + bw < n - 1 ? k + bw : n - 1;
            for (long m = k + 1; m <= mmax; m++) sm -= LB(m, m - k) * c[m];
            c[k] = sm / LB(k, 0);
        }
        memcpy(u, c, sizeof(double) * n);
        memset(rho, 0, sizeof(double) * n);
        o_advance_boris(f, s, P, np, NULL);
        o_advance_boundary(f, s, P, np, NULL);
        s->niter++;
    }
#undef LB
#undef DIR
    s->rho = sr;
    free(L); free(c); free(rho);
    return o_mean_energy_eV(P, np, s->mass);
}

/* ======================================================================= */
/* CPU baseline timing                                                     */
/* ======================================================================= */

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

/* Species<D>::advance (particles.hpp:342-349) n_steps times, serial, as
 * plasma2d runs it: boris + MCC with the shared SHR3 stream, then the
 * boundary pass with deposit when selfconsistent. */
double o_bench_steps(const o_fields *f, o_species *s, o_particle *p, long n, int n_steps)
{
    double t0 = now_s();
    for (int st = 0; st < n_steps; st++) {
        o_advance_boris(f, s, p, n, NULL);
        o_advance_boundary(f, s, p, n, NULL);
        s->niter++;
    }
    return now_s() - t0;
}

/* OpenMP variant (the tests/test_particles_omp.cpp:70 idiom): parallel for
 * over particles, Philox per particle, private rho per thread summed at the
 * end of each step. */
double o_bench_steps_omp(const o_fields *f, o_species *s, o_particle *P, long n, int n_steps,
                         int threads)
{
#ifdef _OPENMP
    const long nc = (long)f->g.nx * f->g.nz;
    const int T = threads > 0 ? threads : omp_get_max_threads();
    double *priv = NULL;
    if (f->selfconsistent && s->rho) priv = (double *)calloc((size_t)T * nc, sizeof(double));
    double t0 = now_s();
    for (int st = 0; st < n_steps; st++) {
#pragma omp parallel num_threads(T)
        {
            int tid = omp_get_thread_num();
            o_species ls = *s;
            ls.rng_kind = O_RNG_PHILOX;
            ls.rho = priv ? priv + (size_t)tid * nc : NULL;
            long lo = n * tid / T, hi = n * (tid + 1) / T;
            ls.slot0 = s->slot0 + lo;
            o_advance_boris(f, &ls, P + lo, hi - lo, NULL);
            o_advance_boundary(f, &ls, P + lo, hi - lo, NULL);
        }
        if (priv) {
#pragma omp parallel for num_threads(T)
            for (long k = 0; k < nc; k++) {
                double a = 0;
                for (int t = 0; t < T; t++) { a += priv[(size_t)t * nc + k]; priv[(size_t)t * nc + k] = 0; }
                s->rho[k] += a;
            }
        }
        s->niter++;
    }
    double el = now_s() - t0;
    free(priv);
    return el;
#else
    (void)threads;
    return o_bench_steps(f, s, P, n, n_steps);
#endif
}
