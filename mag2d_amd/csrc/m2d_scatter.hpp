// m2d_scatter.hpp -- BaseSpecies::scatter (particles.cpp:208-365) on the
// device, shared by the 2D and 3D particle passes.
#pragma once
#include "m2d_objects.hpp"
#include <cmath>

namespace m2d {

// ------------------------------------------------------------------ MCC
__device__ __forceinline__ double vec_interp(const double *xs, const double *ys, int n, double x)
{
    // vec_interpolate::operator() (tabulate.cpp:124-141)
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

// Interaction::coulomb_sigma (particles.cpp:19-26) at the relative energy
// E (J): pi L^2 ln(lambda_D / L), L = q1 q2 / (4 pi eps0 E)
__host__ __device__ inline double coulomb_sigma(const McInter &it, double E)
{
    const double L = it.qq / E;
    return M_PI * L * L * log(it.lambda_D / L);
}

__device__ __forceinline__ double sigma_v(const McDev &m, const McInter &it, double v)
{
    // Interaction::sigma_v (particles.hpp:61-70)
    if (it.type == M2D_COULOMB) return coulomb_sigma(it, 0.5 * it.mu * v * v) * v;
    if (it.tab_len > 0)
        return vec_interp(m.tabE + it.tab_off, m.tabS + it.tab_off, it.tab_len,
                          0.5 * it.mu * v * v / kQE) * v;
    return it.rate;
}

__device__ __forceinline__ void rot_len(Rng &r, double len, double &x, double &y, double &z)
{
    // t_random::rot(len, x, y, z) (random.cpp:103-115)
    double cs = 1.0 - 2.0 * r.uni();
    x = len * cs;
    cs = sqrt(1.0 - cs * cs);
    double sp, cp;
    sincos(2.0 * M_PI * r.uni(), &sp, &cp);
    y = len * cs * sp;
    z = len * cs * cp;
}

__device__ __forceinline__ double ellint_k(double k)
{
    // complete elliptic integral K(modulus k) by the arithmetic-geometric mean
    double a = 1.0, b = sqrt(1.0 - k * k);
    for (int i = 0; i < 40 && fabs(a - b) > 1e-16 * a; i++) {
        double an = 0.5 * (a + b);
        b = sqrt(a * b);
        a = an;
    }
    return M_PI / (a + b);
}

__device__ inline void deflect(Rng &r, double angle, double &x, double &y, double &z)
{
    // t_random::deflect (random.cpp:133-175)
    double len = tan(angle / 2.0);
    double t = 1.0 - 2.0 * r.uni();
    double x1 = len * t;
    t = sqrt(1.0 - t * t);
    double sp, cp;
    sincos(2.0 * M_PI * r.uni(), &sp, &cp);
    double y1 = len * t * sp, z1 = len * t * cp;
    double tx = y * z1 - z * y1, ty = z * x1 - x * z1, tz = x * y1 - y * x1;
    t = len / sqrt(tx * tx + ty * ty + tz * tz);
    tx *= t; ty *= t; tz *= t;
    double xp = x - y * tz + z * ty, yp = y - z * tx + x * tz, zp = z - x * ty + y * tx;
    t = 2.0 / (1.0 + len * len);
    x1 = tx * t; y1 = ty * t; z1 = tz * t;
    x += -yp * z1 + zp * y1;
    y += -zp * x1 + xp * z1;
    z += -xp * y1 + yp * x1;
}

// BaseSpecies::scatter (particles.cpp:208-365), FP64, same draw order as
// the reference: species pick, partner, interaction pick, angles.  The
// velocity changes only when an interaction fires.
__device__ inline void scatter_d(const McDev &m, Rng &r, double &vx, double &vy, double &vz,
                                 unsigned long long *counters)
{
    double gamma = r.uni() / m.lifetime;
    double tmp = 0.0;
    int sid;
    for (sid = 0; sid < m.n_targets - 1; sid++) {
        tmp += m.t[sid].rate;
        if (tmp > gamma) break;
    }
    const McTarget &T = m.t[sid];
    double vr2, vz2, vt2;
    int64_t partner = -1;  // slot of the live partner (random_particle, particles.hpp:201-211)
    if (T.px == nullptr || T.pn == 0) {
        vr2 = r.normal() * T.vth;
        vz2 = r.normal() * T.vth;
        vt2 = r.normal() * T.vth;
    } else {
        int64_t i = (int64_t)(r.u32() % (uint32_t)T.pn);
        int tries = 0;
        while (is_dead(T.px[i]) && tries++ < 1000) i = (int64_t)(r.u32() % (uint32_t)T.pn);
        if (is_dead(T.px[i])) { raise_flag(counters, M2D_FLAG_PARTNER); return; }
        vr2 = T.pvx[i]; vz2 = T.pvz[i]; vt2 = T.pvy[i];
        partner = i;
    }
    double dvx = vx - vr2, dvz = vz - vz2, dvy = vy - vt2;
    double v_rel = sqrt(dvx * dvx + dvz * dvz + dvy * dvy);
    double m2 = T.mass, mass = m.mass;

    gamma = r.uni() * T.rate;
    tmp = 0.0;
    int k;
    for (k = 0; k < T.n; k++) {
        tmp += sigma_v(m, m.in[T.first + k], v_rel) * T.density;
        if (tmp > gamma) break;
    }
    if (k == T.n) return;  // null collision
    const McInter &it = m.in[T.first + k];
    tmp = 1.0 / (mass + m2);
    switch (it.type) {
    case M2D_SUPERELASTIC: {
        double E = 0.5 * it.mu * v_rel * v_rel + it.DE;
        double vr = sqrt(2 * E / it.mu);
        double rx, rz, ry;
        rot_len(r, vr, rx, rz, ry);
        vx = (rx * m2 + vx * mass + vr2 * m2) * tmp;
        double nvz = (ry * m2 + vz * mass + vz2 * m2) * tmp;
        double nvy = (rz * m2 + vy * mass + vt2 * m2) * tmp;
        vz = nvz; vy = nvy;
        break;
    }
    case M2D_COULOMB:
    case M2D_ELASTIC: {
        // isotropic in the centre of mass (particles.cpp:290-316)
        double cx = (vx * mass + vr2 * m2) * tmp;
        double cz = (vz * mass + vz2 * m2) * tmp;
        double cy = (vy * mass + vt2 * m2) * tmp;
        double rx, ry, rz;
        rot_len(r, v_rel, rx, ry, rz);
        vx = rx * m2 * tmp + cx;
        vz = rz * m2 * tmp + cz;
        vy = ry * m2 * tmp + cy;
        // COULOMB also turns the partner (particles.cpp:309-314); a partner
        // drawn from the continuum has no state to change
        if (it.type == M2D_COULOMB && partner >= 0) {
            T.pvx[partner] = (float)(-rx * mass * tmp + cx);
            T.pvz[partner] = (float)(-rz * mass * tmp + cz);
            T.pvy[partner] = (float)(-ry * mass * tmp + cy);
        }
        break;
    }
    case M2D_CX:
        vx = vr2; vz = vz2; vy = vt2;
        break;
    case M2D_LANGEVIN: {
        double ax = (vx - vr2) * m2 * tmp, ay = (vz - vz2) * m2 * tmp, az = (vy - vt2) * m2 * tmp;
        double beta = sqrt(r.uni()) * it.cutoff;
        if (beta > 1.0) {
            // Nanbu & Kitatani, J. Phys. D 28 (1995) 324 (particles.cpp:337-345)
            double t4 = sqrt(beta * beta * beta * beta - 1.0);
            double xi0 = sqrt(beta * beta - t4), xi1 = sqrt(beta * beta + t4);
            double theta = ellint_k(xi0 / xi1) * M_SQRT2 * beta / xi1;
            deflect(r, M_PI - 2 * theta, ax, ay, az);
        } else {
            double len = sqrt(ax * ax + ay * ay + az * az);
            rot_len(r, len, ax, ay, az);
        }
        double nvx = ax + (vx * mass + vr2 * m2) * tmp;
        double nvz = ay + (vz * mass + vz2 * m2) * tmp;
        double nvy = az + (vy * mass + vt2 * m2) * tmp;
        vx = nvx; vz = nvz; vy = nvy;
        break;
    }
    default:
        return;
    }
}

// the same on an FP32 particle state
__device__ inline void scatter(const McDev &m, Rng &r, float &vxf, float &vyf, float &vzf,
                               unsigned long long *counters)
{
    double vx = vxf, vy = vyf, vz = vzf;
    scatter_d(m, r, vx, vy, vz, counters);
    vxf = (float)vx; vyf = (float)vy; vzf = (float)vz;
}

}  // namespace m2d
