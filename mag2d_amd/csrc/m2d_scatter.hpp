// m2d_scatter.hpp -- BaseSpecies::scatter (particles.cpp:208-365) on the
// device, shared by the 2D and 3D particle passes and the adjoint.
#pragma once
#include "m2d_objects.hpp"
#include <cmath>

namespace m2d {

// ------------------------------------------------------------------ derivatives
// A value with its derivatives with respect to three inputs (forward mode).
// The scatter below is written once over its velocity type: as double it is
// the collision itself, as Dual3 seeded with the unit vectors it also gives
// the 3 x 3 Jacobian of the new velocity with respect to the old one for the
// same draws (the random numbers, and every branch taken on a value, are the
// same in both).
struct Dual3 {
    double v, d[3];
};
__device__ __forceinline__ Dual3 mk(double v, double d0, double d1, double d2) { return Dual3{v, {d0, d1, d2}}; }
__device__ __forceinline__ Dual3 operator+(Dual3 a, Dual3 b)
{
    return mk(a.v + b.v, a.d[0] + b.d[0], a.d[1] + b.d[1], a.d[2] + b.d[2]);
}
__device__ __forceinline__ Dual3 operator-(Dual3 a, Dual3 b)
{
    return mk(a.v - b.v, a.d[0] - b.d[0], a.d[1] - b.d[1], a.d[2] - b.d[2]);
}
__device__ __forceinline__ Dual3 operator-(Dual3 a) { return mk(-a.v, -a.d[0], -a.d[1], -a.d[2]); }
__device__ __forceinline__ Dual3 operator+(Dual3 a, double b) { return mk(a.v + b, a.d[0], a.d[1], a.d[2]); }
__device__ __forceinline__ Dual3 operator+(double a, Dual3 b) { return b + a; }
__device__ __forceinline__ Dual3 operator-(Dual3 a, double b) { return mk(a.v - b, a.d[0], a.d[1], a.d[2]); }
__device__ __forceinline__ Dual3 operator-(double a, Dual3 b) { return mk(a - b.v, -b.d[0], -b.d[1], -b.d[2]); }
__device__ __forceinline__ Dual3 operator*(Dual3 a, Dual3 b)
{
    return mk(a.v * b.v, a.d[0] * b.v + a.v * b.d[0], a.d[1] * b.v + a.v * b.d[1], a.d[2] * b.v + a.v * b.d[2]);
}
__device__ __forceinline__ Dual3 operator*(Dual3 a, double b) { return mk(a.v * b, a.d[0] * b, a.d[1] * b, a.d[2] * b); }
__device__ __forceinline__ Dual3 operator*(double a, Dual3 b) { return b * a; }
__device__ __forceinline__ Dual3 operator/(Dual3 a, Dual3 b)
{
    const double q = a.v / b.v;
    return mk(q, (a.d[0] - q * b.d[0]) / b.v, (a.d[1] - q * b.d[1]) / b.v, (a.d[2] - q * b.d[2]) / b.v);
}
__device__ __forceinline__ Dual3 operator/(Dual3 a, double b) { return mk(a.v / b, a.d[0] / b, a.d[1] / b, a.d[2] / b); }
__device__ __forceinline__ Dual3 operator/(double a, Dual3 b)
{
    const double q = a / b.v;
    return mk(q, -q * b.d[0] / b.v, -q * b.d[1] / b.v, -q * b.d[2] / b.v);
}
__device__ __forceinline__ double vsqrt(double a) { return sqrt(a); }
__device__ __forceinline__ Dual3 vsqrt(Dual3 a)
{
    const double r = sqrt(a.v), h = 0.5 / r;
    return mk(r, a.d[0] * h, a.d[1] * h, a.d[2] * h);
}
__device__ __forceinline__ double val(double x) { return x; }
__device__ __forceinline__ double val(const Dual3 &x) { return x.v; }
// a constant of the velocity type
template <class V> __device__ __forceinline__ V vconst(double x);
template <> __device__ __forceinline__ double vconst<double>(double x) { return x; }
template <> __device__ __forceinline__ Dual3 vconst<Dual3>(double x) { return mk(x, 0.0, 0.0, 0.0); }

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

// E1: the caller knows the interaction is ELASTIC with a constant rate
template <bool E1 = false>
__device__ __forceinline__ double sigma_v(const McDev &m, const McInter &it, double v)
{
    // Interaction::sigma_v (particles.hpp:61-70)
    if (E1) return it.rate;
    if (it.type == M2D_COULOMB) return coulomb_sigma(it, 0.5 * it.mu * v * v) * v;
    if (it.tab_len > 0)
        return vec_interp(m.tabE + it.tab_off, m.tabS + it.tab_off, it.tab_len,
                          0.5 * it.mu * v * v / kQE) * v;
    return it.rate;
}

// t_random::rot(len, x, y, z) (random.cpp:103-115) in two parts: the draws
// (the polar cosine c, its sine s, the azimuth's sine and cosine) and the
// vector of length len they give
struct RotDraw {
    double c, s, sp, cp;
};
__device__ __forceinline__ RotDraw rot_draw(Rng &r)
{
    RotDraw d;
    d.c = 1.0 - 2.0 * r.uni();
    d.s = sqrt(1.0 - d.c * d.c);
    sincos(2.0 * M_PI * r.uni(), &d.sp, &d.cp);
    return d;
}

template <class V>
__device__ __forceinline__ void rot_apply(V len, const RotDraw &d, V &x, V &y, V &z)
{
    x = len * d.c;
    y = len * d.s * d.sp;
    z = len * d.s * d.cp;
}

template <class V>
__device__ __forceinline__ void rot_len(Rng &r, V len, V &x, V &y, V &z)
{
    const RotDraw d = rot_draw(r);
    rot_apply(len, d, x, y, z);
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

template <class V>
__device__ inline void deflect(Rng &r, double angle, V &x, V &y, V &z)
{
    // t_random::deflect (random.cpp:133-175)
    double len = tan(angle / 2.0);
    double t = 1.0 - 2.0 * r.uni();
    double x1 = len * t;
    t = sqrt(1.0 - t * t);
    double sp, cp;
    sincos(2.0 * M_PI * r.uni(), &sp, &cp);
    double y1 = len * t * sp, z1 = len * t * cp;
    V tx = y * z1 - z * y1, ty = z * x1 - x * z1, tz = x * y1 - y * x1;
    const V tn = len / vsqrt(tx * tx + ty * ty + tz * tz);
    tx = tx * tn; ty = ty * tn; tz = tz * tn;
    V xp = x - y * tz + z * ty, yp = y - z * tx + x * tz, zp = z - x * ty + y * tx;
    t = 2.0 / (1.0 + len * len);
    const V x2 = tx * t, y2 = ty * t, z2 = tz * t;
    x = x + (-yp * z2 + zp * y2);
    y = y + (-zp * x2 + xp * z2);
    z = z + (-xp * y2 + yp * x2);
}

// BaseSpecies::scatter (particles.cpp:208-365), FP64, same draw order as
// the reference: species pick, partner, interaction pick, angles.  The
// velocity changes only when an interaction fires.  V is double, or Dual3
// for the Jacobian (the picks compare values only).
// E1: the caller has checked elastic1(m).  The same code then takes no
// runtime branch before the null-collision test, so the word each draw
// reads is known when it is compiled: the stream's Philox blocks, the three
// normals' functions and the rotation's angles become independent chains
// whose latencies overlap (one thread's chain of dependent FP64 operations
// is what a scatter costs).  The angles are drawn before that test, from
// the words the reference draws them from after it; a null collision thus
// leaves the stream two words further on, so E1 is for callers that draw
// nothing from the stream after the scatter.
template <class V, bool E1 = false>
__device__ inline void scatter_d(const McDev &m, Rng &r, V &vx, V &vy, V &vz, unsigned long long *counters)
{
    double gamma = r.uni() / m.lifetime;
    double tmp = 0.0;
    int sid;
    const int n_targets = E1 ? 1 : m.n_targets;
    for (sid = 0; sid < n_targets - 1; sid++) {
        tmp += m.t[sid].rate;
        if (tmp > gamma) break;
    }
    const McTarget &T = m.t[sid];
    double vr2, vz2, vt2;
    int64_t partner = -1;  // slot of the live partner (random_particle, particles.hpp:201-211)
    if (E1 || T.px == nullptr || T.pn == 0) {
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
    V dvx = vx - vr2, dvz = vz - vz2, dvy = vy - vt2;
    V v_rel = vsqrt(dvx * dvx + dvz * dvz + dvy * dvy);
    double m2 = T.mass, mass = m.mass;

    gamma = r.uni() * T.rate;
    RotDraw rd{};
    if constexpr (E1) rd = rot_draw(r);
    tmp = 0.0;
    int k;
    const int n_inter = E1 ? 1 : T.n;
    for (k = 0; k < n_inter; k++) {
        tmp += sigma_v<E1>(m, m.in[T.first + k], val(v_rel)) * T.density;
        if (tmp > gamma) break;
    }
    if (k == n_inter) return;  // null collision
    const McInter &it = m.in[T.first + k];
    tmp = 1.0 / (mass + m2);
    switch (E1 ? (int)M2D_ELASTIC : it.type) {
    case M2D_SUPERELASTIC: {
        V E = 0.5 * it.mu * v_rel * v_rel + it.DE;
        V vr = vsqrt(2 * E / it.mu);
        V rx, rz, ry;
        rot_len(r, vr, rx, rz, ry);
        vx = (rx * m2 + vx * mass + vr2 * m2) * tmp;
        V nvz = (ry * m2 + vz * mass + vz2 * m2) * tmp;
        V nvy = (rz * m2 + vy * mass + vt2 * m2) * tmp;
        vz = nvz; vy = nvy;
        break;
    }
    case M2D_COULOMB:
    case M2D_ELASTIC: {
        // isotropic in the centre of mass (particles.cpp:290-316)
        V cx = (vx * mass + vr2 * m2) * tmp;
        V cz = (vz * mass + vz2 * m2) * tmp;
        V cy = (vy * mass + vt2 * m2) * tmp;
        V rx, ry, rz;
        if constexpr (E1) rot_apply(v_rel, rd, rx, ry, rz);
        else rot_len(r, v_rel, rx, ry, rz);
        vx = rx * m2 * tmp + cx;
        vz = rz * m2 * tmp + cz;
        vy = ry * m2 * tmp + cy;
        // COULOMB also turns the partner (particles.cpp:309-314); a partner
        // drawn from the continuum has no state to change
        if (it.type == M2D_COULOMB && partner >= 0) {
            T.pvx[partner] = (float)val(-rx * mass * tmp + cx);
            T.pvz[partner] = (float)val(-rz * mass * tmp + cz);
            T.pvy[partner] = (float)val(-ry * mass * tmp + cy);
        }
        break;
    }
    case M2D_CX:
        vx = vconst<V>(vr2); vz = vconst<V>(vz2); vy = vconst<V>(vt2);
        break;
    case M2D_LANGEVIN: {
        V ax = (vx - vr2) * m2 * tmp, ay = (vz - vz2) * m2 * tmp, az = (vy - vt2) * m2 * tmp;
        double beta = sqrt(r.uni()) * it.cutoff;
        if (beta > 1.0) {
            // Nanbu & Kitatani, J. Phys. D 28 (1995) 324 (particles.cpp:337-345)
            double t4 = sqrt(beta * beta * beta * beta - 1.0);
            double xi0 = sqrt(beta * beta - t4), xi1 = sqrt(beta * beta + t4);
            double theta = ellint_k(xi0 / xi1) * M_SQRT2 * beta / xi1;
            deflect(r, M_PI - 2 * theta, ax, ay, az);
        } else {
            V len = vsqrt(ax * ax + ay * ay + az * az);
            rot_len(r, len, ax, ay, az);
        }
        V nvx = ax + (vx * mass + vr2 * m2) * tmp;
        V nvz = ay + (vz * mass + vz2 * m2) * tmp;
        V nvy = az + (vy * mass + vt2 * m2) * tmp;
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

// one target, a continuum partner and one ELASTIC interaction of constant
// rate: the collision set scatter_d<V, true> is compiled for
__device__ __forceinline__ bool elastic1(const McDev &m)
{
    const McTarget &T = m.t[0];
    return m.n_targets == 1 && (T.px == nullptr || T.pn == 0) && T.n == 1 && m.in[T.first].type == M2D_ELASTIC &&
           m.in[T.first].tab_len == 0;
}

// scatter() of a slot whose stream nothing draws from afterwards (the
// collision queue): the same velocities, through scatter_d<double, true>
// where elastic1(m) holds
__device__ inline void scatter_last(const McDev &m, Rng &r, float &vxf, float &vyf, float &vzf,
                                    unsigned long long *counters)
{
    double vx = vxf, vy = vyf, vz = vzf;
    if (elastic1(m)) scatter_d<double, true>(m, r, vx, vy, vz, counters);
    else scatter_d(m, r, vx, vy, vz, counters);
    vxf = (float)vx; vyf = (float)vy; vzf = (float)vz;
}

}  // namespace m2d
