// m2d_gather2d.hpp -- Fields::E at a particle (fields.hpp:124-150) through
// the padded FP32 edge-gradient tables, shared by the single-run and the
// batched 2D passes so both compute the same bits.
#pragma once
#include "m2d_internal.hpp"

namespace m2d {

// ------------------------------------------------------------------ gather
// Fields::E through the padded edge-gradient tables.  Index arithmetic is
// FP64 on the FP32 position (exactly the reference's x*idx), weights FP32.
// A is any argument block with the table pointers gx, gz, the grid nx, nz,
// idx, idz, xmin, zmin and the constant-field switch econst/ecx/ecz.
// ORIGIN0: the caller guarantees xmin = zmin = 0 and no constant field
// (x - 0 is x exactly, so the result is the same).  Returns true when the
// point lies outside the grid (a NaN included): the caller raises
// M2D_FLAG_OUT_OF_RANGE, and (generic path) the point is clamped onto the
// grid.  Every table access stays inside the tables for any input, so the
// gather may run ahead of the test that discards its result.
// x truncated to an int as v_cvt_i32_f64 defines it: out-of-range values
// saturate and a NaN gives 0 (a C++ conversion of those is undefined)
__device__ __forceinline__ int cvt_i32_sat(double x)
{
    int r;
    asm("v_cvt_i32_f64 %0, %1" : "=v"(r) : "v"(x));
    return r;
}

__device__ __forceinline__ int cvt_i32_sat_f(float x)
{
    int r;
    asm("v_cvt_i32_f32 %0, %1" : "=v"(r) : "v"(x));
    return r;
}

// Cell indices and in-cell weights of the gather at one point.  Index
// arithmetic is FP64 on the FP32 position, as documented above.
struct GatherAt {
    int ih, jx;        // x component: edge row ih, column jx (and their +1 neighbours)
    int i0, jh;        // z component: row i0, edge column jh
    float fx, fyx;     // x component weights
    float fxz, fz;     // z component weights
};

template <bool ORIGIN0, class A>
__device__ __forceinline__ bool gather_at(const A &a, double xd, double zd, GatherAt &g)
{
    const int nx = a.nx, nz = a.nz;
    double X = ORIGIN0 ? xd * a.idx : (xd - a.xmin) * a.idx;
    double Z = ORIGIN0 ? zd * a.idz : (zd - a.zmin) * a.idz;
    const bool bad = !(X > -1.0 && X < nx - 0.5 && Z > -1.0 && Z < nz - 0.5);
    // Inside that range every index below is already in the tables (x > -1
    // truncates to 0).  Outside it the generic path moves the point onto the
    // grid (fmax of a NaN and 0 is 0).  The ORIGIN0 pass does not: such a
    // point is outside the box, so its particle is removed after the push;
    // the unsigned clamps alone keep the table reads in bounds (a NaN
    // converts to 0, a negative index wraps to a large one).
    if (!ORIGIN0 && bad) {
        X = fmin(fmax(X, 0.0), (double)(nx - 1));
        Z = fmin(fmax(Z, 0.0), (double)(nz - 1));
    }
    auto trunc = [](double t) { return ORIGIN0 ? cvt_i32_sat(t) : (int)t; };
    // x component: i = (int)(x idx + .5), j = min((int)(z idz), lmax-2)
    g.ih = (int)min((unsigned)trunc(X + 0.5), (unsigned)(nx - 1));
    g.jx = (int)min((unsigned)trunc(Z), (unsigned)(nz - 2));
    // z component: i = min((int)(x idx), jmax-2), j = (int)(z idz + .5)
    g.i0 = (int)min((unsigned)trunc(X), (unsigned)(nx - 2));
    g.jh = (int)min((unsigned)trunc(Z + 0.5), (unsigned)(nz - 1));
    g.fx = (float)(X - g.ih + 0.5);
    g.fyx = (float)(Z - g.jx);
    g.fxz = (float)(X - g.i0);
    g.fz = (float)(Z - g.jh + 0.5);
    return bad;
}

// The same for a grid with its origin at 0 and positions given as FP32
// pairs (x + xl; xl = 0 for FP32 positions), in FP32 arithmetic: x idx is
// carried as p + e, p the rounded product with idx's leading FP32 part and e
// its rounding error (fma) plus the products with the remainders, which
// holds x idx to about 2^-46 of its value.  The cell indices come from p;
// the weights are (p - index) + 1/2 (exact: p and the index are close) plus
// e.  A p that rounds across a cell face can pick the neighbouring index,
// whose weight then lies a rounding outside [0, 1]: the interpolation is
// continuous across faces, so the gathered field changes only by that
// rounding.  A.idxh/idxl and idzh/idzl are idx and idz as FP32 pairs.
template <class A>
__device__ __forceinline__ bool gather_at_f(const A &a, float x, float xl, float z, float zl, GatherAt &g)
{
    const int nx = a.nx, nz = a.nz;
    const float px = x * a.idxh, pz = z * a.idzh;
    const float ex = fmaf(x, a.idxh, -px) + fmaf(xl, a.idxh, x * a.idxl);
    const float ez = fmaf(z, a.idzh, -pz) + fmaf(zl, a.idzh, z * a.idzl);
    const bool bad = !(px > -1.f && px < nx - 0.5f && pz > -1.f && pz < nz - 0.5f);
    // saturating conversions (a NaN gives 0) and unsigned clamps keep every
    // table read in bounds for any input, as in gather_at
    g.ih = (int)min((unsigned)cvt_i32_sat_f(px + 0.5f), (unsigned)(nx - 1));
    g.jx = (int)min((unsigned)cvt_i32_sat_f(pz), (unsigned)(nz - 2));
    g.i0 = (int)min((unsigned)cvt_i32_sat_f(px), (unsigned)(nx - 2));
    g.jh = (int)min((unsigned)cvt_i32_sat_f(pz + 0.5f), (unsigned)(nz - 1));
    g.fx = ((px - (float)g.ih) + 0.5f) + ex;
    g.fyx = (pz - (float)g.jx) + ez;
    g.fxz = (px - (float)g.i0) + ex;
    g.fz = ((pz - (float)g.jh) + 0.5f) + ez;
    return bad;
}

// The cell k and the offsets u, 1 - u of a non-negative position x + xl (an
// FP32 pair; xl = 0 for FP32 positions) on a grid with its origin at 0, in
// FP32 arithmetic: x idx is carried as p + e (see gather_at_f), k = trunc(p)
// moved by one where e carries p + e across a node, and u = (p - k) + e,
// ou = ((k + 1) - p) - e, each difference exact (p and the node are close),
// so u and ou carry one rounding each, as the FP64 offsets do.
__device__ __forceinline__ void cell_f(float x, float xl, float sh, float sl, int &k, float &u, float &ou)
{
    const float p = x * sh;
    const float e = fmaf(x, sh, -p) + fmaf(xl, sh, x * sl);
    k = cvt_i32_sat_f(p);
    float r = (p - (float)k) + e;
    if (r < 0.f) k -= 1;
    else if (r >= 1.f) k += 1;
    u = (p - (float)k) + e;
    ou = ((float)(k + 1) - p) - e;
}

// the bilinear interpolation of the four edge samples of each component,
// E = -grad
__device__ __forceinline__ void gather_interp(const GatherAt &g, float g00, float g01, float g10, float g11, float h00,
                                              float h01, float h10, float h11, float &ex, float &ez)
{
    const float fx = g.fx, fyx = g.fyx, fxz = g.fxz, fz = g.fz;
    float gxv = (1.f - fx) * ((1.f - fyx) * g00 + fyx * g01) + fx * ((1.f - fyx) * g10 + fyx * g11);
    float gzv = (1.f - fxz) * ((1.f - fz) * h00 + fz * h01) + fxz * ((1.f - fz) * h10 + fz * h11);
    ex = -gxv;
    ez = -gzv;
}

// the eight edge samples from the tables in memory: gx, gz with the phase
// ph applied, or (PHT) the phase-applied tables phx, phz (one float per
// edge: the same fmaf values, so the same bits)
template <bool PHT, class A>
__device__ __forceinline__ void gather_fetch(const A &a, const GatherAt &g, float ph, float &ex, float &ez)
{
    const int nz = a.nz;
    // indices are clamped non-negative and the tables hold fewer than 2^24
    // rows: 24-bit products are exact (full-rate multiplies)
    const int ox = __umul24(g.ih, nz) + g.jx, oz = __umul24(g.i0, nz + 1) + g.jh;
    float g00, g01, g10, g11, h00, h01, h10, h11;
    if constexpr (PHT) {
        // 32-bit byte offsets from the table bases (m2d_fields2d_create
        // keeps (nx+1) (nz+1) below 2^30, so the tables' (nx+1) nz and
        // nx (nz+1) floats span less than 2^32 bytes): the loads take the
        // base from scalar registers and one 32-bit offset each
        const unsigned bx = (unsigned)ox * 4u, bz = (unsigned)oz * 4u;
        const unsigned bx1 = bx + (unsigned)nz * 4u, bz1 = bz + (unsigned)(nz + 1) * 4u;
        const char *tx = (const char *)a.phx, *tz = (const char *)a.phz;
        const float *r0 = (const float *)(tx + bx), *r1 = (const float *)(tx + bx1);
        const float *q0 = (const float *)(tz + bz), *q1 = (const float *)(tz + bz1);
        g00 = r0[0]; g01 = r0[1]; g10 = r1[0]; g11 = r1[1];
        h00 = q0[0]; h01 = q0[1]; h10 = q1[0]; h11 = q1[1];
    } else {
        const float2 *r0 = a.gx + ox, *r1 = r0 + nz;
        const float2 p00 = r0[0], p01 = r0[1], p10 = r1[0], p11 = r1[1];
        const float2 *q0 = a.gz + oz, *q1 = q0 + (nz + 1);
        const float2 s00 = q0[0], s01 = q0[1], s10 = q1[0], s11 = q1[1];
        g00 = fmaf(ph, p00.y, p00.x); g01 = fmaf(ph, p01.y, p01.x);
        g10 = fmaf(ph, p10.y, p10.x); g11 = fmaf(ph, p11.y, p11.x);
        h00 = fmaf(ph, s00.y, s00.x); h01 = fmaf(ph, s01.y, s01.x);
        h10 = fmaf(ph, s10.y, s10.x); h11 = fmaf(ph, s11.y, s11.x);
    }
    gather_interp(g, g00, g01, g10, g11, h00, h01, h10, h11, ex, ez);
}

// PHT: read the tables with the phase already applied (a.phx, a.phz)
// instead of gx, gz.
template <bool ORIGIN0 = false, bool PHT = false, class A>
__device__ __forceinline__ bool gather_E_clamped(const A &a, double xd, double zd, float ph, float &ex,
                                                 float &ez)
{
    if (!ORIGIN0 && a.econst) { ex = a.ecx; ez = a.ecz; return false; }
    GatherAt g;
    const bool bad = gather_at<ORIGIN0>(a, xd, zd, g);
    gather_fetch<PHT>(a, g, ph, ex, ez);
    return bad;
}

// the origin-0 gather of FP32 positions (pairs) in FP32 arithmetic
template <bool PHT = false, class A>
__device__ __forceinline__ bool gather_E_f(const A &a, float x, float xl, float z, float zl, float ph, float &ex,
                                           float &ez)
{
    GatherAt g;
    const bool bad = gather_at_f(a, x, xl, z, zl, g);
    gather_fetch<PHT>(a, g, ph, ex, ez);
    return bad;
}

// the same, raising M2D_FLAG_OUT_OF_RANGE in a.counters itself
template <bool ORIGIN0 = false, class A>
__device__ __forceinline__ void gather_E(const A &a, double xd, double zd, float ph, float &ex, float &ez)
{
    if (gather_E_clamped<ORIGIN0>(a, xd, zd, ph, ex, ez)) raise_flag(a.counters, M2D_FLAG_OUT_OF_RANGE);
}

}  // namespace m2d
