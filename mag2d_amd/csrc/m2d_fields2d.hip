// m2d_fields2d.hip -- Fields: device grids, gather tables and the Poisson
// solve that replaces the UMFPACK factorisation of Fields::Fields
// (fields.cpp:115-276) and the back-solves of boundary_solve /
// boundary_solve_rf / solve (fields.cpp:278-353).
//
// Solver: FP64 iterative refinement around an FP32 preconditioned conjugate
// gradient on the Dirichlet-eliminated SPD system 4u_k - sum_free u_n = ...,
// preconditioned by one symmetric multigrid V-cycle (red-black Gauss-Seidel
// on the finest 5-point level, weighted Jacobi on Galerkin 9-point coarse
// levels, dense inverse on the coarsest).  Each outer step computes the FP64
// residual r = b - A u with the reference's own rows, solves A'e = -r in
// FP32, and adds e in FP64, until ||r|| <= tol ||b_eff||.
#include "m2d_objects.hpp"
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>

namespace m2d {

// ------------------------------------------------------------------ tables
// Edge gradients of u and uRF, padded so Field2D::grad's first/last-row
// fallbacks (Field2D.hpp:106-124,148-166) are the bilinear formula itself.
__global__ void k_build_gx(const double *u, const double *urf, int nx, int nz, double idx,
                           float2 *gx)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    int e = blockIdx.y;  // 0..nx
    if (j >= nz || e > nx) return;
    int ee = e < 1 ? 1 : (e > nx - 1 ? nx - 1 : e);
    size_t k1 = (size_t)ee * nz + j, k0 = k1 - nz;
    float2 v;
    v.x = (float)((u[k1] - u[k0]) * idx);
    v.y = (float)((urf[k1] - urf[k0]) * idx);
    gx[(size_t)e * nz + j] = v;
}

__global__ void k_build_gz(const double *u, const double *urf, int nx, int nz, double idz,
                           float2 *gz)
{
    int e = blockIdx.x * blockDim.x + threadIdx.x;  // 0..nz
    int i = blockIdx.y;
    if (e > nz || i >= nx) return;
    int ee = e < 1 ? 1 : (e > nz - 1 ? nz - 1 : e);
    size_t k1 = (size_t)i * nz + ee, k0 = k1 - 1;
    float2 v;
    v.x = (float)((u[k1] - u[k0]) * idz);
    v.y = (float)((urf[k1] - urf[k0]) * idz);
    gz[(size_t)i * (nz + 1) + e] = v;
}

int fields_build_tables(m2d_fields2d *f)
{
    hipStream_t s = f->ctx->stream;
    dim3 b(256);
    f->ph_valid = false;
    k_build_gx<<<dim3((f->nz + 255) / 256, f->nx + 1), b, 0, s>>>(f->u, f->urf, f->nx, f->nz, f->idx, f->gx);
    k_build_gz<<<dim3((f->nz + 1 + 255) / 256, f->nx), b, 0, s>>>(f->u, f->urf, f->nx, f->nz, f->idz, f->gz);
    M2D_HIP(hipGetLastError());
    return M2D_OK;
}

// ------------------------------------------------------------------ FP64 kernels
__device__ __forceinline__ bool is_dir(uint8_t m) { return m == M2D_FIXED || m == M2D_FIXED_RF; }

// rhs in place (fields.cpp:297-310, 333-346), Dirichlet values into u.
// The node bodies below are shared by the grid kernels and the
// single-workgroup solve, so both paths compute each node the same way.
__device__ __forceinline__ void node_rhs(const uint8_t *mask, const double *voltage, double *rho, double *u,
                                         size_t k, int mode, double scale)
{
    uint8_t m = mask[k];
    double b;
    if (mode == M2D_SOLVE_RAW) b = rho[k];
    else if (m == M2D_FIXED) b = mode == M2D_SOLVE_STATIC ? voltage[k] : 0.0;
    else if (m == M2D_FIXED_RF) b = mode == M2D_SOLVE_RF ? voltage[k] : 0.0;
    else b = rho[k] * scale;
    rho[k] = b;
    if (is_dir(m)) u[k] = b;
}

__global__ void k_rhs(const uint8_t *mask, const double *voltage, double *rho, double *u, size_t n,
                      int mode, double scale)
{
    size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) node_rhs(mask, voltage, rho, u, k, mode, scale);
}

// r = b - A u at node k (0 on Dirichlet nodes), d = -r in FP32; rr = r^2 and
// bb = b_eff^2 (b with the Dirichlet neighbours moved to the right-hand side)
__device__ __forceinline__ void node_residual64(const uint8_t *mask, const double *b, const double *u, int nz,
                                                double *r, float *d, size_t k, double &rr, double &bb)
{
    double rv = 0;
    bb = 0;
    if (!is_dir(mask[k])) {
        // free rows never sit on the frame (checked at creation)
        double au = u[k - nz] + u[k - 1] - 4.0 * u[k] + u[k + 1] + u[k + nz];
        rv = b[k] - au;
        double be = b[k];
        if (is_dir(mask[k - nz])) be -= u[k - nz];
        if (is_dir(mask[k - 1])) be -= u[k - 1];
        if (is_dir(mask[k + 1])) be -= u[k + 1];
        if (is_dir(mask[k + nz])) be -= u[k + nz];
        bb = be * be;
    }
    r[k] = rv;
    d[k] = (float)(-rv);
    rr = rv * rv;
}

// u += e on free nodes
__global__ void k_add_correction(const uint8_t *mask, double *u, const float *e, size_t n, const int *stop)
{
    if (stopped(stop)) return;
    size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n && !is_dir(mask[k])) u[k] += (double)e[k];
}

// ------------------------------------------------------------------ FP32 MG kernels
// smoothing sweeps before and after the coarse correction, and the damping
// of the Jacobi sweeps on the Galerkin levels
constexpr int kPre = 2, kPost = 2;
constexpr float kOmega = 0.8f;


// 9-point operator y = A x (Dirichlet rows give 0)
__device__ __forceinline__ float apply9(const float *st, const float *x, int i, int j, int nx, int nz)
{
    size_t k = (size_t)i * nz + j;
    const float *c = st + 9 * k;
    float s = 0.f;
#pragma unroll
    for (int di = -1; di <= 1; di++) {
        int ii = i + di;
        if (ii < 0 || ii >= nx) continue;
#pragma unroll
        for (int dj = -1; dj <= 1; dj++) {
            int jj = j + dj;
            if (jj < 0 || jj >= nz) continue;
            float a = c[(di + 1) * 3 + (dj + 1)];
            if (a != 0.f) s = fmaf(a, x[(size_t)ii * nz + jj], s);
        }
    }
    return s;
}

// finest level: 5-point operator from the mask, A' = 4I - adjacency(free)
__device__ __forceinline__ float apply5(const uint8_t *dir, const float *x, size_t k, int nz)
{
    float s = 4.f * x[k];
    if (!dir[k - nz]) s -= x[k - nz];
    if (!dir[k - 1]) s -= x[k - 1];
    if (!dir[k + 1]) s -= x[k + 1];
    if (!dir[k + nz]) s -= x[k + nz];
    return s;
}

// red-black Gauss-Seidel update of node (i, j) of colour (i+j)&1 on the
// finest level
__device__ __forceinline__ void node_rbgs5(const uint8_t *dir, float *x, const float *b, int nz, size_t k)
{
    if (dir[k]) return;
    float s = b[k];
    if (!dir[k - nz]) s += x[k - nz];
    if (!dir[k - 1]) s += x[k - 1];
    if (!dir[k + 1]) s += x[k + 1];
    if (!dir[k + nz]) s += x[k + nz];
    x[k] = 0.25f * s;
}

// half sweep (colour = (i+j)&1)
__global__ void k_rbgs5(const uint8_t *dir, float *x, const float *b, int nx, int nz, int colour)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    int i = blockIdx.y;
    if (j >= nz || ((i + j) & 1) != colour) return;
    node_rbgs5(dir, x, b, nz, (size_t)i * nz + j);
}

__device__ __forceinline__ float node_resid5(const uint8_t *dir, const float *x, const float *b, int nz, size_t k)
{
    return dir[k] ? 0.f : b[k] - apply5(dir, x, k, nz);
}

// weighted Jacobi on a 9-point level: xo = xi + w D^-1 (b - A xi)
__device__ __forceinline__ float node_jacobi9(const float *st, const float *dinv, const float *xi, const float *b,
                                              int nx, int nz, float w, int i, int j)
{
    size_t k = (size_t)i * nz + j;
    float dv = dinv[k];
    return dv == 0.f ? 0.f : xi[k] + w * dv * (b[k] - apply9(st, xi, i, j, nx, nz));
}

__global__ void k_jacobi9(const float *st, const float *dinv, const float *xi, float *xo,
                          const float *b, int nx, int nz, float w)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    int i = blockIdx.y;
    if (j >= nz) return;
    xo[(size_t)i * nz + j] = node_jacobi9(st, dinv, xi, b, nx, nz, w, i, j);
}

__device__ __forceinline__ float node_resid9(const float *st, const float *dinv, const float *x, const float *b,
                                             int nx, int nz, int i, int j)
{
    size_t k = (size_t)i * nz + j;
    return dinv[k] == 0.f ? 0.f : b[k] - apply9(st, x, i, j, nx, nz);
}

__global__ void k_resid9(const float *st, const float *dinv, const float *x, const float *b, float *r,
                         int nx, int nz)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    int i = blockIdx.y;
    if (j >= nz) return;
    r[(size_t)i * nz + j] = node_resid9(st, dinv, x, b, nx, nz, i, j);
}

// b_c = P^T r_f with P the bilinear interpolation masked on fine Dirichlet nodes
__device__ __forceinline__ float node_restrict(const float *rf, int nxf, int nzf, const float *dinvc, int nzc,
                                               int I, int J)
{
    size_t K = (size_t)I * nzc + J;
    if (dinvc[K] == 0.f) return 0.f;
    float s = 0.f;
    for (int di = -1; di <= 1; di++) {
        int i = 2 * I + di;
        if (i < 0 || i >= nxf) continue;
        float wi = di == 0 ? 1.f : 0.5f;
        for (int dj = -1; dj <= 1; dj++) {
            int j = 2 * J + dj;
            if (j < 0 || j >= nzf) continue;
            float wj = dj == 0 ? 1.f : 0.5f;
            s = fmaf(wi * wj, rf[(size_t)i * nzf + j], s);  // r is 0 on Dirichlet nodes
        }
    }
    return s;
}

__global__ void k_restrict(const float *rf, int nxf, int nzf, float *bc, const float *dinvc, int nxc,
                           int nzc)
{
    int J = blockIdx.x * blockDim.x + threadIdx.x;
    int I = blockIdx.y;
    if (J >= nzc) return;
    bc[(size_t)I * nzc + J] = node_restrict(rf, nxf, nzf, dinvc, nzc, I, J);
}

// x_f += P x_c on free fine nodes
__device__ __forceinline__ void node_prolong(float *xf, const uint8_t *dirf, int nzf, const float *xc, int nxc,
                                             int nzc, int i, int j)
{
    size_t k = (size_t)i * nzf + j;
    if (dirf[k]) return;
    int I0 = i >> 1, J0 = j >> 1;
    float s = 0.f;
    if ((i & 1) == 0) {
        if ((j & 1) == 0) s = xc[(size_t)I0 * nzc + J0];
        else {
            s = 0.5f * xc[(size_t)I0 * nzc + J0];
            if (J0 + 1 < nzc) s += 0.5f * xc[(size_t)I0 * nzc + J0 + 1];
        }
    } else {
        float a = 0.f, b = 0.f;
        if ((j & 1) == 0) {
            a = xc[(size_t)I0 * nzc + J0];
            if (I0 + 1 < nxc) b = xc[(size_t)(I0 + 1) * nzc + J0];
        } else {
            a = 0.5f * xc[(size_t)I0 * nzc + J0];
            if (J0 + 1 < nzc) a += 0.5f * xc[(size_t)I0 * nzc + J0 + 1];
            if (I0 + 1 < nxc) {
                b = 0.5f * xc[(size_t)(I0 + 1) * nzc + J0];
                if (J0 + 1 < nzc) b += 0.5f * xc[(size_t)(I0 + 1) * nzc + J0 + 1];
            }
        }
        s = 0.5f * (a + b);
    }
    xf[k] += s;
}

__global__ void k_prolong(float *xf, const uint8_t *dirf, int nxf, int nzf, const float *xc, int nxc,
                          int nzc)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    int i = blockIdx.y;
    if (j >= nzf) return;
    node_prolong(xf, dirf, nzf, xc, nxc, nzc, i, j);
}

// dense coarsest solve x = A^-1 b, one block
__global__ void __launch_bounds__(256) k_coarse_solve(const float *inv, const float *b, float *x, int n,
                                                      const int *stop)
{
    if (stopped(stop)) return;
    extern __shared__ float sb[];
    for (int k = threadIdx.x; k < n; k += blockDim.x) sb[k] = b[k];
    __syncthreads();
    for (int row = threadIdx.x; row < n; row += blockDim.x) {
        const float *a = inv + (size_t)row * n;
        float s = 0.f;
        for (int c = 0; c < n; c++) s = fmaf(a[c], sb[c], s);
        x[row] = s;
    }
}


// ------------------------------------------------------------------ fused solver kernels
// The solve is launch-bound on small grids, so its kernels carry their own
// reductions and fuse neighbouring sweeps.  finish_sums: each block sums its
// values in a fixed order and writes a partial; the block that takes the last
// ticket sums the partials in block order, writes out[] and resets the
// ticket, so a reduction costs no extra launch and gives the same bits on
// every run.
// out = a.b (FP64 accumulation)
__global__ void __launch_bounds__(kRedBlock) k_dot_fin(const float *a, const float *b, size_t n, double *partial,
                                                      unsigned int *ticket, double *out, int fin, const int *stop)
{
    if (stopped(stop)) return;
    double v[1] = {0};
    for (size_t k = (size_t)blockIdx.x * kRedBlock + threadIdx.x; k < n; k += (size_t)gridDim.x * kRedBlock)
        v[0] += (double)a[k] * b[k];
    finish_sums<1>(v, partial, ticket, out, fin);
}

// q = A' p on the finest level, out = p.q
// The CG direction and the operator in one launch: p' = z + beta p (first:
// p' = z), q = A' p' and out = p'.q.  Each node's p' is recomputed at its
// stencil neighbours from z and p -- the same operations as at the node
// itself, so the same bits -- and written to pn, not p, which other blocks
// are still reading.
__global__ void __launch_bounds__(kRedBlock) k_dir_apply5_dot(const uint8_t *dir, const float *z, const float *p,
                                                             float *pn, float *q, const double *rz_new,
                                                             const double *rz_old, int first, size_t n, int nz,
                                                             double *partial, unsigned int *ticket, double *out,
                                                             int fin, const int *stop)
{
    if (stopped(stop)) return;
    double v[1] = {0};
    const float beta = first ? 0.f : cg_beta(rz_new, rz_old);
    const size_t k = (size_t)blockIdx.x * kRedBlock + threadIdx.x;
    if (k < n) {
        const float pk = cg_dir_at(z, p, k, beta, first);
        pn[k] = pk;
        float y = 0.f;
        if (!dir[k]) {
            // apply5's order
            float t = 4.f * pk;
            if (!dir[k - nz]) t -= cg_dir_at(z, p, k - nz, beta, first);
            if (!dir[k - 1]) t -= cg_dir_at(z, p, k - 1, beta, first);
            if (!dir[k + 1]) t -= cg_dir_at(z, p, k + 1, beta, first);
            if (!dir[k + nz]) t -= cg_dir_at(z, p, k + nz, beta, first);
            y = t;
        }
        q[k] = y;
        v[0] = (double)pk * y;
    }
    finish_sums<1>(v, partial, ticket, out, fin);
}

// e += alpha p, r -= alpha q with alpha = rz / pq; out = r.r
// (first: e starts at 0, so e = alpha p without reading e)
__global__ void __launch_bounds__(kRedBlock) k_cg_update_dot(float *e, float *r, const float *p, const float *q,
                                                            const double *rz, const double *pq, size_t n, int first,
                                                            double *partial, unsigned int *ticket, double *out,
                                                            int fin, const int *stop, CtlCheck post)
{
    if (stopped(stop)) return;
    double v[1] = {0};
    const double den = *pq;
    const float alpha = den != 0.0 ? (float)(*rz / den) : 0.f;
    for (size_t k = (size_t)blockIdx.x * kRedBlock + threadIdx.x; k < n; k += (size_t)gridDim.x * kRedBlock) {
        e[k] = fmaf(alpha, p[k], first ? 0.f : e[k]);
        const float rk = fmaf(-alpha, q[k], r[k]);
        r[k] = rk;
        v[0] += (double)rk * rk;
    }
    finish_sums<1>(v, partial, ticket, out, fin, post);
}

// the FP64 residual with its norms finished in the same launch: out[0] =
// |r|^2, out[1] = |b_eff|^2, out[2] = |d|^2 of the FP32 copy d = -r
__global__ void __launch_bounds__(kRedBlock) k_residual64_fin(const uint8_t *mask, const double *b, const double *u,
                                                             size_t n, int nz, double *r, float *d, double *partial,
                                                             unsigned int *ticket, double *out, int fin,
                                                             const int *stop, CtlEval post)
{
    if (stopped(stop)) return;
    double v[3] = {0, 0, 0};
    const size_t k = (size_t)blockIdx.x * kRedBlock + threadIdx.x;
    if (k < n) {
        node_residual64(mask, b, u, nz, r, d, k, v[0], v[1]);
        v[2] = (double)d[k] * d[k];
    }
    finish_sums<3>(v, partial, ticket, out, fin, post);
}

// prolongation value at fine node (i, j) (the increment of node_prolong)
__device__ __forceinline__ float prolong_at(const float *xc, int nxc, int nzc, int i, int j)
{
    int I0 = i >> 1, J0 = j >> 1;
    float s = 0.f;
    if ((i & 1) == 0) {
        if ((j & 1) == 0) s = xc[(size_t)I0 * nzc + J0];
        else {
            s = 0.5f * xc[(size_t)I0 * nzc + J0];
            if (J0 + 1 < nzc) s += 0.5f * xc[(size_t)I0 * nzc + J0 + 1];
        }
    } else {
        float a = 0.f, c = 0.f;
        if ((j & 1) == 0) {
            a = xc[(size_t)I0 * nzc + J0];
            if (I0 + 1 < nxc) c = xc[(size_t)(I0 + 1) * nzc + J0];
        } else {
            a = 0.5f * xc[(size_t)I0 * nzc + J0];
            if (J0 + 1 < nzc) a += 0.5f * xc[(size_t)I0 * nzc + J0 + 1];
            if (I0 + 1 < nxc) {
                c = 0.5f * xc[(size_t)(I0 + 1) * nzc + J0];
                if (J0 + 1 < nzc) c += 0.5f * xc[(size_t)(I0 + 1) * nzc + J0 + 1];
            }
        }
        s = 0.5f * (a + c);
    }
    return s;
}

// ---- the finest level in LDS tiles.  A block holds a tile of kT0 x kT0
// nodes (16, 8 or 4 by the grid's size, see l0_tile) plus a halo and runs several red-black half sweeps on it with
// workgroup barriers in between.  A node at distance d from the halo's edge
// keeps the exact global value for d half sweeps after the first, so the
// halo is as wide as the sweeps (and the residual stencil) need.  Updates
// use the node_rbgs5 order of additions, so the result is the same as the
// half sweeps run one launch each.
constexpr int kH0d = 2 * kPre + 1;   // descent: 2 kPre - 1 sweeps, residual, restriction
constexpr int kH0u = 2 * kPost;      // ascent: 2 kPost sweeps
constexpr int kT0Threads = 256;

// one red-black half sweep over the held region but its outermost ring
template <int S>
__device__ __forceinline__ void tile_rbgs(float *sx, const float *sb, const uint8_t *sd, int i0, int j0, int colour)
{
    for (int t = threadIdx.x; t < S * S; t += kT0Threads) {
        const int li = t / S, lj = t % S;
        if (li == 0 || li == S - 1 || lj == 0 || lj == S - 1) continue;
        if (((i0 + li + j0 + lj) & 1) != colour || sd[t]) continue;
        float s = sb[t];
        if (!sd[t - S]) s += sx[t - S];
        if (!sd[t - 1]) s += sx[t - 1];
        if (!sd[t + 1]) s += sx[t + 1];
        if (!sd[t + S]) s += sx[t + S];
        sx[t] = 0.25f * s;
    }
}

// descent on the finest level: pre-smoothing from x = 0, the residual and its
// restriction to the coarse nodes whose fine node (2I, 2J) lies in the tile;
// writes the smoothed x of the tile and those coarse right-hand sides
template <int kT0>
__global__ void __launch_bounds__(kT0Threads) k_l0_down(const uint8_t *dir, const float *b, float *x, int nx, int nz,
                                                      float *bc, const float *dinvc, int nxc, int nzc,
                                                      const int *stop)
{
    if (stopped(stop)) return;
    constexpr int S = kT0 + 2 * kH0d;
    __shared__ float sx[S * S], sb[S * S], sr[S * S];
    __shared__ uint8_t sd[S * S];
    const int gi0 = blockIdx.y * kT0, gj0 = blockIdx.x * kT0;
    const int i0 = gi0 - kH0d, j0 = gj0 - kH0d;
    for (int t = threadIdx.x; t < S * S; t += kT0Threads) {
        const int i = i0 + t / S, j = j0 + t % S;
        const bool in = i >= 0 && i < nx && j >= 0 && j < nz;
        const size_t k = (size_t)i * nz + j;
        const uint8_t d = in ? dir[k] : (uint8_t)1;
        const float bv = in ? b[k] : 0.f;
        sd[t] = d;
        sb[t] = bv;
        // the first red half sweep from zero reads no neighbour
        sx[t] = (((i + j) & 1) == 0 && !d) ? 0.25f * (bv + 0.f) : 0.f;
    }
    __syncthreads();
    for (int h = 1; h < 2 * kPre; h++) {
        tile_rbgs<S>(sx, sb, sd, i0, j0, h & 1);
        __syncthreads();
    }
    // residual (node_resid5) where it reads only exact x: the 2 kPre - 1
    // sweeps after the first left the outer 2 kPre - 1 rings inexact
    constexpr int R = 2 * kPre;
    for (int t = threadIdx.x; t < S * S; t += kT0Threads) {
        const int li = t / S, lj = t % S;
        if (li < R || li >= S - R || lj < R || lj >= S - R) continue;
        float r = 0.f;
        if (!sd[t]) {
            float s = 4.f * sx[t];
            if (!sd[t - S]) s -= sx[t - S];
            if (!sd[t - 1]) s -= sx[t - 1];
            if (!sd[t + 1]) s -= sx[t + 1];
            if (!sd[t + S]) s -= sx[t + S];
            r = sb[t] - s;
        }
        sr[t] = r;
    }
    __syncthreads();
    // restriction (node_restrict) for the coarse nodes of this tile
    for (int t = threadIdx.x; t < (kT0 / 2) * (kT0 / 2); t += kT0Threads) {
        const int I = gi0 / 2 + t / (kT0 / 2), J = gj0 / 2 + t % (kT0 / 2);
        if (I >= nxc || J >= nzc) continue;
        const size_t K = (size_t)I * nzc + J;
        if (dinvc[K] == 0.f) { bc[K] = 0.f; continue; }
        float s = 0.f;
        for (int di = -1; di <= 1; di++) {
            const int i = 2 * I + di;
            if (i < 0 || i >= nx) continue;
            const float wi = di == 0 ? 1.f : 0.5f;
            for (int dj = -1; dj <= 1; dj++) {
                const int j = 2 * J + dj;
                if (j < 0 || j >= nz) continue;
                const float wj = dj == 0 ? 1.f : 0.5f;
                s = fmaf(wi * wj, sr[(i - i0) * S + (j - j0)], s);
            }
        }
        bc[K] = s;
    }
    for (int t = threadIdx.x; t < kT0 * kT0; t += kT0Threads) {
        const int i = gi0 + t / kT0, j = gj0 + t % kT0;
        if (i < nx && j < nz) x[(size_t)i * nz + j] = sx[(i - i0) * S + (j - j0)];
    }
}

// ascent on the finest level: x = xd + P x_c, then the post-smoothing
// sweeps.  xd (the descent's x) and x are different arrays: a block reads
// its halo from xd, and other blocks of the launch, not all resident at
// once on a large grid, write x.
template <int kT0>
__global__ void __launch_bounds__(kT0Threads) k_l0_up(const uint8_t *dir, const float *b, const float *xd, float *x,
                                                    int nx, int nz, const float *xc, int nxc, int nzc,
                                                    const int *stop)
{
    if (stopped(stop)) return;
    constexpr int S = kT0 + 2 * kH0u;
    __shared__ float sx[S * S], sb[S * S];
    __shared__ uint8_t sd[S * S];
    const int gi0 = blockIdx.y * kT0, gj0 = blockIdx.x * kT0;
    const int i0 = gi0 - kH0u, j0 = gj0 - kH0u;
    for (int t = threadIdx.x; t < S * S; t += kT0Threads) {
        const int i = i0 + t / S, j = j0 + t % S;
        const bool in = i >= 0 && i < nx && j >= 0 && j < nz;
        const size_t k = (size_t)i * nz + j;
        const uint8_t d = in ? dir[k] : (uint8_t)1;
        float xv = in ? xd[k] : 0.f;
        if (!d) xv += prolong_at(xc, nxc, nzc, i, j);  // node_prolong
        sd[t] = d;
        sb[t] = in ? b[k] : 0.f;
        sx[t] = xv;
    }
    __syncthreads();
    for (int h = 0; h < 2 * kPost; h++) {
        tile_rbgs<S>(sx, sb, sd, i0, j0, (h & 1) ^ 1);
        __syncthreads();
    }
    for (int t = threadIdx.x; t < kT0 * kT0; t += kT0Threads) {
        const int i = gi0 + t / kT0, j = gj0 + t % kT0;
        if (i < nx && j < nz) x[(size_t)i * nz + j] = sx[(i - i0) * S + (j - j0)];
    }
}

// ---- the Galerkin levels in LDS tiles (damped Jacobi, 9-point operator).
// The first sweep starts from x = 0 and reads no neighbour; each later sweep,
// the residual and the restriction widen the inexact border by one ring.
// Galerkin levels' tile edge: 4 nodes, so that even the 64 x 64 level of a
// 128^2 grid spreads over 256 blocks (16 blocks at an edge of 16); the halo
// loads cost less than the idle CUs did (Penning trap step -17%, the
// 512^2 RF basis solve -23%)
constexpr int kTm = 4;
constexpr int kHmd = kPre + 1;  // descent: kPre - 1 sweeps, residual, restriction
constexpr int kHmu = kPost;     // ascent: kPost sweeps

// apply9 on tile-local arrays (same loop order and skips as apply9)
template <int S>
__device__ __forceinline__ float tile_apply9(const float *sst, const float *sx, int t, int i, int j, int nx, int nz)
{
    const float *c = sst + 9 * t;
    float s = 0.f;
#pragma unroll
    for (int di = -1; di <= 1; di++) {
        const int ii = i + di;
        if (ii < 0 || ii >= nx) continue;
#pragma unroll
        for (int dj = -1; dj <= 1; dj++) {
            const int jj = j + dj;
            if (jj < 0 || jj >= nz) continue;
            const float a = c[(di + 1) * 3 + (dj + 1)];
            if (a != 0.f) s = fmaf(a, sx[t + di * S + dj], s);
        }
    }
    return s;
}

// one Jacobi sweep (node_jacobi9) over the held region but its outer ring
template <int S>
__device__ __forceinline__ void tile_jacobi(const float *sst, const float *sdv, const float *sb, const float *xi,
                                            float *xo, int i0, int j0, int nx, int nz, float w)
{
    for (int t = threadIdx.x; t < S * S; t += kT0Threads) {
        const int li = t / S, lj = t % S;
        if (li == 0 || li == S - 1 || lj == 0 || lj == S - 1) continue;
        const float dv = sdv[t];
        xo[t] = dv == 0.f ? 0.f : xi[t] + w * dv * (sb[t] - tile_apply9<S>(sst, xi, t, i0 + li, j0 + lj, nx, nz));
    }
}

// load the coefficients, inverse diagonal and right-hand side of a tile
// (zero outside the level: such nodes act as pinned)
template <int S>
__device__ __forceinline__ void tile_load9(const float *st, const float *dinv, const float *b, int i0, int j0, int nx,
                                           int nz, float *sst, float *sdv, float *sb)
{
    for (int t = threadIdx.x; t < S * S; t += kT0Threads) {
        const int i = i0 + t / S, j = j0 + t % S;
        const bool in = i >= 0 && i < nx && j >= 0 && j < nz;
        const size_t k = (size_t)i * nz + j;
        sdv[t] = in ? dinv[k] : 0.f;
        sb[t] = in ? b[k] : 0.f;
    }
    for (int t = threadIdx.x; t < 9 * S * S; t += kT0Threads) {
        const int node = t / 9, q = t % 9;
        const int i = i0 + node / S, j = j0 + node % S;
        const bool in = i >= 0 && i < nx && j >= 0 && j < nz;
        sst[t] = in ? st[9 * ((size_t)i * nz + j) + q] : 0.f;
    }
}

// descent on a Galerkin level: smoothing from x = 0, residual, restriction
// to the next level's right-hand side; writes the smoothed x of the tile
__global__ void __launch_bounds__(kT0Threads) k_lm_down(const float *st, const float *dinv, const float *b, float *x,
                                                      int nx, int nz, float *bc, const float *dinvc, int nxc, int nzc,
                                                      float w, const int *stop)
{
    if (stopped(stop)) return;
    constexpr int S = kTm + 2 * kHmd;
    __shared__ float sst[9 * S * S], sdv[S * S], sb[S * S], sx[S * S], sx2[S * S];
    const int gi0 = blockIdx.y * kTm, gj0 = blockIdx.x * kTm;
    const int i0 = gi0 - kHmd, j0 = gj0 - kHmd;
    tile_load9<S>(st, dinv, b, i0, j0, nx, nz, sst, sdv, sb);
    for (int t = threadIdx.x; t < S * S; t += kT0Threads) sx[t] = 0.f;
    __syncthreads();
    float *xa = sx, *xb = sx2;
    for (int it = 0; it < kPre; it++) {
        // the first sweep starts from x = +0, where the operator gives +0
        // (every term of apply9 is then a signed zero added to +0): it reads
        // no neighbour and is exact on the outer ring too
        if (it == 0) {
            for (int t = threadIdx.x; t < S * S; t += kT0Threads) {
                const float dv = sdv[t];
                xb[t] = dv == 0.f ? 0.f : xa[t] + w * dv * (sb[t] - 0.f);
            }
        } else {
            tile_jacobi<S>(sst, sdv, sb, xa, xb, i0, j0, nx, nz, w);
        }
        __syncthreads();
        float *tmp = xa; xa = xb; xb = tmp;
    }
    // residual (node_resid9) into xb where xa is exact
    constexpr int R = kPre;
    for (int t = threadIdx.x; t < S * S; t += kT0Threads) {
        const int li = t / S, lj = t % S;
        if (li < R || li >= S - R || lj < R || lj >= S - R) continue;
        xb[t] = sdv[t] == 0.f ? 0.f : sb[t] - tile_apply9<S>(sst, xa, t, i0 + li, j0 + lj, nx, nz);
    }
    __syncthreads();
    for (int t = threadIdx.x; t < (kTm / 2) * (kTm / 2); t += kT0Threads) {
        const int I = gi0 / 2 + t / (kTm / 2), J = gj0 / 2 + t % (kTm / 2);
        if (I >= nxc || J >= nzc) continue;
        const size_t K = (size_t)I * nzc + J;
        if (dinvc[K] == 0.f) { bc[K] = 0.f; continue; }
        float s = 0.f;
        for (int di = -1; di <= 1; di++) {
            const int i = 2 * I + di;
            if (i < 0 || i >= nx) continue;
            const float wi = di == 0 ? 1.f : 0.5f;
            for (int dj = -1; dj <= 1; dj++) {
                const int j = 2 * J + dj;
                if (j < 0 || j >= nz) continue;
                const float wj = dj == 0 ? 1.f : 0.5f;
                s = fmaf(wi * wj, xb[(i - i0) * S + (j - j0)], s);
            }
        }
        bc[K] = s;
    }
    for (int t = threadIdx.x; t < kTm * kTm; t += kT0Threads) {
        const int i = gi0 + t / kTm, j = gj0 + t % kTm;
        if (i < nx && j < nz) x[(size_t)i * nz + j] = xa[(i - i0) * S + (j - j0)];
    }
}

// ascent on a Galerkin level: xo = x + P x_c on free nodes, then the
// sweeps.  x (the descent's) and xo are different arrays: a block reads its
// halo from x while other blocks of the launch write xo.
__global__ void __launch_bounds__(kT0Threads) k_lm_up(const float *st, const float *dinv, const uint8_t *dir,
                                                    const float *b, const float *x, float *xo, int nx, int nz,
                                                    const float *xc, int nxc, int nzc, float w, const int *stop)
{
    if (stopped(stop)) return;
    constexpr int S = kTm + 2 * kHmu;
    __shared__ float sst[9 * S * S], sdv[S * S], sb[S * S], sx[S * S], sx2[S * S];
    const int gi0 = blockIdx.y * kTm, gj0 = blockIdx.x * kTm;
    const int i0 = gi0 - kHmu, j0 = gj0 - kHmu;
    tile_load9<S>(st, dinv, b, i0, j0, nx, nz, sst, sdv, sb);
    for (int t = threadIdx.x; t < S * S; t += kT0Threads) {
        const int i = i0 + t / S, j = j0 + t % S;
        const bool in = i >= 0 && i < nx && j >= 0 && j < nz;
        const size_t k = (size_t)i * nz + j;
        float xv = in ? x[k] : 0.f;
        if (in && !dir[k]) xv += prolong_at(xc, nxc, nzc, i, j);  // node_prolong
        sx[t] = xv;
    }
    __syncthreads();
    float *xa = sx, *xb = sx2;
    for (int it = 0; it < kPost; it++) {
        tile_jacobi<S>(sst, sdv, sb, xa, xb, i0, j0, nx, nz, w);
        __syncthreads();
        float *tmp = xa; xa = xb; xb = tmp;
    }
    for (int t = threadIdx.x; t < kTm * kTm; t += kT0Threads) {
        const int i = gi0 + t / kTm, j = gj0 + t % kTm;
        if (i < nx && j < nz) xo[(size_t)i * nz + j] = xa[(i - i0) * S + (j - j0)];
    }
}

// the V-cycle below level lc in one workgroup: descent from b on level lc,
// dense solve, ascent to x on level lc
constexpr int kMaxLevels = 16;
constexpr size_t kBlockLevelNodes = 1024;  // levels this small run inside the workgroup

struct BLevel {
    int nx, nz, n;
    const float *st, *dinv;
    const uint8_t *dir;
    float *x, *x2, *r, *b;
};

struct CoarseArgs {
    int lc, L;
    BLevel lv[kMaxLevels];

    const float *invT;  // the dense inverse, transposed (row r of A^-1 is column r)
    int coarse_n;
    float omega;
    int pre, post;
    const int *stop;    // the solve's stop flag (see stopped())
};

__device__ void c_jacobi(const BLevel &A, float *x, int sweeps, float w)
{
    float *xa = x, *xb = A.x2;
    for (int it = 0; it < sweeps; it++) {
        for (int k = threadIdx.x; k < A.n; k += kCoarseThreads)
            xb[k] = node_jacobi9(A.st, A.dinv, xa, A.b, A.nx, A.nz, w, k / A.nz, k % A.nz);
        __syncthreads();
        float *t = xa; xa = xb; xb = t;
    }
    if (xa != x) {
        for (int k = threadIdx.x; k < A.n; k += kCoarseThreads) x[k] = xa[k];
        __syncthreads();
    }
}

__global__ void __launch_bounds__(kCoarseThreads) k_vcycle_coarse(CoarseArgs C)
{
    if (stopped(C.stop)) return;
    extern __shared__ float sb[];  // coarse_n right-hand side words, then kDenseParts partial rows
    for (int l = C.lc; l < C.L - 1; l++) {
        const BLevel &A = C.lv[l];
        for (int k = threadIdx.x; k < A.n; k += kCoarseThreads) A.x[k] = 0.f;
        __syncthreads();
        c_jacobi(A, A.x, C.pre, C.omega);
        for (int k = threadIdx.x; k < A.n; k += kCoarseThreads)
            A.r[k] = node_resid9(A.st, A.dinv, A.x, A.b, A.nx, A.nz, k / A.nz, k % A.nz);
        __syncthreads();
        const BLevel &Cn = C.lv[l + 1];
        for (int k = threadIdx.x; k < Cn.n; k += kCoarseThreads)
            Cn.b[k] = node_restrict(A.r, A.nx, A.nz, Cn.dinv, Cn.nz, k / Cn.nz, k % Cn.nz);
        __syncthreads();
    }
    {
        // dense solve x = A^-1 b: kDenseParts threads per row, each summing a
        // contiguous quarter of the columns (8 loads in flight), the quarters
        // added in order
        const BLevel &Z = C.lv[C.L - 1];
        const int n = C.coarse_n;
        float *part = sb + n;
        for (int k = threadIdx.x; k < n; k += kCoarseThreads) sb[k] = Z.b[k];
        __syncthreads();
        const int len = (n + kDenseParts - 1) / kDenseParts;
        for (int t = threadIdx.x; t < n * kDenseParts; t += kCoarseThreads) {
            const int row = t % n, p = t / n;
            const int c0 = p * len, c1 = min(n, c0 + len);
            float s = 0.f;
            int c = c0;
            for (; c + 8 <= c1; c += 8) {
                float a[8];
#pragma unroll
                for (int q = 0; q < 8; q++) a[q] = C.invT[(size_t)(c + q) * n + row];
#pragma unroll
                for (int q = 0; q < 8; q++) s = fmaf(a[q], sb[c + q], s);
            }
            for (; c < c1; c++) s = fmaf(C.invT[(size_t)c * n + row], sb[c], s);
            part[p * n + row] = s;
        }
        __syncthreads();
        for (int row = threadIdx.x; row < n; row += kCoarseThreads) {
            float s = part[row];
            for (int p = 1; p < kDenseParts; p++) s += part[p * n + row];
            Z.x[row] = s;
        }
        __syncthreads();
    }
    for (int l = C.L - 2; l >= C.lc; l--) {
        const BLevel &A = C.lv[l];
        const BLevel &Cn = C.lv[l + 1];
        for (int k = threadIdx.x; k < A.n; k += kCoarseThreads)
            node_prolong(A.x, A.dir, A.nz, Cn.x, Cn.nx, Cn.nz, k / A.nz, k % A.nz);
        __syncthreads();
        c_jacobi(A, A.x, C.post, C.omega);
    }
}

// The common case of k_vcycle_coarse -- one Galerkin level of at most 1024
// nodes above a dense level of at most 256 -- with both levels held in LDS
// for the whole V-cycle (the same node bodies on LDS arrays, so the same
// bits).  Writes x of the Galerkin level.
constexpr int kLdsLevelNodes = 1024, kLdsDenseNodes = 256;

__global__ void __launch_bounds__(kCoarseThreads) k_vcycle_coarse_lds(CoarseArgs C)
{
    if (stopped(C.stop)) return;
    __shared__ float s_st[9 * kLdsLevelNodes], s_dv[kLdsLevelNodes], s_b[kLdsLevelNodes];
    __shared__ float s_x[kLdsLevelNodes], s_x2[kLdsLevelNodes], s_r[kLdsLevelNodes];
    __shared__ uint8_t s_dir[kLdsLevelNodes];
    __shared__ float s_zb[kLdsDenseNodes], s_zx[kLdsDenseNodes], s_part[kDenseParts * kLdsDenseNodes];
    const BLevel &A = C.lv[C.lc];
    const BLevel &Z = C.lv[C.lc + 1];
    const int n = A.n, nx = A.nx, nz = A.nz, cn = C.coarse_n;
    for (int k = threadIdx.x; k < 9 * n; k += kCoarseThreads) s_st[k] = A.st[k];
    for (int k = threadIdx.x; k < n; k += kCoarseThreads) {
        s_dv[k] = A.dinv[k];
        s_b[k] = A.b[k];
        s_dir[k] = A.dir[k];
        s_x[k] = 0.f;
    }
    __syncthreads();
    float *xa = s_x, *xb = s_x2;
    for (int it = 0; it < C.pre; it++) {
        for (int k = threadIdx.x; k < n; k += kCoarseThreads)
            xb[k] = node_jacobi9(s_st, s_dv, xa, s_b, nx, nz, C.omega, k / nz, k % nz);
        __syncthreads();
        float *t = xa; xa = xb; xb = t;
    }
    for (int k = threadIdx.x; k < n; k += kCoarseThreads)
        s_r[k] = node_resid9(s_st, s_dv, xa, s_b, nx, nz, k / nz, k % nz);
    __syncthreads();
    for (int k = threadIdx.x; k < cn; k += kCoarseThreads)
        s_zb[k] = node_restrict(s_r, nx, nz, Z.dinv, Z.nz, k / Z.nz, k % Z.nz);
    __syncthreads();
    {
        const int len = (cn + kDenseParts - 1) / kDenseParts;
        for (int t = threadIdx.x; t < cn * kDenseParts; t += kCoarseThreads) {
            const int row = t % cn, p = t / cn;
            const int c0 = p * len, c1 = min(cn, c0 + len);
            float s = 0.f;
            int c = c0;
            for (; c + 8 <= c1; c += 8) {
                float a[8];
#pragma unroll
                for (int q = 0; q < 8; q++) a[q] = C.invT[(size_t)(c + q) * cn + row];
#pragma unroll
                for (int q = 0; q < 8; q++) s = fmaf(a[q], s_zb[c + q], s);
            }
            for (; c < c1; c++) s = fmaf(C.invT[(size_t)c * cn + row], s_zb[c], s);
            s_part[p * cn + row] = s;
        }
        __syncthreads();
        for (int row = threadIdx.x; row < cn; row += kCoarseThreads) {
            float s = s_part[row];
            for (int p = 1; p < kDenseParts; p++) s += s_part[p * cn + row];
            s_zx[row] = s;
        }
        __syncthreads();
    }
    for (int k = threadIdx.x; k < n; k += kCoarseThreads)
        node_prolong(xa, s_dir, nz, s_zx, Z.nx, Z.nz, k / nz, k % nz);
    __syncthreads();
    for (int it = 0; it < C.post; it++) {
        for (int k = threadIdx.x; k < n; k += kCoarseThreads)
            xb[k] = node_jacobi9(s_st, s_dv, xa, s_b, nx, nz, C.omega, k / nz, k % nz);
        __syncthreads();
        float *t = xa; xa = xb; xb = t;
    }
    for (int k = threadIdx.x; k < n; k += kCoarseThreads) A.x[k] = xa[k];
}

// dense solve x = A^-1 b with the transposed inverse, one workgroup of
// kCoarseThreads: kDenseParts threads per row (the same sums as the dense
// step of k_vcycle_coarse); dynamic LDS of (1 + kDenseParts) n words
__global__ void __launch_bounds__(kCoarseThreads) k_dense_solve(const float *invT, const float *b, float *x, int n,
                                                                 const int *stop)
{
    if (stopped(stop)) return;
    extern __shared__ float sb[];
    float *part = sb + n;
    for (int k = threadIdx.x; k < n; k += kCoarseThreads) sb[k] = b[k];
    __syncthreads();
    const int len = (n + kDenseParts - 1) / kDenseParts;
    for (int t = threadIdx.x; t < n * kDenseParts; t += kCoarseThreads) {
        const int row = t % n, p = t / n;
        const int c0 = p * len, c1 = min(n, c0 + len);
        float s = 0.f;
        int c = c0;
        for (; c + 8 <= c1; c += 8) {
            float a[8];
#pragma unroll
            for (int q = 0; q < 8; q++) a[q] = invT[(size_t)(c + q) * n + row];
#pragma unroll
            for (int q = 0; q < 8; q++) s = fmaf(a[q], sb[c + q], s);
        }
        for (; c < c1; c++) s = fmaf(invT[(size_t)c * n + row], sb[c], s);
        part[p * n + row] = s;
    }
    __syncthreads();
    for (int row = threadIdx.x; row < n; row += kCoarseThreads) {
        float s = part[row];
        for (int p = 1; p < kDenseParts; p++) s += part[p * n + row];
        x[row] = s;
    }
}

// ------------------------------------------------------------------ host: hierarchy
static void p1d(int i, int nc, int I[2], float w[2], int &cnt)
{
    cnt = 0;
    if ((i & 1) == 0) {
        if (i / 2 < nc) { I[cnt] = i / 2; w[cnt++] = 1.f; }
    } else {
        if ((i - 1) / 2 < nc) { I[cnt] = (i - 1) / 2; w[cnt++] = 0.5f; }
        if ((i + 1) / 2 < nc) { I[cnt] = (i + 1) / 2; w[cnt++] = 0.5f; }
    }
}

// Galerkin A_c = P^T A_f P on the structured hierarchy, FP64 on the host
static void galerkin(const std::vector<double> &Af, const std::vector<uint8_t> &dirf, int nxf, int nzf,
                     std::vector<double> &Ac, std::vector<uint8_t> &dirc, int nxc, int nzc)
{
    Ac.assign((size_t)nxc * nzc * 9, 0.0);
    for (int i = 0; i < nxf; i++)
        for (int j = 0; j < nzf; j++) {
            size_t k = (size_t)i * nzf + j;
            if (dirf[k]) continue;
            int Ii[2], Jj[2], ci, cj;
            float wi[2], wj[2];
            p1d(i, nxc, Ii, wi, ci);
            p1d(j, nzc, Jj, wj, cj);
            for (int di = -1; di <= 1; di++)
                for (int dj = -1; dj <= 1; dj++) {
                    double a = Af[9 * k + (di + 1) * 3 + (dj + 1)];
                    if (a == 0.0) continue;
                    int i2 = i + di, j2 = j + dj;
                    if (i2 < 0 || i2 >= nxf || j2 < 0 || j2 >= nzf) continue;
                    if (dirf[(size_t)i2 * nzf + j2]) continue;
                    int Ii2[2], Jj2[2], ci2, cj2;
                    float wi2[2], wj2[2];
                    p1d(i2, nxc, Ii2, wi2, ci2);
                    p1d(j2, nzc, Jj2, wj2, cj2);
                    for (int u1 = 0; u1 < ci; u1++)
                        for (int v1 = 0; v1 < cj; v1++) {
                            double w1 = (double)wi[u1] * wj[v1];
                            size_t K = (size_t)Ii[u1] * nzc + Jj[v1];
                            for (int u2 = 0; u2 < ci2; u2++)
                                for (int v2 = 0; v2 < cj2; v2++) {
                                    int dI = Ii2[u2] - Ii[u1], dJ = Jj2[v2] - Jj[v1];
                                    double w2 = (double)wi2[u2] * wj2[v2];
                                    Ac[9 * K + (dI + 1) * 3 + (dJ + 1)] += w1 * a * w2;
                                }
                        }
                }
        }
    dirc.assign((size_t)nxc * nzc, 0);
    for (size_t K = 0; K < (size_t)nxc * nzc; K++)
        if (!(Ac[9 * K + 4] > 0.0)) {
            dirc[K] = 1;
            for (int q = 0; q < 9; q++) Ac[9 * K + q] = 0.0;
        }
    // drop couplings to pinned coarse nodes
    for (int I = 0; I < nxc; I++)
        for (int J = 0; J < nzc; J++) {
            size_t K = (size_t)I * nzc + J;
            for (int dI = -1; dI <= 1; dI++)
                for (int dJ = -1; dJ <= 1; dJ++) {
                    int I2 = I + dI, J2 = J + dJ;
                    if (I2 < 0 || I2 >= nxc || J2 < 0 || J2 >= nzc || dirc[(size_t)I2 * nzc + J2])
                        Ac[9 * K + (dI + 1) * 3 + (dJ + 1)] = 0.0;
                }
        }
}

// Gauss-Jordan inverse of the SPD coarsest operator restricted to free nodes
static bool dense_inverse(const std::vector<double> &A9, const std::vector<uint8_t> &dir, int nx, int nz,
                          std::vector<float> &inv)
{
    int n = nx * nz;
    std::vector<int> id;
    for (int k = 0; k < n; k++) if (!dir[k]) id.push_back(k);
    int m = (int)id.size();
    std::vector<int> pos(n, -1);
    for (int q = 0; q < m; q++) pos[id[q]] = q;
    std::vector<double> a((size_t)m * 2 * m, 0.0);
    for (int q = 0; q < m; q++) {
        int k = id[q], i = k / nz, j = k % nz;
        for (int di = -1; di <= 1; di++)
            for (int dj = -1; dj <= 1; dj++) {
                int i2 = i + di, j2 = j + dj;
                if (i2 < 0 || i2 >= nx || j2 < 0 || j2 >= nz) continue;
                int p = pos[i2 * nz + j2];
                if (p < 0) continue;
                a[(size_t)q * 2 * m + p] = A9[9 * (size_t)k + (di + 1) * 3 + (dj + 1)];
            }
        a[(size_t)q * 2 * m + m + q] = 1.0;
    }
    for (int c = 0; c < m; c++) {
        int piv = c;
        for (int r = c + 1; r < m; r++)
            if (fabs(a[(size_t)r * 2 * m + c]) > fabs(a[(size_t)piv * 2 * m + c])) piv = r;
        if (fabs(a[(size_t)piv * 2 * m + c]) < 1e-300) return false;
        if (piv != c)
            for (int t = 0; t < 2 * m; t++) std::swap(a[(size_t)c * 2 * m + t], a[(size_t)piv * 2 * m + t]);
        double d = 1.0 / a[(size_t)c * 2 * m + c];
        for (int t = 0; t < 2 * m; t++) a[(size_t)c * 2 * m + t] *= d;
        for (int r = 0; r < m; r++) {
            if (r == c) continue;
            double f = a[(size_t)r * 2 * m + c];
            if (f == 0.0) continue;
            for (int t = 0; t < 2 * m; t++) a[(size_t)r * 2 * m + t] -= f * a[(size_t)c * 2 * m + t];
        }
    }
    inv.assign((size_t)n * n, 0.f);
    for (int q = 0; q < m; q++)
        for (int p = 0; p < m; p++) inv[(size_t)id[q] * n + id[p]] = (float)a[(size_t)q * 2 * m + m + p];
    return true;
}

static int dmalloc(void **p, size_t bytes)
{
    hipError_t e = hipMalloc(p, bytes < 16 ? 16 : bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        return set_error(M2D_ERR_NO_MEMORY, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    }
    return M2D_OK;
}

static int mg_setup(m2d_fields2d *f)
{
    Mg &mg = f->mg;
    const int nx = f->nx, nz = f->nz;
    // finest level operator from the mask
    std::vector<uint8_t> dir((size_t)nx * nz);
    for (size_t k = 0; k < dir.size(); k++) dir[k] = f->h_mask[k] == M2D_FIXED || f->h_mask[k] == M2D_FIXED_RF;
    std::vector<double> A((size_t)nx * nz * 9, 0.0);
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < nz; j++) {
            size_t k = (size_t)i * nz + j;
            if (dir[k]) continue;
            A[9 * k + 4] = 4.0;
            const int di[4] = {-1, 1, 0, 0}, dj[4] = {0, 0, -1, 1};
            for (int q = 0; q < 4; q++) {
                size_t k2 = (size_t)(i + di[q]) * nz + (j + dj[q]);
                if (!dir[k2]) A[9 * k + (di[q] + 1) * 3 + (dj[q] + 1)] = -1.0;
            }
        }
    std::vector<std::vector<double>> ops{A};
    std::vector<std::vector<uint8_t>> dirs{dir};
    std::vector<std::pair<int, int>> dims{{nx, nz}};
    while (true) {
        int cx = dims.back().first, cz = dims.back().second;
        if ((size_t)cx * cz <= 256 || cx < 5 || cz < 5) break;
        int nxc = (cx + 1) / 2, nzc = (cz + 1) / 2;
        std::vector<double> Ac;
        std::vector<uint8_t> dc;
        galerkin(ops.back(), dirs.back(), cx, cz, Ac, dc, nxc, nzc);
        ops.push_back(std::move(Ac));
        dirs.push_back(std::move(dc));
        dims.push_back({nxc, nzc});
    }
    mg.lv.resize(ops.size());
    for (size_t l = 0; l < ops.size(); l++) {
        MgLevel &L = mg.lv[l];
        L.nx = dims[l].first;
        L.nz = dims[l].second;
        L.n = (size_t)L.nx * L.nz;
        std::vector<float> st(9 * L.n), dinv(L.n);
        for (size_t k = 0; k < L.n; k++) {
            for (int q = 0; q < 9; q++) st[9 * k + q] = (float)ops[l][9 * k + q];
            double d = ops[l][9 * k + 4];
            dinv[k] = dirs[l][k] || !(d > 0) ? 0.f : (float)(1.0 / d);
        }
        M2D_TRY(dmalloc((void **)&L.st, sizeof(float) * 9 * L.n));
        M2D_TRY(dmalloc((void **)&L.dinv, sizeof(float) * L.n));
        M2D_TRY(dmalloc((void **)&L.dir, L.n));
        M2D_TRY(dmalloc((void **)&L.x, sizeof(float) * L.n));
        M2D_TRY(dmalloc((void **)&L.x2, sizeof(float) * L.n));
        M2D_TRY(dmalloc((void **)&L.r, sizeof(float) * L.n));
        M2D_TRY(dmalloc((void **)&L.b, sizeof(float) * L.n));
        M2D_HIP(hipMemcpy(L.st, st.data(), sizeof(float) * 9 * L.n, hipMemcpyHostToDevice));
        M2D_HIP(hipMemcpy(L.dinv, dinv.data(), sizeof(float) * L.n, hipMemcpyHostToDevice));
        M2D_HIP(hipMemcpy(L.dir, dirs[l].data(), L.n, hipMemcpyHostToDevice));
    }
    std::vector<float> inv;
    const MgLevel &C = mg.lv.back();
    M2D_CHECK(dense_inverse(ops.back(), dirs.back(), C.nx, C.nz, inv), M2D_ERR_BAD_ARG,
              "Fields: coarsest multigrid operator is singular");
    mg.coarse_n = (int)C.n;
    M2D_TRY(dmalloc((void **)&mg.coarse_inv, sizeof(float) * inv.size()));
    M2D_HIP(hipMemcpy(mg.coarse_inv, inv.data(), sizeof(float) * inv.size(), hipMemcpyHostToDevice));
    {
        const size_t cn = C.n;
        std::vector<float> invT(cn * cn);
        for (size_t r = 0; r < cn; r++)
            for (size_t c = 0; c < cn; c++) invT[c * cn + r] = inv[r * cn + c];
        M2D_TRY(dmalloc((void **)&mg.coarse_invT, sizeof(float) * invT.size()));
        M2D_HIP(hipMemcpy(mg.coarse_invT, invT.data(), sizeof(float) * invT.size(), hipMemcpyHostToDevice));
    }
    const size_t n = f->n;
    M2D_TRY(dmalloc((void **)&mg.pr, sizeof(float) * n));
    M2D_TRY(dmalloc((void **)&mg.pz, sizeof(float) * n));
    M2D_TRY(dmalloc((void **)&mg.pp, sizeof(float) * n));
    M2D_TRY(dmalloc((void **)&mg.pp2, sizeof(float) * n));
    M2D_TRY(dmalloc((void **)&mg.pq, sizeof(float) * n));
    M2D_TRY(dmalloc((void **)&mg.pe, sizeof(float) * n));
    M2D_TRY(dmalloc((void **)&mg.r64, sizeof(double) * n));
    // 16 device scalars, then the solve's control block (SolveCtl)
    M2D_TRY(dmalloc((void **)&mg.dots, sizeof(double) * 32));
    M2D_TRY(dmalloc((void **)&mg.ticket, sizeof(unsigned int)));
    M2D_HIP(hipMemset(mg.ticket, 0, sizeof(unsigned int)));
    // the coarse part of the V-cycle runs in one workgroup from the first
    // level of at most kBlockLevelNodes nodes (the dense level at the latest)
    mg.lc = (int)mg.lv.size() - 1;
    for (int l = 1; l < (int)mg.lv.size(); l++)
        if (mg.lv[l].n <= kBlockLevelNodes) { mg.lc = l; break; }
    M2D_CHECK(mg.lv.size() <= (size_t)kMaxLevels && mg.coarse_n <= 3200, M2D_ERR_BAD_ARG,
              "Fields: multigrid hierarchy of %d levels, coarsest %d nodes", (int)mg.lv.size(), mg.coarse_n);
    M2D_HIP(hipHostMalloc((void **)&mg.h_dots, sizeof(double) * 32));
    mg.npartial = (int)std::min<size_t>((n + 255) / 256, 65536);
    // up to three sums per block of the fused reductions
    M2D_TRY(dmalloc((void **)&mg.partial, sizeof(double) * 3 * std::max<size_t>(mg.npartial, (n + 255) / 256)));
    return M2D_OK;
}

static void mg_free(Mg &mg)
{
    for (MgLevel &L : mg.lv)
        for (void *p : {(void *)L.st, (void *)L.dinv, (void *)L.dir, (void *)L.x, (void *)L.x2,
                        (void *)L.r, (void *)L.b})
            if (p) (void)hipFree(p);
    mg.lv.clear();
    for (void *p : {(void *)mg.coarse_inv, (void *)mg.pr, (void *)mg.pz, (void *)mg.pp, (void *)mg.pp2, (void *)mg.pq,
                    (void *)mg.pe, (void *)mg.r64, (void *)mg.dots, (void *)mg.partial, (void *)mg.ticket, (void *)mg.coarse_invT})
        if (p) (void)hipFree(p);
    if (mg.h_dots) (void)hipHostFree(mg.h_dots);
    mg = Mg();
}

// ------------------------------------------------------------------ host: V-cycle

// one V-cycle: the finest level in grid kernels (first sweep from zero,
// fused residual + restriction, fused prolongation + black sweep), middle
// levels in grid kernels, and every level from mg.lc down in one workgroup
// edge of the finest level's LDS tiles: more, smaller tiles on smaller grids,
// whose V-cycle is bound by latency rather than bytes (with edges of 16 and
// 32 the 128^2 and 512^2 solves kept most CUs idle)
constexpr int kT0Tiny = 4, kT0Small = 8, kT0Large = 16;
static int l0_tile(size_t n) { return n <= 16384 ? kT0Tiny : n <= 65536 ? kT0Small : kT0Large; }

static int vcycle(m2d_fields2d *f, const float *rhs, float *out, const int *stop)
{
    Mg &mg = f->mg;
    hipStream_t s = f->ctx->stream;
    const int L = (int)mg.lv.size();
    MgLevel &A0 = mg.lv[0];
    if (L == 1) {
        k_coarse_solve<<<1, 256, sizeof(float) * mg.coarse_n, s>>>(mg.coarse_inv, rhs, out, mg.coarse_n, stop);
        M2D_HIP(hipGetLastError());
        return M2D_OK;
    }
    // ---- level 0 descent (LDS tiles: smoothing, residual, restriction)
    const int T0 = l0_tile(A0.n);
    const dim3 gt((A0.nz + T0 - 1) / T0, (A0.nx + T0 - 1) / T0);
    {
        MgLevel &C = mg.lv[1];
        auto down = [&](auto kern) {
            kern<<<gt, kT0Threads, 0, s>>>(A0.dir, rhs, A0.x, A0.nx, A0.nz, C.b, C.dinv, C.nx, C.nz, stop);
        };
        switch (T0) {
        case 4: down(k_l0_down<4>); break;
        case 8: down(k_l0_down<8>); break;
        case 16: down(k_l0_down<16>); break;
        default: down(k_l0_down<32>); break;
        }
    }
    // ---- middle levels [1, lc) in LDS tiles
    for (int l = 1; l < mg.lc; l++) {
        MgLevel &A = mg.lv[l];
        MgLevel &C = mg.lv[l + 1];
        const dim3 g((A.nz + kTm - 1) / kTm, (A.nx + kTm - 1) / kTm);
        k_lm_down<<<g, kT0Threads, 0, s>>>(A.st, A.dinv, A.b, A.x, A.nx, A.nz, C.b, C.dinv, C.nx, C.nz, kOmega,
                                           stop);
    }
    // ---- levels [lc, L) in one workgroup
    {
        CoarseArgs C;
        C.lc = mg.lc;
        C.L = L;
        for (int l = 0; l < L; l++) {
            const MgLevel &A = mg.lv[l];
            C.lv[l] = BLevel{A.nx, A.nz, (int)A.n, A.st, A.dinv, A.dir, A.x, A.x2, A.r, A.b};
        }
        C.invT = mg.coarse_invT;
        C.coarse_n = mg.coarse_n;
        C.omega = kOmega;
        C.pre = kPre;
        C.post = kPost;
        C.stop = stop;
        if (L - 1 - mg.lc == 1 && mg.lv[mg.lc].n <= (size_t)kLdsLevelNodes && mg.coarse_n <= kLdsDenseNodes)
            k_vcycle_coarse_lds<<<1, kCoarseThreads, 0, s>>>(C);
        else
            k_vcycle_coarse<<<1, kCoarseThreads, sizeof(float) * mg.coarse_n * (1 + kDenseParts), s>>>(C);
    }
    // ---- middle levels ascent: the descent's x in lv[l].x, the ascent's in
    // lv[l].x2 (level lc's from the one-workgroup kernel in lv[lc].x)
    auto ascent_x = [&](int l) -> const float * { return l >= mg.lc ? mg.lv[l].x : mg.lv[l].x2; };
    for (int l = mg.lc - 1; l >= 1; l--) {
        MgLevel &A = mg.lv[l];
        MgLevel &Cc = mg.lv[l + 1];
        const dim3 g((A.nz + kTm - 1) / kTm, (A.nx + kTm - 1) / kTm);
        k_lm_up<<<g, kT0Threads, 0, s>>>(A.st, A.dinv, A.dir, A.b, A.x, A.x2, A.nx, A.nz, ascent_x(l + 1), Cc.nx,
                                         Cc.nz, kOmega, stop);
    }
    // ---- level 0 ascent (LDS tiles: prolongation, smoothing)
    {
        MgLevel &C = mg.lv[1];
        auto up = [&](auto kern) {
            kern<<<gt, kT0Threads, 0, s>>>(A0.dir, rhs, A0.x, out, A0.nx, A0.nz, ascent_x(1), C.nx, C.nz, stop);
        };
        switch (T0) {
        case 4: up(k_l0_up<4>); break;
        case 8: up(k_l0_up<8>); break;
        case 16: up(k_l0_up<16>); break;
        default: up(k_l0_up<32>); break;
        }
    }
    M2D_HIP(hipGetLastError());
    return M2D_OK;
}

// fused-reduction launch shape over the finest level
static unsigned red_blocks(size_t n) { return (unsigned)((n + kRedBlock - 1) / kRedBlock); }

__global__ void k_ctl_init(SolveCtl *c)
{
    if (threadIdx.x) return;
    SolveCtl z{};
    z.stop_pcg = 1;
    z.no_add = 1;
    z.rel = INFINITY;
    *c = z;
}

// opens a CG pass on the FP32 residual d (|d|^2 = dots[10])
__global__ void k_ctl_open(SolveCtl *c, const double *dots, double tol)
{
    if (threadIdx.x || c->halted) return;
    c->pass_iters = 0;
    const double r0 = sqrt(dots[10]);
    if (c->stop_outer || r0 == 0.0) {
        // nothing to solve: no iteration runs and nothing is added
        c->stop_pcg = 1;
        c->no_add = 1;
        return;
    }
    c->thr = inner_rtol(c->rel, tol) * r0;
    c->stop_pcg = 0;
    c->no_add = 0;
}

// after the last iteration queued for a pass: a pass that has neither
// converged nor reached max_iters is suspended -- every launch queued after
// it returns at once -- until the host queues more of its iterations
// (k_ctl_resume); at max_iters it ends, as a loop with that cap would end it
__global__ void k_ctl_end(SolveCtl *c, int max_iters)
{
    if (threadIdx.x || c->stop_pcg) return;
    c->stop_pcg = 1;
    if (c->pass_iters >= max_iters) return;
    c->halted = 1;
    c->stop_outer = 1;
    c->no_add = 1;
}

// the suspended pass continues: its CG state is where it stopped
__global__ void k_ctl_resume(SolveCtl *c)
{
    if (threadIdx.x || !c->halted) return;
    c->halted = 0;
    c->stop_outer = 0;
    c->stop_pcg = 0;
    c->no_add = 0;
}

// after the pass's correction
__global__ void k_ctl_close(SolveCtl *c)
{
    if (threadIdx.x || c->no_add) return;
    c->outer++;
    c->no_add = 1;
}

static SolveCtl *solve_ctl(Mg &mg) { return (SolveCtl *)(mg.dots + 16); }

// Iterations from + 1 .. to of the open CG pass on A' e = d (d in mg.pr, e
// in mg.pe), from 0: the pass's start; every launch stops with the pass
// (SolveCtl::stop_pcg).  cur: which of dots[0], dots[4] holds the current
// r.z, (iterations run - 1) & 1.  dots: [1] p.q, [2] r.r
static int queue_pcg(m2d_fields2d *f, int from, int to, int &cur)
{
    Mg &mg = f->mg;
    hipStream_t s = f->ctx->stream;
    const size_t n = f->n;
    const unsigned rb = red_blocks(n);
    const MgLevel &L0 = mg.lv[0];
    SolveCtl *ctl = solve_ctl(mg);
    const int *stop = &ctl->stop_pcg;
    // iteration it's direction goes to P[it & 1], the previous one's is read from the other
    float *const P[2] = {mg.pp, mg.pp2};
    if (from == 0) {
        M2D_TRY(vcycle(f, mg.pr, mg.pz, stop));
        reduce_launch<1>(rb, mg.partial, mg.dots, s, [&](int fin) {
            k_dot_fin<<<rb, kRedBlock, 0, s>>>(mg.pr, mg.pz, n, mg.partial, mg.ticket, mg.dots, fin, stop);
        }, stop);
        cur = 0;
    }
    for (int it = from + 1; it <= to; it++) {
        if (it > 1) {
            M2D_TRY(vcycle(f, mg.pr, mg.pz, stop));
            reduce_launch<1>(rb, mg.partial, mg.dots + 4 * (cur ^ 1), s, [&](int fin) {
                k_dot_fin<<<rb, kRedBlock, 0, s>>>(mg.pr, mg.pz, n, mg.partial, mg.ticket, mg.dots + 4 * (cur ^ 1),
                                                   fin, stop);
            }, stop);
            cur ^= 1;
        }
        const float *pold = P[(it - 1) & 1];
        float *pnew = P[it & 1];
        reduce_launch<1>(rb, mg.partial, mg.dots + 1, s, [&](int fin) {
            k_dir_apply5_dot<<<rb, kRedBlock, 0, s>>>(L0.dir, mg.pz, pold, pnew, mg.pq, mg.dots + 4 * cur,
                                                      mg.dots + 4 * (cur ^ 1), it == 1, n, L0.nz, mg.partial,
                                                      mg.ticket, mg.dots + 1, fin, stop);
        }, stop);
        reduce_launch<1>(rb, mg.partial, mg.dots + 2, s, [&](int fin) {
            k_cg_update_dot<<<rb, kRedBlock, 0, s>>>(mg.pe, mg.pr, pnew, mg.pq, mg.dots + 4 * cur, mg.dots + 1, n,
                                                     it == 1, mg.partial, mg.ticket, mg.dots + 2, fin, stop,
                                                     CtlCheck{ctl});
        }, stop, CtlCheck{ctl});
    }
    M2D_HIP(hipGetLastError());
    return M2D_OK;
}

// the FP64 residual of u, its norms, and the refinement's decision on them
static void queue_residual(m2d_fields2d *f, double *u, double tol, int max_outer)
{
    Mg &mg = f->mg;
    hipStream_t s = f->ctx->stream;
    const size_t n = f->n;
    const unsigned rb = red_blocks(n);
    SolveCtl *ctl = solve_ctl(mg);
    const int *stop = &ctl->stop_outer;
    const CtlEval post{ctl, tol, max_outer};
    reduce_launch<3>(rb, mg.partial, mg.dots + 8, s, [&](int fin) {
        k_residual64_fin<<<rb, kRedBlock, 0, s>>>(f->mask, f->rho, u, n, f->nz, mg.r64, mg.pr, mg.partial, mg.ticket,
                                                  mg.dots + 8, fin, stop, post);
    }, stop, post);
}

// FP64 iterative refinement of u over FP32 multigrid-preconditioned CG
// passes, queued whole (queue_refinement)
static int solve(m2d_fields2d *f, int mode, double tol, m2d_solve_info *info)
{
    hipStream_t s = f->ctx->stream;
    Mg &mg = f->mg;
    const size_t n = f->n;
    const unsigned gb = (unsigned)((n + 255) / 256);
    double *u = mode == M2D_SOLVE_RF ? f->urf : f->u;
    double scale = -(f->p.grid.dx * f->p.grid.dx) / kEps0 / f->p.dV;  // fields.cpp:307
    if (tol <= 0) tol = 1e-10;
    SolveCtl *ctl = solve_ctl(mg);
    SolveCtl *h = (SolveCtl *)(mg.h_dots + 16);
    k_rhs<<<gb, 256, 0, s>>>(f->mask, f->voltage, f->rho, u, n, mode, scale);
    k_ctl_init<<<1, 64, 0, s>>>(ctl);
    queue_residual(f, u, tol, kMaxOuter);
    M2D_TRY(queue_refinement(
        s, ctl, h, mg.dots, mg.est, tol, 60, [&](int from, int to, int &cur) { return queue_pcg(f, from, to, cur); },
        [&] { k_add_correction<<<gb, 256, 0, s>>>(f->mask, u, mg.pe, n, &ctl->no_add); },
        [&] { queue_residual(f, u, tol, kMaxOuter); }));
    if (info) { info->outer_iters = h->outer; info->inner_iters = h->iters; info->rel_residual = h->rel; }
    if (h->nan) return set_error(M2D_ERR_NAN, "Fields solve: residual is not finite");
    M2D_TRY(fields_build_tables(f));
    if (h->rel > tol)
        return set_error(M2D_ERR_NOT_CONVERGED, "Fields solve: relative residual %.3e > %.3e after %d steps",
                         h->rel, tol, h->outer);
    return M2D_OK;
}

}  // namespace m2d

using namespace m2d;

extern "C" {

int m2d_fields2d_create(m2d_ctx *ctx, const m2d_fields2d_params *p, const uint8_t *mask,
                        const double *voltage, m2d_fields2d **out)
{
    M2D_CHECK(ctx && p && mask && voltage && out, M2D_ERR_BAD_ARG, "m2d_fields2d_create: NULL argument");
    *out = nullptr;
    const m2d_grid2d &g = p->grid;
    M2D_CHECK(g.nx >= 3 && g.nz >= 3 && g.dx > 0 && g.dz > 0, M2D_ERR_BAD_ARG,
              "m2d_fields2d_create: grid %dx%d spacing %g,%g", g.nx, g.nz, g.dx, g.dz);
    // node indices are 32-bit, the gathers multiply in 24 bits and address
    // the FP32 edge tables ((nx+1) (nz+1) floats at most) by 32-bit byte offsets
    M2D_CHECK(g.nx < (1 << 24) && g.nz < (1 << 24) && (size_t)(g.nx + 1) * (g.nz + 1) < (1ull << 30),
              M2D_ERR_BAD_ARG, "m2d_fields2d_create: grid %dx%d too large", g.nx, g.nz);
    M2D_CHECK(p->dV > 0, M2D_ERR_BAD_ARG, "m2d_fields2d_create: dV must be > 0");
    const size_t n = (size_t)g.nx * g.nz;
    // a free 5-point row on the frame would reach outside the grid (fields.cpp:229-237)
    for (int i = 0; i < g.nx; i++)
        for (int j = 0; j < g.nz; j++) {
            uint8_t m = mask[(size_t)i * g.nz + j];
            M2D_CHECK(m <= M2D_BOUNDARY, M2D_ERR_BAD_ARG, "mask value %d at (%d,%d)", (int)m, i, j);
            bool edge = i == 0 || j == 0 || i == g.nx - 1 || j == g.nz - 1;
            M2D_CHECK(!edge || m == M2D_FIXED || m == M2D_FIXED_RF, M2D_ERR_BAD_ARG,
                      "m2d_fields2d_create: node (%d,%d) on the grid edge is not Dirichlet; its "
                      "Cartesian 5-point row would leave the grid", i, j);
        }
    M2D_HIP(hipSetDevice(ctx->device));
    m2d_fields2d *f = new (std::nothrow) m2d_fields2d();
    M2D_CHECK(f, M2D_ERR_NO_MEMORY, "m2d_fields2d_create: host allocation");
    f->ctx = ctx;
    ctx_retain(ctx);
    f->p = *p;
    f->nx = g.nx;
    f->nz = g.nz;
    f->n = n;
    f->idx = 1.0 / g.dx;
    f->idz = 1.0 / g.dz;
    f->h_mask.assign(mask, mask + n);
    int r = M2D_OK;
    auto fail = [&](int code) { m2d_fields2d_destroy(f); return code; };
    if ((r = dmalloc((void **)&f->mask, n)) || (r = dmalloc((void **)&f->cellfree, n)) ||
        (r = dmalloc((void **)&f->voltage, 8 * n)) || (r = dmalloc((void **)&f->u, 8 * n)) ||
        (r = dmalloc((void **)&f->urf, 8 * n)) || (r = dmalloc((void **)&f->rho, 8 * n)) ||
        (r = dmalloc((void **)&f->gx, sizeof(float2) * (size_t)(g.nx + 1) * g.nz)) ||
        (r = dmalloc((void **)&f->gz, sizeof(float2) * (size_t)g.nx * (g.nz + 1))) ||
        (r = dmalloc((void **)&f->phx, sizeof(float) * (size_t)(g.nx + 1) * g.nz)) ||
        (r = dmalloc((void **)&f->phz, sizeof(float) * (size_t)g.nx * (g.nz + 1))))
        return fail(r);
    // sort tiles of 8 x 8 cells; the deposit window of tile (ti, tj) spans
    // cells [8 ti - 8, 8 ti + 16) x [8 tj - 8, 8 tj + 16)
    const int ntx = (g.nx - 2) / kTile + 1, ntz = (g.nz - 2) / kTile + 1;
    if ((r = dmalloc((void **)&f->winfree, (size_t)ntx * ntz))) return fail(r);
    std::vector<uint8_t> cf(n, 0);
    for (int i = 0; i + 1 < g.nx; i++)
        for (int j = 0; j + 1 < g.nz; j++) {
            size_t k = (size_t)i * g.nz + j;
            cf[k] = mask[k] == M2D_FREE || mask[k + g.nz] == M2D_FREE || mask[k + 1] == M2D_FREE ||
                    mask[k + g.nz + 1] == M2D_FREE;
        }
    std::vector<uint8_t> wf((size_t)ntx * ntz, 0);
    for (int ti = 0; ti < ntx; ti++)
        for (int tj = 0; tj < ntz; tj++) {
            const int i0 = ti * kTile - kHalo, j0 = tj * kTile - kHalo;
            bool all = i0 >= 0 && j0 >= 0 && i0 + kWin - 2 <= g.nx - 2 && j0 + kWin - 2 <= g.nz - 2;
            for (int i = i0; all && i < i0 + kWin - 1; i++)
                for (int j = j0; all && j < j0 + kWin - 1; j++) all = cf[(size_t)i * g.nz + j] != 0;
            wf[(size_t)ti * ntz + tj] = all;
        }
    if (hipMemcpy(f->mask, mask, n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(f->winfree, wf.data(), wf.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(f->cellfree, cf.data(), n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(f->voltage, voltage, 8 * n, hipMemcpyHostToDevice) != hipSuccess)
        return fail(set_error(M2D_ERR_HIP, "m2d_fields2d_create: upload failed"));
    if ((r = fill_bytes(ctx, f->u, 0, 8 * n)) || (r = fill_bytes(ctx, f->urf, 0, 8 * n)) ||
        (r = fill_bytes(ctx, f->rho, 0, 8 * n)))
        return fail(r);
    if ((r = mg_setup(f))) return fail(r);
    if ((r = fields_build_tables(f))) return fail(r);
    if (hipStreamSynchronize(ctx->stream) != hipSuccess)
        return fail(set_error(M2D_ERR_HIP, "m2d_fields2d_create: synchronise failed"));
    *out = f;
    return M2D_OK;
}

int m2d_fields2d_destroy(m2d_fields2d *f)
{
    if (!f) return M2D_OK;
    if (f->ctx) (void)hipStreamSynchronize(f->ctx->stream);
    for (void *p : {(void *)f->mask, (void *)f->cellfree, (void *)f->winfree, (void *)f->voltage, (void *)f->u,
                    (void *)f->urf, (void *)f->rho, (void *)f->gx, (void *)f->gz, (void *)f->phx, (void *)f->phz, (void *)f->bt,
                    (void *)f->smooth_tmp})
        if (p) (void)hipFree(p);
    mg_free(f->mg);
    if (f->ctx) ctx_release(f->ctx);
    delete f;
    return M2D_OK;
}

int m2d_fields2d_set_magnetic_field(m2d_fields2d *f, const m2d_grid2d *g, const double *Br, const double *Bz)
{
    M2D_CHECK(f, M2D_ERR_BAD_ARG, "m2d_fields2d_set_magnetic_field: NULL fields");
    M2D_HIP(hipStreamSynchronize(f->ctx->stream));
    if (f->bt) {
        M2D_HIP(hipFree(f->bt));
        f->bt = nullptr;
    }
    if (!Br) return M2D_OK;  // back to the constant B of the params
    M2D_CHECK(g && Bz, M2D_ERR_BAD_ARG, "m2d_fields2d_set_magnetic_field: NULL argument");
    M2D_CHECK(g->nx >= 2 && g->nz >= 2 && g->dx > 0 && g->dz > 0 && (int64_t)g->nx * g->nz < (1ll << 30),
              M2D_ERR_BAD_ARG, "m2d_fields2d_set_magnetic_field: bad grid %d x %d", g->nx, g->nz);
    const size_t n = (size_t)g->nx * g->nz;
    std::vector<float2> h(n);
    for (size_t k = 0; k < n; k++) {
        // Fields::load_magnetic_field (fields.cpp:954-958)
        M2D_CHECK(!std::isnan(Br[k]) && !std::isnan(Bz[k]), M2D_ERR_BAD_ARG,
                  "Fields::load_magnetic_field() garbage loaded");
        h[k] = make_float2((float)Br[k], (float)Bz[k]);
    }
    M2D_TRY(dmalloc((void **)&f->bt, n * sizeof(float2)));
    M2D_HIP(hipMemcpy(f->bt, h.data(), n * sizeof(float2), hipMemcpyHostToDevice));
    f->bgrid = *g;
    return M2D_OK;
}

// Fields::u_smooth, symmetry part (fields.cpp:38-72): thread (i, j) of the
// triangle 0 <= j <= i <= ic/2 averages the 8 mirror images of the node on
// the square grid, in the reference's order, and writes the average to all
// of them.  The triangle is a fundamental domain of the square's symmetries,
// so no two threads touch the same node.
__global__ void k_smooth_sym(double *u, int lmax, int ic, int jc)
{
    const int i = blockIdx.y, j = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > ic / 2 || j > i) return;
    double *n[8] = {u + (size_t)i * lmax + j,        u + (size_t)(ic - i) * lmax + j,
                    u + (size_t)i * lmax + (jc - j), u + (size_t)(ic - i) * lmax + (jc - j),
                    u + (size_t)(jc - j) * lmax + (ic - i), u + (size_t)j * lmax + i,
                    u + (size_t)(jc - j) * lmax + i, u + (size_t)j * lmax + (ic - i)};
    double sum = 0;
#pragma unroll
    for (int k = 0; k < 8; k++) sum += *n[k];
    sum /= 8.0;
#pragma unroll
    for (int k = 0; k < 8; k++) *n[k] = sum;
}

// Fields::u_smooth, filter part (fields.cpp:74-111): the 2dx x 2dx
// square-particle stencil on a copy t of the field, at the interior nodes
// within r2 (squared, in cells) of the centre (r2 <= 0: everywhere).  Every
// product is by a power of two, so the sum has the reference's bits whether
// or not it is fused.
__global__ void k_smooth9(const double *t, double *u, int jmax, int lmax, double icf, double jcf, double r2)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x + 1, i = blockIdx.y + 1;
    if (i > jmax - 2 || j > lmax - 2) return;
    if (r2 > 0) {
        const double r = (i - icf) * (i - icf) + (j - jcf) * (j - jcf);
        if (r > r2) return;
    }
    const double *c = t + (size_t)i * lmax + j;
    const double sum = c[0] + c[-lmax] * 0.5 + c[lmax] * 0.5 + c[-1] * 0.5 + c[1] * 0.5 + c[-lmax - 1] * 0.25 +
                       c[lmax - 1] * 0.25 + c[-lmax + 1] * 0.25 + c[lmax + 1] * 0.25;
    u[(size_t)i * lmax + j] = sum / 4.0;
}

static double *field_ptr(m2d_fields2d *f, int which)
{
    switch (which) {
    case M2D_FIELD_U: return f->u;
    case M2D_FIELD_URF: return f->urf;
    case M2D_FIELD_RHO: return f->rho;
    case M2D_FIELD_VOLTAGE: return f->voltage;
    default: return nullptr;
    }
}

double *m2d_fields2d_data(m2d_fields2d *f, int which) { return f ? field_ptr(f, which) : nullptr; }

int m2d_fields2d_upload(m2d_fields2d *f, int which, const double *host)
{
    M2D_CHECK(f && host, M2D_ERR_BAD_ARG, "m2d_fields2d_upload: NULL argument");
    double *d = field_ptr(f, which);
    M2D_CHECK(d, M2D_ERR_BAD_ARG, "m2d_fields2d_upload: unknown field %d", which);
    M2D_HIP(hipMemcpyAsync(d, host, 8 * f->n, hipMemcpyHostToDevice, f->ctx->stream));
    if (which == M2D_FIELD_U || which == M2D_FIELD_URF) M2D_TRY(fields_build_tables(f));
    M2D_HIP(hipStreamSynchronize(f->ctx->stream));
    return M2D_OK;
}

int m2d_fields2d_download(m2d_fields2d *f, int which, double *host)
{
    M2D_CHECK(f && host, M2D_ERR_BAD_ARG, "m2d_fields2d_download: NULL argument");
    double *d = field_ptr(f, which);
    M2D_CHECK(d, M2D_ERR_BAD_ARG, "m2d_fields2d_download: unknown field %d", which);
    M2D_HIP(hipMemcpyAsync(host, d, 8 * f->n, hipMemcpyDeviceToHost, f->ctx->stream));
    M2D_HIP(hipStreamSynchronize(f->ctx->stream));
    return M2D_OK;
}

int m2d_fields2d_update(m2d_fields2d *f)
{
    M2D_CHECK(f, M2D_ERR_BAD_ARG, "m2d_fields2d_update: NULL argument");
    return fields_build_tables(f);
}

int m2d_fields2d_u_smooth(m2d_fields2d *f, int which, int symmetry, double radius)
{
    M2D_CHECK(f, M2D_ERR_BAD_ARG, "m2d_fields2d_u_smooth: NULL argument");
    M2D_CHECK(which == M2D_FIELD_U || which == M2D_FIELD_URF, M2D_ERR_BAD_ARG,
              "m2d_fields2d_u_smooth: smooths u or uRF");
    double *u = field_ptr(f, which);
    const int jmax = f->nx, lmax = f->nz, ic = jmax - 1, jc = lmax - 1;
    hipStream_t st = f->ctx->stream;
    if (symmetry) {
        M2D_CHECK(ic == jc, M2D_ERR_UNSUPPORTED, "Fields::u_smooth: smoothing of non-square matrix not implemented");
        const int h = ic / 2 + 1;
        k_smooth_sym<<<dim3((unsigned)((h + 127) / 128), (unsigned)h), 128, 0, st>>>(u, lmax, ic, jc);
        M2D_HIP(hipGetLastError());
    }
    if (!f->smooth_tmp) M2D_TRY(dmalloc((void **)&f->smooth_tmp, 8 * f->n));
    M2D_HIP(hipMemcpyAsync(f->smooth_tmp, u, 8 * f->n, hipMemcpyDeviceToDevice, st));
    const double r2 = radius > 0 ? (radius / f->p.grid.dx) * (radius / f->p.grid.dx) : radius;
    if (jmax > 2 && lmax > 2) {
        k_smooth9<<<dim3((unsigned)((lmax - 2 + 127) / 128), (unsigned)(jmax - 2)), 128, 0, st>>>(
            f->smooth_tmp, u, jmax, lmax, ic / 2.0, jc / 2.0, r2);
        M2D_HIP(hipGetLastError());
    }
    return fields_build_tables(f);
}

int m2d_fields2d_reset(m2d_fields2d *f)
{
    M2D_CHECK(f, M2D_ERR_BAD_ARG, "m2d_fields2d_reset: NULL argument");
    return fill_bytes(f->ctx, f->rho, 0, 8 * f->n);
}

int m2d_fields2d_solve(m2d_fields2d *f, int mode, double tol, m2d_solve_info *info)
{
    M2D_CHECK(f, M2D_ERR_BAD_ARG, "m2d_fields2d_solve: NULL argument");
    M2D_CHECK(mode == M2D_SOLVE_STATIC || mode == M2D_SOLVE_RF || mode == M2D_SOLVE_RAW, M2D_ERR_BAD_ARG,
              "m2d_fields2d_solve: unknown mode %d", mode);
    return solve(f, mode, tol, info);
}

}  // extern "C"
