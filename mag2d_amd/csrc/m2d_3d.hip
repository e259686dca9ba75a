// m2d_3d.hip -- the 3D path: Field3D gather/deposit (Field3D.hpp:40-163),
// the Species<CARTESIAN3D> step (species3d.cpp:3-94) as one pass, its cell
// sort, and the 7-point Poisson solve that replaces Solver (fields3d.cpp:35-101).
//
// Solver: FP64 iterative refinement around FP32 conjugate gradients on the
// Dirichlet-eliminated SPD system 6u - sum_free u_nb = ..., preconditioned by
// one symmetric geometric multigrid V-cycle: coarse node I sits on fine node
// 2I, every level carries the 7-point operator of its own mask, restriction
// is P^T / 2 (full weighting with the h^2 of the coarser spacing), red-black
// Gauss-Seidel smooths, and the coarsest level is solved by a dense inverse.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <rocprim/device/device_radix_sort.hpp>
#include "m2d_objects.hpp"
#include "m2d_scatter.hpp"

namespace m2d {

struct Mg3Level {
    int nx = 0, ny = 0, nz = 0;
    size_t n = 0;
    uint8_t *dir = nullptr;   // 1 on Dirichlet nodes (and the coarse faces)
    uint8_t *nbm = nullptr;   // per node: bits 0-5 set for free neighbours -x,-y,-z,+z,+y,+x; 0x80 if pinned
    float *x = nullptr, *b = nullptr, *r = nullptr;
};

}  // namespace m2d

struct m2d_fields3d {
    m2d_ctx *ctx = nullptr;
    m2d_fields3d_params p{};
    int nx = 0, ny = 0, nz = 0;
    size_t n = 0;
    double idx = 0, idy = 0, idz = 0;
    uint8_t *mask = nullptr;
    uint8_t *cellfree = nullptr;   // cell (i,j,k) at its lower corner: any of 8 corners FREE
    double *voltage = nullptr, *u = nullptr, *rho = nullptr;
    float *tx = nullptr, *ty = nullptr, *tz = nullptr;  // (u_m - u_{m-e}) * 1/h per axis, FP32
    std::vector<m2d::Mg3Level> lv;
    float *coarse_inv = nullptr;
    float *coarse_invT = nullptr;  // the same, transposed (read by k_dense_solve)
    unsigned int *ticket = nullptr;  // block counter of the fused reductions
    int coarse_n = 0;
    bool coarse_dense = true;
    int lc3 = 1;                   // first level of the one-workgroup coarse V-cycle
    float *pr = nullptr, *pz = nullptr, *pp = nullptr, *pp2 = nullptr, *pq = nullptr, *pe = nullptr;
    double *r64 = nullptr, *dots = nullptr, *h_dots = nullptr, *partial = nullptr;
    int npartial = 0;
    m2d::SolveEst est;  // the last solve's passes and iterations: what the next queues
};

namespace m2d {

__device__ __forceinline__ bool dir3(uint8_t m) { return m == M2D_FIXED || m == M2D_FIXED_RF; }

__device__ __forceinline__ void node3(size_t m, int ny, int nz, int &i, int &j, int &k)
{
    k = (int)(m % nz);
    size_t t = m / nz;
    j = (int)(t % ny);
    i = (int)(t / ny);
}

// ------------------------------------------------------------------ tables and solve kernels
__global__ void k3_tables(const double *u, int nx, int ny, int nz, double idx, double idy, double idz,
                          float *tx, float *ty, float *tz)
{
    const size_t n = (size_t)nx * ny * nz;
    size_t m = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    int i, j, k;
    node3(m, ny, nz, i, j, k);
    const size_t sx = (size_t)ny * nz;
    tx[m] = i >= 1 ? (float)((u[m] - u[m - sx]) * idx) : 0.f;
    ty[m] = j >= 1 ? (float)((u[m] - u[m - nz]) * idy) : 0.f;
    tz[m] = k >= 1 ? (float)((u[m] - u[m - 1]) * idz) : 0.f;
}

// Solver::solve rows (fields3d.cpp:86-93); Dirichlet values into u
__global__ void k3_rhs(const uint8_t *mask, const double *voltage, double *rho, double *u, size_t n,
                       double scale)
{
    size_t m = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    const bool d = dir3(mask[m]);
    const double b = d ? voltage[m] : rho[m] * -scale;
    rho[m] = b;
    if (d) u[m] = b;
}

// r = b - A u on free rows (matrix_init, fields3d.cpp:35-76); d = -r in FP32;
__global__ void k3_add_correction(const uint8_t *mask, double *u, const float *e, size_t n, const int *stop)
{
    if (stopped(stop)) return;
    size_t m = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m < n && !dir3(mask[m])) u[m] += (double)e[m];
}

// r = b - A u on free nodes (0 on Dirichlet nodes) with its norms finished in
// the same launch: out[0] = |r|^2, out[1] = |b_eff|^2, out[2] = |d|^2, d = -r in FP32
__global__ void __launch_bounds__(kRedBlock) k3_residual64_fin(const uint8_t *mask, const double *b, const double *u,
                                                              int nx, int ny, int nz, double *r, float *d,
                                                              double *partial, unsigned int *ticket, double *out,
                                                              int fin, const int *stop, CtlEval post)
{
    if (stopped(stop)) return;
    const size_t n = (size_t)nx * ny * nz, sx = (size_t)ny * nz;
    double v[3] = {0, 0, 0};
    for (size_t m = (size_t)blockIdx.x * kRedBlock + threadIdx.x; m < n; m += (size_t)gridDim.x * kRedBlock) {
        double rv = 0;
        if (!dir3(mask[m])) {
            // free rows never touch the grid faces (checked at creation)
            const size_t nb[6] = {m - sx, m - nz, m - 1, m + 1, m + nz, m + sx};
            double au = -6.0 * u[m], be = b[m];
#pragma unroll
            for (int q = 0; q < 6; q++) {
                au += u[nb[q]];
                if (dir3(mask[nb[q]])) be -= u[nb[q]];
            }
            rv = b[m] - au;
            v[1] += be * be;
        }
        r[m] = rv;
        const float dm = (float)(-rv);
        d[m] = dm;
        v[0] += rv * rv;
        v[2] += (double)dm * dm;
    }
    finish_sums<3>(v, partial, ticket, out, fin, post);
}

// Neighbour masks of the 7-point levels: one byte per node replaces the
// seven Dirichlet flags a stencil update reads.  Bits 0-5: neighbours
// -x, -y, -z, +z, +y, +x free; kNbSelf: the node itself is pinned.
constexpr uint8_t kNbSelf = 0x80;

// A' x = 6 x - sum over free neighbours (Dirichlet rows: 0)
__device__ __forceinline__ float apply7(uint8_t nb, const float *x, size_t m, size_t sx, int nz)
{
    float s = 6.f * x[m];
    if (nb & 1) s -= x[m - sx];
    if (nb & 2) s -= x[m - nz];
    if (nb & 4) s -= x[m - 1];
    if (nb & 8) s -= x[m + 1];
    if (nb & 16) s -= x[m + nz];
    if (nb & 32) s -= x[m + sx];
    return s;
}

// The CG direction and the operator in one launch: p' = z + beta p (first:
// p' = z), q = A' p' and out = p'.q.  Each node's p' is recomputed at its
// stencil neighbours from z and p (the same operations, so the same bits) and
// written to pn, not p, which other blocks are still reading.
__global__ void __launch_bounds__(kRedBlock) k3_dir_apply_dot(const uint8_t *nbm, const float *z, const float *p,
                                                             float *pn, float *q, const double *rz_new,
                                                             const double *rz_old, int first, int nx, int ny, int nz,
                                                             double *partial, unsigned int *ticket, double *out,
                                                             int fin, const int *stop)
{
    if (stopped(stop)) return;
    const size_t n = (size_t)nx * ny * nz, sx = (size_t)ny * nz;
    const float beta = first ? 0.f : cg_beta(rz_new, rz_old);
    double v[1] = {0};
    for (size_t m = (size_t)blockIdx.x * kRedBlock + threadIdx.x; m < n; m += (size_t)gridDim.x * kRedBlock) {
        const uint8_t nb = nbm[m];
        const float pk = cg_dir_at(z, p, m, beta, first);
        pn[m] = pk;
        float ym = 0.f;
        if (!(nb & kNbSelf)) {
            // apply7's order
            float t = 6.f * pk;
            if (nb & 1) t -= cg_dir_at(z, p, m - sx, beta, first);
            if (nb & 2) t -= cg_dir_at(z, p, m - nz, beta, first);
            if (nb & 4) t -= cg_dir_at(z, p, m - 1, beta, first);
            if (nb & 8) t -= cg_dir_at(z, p, m + 1, beta, first);
            if (nb & 16) t -= cg_dir_at(z, p, m + nz, beta, first);
            if (nb & 32) t -= cg_dir_at(z, p, m + sx, beta, first);
            ym = t;
        }
        q[m] = ym;
        v[0] += (double)pk * ym;
    }
    finish_sums<1>(v, partial, ticket, out, fin);
}

// Gauss-Seidel update of node m (the node bodies here are shared by the grid
// kernels and the one-workgroup coarse V-cycle)
__device__ __forceinline__ void node3_rbgs(const uint8_t *nbm, float *x, const float *b, size_t m, size_t sx, int nz)
{
    const uint8_t nb = nbm[m];
    if (nb & kNbSelf) return;
    float s = b[m];
    if (nb & 1) s += x[m - sx];
    if (nb & 2) s += x[m - nz];
    if (nb & 4) s += x[m - 1];
    if (nb & 8) s += x[m + 1];
    if (nb & 16) s += x[m + nz];
    if (nb & 32) s += x[m + sx];
    x[m] = s * (1.f / 6.f);
}

// The six neighbours of node (i, j, k) = m, read beside its mask byte: the
// indices come from the coordinates, not from the mask (a face's missing
// neighbour reads the node itself, which the mask never selects), so the
// loads do not wait for the mask.  Order: -x, -y, -z, +z, +y, +x (mask bits
// 1, 2, 4, 8, 16, 32).
struct Nb7 {
    float v[6];
};
__device__ __forceinline__ Nb7 neighbours7(const float *x, size_t m, int i, int j, int k, int nx, int ny, int nz)
{
    const size_t sx = (size_t)ny * nz;
    Nb7 q;
    q.v[0] = x[i > 0 ? m - sx : m];
    q.v[1] = x[j > 0 ? m - nz : m];
    q.v[2] = x[k > 0 ? m - 1 : m];
    q.v[3] = x[k < nz - 1 ? m + 1 : m];
    q.v[4] = x[j < ny - 1 ? m + nz : m];
    q.v[5] = x[i < nx - 1 ? m + sx : m];
    return q;
}

// red-black Gauss-Seidel half sweep, colour = (i+j+k) & 1.  Threads cover
// only the nodes of the colour: thread (kk, j, i) updates k = 2 kk + parity,
// blocks of 64 x 4 (kk, j) and a grid of (nz/2 / 64, ny / 4, nx).  The node
// body is node3_rbgs's, in the same order, on neighbours read early.
constexpr int kRbgsX = 64, kRbgsY = 4;
__global__ void __launch_bounds__(kRbgsX * kRbgsY) k3_rbgs(const uint8_t *nbm, float *x, const float *b, int nx,
                                                           int ny, int nz, int colour, const int *stop)
{
    if (stopped(stop)) return;
    const size_t sx = (size_t)ny * nz;
    const int i = blockIdx.z, j = blockIdx.y * kRbgsY + threadIdx.y;
    const int k = 2 * (blockIdx.x * kRbgsX + threadIdx.x) + ((i + j + colour) & 1);
    if (j >= ny || k >= nz) return;
    const size_t m = (size_t)i * sx + (size_t)j * nz + k;
    const uint8_t nb = nbm[m];
    const float bm = b[m];
    const Nb7 q = neighbours7(x, m, i, j, k, nx, ny, nz);
    if (nb & kNbSelf) return;
    float s = bm;
    if (nb & 1) s += q.v[0];
    if (nb & 2) s += q.v[1];
    if (nb & 4) s += q.v[2];
    if (nb & 8) s += q.v[3];
    if (nb & 16) s += q.v[4];
    if (nb & 32) s += q.v[5];
    x[m] = s * (1.f / 6.f);
}

// x = 0, then the colour-0 half sweep, in one pass over every node: from
// zero, every neighbour a colour-0 node reads is +0, so its update is
// (b + 0) / 6 -- the +0 once, which turns a -0 right-hand side into +0 as
// the sweep's sums do, and not at all for a node with no free neighbour.
// Blocks of 64 x 4 nodes (k, j) and a grid of (nz / 64, ny / 4, nx).
__global__ void __launch_bounds__(256) k3_rbgs_from_zero(const uint8_t *nbm, float *x, const float *b, int nx,
                                                         int ny, int nz, const int *stop)
{
    if (stopped(stop)) return;
    const int i = blockIdx.z, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.x * 64 + threadIdx.x;
    if (j >= ny || k >= nz) return;
    const size_t m = ((size_t)i * ny + j) * nz + k;
    const uint8_t nb = nbm[m];
    float v = 0.f;
    if (((i + j + k) & 1) == 0 && !(nb & kNbSelf)) {
        float t = b[m];
        if (nb & 63) t += 0.f;
        v = t * (1.f / 6.f);
    }
    x[m] = v;
}

// r = b - A' x; blocks of 64 x 4 nodes (k, j) and a grid of (nz / 64, ny / 4, nx)
__global__ void __launch_bounds__(256) k3_resid(const uint8_t *nbm, const float *x, const float *b, float *r, int nx,
                                                int ny, int nz, const int *stop)
{
    if (stopped(stop)) return;
    const int i = blockIdx.z, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.x * 64 + threadIdx.x;
    if (j >= ny || k >= nz) return;
    const size_t m = ((size_t)i * ny + j) * nz + k;
    const uint8_t nb = nbm[m];
    const float bm = b[m], xm = x[m];
    const Nb7 q = neighbours7(x, m, i, j, k, nx, ny, nz);
    // apply7's body, in its order, on the neighbours read early
    float s = 6.f * xm;
    if (nb & 1) s -= q.v[0];
    if (nb & 2) s -= q.v[1];
    if (nb & 4) s -= q.v[2];
    if (nb & 8) s -= q.v[3];
    if (nb & 16) s -= q.v[4];
    if (nb & 32) s -= q.v[5];
    r[m] = (nb & kNbSelf) ? 0.f : bm - s;
}


// 1D weights of the trilinear prolongation: fine i from coarse I
__device__ __forceinline__ float w1(int i, int I) { int d = i - 2 * I; return d == 0 ? 1.f : (d == 1 || d == -1 ? 0.5f : 0.f); }

// b_c = P^T r_f / 2 at free coarse node M = (I, J, K).  The 27 fine nodes
// are summed in (i, j, k) order; one outside the grid reads a node inside
// with weight 0, which adds +0 to a sum that never holds -0: the same bits
// as skipping it, with every load issued at once.
__device__ __forceinline__ float node3_restrict_at(const float *rf, int fx, int fy, int fz, int I, int J, int K)
{
    float w[3][3];
    int ix[3][3];
    const int c[3] = {I, J, K}, nf[3] = {fx, fy, fz};
#pragma unroll
    for (int d = 0; d < 3; d++) {
#pragma unroll
        for (int t = 0; t < 3; t++) {
            const int f = 2 * c[d] - 1 + t;
            const bool in = f >= 0 && f < nf[d];
            ix[d][t] = in ? f : 2 * c[d];
            w[d][t] = in ? w1(f, c[d]) : 0.f;
        }
    }
    float v[27];
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 3; b++)
#pragma unroll
            for (int e = 0; e < 3; e++)
                v[(a * 3 + b) * 3 + e] = rf[((size_t)ix[0][a] * fy + ix[1][b]) * fz + ix[2][e]];
    float s = 0.f;
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 3; b++) {
            const float wij = w[0][a] * w[1][b];
#pragma unroll
            for (int e = 0; e < 3; e++) s = fmaf(wij * w[2][e], v[(a * 3 + b) * 3 + e], s);  // r is 0 on Dirichlet nodes
        }
    return 0.5f * s;
}

__device__ __forceinline__ float node3_restrict(const float *rf, int fx, int fy, int fz, const uint8_t *dirc, int cy,
                                                int cz, size_t M)
{
    if (dirc[M]) return 0.f;
    int I, J, K;
    node3(M, cy, cz, I, J, K);
    return node3_restrict_at(rf, fx, fy, fz, I, J, K);
}

// blocks of 64 x 4 coarse nodes (K, J) and a grid of (cz / 64, cy / 4, cx):
// the coordinates come from the launch, not from divisions
__global__ void __launch_bounds__(256) k3_restrict(const float *rf, int fx, int fy, int fz, float *bc,
                                                   const uint8_t *dirc, int cx, int cy, int cz, const int *stop)
{
    if (stopped(stop)) return;
    const int I = blockIdx.z, J = blockIdx.y * 4 + threadIdx.y, K = blockIdx.x * 64 + threadIdx.x;
    if (J >= cy || K >= cz) return;
    const size_t M = ((size_t)I * cy + J) * cz + K;
    bc[M] = dirc[M] ? 0.f : node3_restrict_at(rf, fx, fy, fz, I, J, K);
}

// x_f += P x_c at free fine node m = (i, j, k)
__device__ __forceinline__ void node3_prolong_at(float *xf, int i, int j, int k, const float *xc, int cx, int cy,
                                                 int cz, size_t m)
{
    // per axis the coarse nodes of fine index f: f / 2 (weight 1) when f is
    // even, else (f - 1) / 2 and (f + 1) / 2 (0.5 each; the second can lie
    // past the coarse grid).  All 8 combinations are summed in (a, b, c)
    // order; an absent node is read at the first with weight 0, which adds
    // +0 to a sum that never holds -0 -- the same bits as skipping it, with
    // the loads issued together.
    int Ia[2], Ja[2], Ka[2];
    float wa[2], wb[2], wc[2];
    auto split = [](int f, int nc, int *I, float *w) {
        const bool odd = f & 1, second = odd && (f + 1) / 2 < nc;
        I[0] = f >> 1;
        w[0] = odd ? 0.5f : 1.f;
        I[1] = second ? (f + 1) / 2 : I[0];
        w[1] = second ? 0.5f : 0.f;
    };
    split(i, cx, Ia, wa);
    split(j, cy, Ja, wb);
    split(k, cz, Ka, wc);
    float v[8];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int c = 0; c < 2; c++) v[(a * 2 + b) * 2 + c] = xc[((size_t)Ia[a] * cy + Ja[b]) * cz + Ka[c]];
    float s = 0.f;
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int c = 0; c < 2; c++) s = fmaf(wa[a] * wb[b] * wc[c], v[(a * 2 + b) * 2 + c], s);
    xf[m] += s;
}

// x_f += P x_c at fine node m (a free node)
__device__ __forceinline__ void node3_prolong_free(float *xf, int fy, int fz, const float *xc, int cx, int cy, int cz,
                                                   size_t m)
{
    int i, j, k;
    node3(m, fy, fz, i, j, k);
    node3_prolong_at(xf, i, j, k, xc, cx, cy, cz, m);
}

// x_f += P x_c at fine node m when it is free
__device__ __forceinline__ void node3_prolong(float *xf, const uint8_t *dirf, int fy, int fz, const float *xc, int cx,
                                              int cy, int cz, size_t m)
{
    if (!dirf[m]) node3_prolong_free(xf, fy, fz, xc, cx, cy, cz, m);
}

// x_f += P x_c on free fine nodes; blocks of 64 x 4 nodes (k, j) and a grid
// of (fz / 64, fy / 4, fx)
__global__ void __launch_bounds__(256) k3_prolong(float *xf, const uint8_t *dirf, int fx, int fy, int fz,
                                                  const float *xc, int cx, int cy, int cz, const int *stop)
{
    if (stopped(stop)) return;
    const int i = blockIdx.z, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.x * 64 + threadIdx.x;
    if (j >= fy || k >= fz) return;
    const size_t m = ((size_t)i * fy + j) * fz + k;
    if (!dirf[m]) node3_prolong_at(xf, i, j, k, xc, cx, cy, cz, m);
}

// ---- the coarse levels [lc, L) of the V-cycle in one workgroup: the same
// node bodies and sweep order as the grid kernels, a workgroup barrier
// between sweeps instead of a launch
constexpr int kMaxLevels3 = 16;
constexpr size_t kBlockLevelNodes3 = 4096;  // levels this small run inside the workgroup

struct BLevel3 {
    int nx, ny, nz, n;
    const uint8_t *dir, *nbm;
    float *x, *b, *r;
};

struct Coarse3Args {
    int lc, L;
    BLevel3 lv[kMaxLevels3];
    int dense;          // dense coarsest solve (else sweeps symmetric RB-GS sweeps)
    const float *invT;  // the dense inverse, transposed
    int coarse_n, sweeps, pre, post;
    const int *stop;    // the solve's stop flag (see stopped())
};

__device__ void c3_smooth(const BLevel3 &V, float *x, const float *b, int sweeps, bool forward)
{
    const size_t sx = (size_t)V.ny * V.nz;
    for (int it = 0; it < sweeps; it++) {
        for (int h = 0; h < 2; h++) {
            const int colour = forward ? h : 1 - h;
            for (int m = threadIdx.x; m < V.n; m += kCoarseThreads) {
                int i, j, k;
                node3((size_t)m, V.ny, V.nz, i, j, k);
                if (((i + j + k) & 1) == colour) node3_rbgs(V.nbm, x, b, (size_t)m, sx, V.nz);
            }
            __syncthreads();
        }
    }
}

__global__ void __launch_bounds__(kCoarseThreads) k3_vcycle_coarse(Coarse3Args C)
{
    if (stopped(C.stop)) return;
    extern __shared__ float sb[];  // dense solve: coarse_n words, then kDenseParts partial rows
    for (int l = C.lc; l < C.L - 1; l++) {
        const BLevel3 &V = C.lv[l];
        for (int m = threadIdx.x; m < V.n; m += kCoarseThreads) V.x[m] = 0.f;
        __syncthreads();
        c3_smooth(V, V.x, V.b, C.pre, true);
        const size_t sx = (size_t)V.ny * V.nz;
        for (int m = threadIdx.x; m < V.n; m += kCoarseThreads) {
            const uint8_t nb = V.nbm[m];
            V.r[m] = (nb & kNbSelf) ? 0.f : V.b[m] - apply7(nb, V.x, (size_t)m, sx, V.nz);
        }
        __syncthreads();
        const BLevel3 &W = C.lv[l + 1];
        for (int M = threadIdx.x; M < W.n; M += kCoarseThreads)
            W.b[M] = node3_restrict(V.r, V.nx, V.ny, V.nz, W.dir, W.ny, W.nz, (size_t)M);
        __syncthreads();
    }
    {
        const BLevel3 &Z = C.lv[C.L - 1];
        if (C.dense) {
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
        } else {
            for (int m = threadIdx.x; m < Z.n; m += kCoarseThreads) Z.x[m] = 0.f;
            __syncthreads();
            c3_smooth(Z, Z.x, Z.b, C.sweeps / 2, true);
            c3_smooth(Z, Z.x, Z.b, C.sweeps / 2, false);
        }
    }
    for (int l = C.L - 2; l >= C.lc; l--) {
        const BLevel3 &V = C.lv[l];
        const BLevel3 &W = C.lv[l + 1];
        for (int m = threadIdx.x; m < V.n; m += kCoarseThreads)
            node3_prolong(V.x, V.dir, V.ny, V.nz, W.x, W.nx, W.ny, W.nz, (size_t)m);
        __syncthreads();
        c3_smooth(V, V.x, V.b, C.post, false);
    }
}

// The common case -- one red-black level of at most 4096 nodes above a dense
// level of at most 512 -- with both levels in LDS and each thread's nodes and
// their colours computed once.  Same node updates and order, same bits.
constexpr int kLds3Nodes = 4096, kLds3Dense = 512, kLds3PerThread = kLds3Nodes / kCoarseThreads;

__global__ void __launch_bounds__(kCoarseThreads) k3_vcycle_coarse_lds(Coarse3Args C)
{
    if (stopped(C.stop)) return;
    __shared__ float s_x[kLds3Nodes], s_b[kLds3Nodes], s_r[kLds3Nodes];
    __shared__ uint8_t s_nb[kLds3Nodes];
    __shared__ float s_zb[kLds3Dense], s_zx[kLds3Dense];
    const BLevel3 &V = C.lv[C.lc];
    const BLevel3 &Z = C.lv[C.lc + 1];
    const int n = V.n, ny = V.ny, nz = V.nz, cn = C.coarse_n;
    const size_t sx = (size_t)ny * nz;
    int par[kLds3PerThread];
#pragma unroll
    for (int q = 0; q < kLds3PerThread; q++) {
        const int m = threadIdx.x + q * kCoarseThreads;
        par[q] = 2;
        if (m < n) {
            int i, j, k;
            node3((size_t)m, ny, nz, i, j, k);
            par[q] = (i + j + k) & 1;
            s_nb[m] = V.nbm[m];
            s_b[m] = V.b[m];
            s_x[m] = 0.f;
        }
    }
    __syncthreads();
    auto sweep = [&](int colour) {
#pragma unroll
        for (int q = 0; q < kLds3PerThread; q++)
            if (par[q] == colour) node3_rbgs(s_nb, s_x, s_b, (size_t)(threadIdx.x + q * kCoarseThreads), sx, nz);
        __syncthreads();
    };
    for (int it = 0; it < C.pre; it++) { sweep(0); sweep(1); }
#pragma unroll
    for (int q = 0; q < kLds3PerThread; q++) {
        const int m = threadIdx.x + q * kCoarseThreads;
        if (m < n) {
            const uint8_t nb = s_nb[m];
            s_r[m] = (nb & kNbSelf) ? 0.f : s_b[m] - apply7(nb, s_x, (size_t)m, sx, nz);
        }
    }
    __syncthreads();
    for (int M = threadIdx.x; M < cn; M += kCoarseThreads)
        s_zb[M] = node3_restrict(s_r, V.nx, ny, nz, Z.dir, Z.ny, Z.nz, (size_t)M);
    __syncthreads();
    {
        // the dense solve of k3_vcycle_coarse; its partial rows reuse s_r
        float *part = s_r;
        const int len = (cn + kDenseParts - 1) / kDenseParts;
        for (int t = threadIdx.x; t < cn * kDenseParts; t += kCoarseThreads) {
            const int row = t % cn, p = t / cn;
            const int c0 = p * len, c1 = min(cn, c0 + len);
            float s = 0.f;
            int c = c0;
            for (; c + 32 <= c1; c += 32) {  // 32 loads in flight: the loop waits on L2
                float a[32];
#pragma unroll
                for (int q = 0; q < 32; q++) a[q] = C.invT[(size_t)(c + q) * cn + row];
#pragma unroll
                for (int q = 0; q < 32; q++) s = fmaf(a[q], s_zb[c + q], s);
            }
            for (; c < c1; c++) s = fmaf(C.invT[(size_t)c * cn + row], s_zb[c], s);
            part[p * cn + row] = s;
        }
        __syncthreads();
        for (int row = threadIdx.x; row < cn; row += kCoarseThreads) {
            float s = part[row];
            for (int p = 1; p < kDenseParts; p++) s += part[p * cn + row];
            s_zx[row] = s;
        }
        __syncthreads();
    }
    // node3_prolong with the pinned flag from the neighbour mask
#pragma unroll
    for (int q = 0; q < kLds3PerThread; q++) {
        const int m = threadIdx.x + q * kCoarseThreads;
        if (m < n && !(s_nb[m] & kNbSelf)) node3_prolong_free(s_x, ny, nz, s_zx, Z.nx, Z.ny, Z.nz, (size_t)m);
    }
    __syncthreads();
    for (int it = 0; it < C.post; it++) { sweep(1); sweep(0); }
#pragma unroll
    for (int q = 0; q < kLds3PerThread; q++) {
        const int m = threadIdx.x + q * kCoarseThreads;
        if (m < n) V.x[m] = s_x[m];
    }
}

// ------------------------------------------------------------------ point kernels (FP64)
struct Grid3 {
    int nx, ny, nz;
    double idx, idy, idz, hx, hy, hz;  // h = (n-1) d idx: the clamp limits of grad
};

__device__ __forceinline__ double d3at(const double *d, const Grid3 &g, int i, int j, int k)
{
    return d[((size_t)i * g.ny + j) * g.nz + k];
}

// Field3D::interpolate (Field3D.hpp:68-96); out of range gives NaN
__global__ void k_interp3d(const double *d, Grid3 g, const double *xs, const double *ys, const double *zs,
                           int64_t n, double *out)
{
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    double x = xs[p], y = ys[p], z = zs[p];
    const double X = x * g.idx, Y = y * g.idy, Z = z * g.idz;
    if (!(X >= 0 && X < g.nx - 1.0 && Y >= 0 && Y < g.ny - 1.0 && Z >= 0 && Z < g.nz - 1.0)) {
        out[p] = NAN;
        return;
    }
    const int i = (int)X, j = (int)Y, k = (int)Z;
    double u = x * g.idx - i, v = y * g.idy - j, w = z * g.idz - k;
    double r = 0;
    r += (1 - u) * (1 - v) * (1 - w) * d3at(d, g, i, j, k);
    r += u * (1 - v) * (1 - w) * d3at(d, g, i + 1, j, k);
    r += (1 - u) * v * (1 - w) * d3at(d, g, i, j + 1, k);
    r += u * v * (1 - w) * d3at(d, g, i + 1, j + 1, k);
    r += (1 - u) * (1 - v) * w * d3at(d, g, i, j, k + 1);
    r += u * (1 - v) * w * d3at(d, g, i + 1, j, k + 1);
    r += (1 - u) * v * w * d3at(d, g, i, j + 1, k + 1);
    r += u * v * w * d3at(d, g, i + 1, j + 1, k + 1);
    out[p] = r;
}

// Field3D::grad_component (Field3D.hpp:112-141), FP64: clamped half a cell
// inside along the differenced axis, staggered differences, trilinear weights
__device__ double grad3_component(const double *d, const Grid3 &g, double x, double y, double z, int ex, int ey,
                                  int ez)
{
    double X = fmin(fmax(x * g.idx, 0.5 * ex), g.hx - 0.5 * ex) + 0.5 * ex;
    double Y = fmin(fmax(y * g.idy, 0.5 * ey), g.hy - 0.5 * ey) + 0.5 * ey;
    double Z = fmin(fmax(z * g.idz, 0.5 * ez), g.hz - 0.5 * ez) + 0.5 * ez;
    int i = min((int)X, g.nx - 2), j = min((int)Y, g.ny - 2), k = min((int)Z, g.nz - 2);
    double u = X - i, v = Y - j, w = Z - k;
    double gg[8];
    for (int c = 0; c < 8; c++) {
        int a = i + (c & 1), b = j + ((c >> 1) & 1), e = k + ((c >> 2) & 1);
        gg[c] = d3at(d, g, a, b, e) - d3at(d, g, a - ex, b - ey, e - ez);
    }
    double r = (1 - u) * (1 - v) * (1 - w) * gg[0] + u * (1 - v) * (1 - w) * gg[1] +
               (1 - u) * v * (1 - w) * gg[2] + u * v * (1 - w) * gg[3] + (1 - u) * (1 - v) * w * gg[4] +
               u * (1 - v) * w * gg[5] + (1 - u) * v * w * gg[6] + u * v * w * gg[7];
    return r * (g.idx * ex + g.idy * ey + g.idz * ez);
}

__global__ void k_grad3d(const double *d, Grid3 g, const double *xs, const double *ys, const double *zs,
                         int64_t n, double *gx, double *gy, double *gz)
{
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    gx[p] = grad3_component(d, g, xs[p], ys[p], zs[p], 1, 0, 0);
    gy[p] = grad3_component(d, g, xs[p], ys[p], zs[p], 0, 1, 0);
    gz[p] = grad3_component(d, g, xs[p], ys[p], zs[p], 0, 0, 1);
}

// ------------------------------------------------------------------ the 3D pass
struct PushArgs3 {
    float *x, *y, *z, *vx, *vy, *vz;
    int64_t n;
    const float *tx, *ty, *tz;
    const uint8_t *cellfree;
    Grid3 g;
    float bx1, by1, bz1;           // largest floats <= x_max, y_max, z_max
    double xper, yper, zper;
    float qmdt2, tx_, ty_, tz_, sx, sy, sz, dt;
    int rotate, push, boundary, periodic;
    uint64_t trial_thr;
    uint64_t seed;
    uint32_t stream;
    uint64_t step;
    const McDev *mcc;
    uint32_t *cq;
    unsigned long long *cq_len;
    unsigned long long *acc;
    unsigned long long *counters;
};

// one E component through the FP32 difference table: the clamp and the
// cell index in FP64 on the FP32 position, weights FP32
__device__ __forceinline__ float gather3_component(const float *t, const Grid3 &g, double x, double y, double z,
                                                   int ex, int ey, int ez)
{
    double X = fmin(fmax(x * g.idx, 0.5 * ex), g.hx - 0.5 * ex) + 0.5 * ex;
    double Y = fmin(fmax(y * g.idy, 0.5 * ey), g.hy - 0.5 * ey) + 0.5 * ey;
    double Z = fmin(fmax(z * g.idz, 0.5 * ez), g.hz - 0.5 * ez) + 0.5 * ez;
    int i = min(max((int)X, 0), g.nx - 2), j = min(max((int)Y, 0), g.ny - 2), k = min(max((int)Z, 0), g.nz - 2);
    const float u = (float)(X - i), v = (float)(Y - j), w = (float)(Z - k);
    const size_t sx = (size_t)g.ny * g.nz;
    const float *c = t + ((size_t)i * g.ny + j) * g.nz + k;
    float a0 = c[0], a1 = c[sx], a2 = c[g.nz], a3 = c[sx + g.nz];
    float a4 = c[1], a5 = c[sx + 1], a6 = c[g.nz + 1], a7 = c[sx + g.nz + 1];
    float lo = (1.f - v) * ((1.f - u) * a0 + u * a1) + v * ((1.f - u) * a2 + u * a3);
    float hi = (1.f - v) * ((1.f - u) * a4 + u * a5) + v * ((1.f - u) * a6 + u * a7);
    return (1.f - w) * lo + w * hi;
}

// eight CIC weights in 2^-31 fixed point, corrected on the largest so a
// particle deposits exactly 2^31 (order-independent integer sums)
__device__ __forceinline__ bool cic3(double X, double Y, double Z, const Grid3 &g, int &cell, uint32_t w[8])
{
    // range in FP64 before any conversion: cells [0, n-2] on every axis
    if (!(X >= 0.0 && X < g.nx - 1.0 && Y >= 0.0 && Y < g.ny - 1.0 && Z >= 0.0 && Z < g.nz - 1.0)) return false;
    const int i = (int)X, j = (int)Y, k = (int)Z;
    // offsets and their complements from FP64, each rounded once (see the
    // 2D cic_weights)
    const double du = X - i, dv = Y - j, dw = Z - k;
    const float S = 2147483648.0f;
    const float a[2] = {(float)(1.0 - du), (float)du}, b[2] = {(float)(1.0 - dv), (float)dv},
                c[2] = {(float)(1.0 - dw), (float)dw};
    uint32_t s = 0;
#pragma unroll
    for (int q = 0; q < 8; q++) {
        w[q] = __float2uint_rn(a[q & 1] * b[(q >> 1) & 1] * c[q >> 2] * S);
        s += w[q];
    }
    // the first largest weight takes the rounding's remainder (selects, not
    // an index: a computed index would put w in memory)
    int km = 0;
    uint32_t wm = w[0];
#pragma unroll
    for (int q = 1; q < 8; q++)
        if (w[q] > wm) { km = q; wm = w[q]; }
#pragma unroll
    for (int q = 0; q < 8; q++)
        if (q == km) w[q] += 0x80000000u - s;
    cell = (int)(((size_t)i * g.ny + j) * g.nz + k);
    return true;
}

// Segmented sum across the wavefront of the lanes' weights keyed by their
// cell (a sorted store puts a cell's particles on neighbouring lanes); the
// last lane of each segment adds its sums to the 64-bit grid.  All 64
// lanes must call it together.
__device__ __forceinline__ void wave_deposit3(int cell, const uint32_t w8[8], const Grid3 &g,
                                              unsigned long long *acc)
{
    const int lane = __lane_id();
    const bool valid = cell >= 0;
    const int key = valid ? cell : -1 - lane;
    const int prev = __shfl_up(key, 1);
    const unsigned long long heads = __ballot(lane == 0 || prev != key);
    const int seg = 63 - __clzll(heads & ((2ull << lane) - 1ull));
    unsigned long long s[8];
#pragma unroll
    for (int q = 0; q < 8; q++) s[q] = valid ? w8[q] : 0ull;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
#pragma unroll
        for (int q = 0; q < 8; q++) {
            unsigned long long t = __shfl_up(s[q], off);
            if (lane - off >= seg) s[q] += t;
        }
    }
    const bool tail = lane == 63 || ((heads >> (lane + 1)) & 1ull);
    if (valid && tail) {
        const size_t sx = (size_t)g.ny * g.nz;
#pragma unroll
        for (int q = 0; q < 8; q++) {
            const size_t m = (size_t)cell + ((q & 1) ? sx : 0) + (((q >> 1) & 1) ? (size_t)g.nz : 0) + (q >> 2);
            if (s[q]) atomicAdd(acc + m, s[q]);
        }
    }
}

// mod (mymath.cpp:62-67)
__device__ __forceinline__ double mod3(double x, double y)
{
    if (x >= 0.0 && x <= y) return x;
    return x - y * (int)(x / y) + (x < 0 ? y : 0);
}

// 3D sort tiles: 4 x 4 x 4 cells, key = tile << 6 | cell in tile; the deposit
// window of a tile spans its cells plus a halo of 2 cells: 9 nodes per axis
constexpr int kTile3 = 4, kHalo3 = 2, kWin3 = kTile3 + 1 + 2 * kHalo3, kWinN3 = kWin3 * kWin3 * kWin3;

struct Step3 {
    bool removed = false, coll = false, nan_pos = false, bad = false;
    int cell = -1;        // deposit cell (lower corner node), -1: none
    int ci = 0, cj = 0, ck = 0;
    uint32_t w8[8];
};

// Species<CARTESIAN3D>::advance for the particle in slot p (species3d.cpp:10-92):
// gather, Boris, drift, MCC trial, box / mod, is_free, deposit weights
template <bool DEPOSIT, bool MCC>
__device__ __forceinline__ void step3(const PushArgs3 &a, int64_t p, Step3 &o)
{
    float x = a.x[p];
    if (is_dead(x)) return;
    float y = a.y[p], z = a.z[p];
    float vx = a.vx[p], vy = a.vy[p], vz = a.vz[p];
    if (a.push) {
        // E = -grad u (ElMag3D::E, fields3d.hpp:91-97)
        const double xd = x, yd = y, zd = z;
        const float ex = -gather3_component(a.tx, a.g, xd, yd, zd, 1, 0, 0);
        const float ey = -gather3_component(a.ty, a.g, xd, yd, zd, 0, 1, 0);
        const float ez = -gather3_component(a.tz, a.g, xd, yd, zd, 0, 0, 1);
        // species3d.cpp:24-61
        vx = fmaf(ex, a.qmdt2, vx);
        vy = fmaf(ey, a.qmdt2, vy);
        vz = fmaf(ez, a.qmdt2, vz);
        if (a.rotate) {
            float vpx = vx - vy * a.tz_ + vz * a.ty_;
            float vpy = vy - vz * a.tx_ + vx * a.tz_;
            float vpz = vz - vx * a.ty_ + vy * a.tx_;
            vx = vx - vpy * a.sz + vpz * a.sy;
            vy = vy - vpz * a.sx + vpx * a.sz;
            vz = vz - vpx * a.sy + vpy * a.sx;
        }
        vx = fmaf(ex, a.qmdt2, vx);
        vy = fmaf(ey, a.qmdt2, vy);
        vz = fmaf(ez, a.qmdt2, vz);
        x = fmaf(vx, a.dt, x);
        y = fmaf(vy, a.dt, y);
        z = fmaf(vz, a.dt, z);
        // MCC trial (species3d.cpp:64-67); the velocity change is applied after the pass
        if (MCC) {
            uint32_t c0 = (uint32_t)(p >> 2), c1 = (uint32_t)a.step, c2 = 0u;
            uint32_t c3 = (kPurposeTrial << 24) | ((uint32_t)(a.step >> 32) & 0xFFFFFFu);
            philox4x32_10(c0, c1, c2, c3, (uint32_t)a.seed, (uint32_t)(a.seed >> 32) ^ a.stream);
            const int q = (int)(p & 3);
            const uint32_t tw = q == 0 ? c0 : q == 1 ? c1 : q == 2 ? c2 : c3;
            if ((uint64_t)tw < a.trial_thr) {
                unsigned long long at = atomicAdd(a.cq_len, 1ull);
                a.cq[at] = (uint32_t)p;
                o.coll = true;
            }
        }
    }
    if (a.boundary) {
        // species3d.cpp:69-85
        if (x > a.bx1 || x < 0.f || y > a.by1 || y < 0.f || z > a.bz1 || z < 0.f || x != x || y != y || z != z) {
            if (x != x || y != y || z != z) { o.removed = true; o.nan_pos = true; }
            else if (!a.periodic) o.removed = true;
            else {
                x = (float)mod3(x, a.xper);
                y = (float)mod3(y, a.yper);
                z = (float)mod3(z, a.zper);
            }
        }
        if (!o.removed) {
            // Geometry::is_free (fields3d.hpp:48-55)
            const double X = (double)x * a.g.idx, Y = (double)y * a.g.idy, Z = (double)z * a.g.idz;
            bool fr = false;
            if (X >= 0.0 && X < a.g.nx - 1.0 && Y >= 0.0 && Y < a.g.ny - 1.0 && Z >= 0.0 && Z < a.g.nz - 1.0)
                fr = a.cellfree[((size_t)(int)X * a.g.ny + (int)Y) * a.g.nz + (int)Z] != 0;
            o.removed = !fr;
        }
    }
    if (DEPOSIT && !o.removed) {
        const double X = (double)x * a.g.idx, Y = (double)y * a.g.idy, Z = (double)z * a.g.idz;
        if (cic3(X, Y, Z, a.g, o.cell, o.w8)) {
            o.ci = (int)X; o.cj = (int)Y; o.ck = (int)Z;
        } else {
            o.cell = -1;
            o.bad = true;
        }
    }
    if (o.removed) {
        a.x[p] = dead_value();
    } else {
        a.x[p] = x; a.y[p] = y; a.z[p] = z;
        a.vx[p] = vx; a.vy[p] = vy; a.vz[p] = vz;
    }
}

__device__ __forceinline__ void deposit8_global(unsigned long long *acc, const Grid3 &g, int cell,
                                                const uint32_t w8[8])
{
    const size_t sx = (size_t)g.ny * g.nz;
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const size_t m = (size_t)cell + ((q & 1) ? sx : 0) + (((q >> 1) & 1) ? (size_t)g.nz : 0) + (q >> 2);
        if (w8[q]) atomicAdd(acc + m, (unsigned long long)w8[q]);
    }
}

// untiled: one particle per thread, deposit summed across the wavefront
template <bool DEPOSIT, bool MCC>
__global__ void __launch_bounds__(256) k_advance3d(PushArgs3 a)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    Step3 o;
    if (p < a.n) step3<DEPOSIT, MCC>(a, p, o);
    if (DEPOSIT) wave_deposit3(o.cell, o.w8, a.g, a.acc);
    wave_count(&a.counters[CNT_REMOVED], o.removed);
    if (MCC) wave_count(&a.counters[CNT_COLL], o.coll);
    if (__ballot(o.nan_pos) && __lane_id() == 0) raise_flag(a.counters, M2D_FLAG_NAN);
    if (__ballot(o.bad) && __lane_id() == 0) raise_flag(a.counters, M2D_FLAG_OUT_OF_RANGE);
}

// tiled (store sorted by m2d_sort_3d): block b runs chunk sched[b] of one
// sort tile and sums its deposits in an LDS window of the tile + halo, then
// adds the window's non-zero nodes to the 64-bit grid with atomics (integer
// sums: the grid does not depend on their order); deposits that leave the
// window go to the grid directly.  Six waves per SIMD (at most 80 VGPRs):
// the pass waits on its 24 gathers, and 4 waves (111 VGPRs unconstrained)
// ran it 12% slower.
template <bool MCC>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(6))) k_advance3d_tiled(PushArgs3 a, const int4 *sched,
                                                         const unsigned int *n_chunks, int nty3, int ntz3)
{
    __shared__ unsigned long long win[kWinN3];
    // plain block order: grouping chunks per XCD (contiguous ranges or runs
    // of 8 to 128 chunks) made this pass 3-27% slower
    const unsigned ch = blockIdx.x;
    if (ch >= *n_chunks) return;  // uniform, before any barrier
    const int4 s = sched[ch];
    const int t = s.x;
    const int wi0 = (t / (nty3 * ntz3)) * kTile3 - kHalo3, wj0 = ((t / ntz3) % nty3) * kTile3 - kHalo3,
              wk0 = (t % ntz3) * kTile3 - kHalo3;
    for (int k = threadIdx.x; k < kWinN3; k += 256) win[k] = 0ull;
    __syncthreads();
    bool removed_any = false, coll_any = false, nan_any = false, bad_any = false;
    int n_removed = 0, n_coll = 0, n_spill = 0;
    for (int64_t p = s.y + threadIdx.x; p < s.z; p += 256) {
        Step3 o;
        step3<true, MCC>(a, p, o);
        n_removed += o.removed;
        n_coll += o.coll;
        nan_any |= o.nan_pos;
        bad_any |= o.bad;
        if (o.cell >= 0) {
            const int da = o.ci - wi0, db = o.cj - wj0, dc = o.ck - wk0;
            if ((unsigned)da < (unsigned)(kWin3 - 1) && (unsigned)db < (unsigned)(kWin3 - 1) &&
                (unsigned)dc < (unsigned)(kWin3 - 1)) {
                unsigned long long *w = win + (da * kWin3 + db) * kWin3 + dc;
#pragma unroll
                for (int q = 0; q < 8; q++) {
                    const int off = ((q & 1) ? kWin3 * kWin3 : 0) + (((q >> 1) & 1) ? kWin3 : 0) + (q >> 2);
                    atomicAdd(w + off, (unsigned long long)o.w8[q]);  // a zero weight adds 0: no branch
                }
            } else {
                deposit8_global(a.acc, a.g, o.cell, o.w8);
                n_spill++;
            }
        }
    }
    (void)removed_any; (void)coll_any;
    __syncthreads();
    for (int k = threadIdx.x; k < kWinN3; k += 256) {
        const unsigned long long v = win[k];
        if (!v) continue;
        // window node (da, db, dc) of the tile's window: only nodes of the
        // grid receive deposits, so a non-zero entry lies on the grid
        const int i = wi0 + k / (kWin3 * kWin3), j = wj0 + (k / kWin3) % kWin3, l = wk0 + k % kWin3;
        if ((unsigned)i < (unsigned)a.g.nx && (unsigned)j < (unsigned)a.g.ny && (unsigned)l < (unsigned)a.g.nz)
            atomicAdd(a.acc + ((size_t)i * a.g.ny + j) * a.g.nz + l, v);
    }
    int sum = n_removed;
    for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off);
    if (__lane_id() == 0 && sum) atomicAdd(&a.counters[CNT_REMOVED], (unsigned long long)sum);
    int sp = n_spill;
    for (int off = 32; off > 0; off >>= 1) sp += __shfl_xor(sp, off);
    if (__lane_id() == 0 && sp) atomicAdd(&a.counters[CNT_SPILL], (unsigned long long)sp);
    if (MCC) {
        int c = n_coll;
        for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off);
        if (__lane_id() == 0 && c) atomicAdd(&a.counters[CNT_COLL], (unsigned long long)c);
    }
    if (__ballot(nan_any) && __lane_id() == 0) raise_flag(a.counters, M2D_FLAG_NAN);
    if (__ballot(bad_any) && __lane_id() == 0) raise_flag(a.counters, M2D_FLAG_OUT_OF_RANGE);
}

// BaseSpecies::scatter for the queued slots (velocity only, after the pass)
__global__ void __launch_bounds__(256) k_scatter3d(PushArgs3 a, unsigned long long *next_len)
{
    const unsigned long long len = *a.cq_len;
    const unsigned long long start = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x;
    const unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;
    if (start == 0) *next_len = 0ull;
    for (unsigned long long k = start; k < len; k += stride) {
        const uint32_t p = a.cq[k];
        if (is_dead(a.x[p])) continue;
        float vx = a.vx[p], vy = a.vy[p], vz = a.vz[p];
        Rng rg;
        rg.begin(a.seed, a.stream, p, a.step, kPurposeMcc);
        scatter(*a.mcc, rg, vx, vy, vz, a.counters);
        a.vx[p] = vx; a.vy[p] = vy; a.vz[p] = vz;
    }
}

// leapfrog start: rotation by -dt/2, then a backward half kick
__global__ void __launch_bounds__(256) k_advance_init3d(PushArgs3 a)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= a.n) return;
    const float x = a.x[p];
    if (is_dead(x)) return;
    const float y = a.y[p], z = a.z[p];
    float vx = a.vx[p], vy = a.vy[p], vz = a.vz[p];
    const double xd = x, yd = y, zd = z;
    const float ex = -gather3_component(a.tx, a.g, xd, yd, zd, 1, 0, 0);
    const float ey = -gather3_component(a.ty, a.g, xd, yd, zd, 0, 1, 0);
    const float ez = -gather3_component(a.tz, a.g, xd, yd, zd, 0, 0, 1);
    if (a.rotate) {
        float vpx = vx - vy * a.tz_ + vz * a.ty_;
        float vpy = vy - vz * a.tx_ + vx * a.tz_;
        float vpz = vz - vx * a.ty_ + vy * a.tx_;
        vx = vx - vpy * a.sz + vpz * a.sy;
        vy = vy - vpz * a.sx + vpx * a.sz;
        vz = vz - vpx * a.sy + vpy * a.sx;
    }
    a.vx[p] = fmaf(ex, a.qmdt2, vx);
    a.vy[p] = fmaf(ey, a.qmdt2, vy);
    a.vz[p] = fmaf(ez, a.qmdt2, vz);
}

// Field3D::accumulate for every live particle
__global__ void __launch_bounds__(256) k_accumulate3d(const float *xs, const float *ys, const float *zs,
                                                      int64_t n, Grid3 g, unsigned long long *acc,
                                                      unsigned long long *counters)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int cell = -1;
    uint32_t w8[8];
    bool bad = false;
    if (p < n) {
        const float x = xs[p];
        if (!is_dead(x) &&
            !cic3((double)x * g.idx, (double)ys[p] * g.idy, (double)zs[p] * g.idz, g, cell, w8)) {
            cell = -1;
            bad = true;
        }
    }
    wave_deposit3(cell, w8, g, acc);
    if (__ballot(bad) && __lane_id() == 0) raise_flag(counters, M2D_FLAG_OUT_OF_RANGE);
}

// ------------------------------------------------------------------ sort
// tile-major cell key: (tile << 6) | cell in its 4 x 4 x 4 tile
__global__ void k_sort_keys3d(const float *xs, const float *ys, const float *zs, int64_t n, Grid3 g, int nty3,
                              int ntz3, uint32_t dead_key, uint32_t *keys, uint32_t *vals)
{
    int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const float x = xs[p];
    uint32_t key = dead_key;
    if (!is_dead(x)) {
        auto idx = [](double X, int nmax) { return X >= 0.0 ? (int)fmin(X, (double)nmax) : 0; };
        const int i = idx((double)x * g.idx, g.nx - 2), j = idx((double)ys[p] * g.idy, g.ny - 2),
                  k = idx((double)zs[p] * g.idz, g.nz - 2);
        const uint32_t tile = (uint32_t)(((i / kTile3) * nty3 + j / kTile3) * ntz3 + k / kTile3);
        key = (tile << 6) | (uint32_t)((i % kTile3) * 16 + (j % kTile3) * 4 + k % kTile3);
    }
    keys[p] = key;
    vals[p] = (uint32_t)p;
}

__global__ void k_sort_gather3d(m2d_particles3d src, m2d_particles3d dst, const uint32_t *keys,
                                const uint32_t *vals, uint32_t dead_key, unsigned long long *n_live)
{
    // destination ranges in XCD-contiguous order (as k_sort_gather)
    const unsigned nb = (unsigned)((src.n + blockDim.x - 1) / blockDim.x);
    const unsigned b = xcd_block(blockIdx.x, nb);
    if (b >= nb) return;
    int64_t i = (int64_t)b * blockDim.x + threadIdx.x;
    if (i >= src.n) return;
    const uint32_t k = keys[i];
    if (k == dead_key) {
        dst.x[i] = dead_value();
        return;
    }
    if (i == src.n - 1 || keys[i + 1] == dead_key) *n_live = (unsigned long long)(i + 1);
    const int64_t s = vals[i];
    dst.x[i] = src.x[s];
    dst.y[i] = src.y[s];
    dst.z[i] = src.z[s];
    dst.vx[i] = src.vx[s];
    dst.vy[i] = src.vy[s];
    dst.vz[i] = src.vz[s];
}

// ------------------------------------------------------------------ host helpers
static int dmalloc3(void **p, size_t bytes)
{
    hipError_t e = hipMalloc(p, bytes < 16 ? 16 : bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        return set_error(M2D_ERR_NO_MEMORY, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    }
    return M2D_OK;
}

static Grid3 grid3_of(const m2d_grid3d &g)
{
    Grid3 r;
    r.nx = g.nx; r.ny = g.ny; r.nz = g.nz;
    r.idx = 1.0 / g.dx; r.idy = 1.0 / g.dy; r.idz = 1.0 / g.dz;
    r.hx = (g.nx - 1) * g.dx * r.idx;
    r.hy = (g.ny - 1) * g.dy * r.idy;
    r.hz = (g.nz - 1) * g.dz * r.idz;
    return r;
}

static int check_grid3(const m2d_grid3d *g)
{
    M2D_CHECK(g, M2D_ERR_BAD_ARG, "3D grid is NULL");
    M2D_CHECK(g->nx >= 3 && g->ny >= 3 && g->nz >= 3 && g->dx > 0 && g->dy > 0 && g->dz > 0, M2D_ERR_BAD_ARG,
              "3D grid %dx%dx%d spacing %g,%g,%g", g->nx, g->ny, g->nz, g->dx, g->dy, g->dz);
    M2D_CHECK((size_t)g->nx * g->ny * g->nz < (1ull << 31), M2D_ERR_BAD_ARG, "3D grid too large");
    return M2D_OK;
}

static int check_parts3(const m2d_particles3d *p)
{
    M2D_CHECK(p, M2D_ERR_BAD_ARG, "3D particle store is NULL");
    M2D_CHECK(p->n >= 0 && p->n < (1ll << 31), M2D_ERR_BAD_ARG, "3D store of %lld slots", (long long)p->n);
    M2D_CHECK(p->n == 0 || (p->x && p->y && p->z && p->vx && p->vy && p->vz), M2D_ERR_BAD_ARG,
              "3D particle arrays must be non-NULL");
    return M2D_OK;
}

// Gauss-Jordan inverse of the coarsest 7-point operator on its free nodes
static bool dense_inverse7(const std::vector<uint8_t> &dir, int nx, int ny, int nz, std::vector<float> &inv)
{
    const int n = nx * ny * nz;
    std::vector<int> id, pos(n, -1);
    for (int m = 0; m < n; m++)
        if (!dir[m]) { pos[m] = (int)id.size(); id.push_back(m); }
    const int q = (int)id.size();
    std::vector<double> a((size_t)q * 2 * q, 0.0);
    const int sx = ny * nz;
    for (int r = 0; r < q; r++) {
        const int m = id[r];
        a[(size_t)r * 2 * q + r] = 6.0;
        for (int nb : {m - sx, m - nz, m - 1, m + 1, m + nz, m + sx}) {
            if (nb < 0 || nb >= n || pos[nb] < 0) continue;
            a[(size_t)r * 2 * q + pos[nb]] = -1.0;
        }
        a[(size_t)r * 2 * q + q + r] = 1.0;
    }
    for (int c = 0; c < q; c++) {
        int piv = c;
        for (int r = c + 1; r < q; r++)
            if (fabs(a[(size_t)r * 2 * q + c]) > fabs(a[(size_t)piv * 2 * q + c])) piv = r;
        if (fabs(a[(size_t)piv * 2 * q + c]) < 1e-300) return false;
        if (piv != c)
            for (int t = 0; t < 2 * q; t++) std::swap(a[(size_t)c * 2 * q + t], a[(size_t)piv * 2 * q + t]);
        const double d = 1.0 / a[(size_t)c * 2 * q + c];
        for (int t = 0; t < 2 * q; t++) a[(size_t)c * 2 * q + t] *= d;
        for (int r = 0; r < q; r++) {
            if (r == c) continue;
            const double f = a[(size_t)r * 2 * q + c];
            if (f == 0.0) continue;
            for (int t = 0; t < 2 * q; t++) a[(size_t)r * 2 * q + t] -= f * a[(size_t)c * 2 * q + t];
        }
    }
    inv.assign((size_t)n * n, 0.f);
    for (int r = 0; r < q; r++)
        for (int c = 0; c < q; c++) inv[(size_t)id[r] * n + id[c]] = (float)a[(size_t)r * 2 * q + q + c];
    return true;
}

constexpr int kCoarseDenseMax = 1024;   // free nodes for the dense coarsest solve
constexpr int kCoarseSweeps = 40;       // symmetric RB-GS sweeps otherwise

static int mg3_setup(m2d_fields3d *f, const std::vector<uint8_t> &mask)
{
    std::vector<std::vector<uint8_t>> dirs;
    std::vector<int> dims;  // nx, ny, nz per level
    std::vector<uint8_t> d0(f->n);
    for (size_t m = 0; m < f->n; m++) d0[m] = mask[m] == M2D_FIXED || mask[m] == M2D_FIXED_RF;
    dirs.push_back(std::move(d0));
    dims.insert(dims.end(), {f->nx, f->ny, f->nz});
    for (;;) {
        const int fx = dims[dims.size() - 3], fy = dims[dims.size() - 2], fz = dims[dims.size() - 1];
        if ((size_t)fx * fy * fz <= 512 || fx < 5 || fy < 5 || fz < 5) break;
        const int cx = (fx - 1) / 2 + 1, cy = (fy - 1) / 2 + 1, cz = (fz - 1) / 2 + 1;
        const std::vector<uint8_t> &df = dirs.back();
        std::vector<uint8_t> dc((size_t)cx * cy * cz);
        for (int I = 0; I < cx; I++)
            for (int J = 0; J < cy; J++)
                for (int K = 0; K < cz; K++) {
                    const bool face = I == 0 || J == 0 || K == 0 || I == cx - 1 || J == cy - 1 || K == cz - 1;
                    dc[((size_t)I * cy + J) * cz + K] =
                        face || df[((size_t)(2 * I) * fy + 2 * J) * fz + 2 * K];
                }
        dirs.push_back(std::move(dc));
        dims.insert(dims.end(), {cx, cy, cz});
    }
    const size_t L = dirs.size();
    f->lv.resize(L);
    for (size_t l = 0; l < L; l++) {
        Mg3Level &V = f->lv[l];
        V.nx = dims[3 * l]; V.ny = dims[3 * l + 1]; V.nz = dims[3 * l + 2];
        V.n = (size_t)V.nx * V.ny * V.nz;
        M2D_TRY(dmalloc3((void **)&V.dir, V.n));
        M2D_TRY(dmalloc3((void **)&V.nbm, V.n));
        M2D_TRY(dmalloc3((void **)&V.x, sizeof(float) * V.n));
        M2D_TRY(dmalloc3((void **)&V.b, sizeof(float) * V.n));
        M2D_TRY(dmalloc3((void **)&V.r, sizeof(float) * V.n));
        M2D_HIP(hipMemcpy(V.dir, dirs[l].data(), V.n, hipMemcpyHostToDevice));
        // neighbour masks (a free node never sits on a face)
        const std::vector<uint8_t> &d = dirs[l];
        std::vector<uint8_t> nbm(V.n);
        const size_t sx = (size_t)V.ny * V.nz;
        for (int i = 0; i < V.nx; i++)
            for (int j = 0; j < V.ny; j++)
                for (int k = 0; k < V.nz; k++) {
                    const size_t m = (size_t)i * sx + (size_t)j * V.nz + k;
                    if (d[m]) { nbm[m] = kNbSelf; continue; }
                    const bool inside = i > 0 && i < V.nx - 1 && j > 0 && j < V.ny - 1 && k > 0 && k < V.nz - 1;
                    M2D_CHECK(inside, M2D_ERR_BAD_ARG, "Fields3D: free node on a face of level %d", (int)l);
                    const size_t nb[6] = {m - sx, m - (size_t)V.nz, m - 1, m + 1, m + (size_t)V.nz, m + sx};
                    uint8_t bits = 0;
                    for (int q = 0; q < 6; q++)
                        if (!d[nb[q]]) bits |= (uint8_t)(1u << q);
                    nbm[m] = bits;
                }
        M2D_HIP(hipMemcpy(V.nbm, nbm.data(), V.n, hipMemcpyHostToDevice));
    }
    M2D_CHECK(L <= (size_t)kMaxLevels3, M2D_ERR_BAD_ARG, "Fields3D: %d multigrid levels", (int)L);
    f->lc3 = (int)L - 1;
    for (size_t l = 1; l < L; l++)
        if (f->lv[l].n <= kBlockLevelNodes3) { f->lc3 = (int)l; break; }
    const Mg3Level &C = f->lv.back();
    size_t nfree = 0;
    for (uint8_t d : dirs.back()) nfree += !d;
    f->coarse_dense = nfree <= (size_t)kCoarseDenseMax && C.n <= 1024;
    if (f->coarse_dense) {
        std::vector<float> inv;
        M2D_CHECK(dense_inverse7(dirs.back(), C.nx, C.ny, C.nz, inv), M2D_ERR_BAD_ARG,
                  "Fields3D: coarsest multigrid operator is singular");
        f->coarse_n = (int)C.n;
        M2D_TRY(dmalloc3((void **)&f->coarse_inv, sizeof(float) * inv.size()));
        M2D_HIP(hipMemcpy(f->coarse_inv, inv.data(), sizeof(float) * inv.size(), hipMemcpyHostToDevice));
        const size_t cn = (size_t)f->coarse_n;
        std::vector<float> invT(cn * cn);
        for (size_t r = 0; r < cn; r++)
            for (size_t c = 0; c < cn; c++) invT[c * cn + r] = inv[r * cn + c];
        M2D_TRY(dmalloc3((void **)&f->coarse_invT, sizeof(float) * invT.size()));
        M2D_HIP(hipMemcpy(f->coarse_invT, invT.data(), sizeof(float) * invT.size(), hipMemcpyHostToDevice));
    }
    const size_t n = f->n;
    for (float **q : {&f->pr, &f->pz, &f->pp, &f->pp2, &f->pq, &f->pe}) M2D_TRY(dmalloc3((void **)q, sizeof(float) * n));
    M2D_TRY(dmalloc3((void **)&f->r64, sizeof(double) * n));
    // 16 scalars, then the solve's control block
    M2D_TRY(dmalloc3((void **)&f->dots, sizeof(double) * 32));
    M2D_HIP(hipHostMalloc((void **)&f->h_dots, sizeof(double) * 32));
    f->npartial = (int)std::min<size_t>((n + 255) / 256, 65536);
    // up to three sums per block of the fused reductions
    M2D_TRY(dmalloc3((void **)&f->partial, sizeof(double) * 3 * std::max<size_t>(f->npartial, (n + 255) / 256)));
    M2D_TRY(dmalloc3((void **)&f->ticket, sizeof(unsigned int)));
    M2D_HIP(hipMemset(f->ticket, 0, sizeof(unsigned int)));
    return M2D_OK;
}

constexpr int kPre3 = 2, kPost3 = 2;

__global__ void k3_zero(float *x, size_t n, const int *stop)
{
    if (stopped(stop)) return;
    size_t m = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m < n) x[m] = 0.f;
}

static unsigned blocks_for(size_t n) { return (unsigned)((n + 255) / 256); }

static int vcycle3(m2d_fields3d *f, const float *rhs, float *out, const int *stop)
{
    hipStream_t s = f->ctx->stream;
    const int L = (int)f->lv.size();
    // from_zero: x starts at 0 (the first half sweep also clears x)
    auto smooth = [&](Mg3Level &V, float *x, const float *b, int sweeps, bool forward, bool from_zero = false) {
        const dim3 g((unsigned)(((V.nz + 1) / 2 + kRbgsX - 1) / kRbgsX), (unsigned)((V.ny + kRbgsY - 1) / kRbgsY),
                     (unsigned)V.nx);
        const dim3 blk(kRbgsX, kRbgsY);
        for (int it = 0; it < sweeps; it++) {
            if (it == 0 && from_zero && forward) {
                k3_rbgs_from_zero<<<dim3((unsigned)((V.nz + 63) / 64), (unsigned)((V.ny + 3) / 4), (unsigned)V.nx),
                                    dim3(64, 4), 0, s>>>(V.nbm, x, b, V.nx, V.ny, V.nz, stop);
            } else {
                k3_rbgs<<<g, blk, 0, s>>>(V.nbm, x, b, V.nx, V.ny, V.nz, forward ? 0 : 1, stop);
            }
            k3_rbgs<<<g, blk, 0, s>>>(V.nbm, x, b, V.nx, V.ny, V.nz, forward ? 1 : 0, stop);
        }
    };
    // levels [lc, L) run in one workgroup (k3_vcycle_coarse) when lc > 0
    const int lc = L > 1 ? f->lc3 : L;
    for (int l = 0; l < std::min(lc, L - 1); l++) {
        Mg3Level &V = f->lv[l];
        float *x = l == 0 ? out : V.x;
        const float *b = l == 0 ? rhs : V.b;
        smooth(V, x, b, kPre3, true, true);
        k3_resid<<<dim3((unsigned)((V.nz + 63) / 64), (unsigned)((V.ny + 3) / 4), (unsigned)V.nx), dim3(64, 4), 0,
                   s>>>(V.nbm, x, b, V.r, V.nx, V.ny, V.nz, stop);
        Mg3Level &C = f->lv[l + 1];
        k3_restrict<<<dim3((unsigned)((C.nz + 63) / 64), (unsigned)((C.ny + 3) / 4), (unsigned)C.nx), dim3(64, 4), 0,
                      s>>>(V.r, V.nx, V.ny, V.nz, C.b, C.dir, C.nx, C.ny, C.nz, stop);
    }
    if (lc < L) {
        Coarse3Args C;
        C.lc = lc;
        C.L = L;
        for (int l = 0; l < L; l++) {
            const Mg3Level &V = f->lv[l];
            C.lv[l] = BLevel3{V.nx, V.ny, V.nz, (int)V.n, V.dir, V.nbm, V.x, V.b, V.r};
        }
        C.dense = f->coarse_dense;
        C.invT = f->coarse_invT;
        C.coarse_n = f->coarse_n;
        C.sweeps = kCoarseSweeps;
        C.pre = kPre3;
        C.post = kPost3;
        C.stop = stop;
        if (L - 1 - lc == 1 && f->lv[lc].n <= (size_t)kLds3Nodes && C.dense && f->coarse_n <= kLds3Dense) {
            k3_vcycle_coarse_lds<<<1, kCoarseThreads, 0, s>>>(C);
        } else {
            const size_t shm = C.dense ? sizeof(float) * f->coarse_n * (1 + kDenseParts) : 0;
            k3_vcycle_coarse<<<1, kCoarseThreads, shm, s>>>(C);
        }
    } else {
        Mg3Level &C = f->lv[L - 1];
        float *x = L == 1 ? out : C.x;
        const float *b = L == 1 ? rhs : C.b;
        if (f->coarse_dense) {
            k_dense_solve<<<1, kCoarseThreads, sizeof(float) * f->coarse_n * (1 + kDenseParts), s>>>(
                f->coarse_invT, b, x, f->coarse_n, stop);
        } else {
            k3_zero<<<blocks_for(C.n), 256, 0, s>>>(x, C.n, stop);
            smooth(C, x, b, kCoarseSweeps / 2, true);
            smooth(C, x, b, kCoarseSweeps / 2, false);
        }
    }
    for (int l = std::min(lc, L - 1) - 1; l >= 0; l--) {
        Mg3Level &V = f->lv[l];
        Mg3Level &C = f->lv[l + 1];
        float *x = l == 0 ? out : V.x;
        const float *b = l == 0 ? rhs : V.b;
        k3_prolong<<<dim3((unsigned)((V.nz + 63) / 64), (unsigned)((V.ny + 3) / 4), (unsigned)V.nx), dim3(64, 4), 0,
                     s>>>(x, V.dir, V.nx, V.ny, V.nz, C.x, C.nx, C.ny, C.nz, stop);
        smooth(V, x, b, kPost3, false);
    }
    M2D_HIP(hipGetLastError());
    return M2D_OK;
}

static SolveCtl *solve_ctl3(m2d_fields3d *f) { return (SolveCtl *)(f->dots + 16); }

// blocks of the fused reductions over the finest level: one node a thread up
// to kRed3Blocks blocks (524288 nodes), then that many, whose threads stride
// over several nodes (each block's reduction and partial sum then serve more
// nodes; the FP64 sums' order changes with the count, and with it their last
// bits, which the FP32 CG coefficients round away in practice).  At 128^3,
// 2048 blocks made the step 3% faster than one node a thread, 1024 and 4096
// less so.
constexpr size_t kRed3Blocks = 2048;
static unsigned red_blocks3(size_t n)
{
    return (unsigned)std::min<size_t>((n + kRedBlock - 1) / kRedBlock, kRed3Blocks);
}

// Iterations from + 1 .. to of the open FP32 PCG pass on A' e = d (d in pr,
// e in pe), from 0: the pass's start; every launch stops with the pass
// (SolveCtl::stop_pcg).  cur: which of dots[0], dots[4] holds the current
// r.z, (iterations run - 1) & 1.  dots: [1] p.q, [2] r.r
static int queue_pcg3(m2d_fields3d *f, int from, int to, int &cur)
{
    hipStream_t s = f->ctx->stream;
    const size_t n = f->n;
    const unsigned rb = red_blocks3(n);
    const Mg3Level &V = f->lv[0];
    SolveCtl *ctl = solve_ctl3(f);
    const int *stop = &ctl->stop_pcg;
    // iteration it's direction goes to P[it & 1], the previous one's is read from the other
    float *const P[2] = {f->pp, f->pp2};
    if (from == 0) {
        M2D_TRY(vcycle3(f, f->pr, f->pz, stop));
        reduce_launch<1>(rb, f->partial, f->dots, s, [&](int fin) {
            k_dot_fin<<<rb, kRedBlock, 0, s>>>(f->pr, f->pz, n, f->partial, f->ticket, f->dots, fin, stop);
        }, stop);
        cur = 0;
    }
    for (int it = from + 1; it <= to; it++) {
        if (it > 1) {
            M2D_TRY(vcycle3(f, f->pr, f->pz, stop));
            reduce_launch<1>(rb, f->partial, f->dots + 4 * (cur ^ 1), s, [&](int fin) {
                k_dot_fin<<<rb, kRedBlock, 0, s>>>(f->pr, f->pz, n, f->partial, f->ticket, f->dots + 4 * (cur ^ 1),
                                                   fin, stop);
            }, stop);
            cur ^= 1;
        }
        const float *pold = P[(it - 1) & 1];
        float *pnew = P[it & 1];
        reduce_launch<1>(rb, f->partial, f->dots + 1, s, [&](int fin) {
            k3_dir_apply_dot<<<rb, kRedBlock, 0, s>>>(V.nbm, f->pz, pold, pnew, f->pq, f->dots + 4 * cur,
                                                      f->dots + 4 * (cur ^ 1), it == 1, V.nx, V.ny, V.nz, f->partial,
                                                      f->ticket, f->dots + 1, fin, stop);
        }, stop);
        reduce_launch<1>(rb, f->partial, f->dots + 2, s, [&](int fin) {
            k_cg_update_dot<<<rb, kRedBlock, 0, s>>>(f->pe, f->pr, pnew, f->pq, f->dots + 4 * cur, f->dots + 1, n,
                                                     it == 1, f->partial, f->ticket, f->dots + 2, fin, stop,
                                                     CtlCheck{ctl});
        }, stop, CtlCheck{ctl});
    }
    M2D_HIP(hipGetLastError());
    return M2D_OK;
}

// the FP64 residual of u, its norms, and the refinement's decision on them
static void queue_residual3(m2d_fields3d *f, double tol)
{
    hipStream_t s = f->ctx->stream;
    const unsigned rb = red_blocks3(f->n);
    SolveCtl *ctl = solve_ctl3(f);
    const int *stop = &ctl->stop_outer;
    const CtlEval post{ctl, tol, kMaxOuter};
    reduce_launch<3>(rb, f->partial, f->dots + 8, s, [&](int fin) {
        k3_residual64_fin<<<rb, kRedBlock, 0, s>>>(f->mask, f->rho, f->u, f->nx, f->ny, f->nz, f->r64, f->pr,
                                                   f->partial, f->ticket, f->dots + 8, fin, stop, post);
    }, stop, post);
}

static int build_tables3(m2d_fields3d *f)
{
    k3_tables<<<blocks_for(f->n), 256, 0, f->ctx->stream>>>(f->u, f->nx, f->ny, f->nz, f->idx, f->idy, f->idz,
                                                            f->tx, f->ty, f->tz);
    M2D_HIP(hipGetLastError());
    return M2D_OK;
}

static double *field3_ptr(m2d_fields3d *f, int which)
{
    switch (which) {
    case M2D_FIELD_U: return f->u;
    case M2D_FIELD_RHO: return f->rho;
    case M2D_FIELD_VOLTAGE: return f->voltage;
    default: return nullptr;
    }
}

static float f32_below(double v)  // largest float <= v
{
    float r = (float)v;
    if ((double)r > v) r = nextafterf(r, -INFINITY);
    return r;
}

static uint64_t trial_thr3(double prob)
{
    if (!(prob > 0.0)) return 0;
    if (prob >= 1.0) return 1ull << 32;
    return (uint64_t)ceil(prob * 4294967296.0 - 0.5);
}

static void fill_args3(PushArgs3 &a, m2d_fields3d *f, const m2d_species2d *s, const m2d_particles3d *p,
                       double rot_scale, double kick_scale)
{
    a.x = p->x; a.y = p->y; a.z = p->z; a.vx = p->vx; a.vy = p->vy; a.vz = p->vz; a.n = p->n;
    a.tx = f->tx; a.ty = f->ty; a.tz = f->tz;
    a.cellfree = f->cellfree;
    a.g = grid3_of(f->p.grid);
    a.bx1 = f32_below(f->p.x_max); a.by1 = f32_below(f->p.y_max); a.bz1 = f32_below(f->p.z_max);
    a.xper = f->p.x_max; a.yper = f->p.y_max; a.zper = f->p.z_max;
    const double tmp = rot_scale * s->charge * s->dt / (2.0 * s->mass);
    const double tx = f->p.Bx * tmp, ty = f->p.By * tmp, tz = f->p.Bz * tmp;
    const double sc = 2.0 / (1 + tx * tx + ty * ty + tz * tz);
    a.tx_ = (float)tx; a.ty_ = (float)ty; a.tz_ = (float)tz;
    a.sx = (float)(tx * sc); a.sy = (float)(ty * sc); a.sz = (float)(tz * sc);
    a.rotate = tx != 0.0 || ty != 0.0 || tz != 0.0;
    a.qmdt2 = (float)(kick_scale * s->charge / s->mass * s->dt / 2.0);
    a.dt = (float)s->dt;
    a.push = 1;
    a.boundary = 0;
    a.periodic = f->p.boundary == M2D_BC_PERIODIC;
    a.trial_thr = 0;
    a.seed = f->ctx->seed;
    a.stream = s->stream;
    a.step = s->niter;
    a.mcc = nullptr; a.cq = nullptr; a.cq_len = nullptr;
    a.acc = nullptr;
    a.counters = f->ctx->counters;
}

static void free3(m2d_fields3d *f)
{
    for (void *q : {(void *)f->mask, (void *)f->cellfree, (void *)f->voltage, (void *)f->u, (void *)f->rho,
                    (void *)f->tx, (void *)f->ty, (void *)f->tz, (void *)f->coarse_inv, (void *)f->pr,
                    (void *)f->pz, (void *)f->pp, (void *)f->pp2, (void *)f->pq, (void *)f->pe, (void *)f->r64,
                    (void *)f->dots, (void *)f->partial, (void *)f->coarse_invT, (void *)f->ticket})
        if (q) (void)hipFree(q);
    for (Mg3Level &V : f->lv)
        for (void *q : {(void *)V.dir, (void *)V.nbm, (void *)V.x, (void *)V.b, (void *)V.r})
            if (q) (void)hipFree(q);
    if (f->h_dots) (void)hipHostFree(f->h_dots);
}

}  // namespace m2d

using namespace m2d;

extern "C" {

int m2d_fields3d_create(m2d_ctx *ctx, const m2d_fields3d_params *p, const uint8_t *mask, const double *voltage,
                        m2d_fields3d **out)
{
    M2D_CHECK(ctx && p && mask && voltage && out, M2D_ERR_BAD_ARG, "m2d_fields3d_create: NULL argument");
    *out = nullptr;
    M2D_TRY(check_grid3(&p->grid));
    M2D_CHECK(p->x_max > 0 && p->y_max > 0 && p->z_max > 0, M2D_ERR_BAD_ARG, "m2d_fields3d_create: box");
    const m2d_grid3d &g = p->grid;
    const size_t n = (size_t)g.nx * g.ny * g.nz;
    for (int i = 0; i < g.nx; i++)
        for (int j = 0; j < g.ny; j++)
            for (int k = 0; k < g.nz; k++) {
                const uint8_t m = mask[((size_t)i * g.ny + j) * g.nz + k];
                M2D_CHECK(m <= M2D_BOUNDARY, M2D_ERR_BAD_ARG, "mask value %d at (%d,%d,%d)", (int)m, i, j, k);
                const bool face = i == 0 || j == 0 || k == 0 || i == g.nx - 1 || j == g.ny - 1 || k == g.nz - 1;
                M2D_CHECK(!face || m == M2D_FIXED || m == M2D_FIXED_RF, M2D_ERR_BAD_ARG,
                          "m2d_fields3d_create: node (%d,%d,%d) on a grid face is not Dirichlet", i, j, k);
            }
    M2D_HIP(hipSetDevice(ctx->device));
    m2d_fields3d *f = new (std::nothrow) m2d_fields3d();
    M2D_CHECK(f, M2D_ERR_NO_MEMORY, "m2d_fields3d_create: host allocation");
    f->ctx = ctx;
    ctx_retain(ctx);
    f->p = *p;
    f->nx = g.nx; f->ny = g.ny; f->nz = g.nz;
    f->n = n;
    f->idx = 1.0 / g.dx; f->idy = 1.0 / g.dy; f->idz = 1.0 / g.dz;
    int r = M2D_OK;
    auto fail = [&](int code) { m2d_fields3d_destroy(f); return code; };
    if ((r = dmalloc3((void **)&f->mask, n)) || (r = dmalloc3((void **)&f->cellfree, n)) ||
        (r = dmalloc3((void **)&f->voltage, 8 * n)) || (r = dmalloc3((void **)&f->u, 8 * n)) ||
        (r = dmalloc3((void **)&f->rho, 8 * n)) || (r = dmalloc3((void **)&f->tx, 4 * n)) ||
        (r = dmalloc3((void **)&f->ty, 4 * n)) || (r = dmalloc3((void **)&f->tz, 4 * n)))
        return fail(r);
    std::vector<uint8_t> hm(mask, mask + n), cf(n, 0);
    for (int i = 0; i + 1 < g.nx; i++)
        for (int j = 0; j + 1 < g.ny; j++)
            for (int k = 0; k + 1 < g.nz; k++) {
                bool any = false;
                for (int c = 0; c < 8; c++)
                    any = any || hm[((size_t)(i + (c & 1)) * g.ny + j + ((c >> 1) & 1)) * g.nz + k + (c >> 2)] ==
                                     M2D_FREE;
                cf[((size_t)i * g.ny + j) * g.nz + k] = any;
            }
    if (hipMemcpy(f->mask, mask, n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(f->cellfree, cf.data(), n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(f->voltage, voltage, 8 * n, hipMemcpyHostToDevice) != hipSuccess)
        return fail(set_error(M2D_ERR_HIP, "m2d_fields3d_create: upload failed"));
    if ((r = fill_bytes(ctx, f->u, 0, 8 * n)) || (r = fill_bytes(ctx, f->rho, 0, 8 * n))) return fail(r);
    if ((r = mg3_setup(f, hm))) return fail(r);
    if ((r = build_tables3(f))) return fail(r);
    if (hipStreamSynchronize(ctx->stream) != hipSuccess)
        return fail(set_error(M2D_ERR_HIP, "m2d_fields3d_create: synchronise failed"));
    *out = f;
    return M2D_OK;
}

int m2d_fields3d_destroy(m2d_fields3d *f)
{
    if (!f) return M2D_OK;
    if (f->ctx) (void)hipStreamSynchronize(f->ctx->stream);
    free3(f);
    if (f->ctx) ctx_release(f->ctx);
    delete f;
    return M2D_OK;
}

double *m2d_fields3d_data(m2d_fields3d *f, int which) { return f ? field3_ptr(f, which) : nullptr; }

int m2d_fields3d_upload(m2d_fields3d *f, int which, const double *host)
{
    M2D_CHECK(f && host, M2D_ERR_BAD_ARG, "m2d_fields3d_upload: NULL argument");
    double *d = field3_ptr(f, which);
    M2D_CHECK(d, M2D_ERR_BAD_ARG, "m2d_fields3d_upload: unknown field %d", which);
    M2D_HIP(hipMemcpyAsync(d, host, 8 * f->n, hipMemcpyHostToDevice, f->ctx->stream));
    if (which == M2D_FIELD_U) M2D_TRY(build_tables3(f));
    M2D_HIP(hipStreamSynchronize(f->ctx->stream));
    return M2D_OK;
}

int m2d_fields3d_download(m2d_fields3d *f, int which, double *host)
{
    M2D_CHECK(f && host, M2D_ERR_BAD_ARG, "m2d_fields3d_download: NULL argument");
    double *d = field3_ptr(f, which);
    M2D_CHECK(d, M2D_ERR_BAD_ARG, "m2d_fields3d_download: unknown field %d", which);
    M2D_HIP(hipMemcpyAsync(host, d, 8 * f->n, hipMemcpyDeviceToHost, f->ctx->stream));
    M2D_HIP(hipStreamSynchronize(f->ctx->stream));
    return M2D_OK;
}

int m2d_fields3d_reset(m2d_fields3d *f)
{
    M2D_CHECK(f, M2D_ERR_BAD_ARG, "m2d_fields3d_reset: NULL argument");
    return fill_bytes(f->ctx, f->rho, 0, 8 * f->n);
}

int m2d_fields3d_solve(m2d_fields3d *f, double tol, m2d_solve_info *info)
{
    M2D_CHECK(f, M2D_ERR_BAD_ARG, "m2d_fields3d_solve: NULL argument");
    hipStream_t s = f->ctx->stream;
    const size_t n = f->n;
    const unsigned gb = blocks_for(n);
    k3_rhs<<<gb, 256, 0, s>>>(f->mask, f->voltage, f->rho, f->u, n, f->p.rhs_scale);
    M2D_HIP(hipGetLastError());
    if (tol <= 0) tol = 1e-10;
    SolveCtl *ctl = solve_ctl3(f);
    SolveCtl *h = (SolveCtl *)(f->h_dots + 16);
    k_ctl_init<<<1, 64, 0, s>>>(ctl);
    queue_residual3(f, tol);
    M2D_TRY(queue_refinement(
        s, ctl, h, f->dots, f->est, tol, 80,
        [&](int from, int to, int &cur) { return queue_pcg3(f, from, to, cur); },
        [&] { k3_add_correction<<<gb, 256, 0, s>>>(f->mask, f->u, f->pe, n, &ctl->no_add); },
        [&] { queue_residual3(f, tol); }));
    const int outer = h->outer;
    const double rel = h->rel;
    if (info) { info->outer_iters = outer; info->inner_iters = h->iters; info->rel_residual = rel; }
    if (h->nan) return set_error(M2D_ERR_NAN, "Fields3D solve: residual is not finite");
    M2D_TRY(build_tables3(f));
    if (rel > tol)
        return set_error(M2D_ERR_NOT_CONVERGED, "Fields3D solve: relative residual %.3e > %.3e after %d steps",
                         rel, tol, outer);
    return M2D_OK;
}

int m2d_accumulate_3d(m2d_ctx *ctx, const m2d_grid3d *g, const m2d_particles3d *p, double charge, double *rho)
{
    M2D_CHECK(ctx && rho, M2D_ERR_BAD_ARG, "m2d_accumulate_3d: NULL argument");
    M2D_TRY(check_grid3(g));
    M2D_TRY(check_parts3(p));
    const size_t n = (size_t)g->nx * g->ny * g->nz;
    M2D_TRY(ensure_acc(ctx, n));
    if (p->n > 0) {
        k_accumulate3d<<<blocks_for((size_t)p->n), 256, 0, ctx->stream>>>(
            p->x, p->y, p->z, p->n, grid3_of(*g), (unsigned long long *)ctx->acc.ptr, ctx->counters);
        M2D_HIP(hipGetLastError());
    }
    k_flush<<<blocks_for(n), 256, 0, ctx->stream>>>((unsigned long long *)ctx->acc.ptr, rho, charge, n);
    M2D_HIP(hipGetLastError());
    return M2D_OK;
}

int m2d_interpolate_3d(m2d_ctx *ctx, const m2d_grid3d *g, const double *data, const double *x, const double *y,
                       const double *z, int64_t n, double *out)
{
    M2D_CHECK(ctx && data && (n == 0 || (x && y && z && out)), M2D_ERR_BAD_ARG, "m2d_interpolate_3d: NULL argument");
    M2D_TRY(check_grid3(g));
    if (n > 0) {
        k_interp3d<<<blocks_for((size_t)n), 256, 0, ctx->stream>>>(data, grid3_of(*g), x, y, z, n, out);
        M2D_HIP(hipGetLastError());
    }
    return M2D_OK;
}

int m2d_grad_3d(m2d_ctx *ctx, const m2d_grid3d *g, const double *data, const double *x, const double *y,
                const double *z, int64_t n, double *gx, double *gy, double *gz)
{
    M2D_CHECK(ctx && data && (n == 0 || (x && y && z && gx && gy && gz)), M2D_ERR_BAD_ARG,
              "m2d_grad_3d: NULL argument");
    M2D_TRY(check_grid3(g));
    if (n > 0) {
        k_grad3d<<<blocks_for((size_t)n), 256, 0, ctx->stream>>>(data, grid3_of(*g), x, y, z, n, gx, gy, gz);
        M2D_HIP(hipGetLastError());
    }
    return M2D_OK;
}

int m2d_advance_3d(m2d_ctx *ctx, m2d_fields3d *f, const m2d_species2d *s, const m2d_particles3d *p,
                   const m2d_mcc *mcc, double *rho, uint32_t flags, m2d_step_stats *stats)
{
    M2D_CHECK(ctx && f && s, M2D_ERR_BAD_ARG, "m2d_advance_3d: NULL argument");
    M2D_CHECK(f->ctx == ctx, M2D_ERR_BAD_ARG, "m2d_advance_3d: fields belong to another context");
    M2D_TRY(check_parts3(p));
    M2D_CHECK(s->mass > 0 && s->dt > 0, M2D_ERR_BAD_ARG, "m2d_advance_3d: mass and dt must be > 0");
    M2D_CHECK((flags & ~(M2D_STEP_PUSH | M2D_STEP_BOUNDARY | M2D_STEP_DEPOSIT)) == 0, M2D_ERR_BAD_ARG,
              "m2d_advance_3d: unknown flags 0x%x", flags);
    const bool deposit = (flags & M2D_STEP_DEPOSIT) != 0;
    M2D_CHECK(!deposit || rho, M2D_ERR_BAD_ARG, "m2d_advance_3d: DEPOSIT needs rho");
    M2D_CHECK(!mcc || !mcc->h.partner_update, M2D_ERR_UNSUPPORTED,
              "m2d_advance_3d: COULOMB (partner update) runs in m2d_advance_2d only");
    M2D_CHECK(!mcc || !mcc_partners_alias(mcc, p->x), M2D_ERR_UNSUPPORTED,
              "m2d_advance_3d: partners from the colliding species' own store run in m2d_advance_2d only");
    if (stats) {
        M2D_TRY(zero_words(ctx, ctx->counters + CNT_REMOVED, 2));
        M2D_TRY(zero_words(ctx, ctx->counters + CNT_SPILL, 1));
    }
    PushArgs3 a;
    fill_args3(a, f, s, p, 1.0, 1.0);
    a.push = (flags & M2D_STEP_PUSH) != 0;
    a.boundary = (flags & M2D_STEP_BOUNDARY) != 0;
    // an empty store queues nothing (and leaves the alternating counters)
    const bool use_mcc = p->n > 0 && a.push && mcc && std::isfinite(s->lifetime) && s->lifetime > 0;
    unsigned long long *next_len = nullptr;
    if (use_mcc) {
        a.trial_thr = trial_thr3(1.0 - exp(-s->dt / s->lifetime));  // species3d.cpp:8
        a.mcc = mcc->d;
        M2D_TRY(ensure(ctx, ctx->cq, (size_t)p->n * sizeof(uint32_t)));
        a.cq = (uint32_t *)ctx->cq.ptr;
        // a call that failed between its pass and its scatter left the
        // alternating counters unknown: clear both first
        if (ctx->cq_dirty) M2D_TRY(zero_words(ctx, ctx->counters + CNT_QLEN, 2));
        ctx->cq_dirty = true;
        const int slot = (int)(ctx->launch_seq++ & 1);
        a.cq_len = ctx->counters + CNT_QLEN + slot;
        next_len = ctx->counters + CNT_QLEN + (slot ^ 1);
    }
    if (deposit) {
        M2D_TRY(ensure_acc(ctx, f->n));
        a.acc = (unsigned long long *)ctx->acc.ptr;
    }
    // a store sorted by m2d_sort_3d deposits through per-chunk LDS windows
    const int ntx3 = (f->nx - 2) / kTile3 + 1, nty3 = (f->ny - 2) / kTile3 + 1, ntz3 = (f->nz - 2) / kTile3 + 1;
    const m2d_tiling *T = deposit ? find_tiling(ctx, p->x, p->n) : nullptr;
    if (T && (T->ntx != ntx3 || T->nty != nty3 || T->ntz != ntz3)) T = nullptr;
    if (p->n > 0) {
        if (T) {
            launch_timed(ctx, use_mcc ? k_advance3d_tiled<true> : k_advance3d_tiled<false>, dim3(T->max_chunks),
                         dim3(256), a, T->sched, T->n_chunks, nty3, ntz3);
        } else {
            void (*kern)(PushArgs3) = deposit ? (use_mcc ? k_advance3d<true, true> : k_advance3d<true, false>)
                                              : (use_mcc ? k_advance3d<false, true> : k_advance3d<false, false>);
            launch_timed(ctx, kern, dim3(blocks_for((size_t)p->n)), dim3(256), a);
        }
        M2D_HIP(hipGetLastError());
        if (use_mcc) {
            k_scatter3d<<<64, 256, 0, ctx->stream>>>(a, next_len);
            M2D_HIP(hipGetLastError());
            ctx->cq_dirty = false;
        }
    }
    if (deposit) {
        k_flush<<<blocks_for(f->n), 256, 0, ctx->stream>>>(a.acc, rho, s->charge, f->n);
        M2D_HIP(hipGetLastError());
    }
    if (stats) {
        M2D_TRY(read_counters(ctx));
        stats->removed = (int64_t)ctx->h_counters[CNT_REMOVED];
        stats->collisions = (int64_t)ctx->h_counters[CNT_COLL];
        stats->spilled = (int64_t)ctx->h_counters[CNT_SPILL];
    }
    return M2D_OK;
}

int m2d_advance_init_3d(m2d_ctx *ctx, m2d_fields3d *f, const m2d_species2d *s, const m2d_particles3d *p)
{
    M2D_CHECK(ctx && f && s, M2D_ERR_BAD_ARG, "m2d_advance_init_3d: NULL argument");
    M2D_TRY(check_parts3(p));
    M2D_CHECK(s->mass > 0 && s->dt > 0, M2D_ERR_BAD_ARG, "m2d_advance_init_3d: mass and dt must be > 0");
    PushArgs3 a;
    fill_args3(a, f, s, p, -0.5, -1.0);
    if (p->n > 0) {
        k_advance_init3d<<<blocks_for((size_t)p->n), 256, 0, ctx->stream>>>(a);
        M2D_HIP(hipGetLastError());
    }
    return M2D_OK;
}

int m2d_sort_3d(m2d_ctx *ctx, m2d_fields3d *f, const m2d_particles3d *src, const m2d_particles3d *dst,
                int64_t *n_live)
{
    M2D_CHECK(ctx && f && src && dst, M2D_ERR_BAD_ARG, "m2d_sort_3d: NULL argument");
    M2D_TRY(check_parts3(src));
    M2D_TRY(check_parts3(dst));
    M2D_CHECK(dst->n == src->n, M2D_ERR_BAD_ARG, "m2d_sort_3d: dst holds %lld slots, src %lld",
              (long long)dst->n, (long long)src->n);
    M2D_CHECK(src->n == 0 || src->x != dst->x, M2D_ERR_BAD_ARG, "m2d_sort_3d: src and dst must differ");
    const size_t n = (size_t)src->n;
    forget_tiling(ctx, src->x);
    forget_tiling(ctx, dst->x);
    if (n == 0) {
        if (n_live) *n_live = 0;
        return M2D_OK;
    }
    const Grid3 g = grid3_of(f->p.grid);
    const int ntx3 = (g.nx - 2) / kTile3 + 1, nty3 = (g.ny - 2) / kTile3 + 1, ntz3 = (g.nz - 2) / kTile3 + 1;
    const uint64_t keyspace = ((uint64_t)ntx3 * nty3 * ntz3) << 6;
    int bits = 0;
    while ((1ull << bits) <= keyspace) bits++;
    M2D_CHECK(bits <= 32, M2D_ERR_BAD_ARG, "m2d_sort_3d: grid too large for 32-bit keys");
    const uint32_t dead_key = bits == 32 ? 0xFFFFFFFFu : (uint32_t)((1ull << bits) - 1);
    size_t temp = 0;
    // tile-major cell order.  Unlike the 2D store (ordered by tile only), a
    // 3D cell holds few particles (8 on average at the 16M-ion, 128^3
    // workload), so the lanes of a wave spread over several cells anyway, and
    // the cell order's gather locality wins: the tiled pass runs 419 us on a
    // cell-ordered store against 505 us on a tile-ordered one
    hipError_t e = rocprim::radix_sort_pairs(nullptr, temp, (const uint32_t *)nullptr, (uint32_t *)nullptr,
                                             (const uint32_t *)nullptr, (uint32_t *)nullptr, n, 0, bits,
                                             ctx->stream);
    M2D_CHECK(e == hipSuccess, M2D_ERR_HIP, "radix_sort_pairs size query: %s", hipGetErrorString(e));
    const size_t arr = ((n * 4 + 255) / 256) * 256;
    M2D_TRY(ensure(ctx, ctx->sortbuf, 4 * arr + temp + 256));
    uint8_t *b = (uint8_t *)ctx->sortbuf.ptr;
    uint32_t *k0 = (uint32_t *)b, *k1 = (uint32_t *)(b + arr), *v0 = (uint32_t *)(b + 2 * arr),
             *v1 = (uint32_t *)(b + 3 * arr);
    unsigned long long *live = ctx->counters + CNT_LIVE;
    M2D_TRY(zero_words(ctx, live, 1));
    k_sort_keys3d<<<blocks_for(n), 256, 0, ctx->stream>>>(src->x, src->y, src->z, src->n, g, nty3, ntz3, dead_key,
                                                          k0, v0);
    M2D_HIP(hipGetLastError());
    e = rocprim::radix_sort_pairs(b + 4 * arr, temp, k0, k1, v0, v1, n, 0, bits, ctx->stream);
    M2D_CHECK(e == hipSuccess, M2D_ERR_HIP, "radix_sort_pairs: %s", hipGetErrorString(e));
    k_sort_gather3d<<<xcd_grid(blocks_for(n)), 256, 0, ctx->stream>>>(*src, *dst, k1, v1, dead_key, live);
    M2D_HIP(hipGetLastError());
    M2D_TRY(build_tiling(ctx, dst->x, dst->n, k1, dead_key, ntx3, nty3, ntz3, kChunkCap3, 256));
    if (n_live) {
        M2D_TRY(read_counters(ctx));
        *n_live = (int64_t)ctx->h_counters[CNT_LIVE];
    }
    return M2D_OK;
}

int m2d_sort_forget_3d(m2d_ctx *ctx, const m2d_particles3d *p)
{
    M2D_CHECK(ctx && p, M2D_ERR_BAD_ARG, "m2d_sort_forget_3d: NULL argument");
    forget_tiling(ctx, p->x);
    return M2D_OK;
}

int m2d_count_live_3d(m2d_ctx *ctx, const m2d_particles3d *p, int64_t *n_live)
{
    M2D_TRY(check_parts3(p));
    const m2d_particles2d v{p->x, p->z, p->vx, p->vy, p->vz, p->n};
    return m2d_count_live_2d(ctx, &v, n_live);
}

int m2d_mean_energy_3d(m2d_ctx *ctx, const m2d_particles3d *p, double mass, double *mean_eV, int64_t *n_live)
{
    M2D_TRY(check_parts3(p));
    const m2d_particles2d v{p->x, p->z, p->vx, p->vy, p->vz, p->n};
    return m2d_mean_energy_2d(ctx, &v, mass, mean_eV, n_live);
}

}  // extern "C"
